"""Writes tests/golden/config1_reference.npz: outputs of the reference C
library (compiled from source into oracle/_ref by oracle/Makefile) on the
small workload's tree sequence (1000 samples, L=1e6, seed 1), together with
the SHA-256 of the generated input columns so a test can check that the same
input was regenerated.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import tskit_amd as ta  # noqa: E402
from oracle.oracle import Reference  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "config1_reference.npz")


def main():
    tc = ta.simulate_tables(1000, 1e6, 1e4, 1e-8, 1e-8, 1)
    ref = Reference(tc)
    sets2 = ta.random_sample_sets(1000, 2, 500, 1)
    sets4 = ta.random_sample_sets(1000, 4, 250, 2)
    afs_sets = ta.random_sample_sets(1000, 2, 25, 3)
    w = np.linspace(0, 1e6, 11)
    out = {
        "digest": np.frombuffer(tc.column_digest().encode(), dtype=np.uint8),
        "windows": w,
        "sets2": np.concatenate(sets2),
        "sets4": np.concatenate(sets4),
        "afs_sets": np.concatenate(afs_sets),
        "num_trees": np.array([ref.num_trees]),
        "breakpoints": ref.breakpoints(),
    }
    I, O = ref.index()
    out["index_I"], out["index_O"] = I, O
    for mode in ["branch", "site"]:
        out[f"diversity_{mode}"] = ref.stat("diversity", sets2, None, w, mode)
        out[f"divergence_{mode}"] = ref.stat("divergence", sets2, [(0, 1)], w, mode)
        out[f"segsites_{mode}"] = ref.stat("segregating_sites", sets2, None, w, mode)
        out[f"f4_{mode}"] = ref.stat("f4", sets4, [(0, 1, 2, 3), (0, 2, 1, 3)], w, mode)
        out[f"f2_{mode}"] = ref.stat("f2", sets4, [(0, 1), (2, 3)], w, mode)
        out[f"relatedness_{mode}"] = ref.stat(
            "genetic_relatedness", sets4, [(0, 1), (2, 2)], w, mode, True, True)
        out[f"diversity_raw_{mode}"] = ref.stat("diversity", sets2, None, w, mode, False)
        out[f"segsites_raw_{mode}"] = ref.stat("segregating_sites", sets2, None, w, mode, False)
        out[f"afs_{mode}"] = ref.afs(afs_sets, np.linspace(0, 1e6, 5), None, mode)
        out[f"afs_polarised_{mode}"] = ref.afs(afs_sets, np.linspace(0, 1e6, 5), None, mode,
                                               True, True)
    out["afs_tw_branch"] = ref.afs(afs_sets, np.linspace(0, 1e6, 3),
                                   np.array([0, 5000, 20000, np.inf]), "branch")
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
