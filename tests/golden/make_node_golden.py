"""Writes tests/golden/node_reference.npz: node-mode statistics from the
reference C library (compiled from source into oracle/_ref by
oracle/Makefile) on a small generated tree sequence (60 samples, L=5e4,
seed 5) and on the same sequence with gaps cut into 12 edges (several roots,
samples missing from trees), with the SHA-256 of each input's columns.

Run from the repository root: python tests/golden/make_node_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import cases  # noqa: E402
import tskit_amd as ta  # noqa: E402
from oracle.oracle import Reference  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "node_reference.npz")

# (name, stat, indexes, polarised, span_normalise)
CHECKS = [
    ("diversity", "diversity", None, False, True),
    ("segsites_pol", "segregating_sites", None, True, False),
    ("Y1", "Y1", None, False, True),
    ("divergence", "divergence", [(0, 1), (1, 2)], False, True),
    ("f4", "f4", [(0, 1, 2, 3)], False, False),
    ("relatedness", "genetic_relatedness", [(0, 1), (2, 2)], True, True),
]


def inputs():
    return {"generated": cases.generated(60, 5e4, 5), "gapped": cases.gapped(60, 5e4, 5, count=12)}


def main():
    out = {}
    sets = ta.random_sample_sets(60, 4, 12, 7)
    out["sets"] = np.concatenate(sets)
    out["windows"] = np.array([0, 1.1e4, 2.5e4, 5e4])
    for key, tc in inputs().items():
        ref = Reference(tc)
        out[f"{key}_digest"] = np.frombuffer(tc.column_digest().encode(), dtype=np.uint8)
        for name, stat, idx, pol, sn in CHECKS:
            out[f"{key}_{name}"] = ref.stat(stat, sets, idx, out["windows"], "node", sn, pol)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
