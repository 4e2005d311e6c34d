"""Writes tests/golden/relvec_reference.npz: genetic_relatedness_vector
outputs of the reference C library (compiled from source into oracle/_ref by
oracle/Makefile) on a small generated tree sequence (50 samples, L=5e4,
seed 9) and on the same sequence with gaps cut into 10 edges, for random
weights (3 columns, seed 4): windows spanning the sequence and windows that
do not, centred and not, span-normalised and not, and every node as a focal
node; with the SHA-256 of each input's columns.

Run from the repository root: python tests/golden/make_relvec_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import cases  # noqa: E402
from oracle.oracle import Reference  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "relvec_reference.npz")

# (name, windows key, centre, span_normalise, focal key)
CHECKS = [
    ("full_centred", "full", True, True, "samples"),
    ("full_raw", "full", False, False, "samples"),
    ("inner_centred", "inner", True, False, "samples"),
    ("inner_nodes", "inner", False, True, "nodes"),
]


def inputs():
    return {"generated": cases.generated(50, 5e4, 9), "gapped": cases.gapped(50, 5e4, 9, count=10)}


def weights(n):
    return np.random.default_rng(4).normal(size=(n, 3))


def windows():
    return {"full": np.array([0, 1.3e4, 3.1e4, 5e4]), "inner": np.array([4e3, 2.2e4, 4.6e4])}


def focal(tc, key):
    if key == "samples":
        return np.flatnonzero(tc.node_flags & 1).astype(np.int32)
    return np.arange(tc.num_nodes, dtype=np.int32)[::-1].copy()


def main():
    out = {}
    for key, tc in inputs().items():
        ref = Reference(tc)
        out[f"{key}_digest"] = np.frombuffer(tc.column_digest().encode(), dtype=np.uint8)
        W = weights(int(np.sum(tc.node_flags & 1)))
        for name, wk, centre, sn, fk in CHECKS:
            out[f"{key}_{name}"] = ref.relatedness_vector(W, windows()[wk], "branch", sn, centre,
                                                          focal(tc, fk))
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
