"""Synthetic inputs: Hudson coalescent with recombination plus infinite-sites
0/1 mutations (the generator lives in csrc/tsg_sim.cpp, host code).

``sim_ancestry`` plays the role msprime.sim_ancestry plays for the
reference's tests and benchmarks; its output is a valid, sorted, simplified
table collection that the reference C library accepts unchanged.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import ptr
from .exceptions import check
from .tables import TableCollection, _ragged

MUTATION_SEED_OFFSET = 0x9E37


def simulate_tables(num_samples, sequence_length, population_size=1e4, recombination_rate=1e-8,
                    mutation_rate=0.0, random_seed=1):
    """Returns a TableCollection: ancestry from ``random_seed``, mutations
    from ``random_seed + 0x9E37``."""
    L = _lib.lib()
    h = ctypes.c_void_p()
    check(L.tsg_sim_ancestry(int(num_samples), float(sequence_length), float(population_size),
                             float(recombination_rate), int(random_seed), ctypes.byref(h)))
    try:
        if mutation_rate > 0:
            check(L.tsg_sim_mutations(h, float(mutation_rate),
                                      int(random_seed) + MUTATION_SEED_OFFSET))
        sizes = (ctypes.c_uint64 * 4)()
        check(L.tsg_tables_get_sizes(h, sizes))
        N, E, S, Mu = (int(x) for x in sizes)
        flags = np.zeros(N, np.uint32)
        time = np.zeros(N, np.float64)
        left = np.zeros(E, np.float64)
        right = np.zeros(E, np.float64)
        parent = np.zeros(E, np.int32)
        child = np.zeros(E, np.int32)
        pos = np.zeros(S, np.float64)
        msite = np.zeros(Mu, np.int32)
        mnode = np.zeros(Mu, np.int32)
        mparent = np.zeros(Mu, np.int32)
        mtime = np.zeros(Mu, np.float64)
        check(L.tsg_tables_copy_out(
            h, ptr(flags, ctypes.c_uint32), ptr(time, ctypes.c_double),
            ptr(left, ctypes.c_double), ptr(right, ctypes.c_double),
            ptr(parent, ctypes.c_int32), ptr(child, ctypes.c_int32), ptr(pos, ctypes.c_double),
            ptr(msite, ctypes.c_int32), ptr(mnode, ctypes.c_int32),
            ptr(mparent, ctypes.c_int32), ptr(mtime, ctypes.c_double)))
    finally:
        L.tsg_tables_free(h)
    tc = TableCollection(sequence_length)
    tc.node_flags, tc.node_time = flags, time
    tc.edge_left, tc.edge_right, tc.edge_parent, tc.edge_child = left, right, parent, child
    tc.site_position = pos
    tc.ancestral_state = np.full(S, ord("0"), dtype=np.int8)
    tc.ancestral_state_offset = np.arange(S + 1, dtype=np.uint64)
    tc.mutation_site, tc.mutation_node, tc.mutation_parent = msite, mnode, mparent
    tc.derived_state = np.full(Mu, ord("1"), dtype=np.int8)
    tc.derived_state_offset = np.arange(Mu + 1, dtype=np.uint64)
    return tc


def sim_ancestry(samples, sequence_length, population_size=1e4, recombination_rate=1e-8,
                 random_seed=1, mutation_rate=0.0):
    """Simulates a tree sequence (optionally with mutations) on the device."""
    return simulate_tables(samples, sequence_length, population_size, recombination_rate,
                           mutation_rate, random_seed).tree_sequence()


def random_sample_sets(num_samples, num_sets, set_size, seed):
    """Disjoint random sample sets from a seeded permutation of the samples."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(num_samples)
    if num_sets * set_size > num_samples:
        raise ValueError("not enough samples for disjoint sets")
    return [np.sort(perm[k * set_size:(k + 1) * set_size]).astype(np.int32)
            for k in range(num_sets)]


__all__ = ["simulate_tables", "sim_ancestry", "random_sample_sets", "_ragged"]
