"""ctypes binding to the in-tree tsgpu library (``libtsgpu.so``).

The library is the product: every statistic is computed by its HIP kernels.
There is no fallback; importing a statistic without the built library raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtsgpu.so")

_lib = None

# Array arguments are declared as untyped addresses and passed as plain
# integers: a typed ctypes pointer costs one ctypes.cast per argument, which
# at ~10 pointer arguments per statistics call was most of the host time of
# a batch of queries.  The names record the element type each entry expects.
c_i32p = ctypes.c_void_p
c_u32p = ctypes.c_void_p
c_u64p = ctypes.c_void_p
c_f64p = ctypes.c_void_p
c_void = ctypes.c_void_p

# name -> (restype, argtypes); every symbol declared in include/tsgpu.h
ONE_WAY = [c_void, ctypes.c_uint64, c_u64p, c_i32p, ctypes.c_uint64, c_f64p, ctypes.c_uint32, c_f64p]
K_WAY = [
    c_void, ctypes.c_uint64, c_u64p, c_i32p, ctypes.c_uint64, c_i32p,
    ctypes.c_uint64, c_f64p, ctypes.c_uint32, c_f64p,
]
SIGNATURES = {
    "tsg_strerror": (ctypes.c_char_p, [ctypes.c_int]),
    "tsg_treeseq_init": (
        ctypes.c_int,
        [
            ctypes.POINTER(c_void), ctypes.c_double,
            ctypes.c_uint64, c_u32p, c_f64p,
            ctypes.c_uint64, c_f64p, c_f64p, c_i32p, c_i32p,
            ctypes.c_uint64, c_f64p, ctypes.c_char_p, c_u64p,
            ctypes.c_uint64, c_i32p, c_i32p, c_i32p, ctypes.c_char_p, c_u64p,
            c_i32p, c_i32p, ctypes.c_uint32,
        ],
    ),
    "tsg_treeseq_free": (ctypes.c_int, [c_void]),
    "tsg_treeseq_set_num_chunks": (ctypes.c_int, [c_void, ctypes.c_uint32]),
    "tsg_treeseq_upload": (ctypes.c_int, [c_void]),
    "tsg_treeseq_get_num_trees": (ctypes.c_uint64, [c_void]),
    "tsg_treeseq_get_num_samples": (ctypes.c_uint64, [c_void]),
    "tsg_treeseq_get_num_nodes": (ctypes.c_uint64, [c_void]),
    "tsg_treeseq_get_num_edges": (ctypes.c_uint64, [c_void]),
    "tsg_treeseq_get_num_sites": (ctypes.c_uint64, [c_void]),
    "tsg_treeseq_get_num_mutations": (ctypes.c_uint64, [c_void]),
    "tsg_treeseq_get_num_chunks": (ctypes.c_uint64, [c_void]),
    "tsg_treeseq_get_sequence_length": (ctypes.c_double, [c_void]),
    "tsg_treeseq_get_device_bytes": (ctypes.c_uint64, [c_void]),
    "tsg_treeseq_get_max_chunk_nodes": (ctypes.c_uint64, [c_void]),
    "tsg_treeseq_get_breakpoints": (ctypes.c_int, [c_void, c_f64p]),
    "tsg_treeseq_get_samples": (ctypes.c_int, [c_void, c_i32p]),
    "tsg_treeseq_get_index": (ctypes.c_int, [c_void, c_i32p, c_i32p]),
    "tsg_treeseq_get_plan_sizes": (ctypes.c_int, [c_void, c_u64p]),
    "tsg_treeseq_copy_plan": (ctypes.c_int, [c_void] + [c_void] * 17),
    "tsg_treeseq_get_path_size": (ctypes.c_uint64, [c_void]),
    "tsg_treeseq_copy_paths": (ctypes.c_int, [c_void, c_void, c_void]),
    "tsg_treeseq_get_hot_plan_sizes": (ctypes.c_int, [c_void, c_u64p]),
    "tsg_treeseq_copy_hot_plan": (ctypes.c_int, [c_void] + [c_void] * 16),
    "tsg_treeseq_copy_hot_afs_plan": (ctypes.c_int, [c_void] + [c_void] * 9),
    "tsg_treeseq_set_sweep_plan": (ctypes.c_int, [c_void, ctypes.c_int]),
    "tsg_treeseq_get_last_store": (ctypes.c_int, [c_void]),
    "tsg_treeseq_set_profiling": (ctypes.c_int, [c_void, ctypes.c_int]),
    "tsg_treeseq_get_profile": (ctypes.c_int, [c_void, c_f64p, c_u64p]),
    "tsg_treeseq_get_profile_modes": (ctypes.c_int, [c_void, c_f64p, c_u64p]),
    "tsg_treeseq_diversity": (ctypes.c_int, ONE_WAY),
    "tsg_treeseq_segregating_sites": (ctypes.c_int, ONE_WAY),
    "tsg_treeseq_Y1": (ctypes.c_int, ONE_WAY),
    "tsg_treeseq_Tajimas_D": (ctypes.c_int, ONE_WAY),
    "tsg_treeseq_divergence": (ctypes.c_int, K_WAY),
    "tsg_treeseq_Y2": (ctypes.c_int, K_WAY),
    "tsg_treeseq_f2": (ctypes.c_int, K_WAY),
    "tsg_treeseq_Y3": (ctypes.c_int, K_WAY),
    "tsg_treeseq_f3": (ctypes.c_int, K_WAY),
    "tsg_treeseq_f4": (ctypes.c_int, K_WAY),
    "tsg_treeseq_genetic_relatedness": (ctypes.c_int, K_WAY),
    "tsg_treeseq_genetic_relatedness_proportion": (ctypes.c_int, K_WAY),
    "tsg_treeseq_Fst": (ctypes.c_int, K_WAY),
    "tsg_treeseq_allele_frequency_spectrum": (
        ctypes.c_int,
        [
            c_void, ctypes.c_uint64, c_u64p, c_i32p, ctypes.c_uint64, c_f64p,
            ctypes.c_uint64, c_f64p, ctypes.c_uint32, c_f64p,
        ],
    ),
    "tsg_treeseq_genotype_matrix": (ctypes.c_int, [c_void, ctypes.c_uint32, c_i32p]),
    "tsg_treeseq_divergence_matrix": (
        ctypes.c_int,
        [c_void, ctypes.c_uint64, c_u64p, c_i32p, ctypes.c_uint64, c_f64p, ctypes.c_uint32,
         c_f64p],
    ),
    "tsg_treeseq_genetic_relatedness_matrix": (
        ctypes.c_int,
        [c_void, ctypes.c_uint64, c_u64p, c_i32p, ctypes.c_uint64, c_f64p, ctypes.c_uint32,
         c_f64p],
    ),
    "tsg_batch_init": (ctypes.c_int, [ctypes.POINTER(c_void), c_void]),
    "tsg_batch_free": (ctypes.c_int, [c_void]),
    "tsg_batch_clear": (ctypes.c_int, [c_void]),
    "tsg_batch_get_num_queries": (ctypes.c_uint64, [c_void]),
    "tsg_batch_run": (ctypes.c_int, [c_void]),
    "tsg_batch_run_multi": (ctypes.c_int, [c_void, ctypes.c_uint64]),
    "tsg_batch_run_noalloc": (ctypes.c_int, [c_void]),
    "tsg_treeseq_genetic_relatedness_vector": (
        ctypes.c_int,
        [c_void, ctypes.c_uint64, c_f64p, ctypes.c_uint64, c_f64p, ctypes.c_uint64, c_i32p,
         c_f64p, ctypes.c_uint32],
    ),
    "tsg_treeseq_general_stat": (
        ctypes.c_int,
        [c_void, ctypes.c_uint64, c_f64p, ctypes.c_uint64, ctypes.c_int32, c_i32p, c_f64p,
         ctypes.c_uint64, c_f64p, ctypes.c_uint32, c_f64p],
    ),
    "tsg_treeseq_trait_covariance": (
        ctypes.c_int,
        [c_void, ctypes.c_uint64, c_f64p, ctypes.c_uint64, c_f64p, ctypes.c_uint32, c_f64p],
    ),
    "tsg_treeseq_trait_correlation": (
        ctypes.c_int,
        [c_void, ctypes.c_uint64, c_f64p, ctypes.c_uint64, c_f64p, ctypes.c_uint32, c_f64p],
    ),
    "tsg_treeseq_trait_linear_model": (
        ctypes.c_int,
        [c_void, ctypes.c_uint64, c_f64p, ctypes.c_uint64, c_f64p, ctypes.c_uint64, c_f64p,
         ctypes.c_uint32, c_f64p],
    ),
    "tsg_treeseq_genetic_relatedness_weighted": (
        ctypes.c_int,
        [c_void, ctypes.c_uint64, c_f64p, ctypes.c_uint64, c_i32p, ctypes.c_uint64, c_f64p,
         c_f64p, ctypes.c_uint32],
    ),
    "tsg_batch_diversity": (ctypes.c_int, ONE_WAY),
    "tsg_batch_segregating_sites": (ctypes.c_int, ONE_WAY),
    "tsg_batch_Y1": (ctypes.c_int, ONE_WAY),
    "tsg_batch_Tajimas_D": (ctypes.c_int, ONE_WAY),
    "tsg_batch_divergence": (ctypes.c_int, K_WAY),
    "tsg_batch_Y2": (ctypes.c_int, K_WAY),
    "tsg_batch_f2": (ctypes.c_int, K_WAY),
    "tsg_batch_Y3": (ctypes.c_int, K_WAY),
    "tsg_batch_f3": (ctypes.c_int, K_WAY),
    "tsg_batch_f4": (ctypes.c_int, K_WAY),
    "tsg_batch_genetic_relatedness": (ctypes.c_int, K_WAY),
    "tsg_batch_genetic_relatedness_proportion": (ctypes.c_int, K_WAY),
    "tsg_batch_Fst": (ctypes.c_int, K_WAY),
    "tsg_sim_ancestry": (
        ctypes.c_int,
        [ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
         ctypes.POINTER(c_void)],
    ),
    "tsg_sim_mutations": (ctypes.c_int, [c_void, ctypes.c_double, ctypes.c_uint64]),
    "tsg_tables_get_sizes": (ctypes.c_int, [c_void, c_u64p]),
    "tsg_tables_copy_out": (
        ctypes.c_int,
        [c_void, c_u32p, c_f64p, c_f64p, c_f64p, c_i32p, c_i32p, c_f64p, c_i32p, c_i32p, c_i32p,
         c_f64p],
    ),
    "tsg_tables_free": (None, [c_void]),
}


def lib():
    """Returns the loaded library, loading it on first use."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"tskit_amd: {LIB_PATH} is not built; run `python -m tskit_amd.build` "
                "(or __graft_entry__.build())"
            )
        handle = ctypes.CDLL(LIB_PATH)
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = handle
    return _lib


def ptr(array, ctype):
    """Address of the data of a C-contiguous numpy array of element type
    ``ctype`` (None for None).  The caller keeps the array alive for the
    duration of the call it is passed to."""
    if array is None:
        return None
    return array.__array_interface__["data"][0]


def as_c(array, dtype):
    return np.ascontiguousarray(array, dtype=dtype)
