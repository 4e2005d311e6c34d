"""tskit_amd: the tree-sequence statistics hot path of tskit on MI355X.

A drop-in for the ``general_stat``-backed statistics of ``tskit.TreeSequence``
(diversity, divergence, Fst, Tajima's D, f2/f3/f4, allele frequency spectrum,
genetic relatedness, ...), computed by hand-written HIP kernels behind a C ABI
(``libtsgpu.so``, declared in ``include/tsgpu.h``).
"""
from .exceptions import LibraryError, TskitException
from .kastore import load as load_tables
from .sim import random_sample_sets, sim_ancestry, simulate_tables
from .tables import NODE_IS_SAMPLE, TableCollection, parse_text
from .trees import StatBatch, TreeSequence, run_batches

def load(path):
    """A TreeSequence from a ``.trees`` file (as tskit.load)."""
    return load_tables(path).tree_sequence()


__all__ = [
    "load",
    "load_tables",
    "LibraryError",
    "TskitException",
    "TableCollection",
    "TreeSequence",
    "StatBatch",
    "run_batches",
    "NODE_IS_SAMPLE",
    "parse_text",
    "sim_ancestry",
    "simulate_tables",
    "random_sample_sets",
]
