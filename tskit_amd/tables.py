"""Columnar tables: the inputs of a tree sequence.

A deliberately small subset of the reference's ``TableCollection``
(python/tskit/tables.py): the node, edge, site and mutation columns the
statistics path reads, plus ``tree_sequence()`` which hands them to the
device (python/tskit/tables.py:3525-3541).
"""
import numpy as np

NODE_IS_SAMPLE = 1


def _ragged(strings):
    """Encodes a list of byte strings as (data, offsets) columns."""
    enc = [s.encode() if isinstance(s, str) else bytes(s) for s in strings]
    offsets = np.zeros(len(enc) + 1, dtype=np.uint64)
    if enc:
        offsets[1:] = np.cumsum([len(s) for s in enc])
    data = np.frombuffer(b"".join(enc), dtype=np.int8) if enc else np.zeros(0, dtype=np.int8)
    return data.copy(), offsets


class TableCollection:
    """Node, edge, site and mutation columns (numpy arrays)."""

    def __init__(self, sequence_length):
        self.sequence_length = float(sequence_length)
        self.time_units = "unknown"
        self.node_flags = np.zeros(0, dtype=np.uint32)
        self.node_time = np.zeros(0, dtype=np.float64)
        self.edge_left = np.zeros(0, dtype=np.float64)
        self.edge_right = np.zeros(0, dtype=np.float64)
        self.edge_parent = np.zeros(0, dtype=np.int32)
        self.edge_child = np.zeros(0, dtype=np.int32)
        self.site_position = np.zeros(0, dtype=np.float64)
        self.ancestral_state = np.zeros(0, dtype=np.int8)
        self.ancestral_state_offset = np.zeros(1, dtype=np.uint64)
        self.mutation_site = np.zeros(0, dtype=np.int32)
        self.mutation_node = np.zeros(0, dtype=np.int32)
        self.mutation_parent = np.zeros(0, dtype=np.int32)
        self.derived_state = np.zeros(0, dtype=np.int8)
        self.derived_state_offset = np.zeros(1, dtype=np.uint64)

    # -- row-wise construction -------------------------------------------------
    def set_nodes(self, flags, time):
        self.node_flags = np.asarray(flags, dtype=np.uint32).copy()
        self.node_time = np.asarray(time, dtype=np.float64).copy()

    def set_edges(self, left, right, parent, child):
        self.edge_left = np.asarray(left, dtype=np.float64).copy()
        self.edge_right = np.asarray(right, dtype=np.float64).copy()
        self.edge_parent = np.asarray(parent, dtype=np.int32).copy()
        self.edge_child = np.asarray(child, dtype=np.int32).copy()

    def set_sites(self, position, ancestral_state=None):
        self.site_position = np.asarray(position, dtype=np.float64).copy()
        if ancestral_state is None:
            ancestral_state = ["0"] * len(self.site_position)
        self.ancestral_state, self.ancestral_state_offset = _ragged(ancestral_state)

    def set_mutations(self, site, node, derived_state=None, parent=None):
        self.mutation_site = np.asarray(site, dtype=np.int32).copy()
        self.mutation_node = np.asarray(node, dtype=np.int32).copy()
        n = len(self.mutation_site)
        if parent is None:
            parent = np.full(n, -1, dtype=np.int32)
        self.mutation_parent = np.asarray(parent, dtype=np.int32).copy()
        if derived_state is None:
            derived_state = ["1"] * n
        self.derived_state, self.derived_state_offset = _ragged(derived_state)

    @property
    def num_nodes(self):
        return len(self.node_time)

    @property
    def num_edges(self):
        return len(self.edge_left)

    @property
    def num_sites(self):
        return len(self.site_position)

    @property
    def num_mutations(self):
        return len(self.mutation_site)

    def sort_edges(self):
        """Sorts edges by (time[parent], parent, child, left), the order the
        reference requires (c/tskit/tables.c:10466-10490)."""
        tp = self.node_time[self.edge_parent]
        order = np.lexsort((self.edge_left, self.edge_child, self.edge_parent, tp))
        self.edge_left = self.edge_left[order]
        self.edge_right = self.edge_right[order]
        self.edge_parent = self.edge_parent[order]
        self.edge_child = self.edge_child[order]

    def tree_sequence(self):
        from .trees import TreeSequence

        return TreeSequence.load_tables(self)

    def dump(self, path):
        """Writes the tables as a ``.trees`` file (kastore; see kastore.py)."""
        from . import kastore

        kastore.dump(self, path)

    def column_digest(self):
        """SHA-256 over the columns, to check two hosts generated the same input."""
        import hashlib

        h = hashlib.sha256()
        for a in (
            self.node_flags, self.node_time, self.edge_left, self.edge_right,
            self.edge_parent, self.edge_child, self.site_position, self.mutation_site,
            self.mutation_node, self.mutation_parent,
        ):
            h.update(np.ascontiguousarray(a).tobytes())
        h.update(np.float64(self.sequence_length).tobytes())
        return h.hexdigest()


def _lines(text):
    return [ln.split() for ln in text.strip().splitlines() if ln.strip()]


def parse_text(sequence_length, nodes, edges, sites=None, mutations=None):
    """Builds tables from the whitespace text format of the reference's C
    test fixtures (c/tests/testlib.c:25-140): nodes ``flags time population
    individual``; edges ``left right parent child[,child...]``; sites
    ``position ancestral_state``; mutations ``site node derived_state
    [parent]``."""
    tc = TableCollection(sequence_length)
    rows = _lines(nodes)
    tc.set_nodes([int(r[0]) for r in rows], [float(r[1]) for r in rows])
    el, er, ep, ec = [], [], [], []
    for r in _lines(edges):
        for child in r[3].split(","):
            el.append(float(r[0]))
            er.append(float(r[1]))
            ep.append(int(r[2]))
            ec.append(int(child))
    tc.set_edges(el, er, ep, ec)
    if sites:
        rows = _lines(sites)
        tc.set_sites([float(r[0]) for r in rows], [r[1] for r in rows])
    if mutations:
        rows = _lines(mutations)
        tc.set_mutations(
            [int(r[0]) for r in rows],
            [int(r[1]) for r in rows],
            [r[2] for r in rows],
            [int(r[3]) if len(r) > 3 else -1 for r in rows],
        )
    return tc
