"""The .trees (kastore) reader and writer: round trips, files the
reference's own loader accepts, and the reference's data files
(tests/golden/trees, copied from the reference's tests and docs) read to the
same tree sequences as the reference's loader reads them."""
import os

import numpy as np
import pytest

import cases
import tskit_amd as ta
from oracle import oracle as om
from oracle.oracle import Oracle
from tskit_amd import kastore

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = ["basic_tree_seq", "minimal-example", "single-locus-example", "tskit-0.3.3"]
needs_ref = pytest.mark.skipif(not om.reference_available(), reason="oracle/_ref not built")


def _path(name):
    return os.path.join(HERE, "golden", "trees", name + ".trees")


def _same_tables(a, b):
    for col in ["node_flags", "node_time", "edge_left", "edge_right", "edge_parent", "edge_child",
                "site_position", "ancestral_state", "ancestral_state_offset", "mutation_site",
                "mutation_node", "mutation_parent", "derived_state", "derived_state_offset"]:
        x, y = np.asarray(getattr(a, col)), np.asarray(getattr(b, col))
        assert x.shape == y.shape and np.array_equal(x.astype(np.float64), y.astype(np.float64)), col
    assert a.sequence_length == b.sequence_length and a.time_units == b.time_units


def test_roundtrip(tmp_path):
    tc = cases.generated(60, 5e4, 7)
    p = tmp_path / "x.trees"
    tc.dump(p)
    _same_tables(tc, ta.load_tables(p))
    ts = ta.load(p)
    assert ts.num_trees == tc.tree_sequence().num_trees


def test_kastore_items_roundtrip(tmp_path):
    items = {"b/x": np.arange(5, dtype=np.int32), "a": np.array([1.5], np.float64),
             "c": np.zeros(0, np.uint8), "z/k": np.array([3, 4], np.uint64)}
    p = tmp_path / "k.kas"
    kastore.write(p, items)
    got = kastore.read(p)
    assert set(got) == set(items)
    for k in items:
        assert got[k].dtype == items[k].dtype and np.array_equal(got[k], items[k])


def test_bad_file(tmp_path):
    p = tmp_path / "bad.trees"
    p.write_bytes(b"not a kastore file at all" * 4)
    with pytest.raises(kastore.FileFormatError):
        ta.load_tables(p)


def _stats(impl, n, L, has_sites):
    half = [np.arange(n // 2), np.arange(n // 2, n)]
    w = np.linspace(0, L, 3)
    out = []
    for mode in (["branch", "site"] if has_sites else ["branch"]):
        out.append(impl.stat("diversity", half, None, w, mode))
        out.append(impl.stat("divergence", half, [(0, 1)], w, mode))
        out.append(impl.afs(half, w, None, mode, True, False))
    return out


@needs_ref
def test_reference_loads_our_file(tmp_path):
    tc = cases.generated(50, 4e4, 9)
    p = tmp_path / "ours.trees"
    tc.dump(p)
    a = _stats(om.Reference(tc), 50, tc.sequence_length, True)
    b = _stats(om.Reference(tc, path=p), 50, tc.sequence_length, True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@needs_ref
@pytest.mark.parametrize("name", FILES)
def test_reference_files(name):
    """Our reader and the reference's loader give the same statistics (the
    oracle on our tables, the reference on its own load of the file)."""
    tc = ta.load_tables(_path(name))
    ts = tc.tree_sequence()
    n = ts.num_samples
    samples = ts.samples()
    ref = om.Reference(tc, path=_path(name))
    orc = Oracle(tc)
    half = [samples[: n // 2], samples[n // 2:]]
    w = np.linspace(0, tc.sequence_length, 3)
    modes = ["branch", "site"] if tc.num_sites else ["branch"]
    for mode in modes:
        for stat, idx in [("diversity", None), ("divergence", [(0, 1)])]:
            assert np.allclose(orc.stat(stat, half, idx, w, mode),
                               ref.stat(stat, half, idx, w, mode), rtol=1e-12, atol=0)
        assert np.allclose(orc.afs(half, w, None, mode, True, False),
                           ref.afs(half, w, None, mode, True, False), rtol=1e-12, atol=0)
    if tc.num_sites:
        assert np.array_equal(orc.genotype_matrix(), ref.genotype_matrix())
