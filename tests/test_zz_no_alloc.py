"""The no-allocation entry point, tsg_batch_run_noalloc: through the C ABI,
its results equal tsg_batch_run's bit for bit (the same kernels in the same
order over buffers carved from the library's fixed device scratch), on two
calls in a row, the second with a smaller batch; it refuses a handle that is
not resident and a batch larger than its scratch."""
import numpy as np
import pytest

import cases
from tskit_amd import _lib

pytestmark = pytest.mark.gpu

NOT_RESIDENT = -30007
SCRATCH_TOO_SMALL = -30006


def _queue(b, sets, w, quads):
    pairs = [(i, j) for i in range(len(sets)) for j in range(i + 1, len(sets))]
    return [b.diversity(sets, windows=w, mode="branch"),
            b.Fst(sets, indexes=pairs, windows=w, mode="branch"),
            b.Tajimas_D(sets, windows=w, mode="site"),
            b.f4(sets, indexes=quads, windows=w, mode="site"),
            b.divergence(sets, indexes=pairs[:2], windows=w, mode="branch",
                         span_normalise=False)]


def _noalloc(lib, b):
    rc = lib.tsg_batch_run_noalloc(b._ll._b)
    if rc == 0:
        assert lib.tsg_batch_clear(b._ll._b) == 0
        b.done = True
    return rc


def test_noalloc_equals_batch_run(require_gpu):
    lib = _lib.lib()
    tc = cases.generated(3000, 1e6, 13)
    ts = tc.tree_sequence()
    fresh = ts.batch()
    fresh.diversity(cases.sample_sets(3000, 2, 50, 1), mode="branch")
    assert _noalloc(lib, fresh) == NOT_RESIDENT
    ts.ll_tree_sequence.upload()
    L = tc.sequence_length
    for step, (k, size, nw) in enumerate([(6, 300, 40), (4, 50, 7)]):
        sets = cases.sample_sets(3000, k, size, 20 + step)
        w = np.linspace(0, L, nw + 1)
        quads = [(0, 1, 2, 3), (3, 2, 1, 0)]
        b1 = ts.batch()
        p1 = _queue(b1, sets, w, quads)
        b1.run()
        b2 = ts.batch()
        p2 = _queue(b2, sets, w, quads)
        print(f"[case] noalloc call {step + 1}: {k} sets of {size}, {nw} windows", flush=True)
        assert _noalloc(lib, b2) == 0
        for x, y in zip(p1, p2):
            assert np.array_equal(np.asarray(x.result()), np.asarray(y.result()),
                                  equal_nan=True)
    big = ts.batch()
    sets = cases.sample_sets(3000, 40, 20, 3)
    big.diversity(sets, windows=np.linspace(0, L, 200001), mode="site")
    print("[case] noalloc batch beyond the scratch", flush=True)
    assert _noalloc(lib, big) == SCRATCH_TOO_SMALL
