"""nq_count_noalloc (the count of nq_count with every device buffer in the
library's static scratch array) against nq_count through the C ABI: same
weighted counts, node counts and totals, on consecutive calls with a smaller
input second (the array keeps the previous call's contents), for the plain
and the batched kernel; its size limits; and the context's resident set
left untouched."""
import ctypes

import numpy as np
import pytest

import nqueens
from nqueens import constellations as C
from nqueens._lib import Sched

pytestmark = pytest.mark.gpu

SENTINEL = 0x5A5A5A5A5A5A5A5A


def _call(name, ctx, N, rec, sched, nodes=True):
    fn = getattr(nqueens.load_library(), name)
    rec = np.ascontiguousarray(rec, np.uint32)
    w = np.full(len(rec), SENTINEL, np.uint64)
    n = np.full(len(rec), SENTINEL, np.uint64) if nodes else None
    total = ctypes.c_uint64(SENTINEL)
    rc = fn(ctx.h, N, rec, len(rec), ctypes.byref(sched) if sched is not None else None,
            w.ctypes.data_as(ctypes.c_void_p), n.ctypes.data_as(ctypes.c_void_p) if nodes else None,
            ctypes.byref(total))
    return rc, w, n, total.value


def _sched(split, block=128):
    return Sched(split_depth=split, block=block, lds_stack=1, use_graph=1, chunk=64, waves_per_eu=0)


@pytest.mark.parametrize("kind", ["plain", "batched-2", "batched-3"])
def test_noalloc_matches_nq_count_on_consecutive_calls(ctx, oracle, kind):
    sched = None if kind == "plain" else _sched(int(kind[-1]))
    for N, preset in ((15, 5), (11, 4), (13, 5)):      # the second call is smaller than the first
        rec = C.gen_constellations(N, preset)
        want, _, nodes = oracle.count_records(N, rec)
        rc0, w0, n0, t0 = _call("nq_count", ctx, N, rec, sched)
        rc1, w1, n1, t1 = _call("nq_count_noalloc", ctx, N, rec, sched)
        assert rc0 == 0 and rc1 == 0, nqueens.load_library().nq_last_error()
        assert np.array_equal(w1, w0) and np.array_equal(n1, n0) and t1 == t0
        assert np.array_equal(w1, want) and np.array_equal(n1, nodes) and t1 == int(want.sum())
        # without node counts
        rc2, w2, _, t2 = _call("nq_count_noalloc", ctx, N, rec, sched, nodes=False)
        assert rc2 == 0 and np.array_equal(w2, want) and t2 == t1


def test_noalloc_large_set(ctx, digests):
    """N=19, preset 5 (100,706 records) at split depth 1, every record's
    count against the reference fixture's digest."""
    from tests.helpers import solutions_sha256
    N = 19
    rec = C.gen_constellations(N, 5)
    rc, w, n, total = _call("nq_count_noalloc", ctx, N, rec, _sched(1, 256), nodes=False)
    assert rc == 0, nqueens.load_library().nq_last_error()
    assert total == digests[str(N)]["solutions_sum"]
    assert solutions_sha256(rec, w) == digests[str(N)]["solutions_sha256"]


def test_noalloc_limits_and_resident_set(ctx, oracle):
    lib = nqueens.load_library()
    # the resident set of the context survives noalloc calls
    rec12 = C.gen_constellations(12, 5)
    ctx.load(12, rec12)
    # more records than the array holds: refused before any device work
    big = np.zeros(((1 << 18) + 1, 4), np.uint32)
    rc, w, _, total = _call("nq_count_noalloc", ctx, 12, big, None)
    assert rc == -3 and total == SENTINEL and b"at most" in lib.nq_last_error()
    # more work items than the array holds (N=19 split 3: 37.7 M items)
    rec19 = C.gen_constellations(19, 5)
    rc, w, _, total = _call("nq_count_noalloc", ctx, 19, rec19, _sched(3, 256), nodes=False)
    assert rc == -3 and total == SENTINEL and (w == SENTINEL).all()
    assert b"work items" in lib.nq_last_error()
    # bad arguments
    assert _call("nq_count_noalloc", ctx, 40, rec12, None)[0] == -1
    assert _call("nq_count_noalloc", ctx, 12, rec12, _sched(-1))[0] == -1
    rc, _, _, total = _call("nq_count_noalloc", ctx, 12, rec12[:0], None)
    assert rc == 0 and total == 0
    # the loaded N=12 set still counts
    t, _ = ctx.run_batched(_sched(3))
    gw, _ = ctx.results()
    want, _, _ = oracle.count_records(12, rec12)
    assert t == 14200 and np.array_equal(gw, want)
