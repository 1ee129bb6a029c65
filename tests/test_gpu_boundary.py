"""The one-shot C entry point nq_count (what a reference binding calls in
place of exec_solutions, 18Py:1057) through ctypes: plain and batched
schedules against the oracle, its error returns, the CPython snippet of
INTEGRATION.md run as written, and exec_solutions with either use_gpu."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import nqueens
from nqueens import constellations as C
from nqueens._lib import Sched

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
SENTINEL = 0xDEADBEEF12345678


def _nq_count(ctx, N, rec, sched, nodes=True):
    lib = nqueens.load_library()
    rec = np.ascontiguousarray(rec, np.uint32)
    w = np.full(len(rec), SENTINEL, np.uint64)
    n = np.full(len(rec), SENTINEL, np.uint64) if nodes else None
    total = ctypes.c_uint64(SENTINEL)
    sp = ctypes.byref(sched) if sched is not None else None
    rc = lib.nq_count(ctx.h, N, rec, len(rec), sp, w.ctypes.data_as(ctypes.c_void_p),
                      n.ctypes.data_as(ctypes.c_void_p) if nodes else None, ctypes.byref(total))
    return rc, w, n, total.value


@pytest.mark.parametrize("N", [12, 13])
@pytest.mark.parametrize("kind", ["plain", "batched", "batched-graph"])
def test_nq_count_against_oracle(ctx, oracle, N, kind):
    rec = C.gen_constellations(N, 5)
    want, _, nodes = oracle.count_records(N, rec)
    sched = None
    if kind != "plain":
        sched = Sched(split_depth=3, block=128, lds_stack=1, use_graph=int(kind == "batched-graph"), chunk=64,
                      waves_per_eu=0)
    rc, w, n, total = _nq_count(ctx, N, rec, sched)
    assert rc == 0, nqueens.load_library().nq_last_error()
    assert np.array_equal(w, want) and np.array_equal(n, nodes)
    assert total == int(want.sum())
    # without a node buffer the counts are the same
    rc, w2, _, total2 = _nq_count(ctx, N, rec, sched, nodes=False)
    assert rc == 0 and np.array_equal(w2, want) and total2 == total


def test_nq_count_errors_leave_outputs_untouched(ctx):
    lib = nqueens.load_library()
    rec = C.gen_constellations(12, 5)
    for N in (5, 40):                                  # outside the constellation path
        rc, w, n, total = _nq_count(ctx, N, rec, None)
        assert rc == -1 and total == SENTINEL and (w == SENTINEL).all()
        assert b"N=" in lib.nq_last_error()
    # a record whose start row lies below the board: the decode refuses it
    bad = rec[:3].copy()
    bad[1, 3] = (30 << 20) | (int(bad[1, 3]) & 0xFFFFF)
    for sched in (None, Sched(split_depth=3, block=128, lds_stack=1, use_graph=0, chunk=64, waves_per_eu=0)):
        rc, w, n, total = _nq_count(ctx, 12, bad, sched)
        assert rc == -2 and total == SENTINEL and (w == SENTINEL).all() and (n == SENTINEL).all()
        assert b"decode" in lib.nq_last_error()
    # a bad schedule is refused before anything is counted
    rc, w, _, total = _nq_count(ctx, 12, rec, Sched(split_depth=3, block=96, lds_stack=1, use_graph=0, chunk=64,
                                                   waves_per_eu=0))
    assert rc == -1 and total == SENTINEL
    # the context still counts afterwards
    rc, w, _, total = _nq_count(ctx, 12, rec, None, nodes=False)
    assert rc == 0 and total == 14200


def test_integration_ctypes_snippet_as_written(oracle, monkeypatch):
    """The CPython block of INTEGRATION.md, executed verbatim (from the
    repository root, as its library path assumes)."""
    text = (ROOT / "INTEGRATION.md").read_text()
    sect = text.split("## From CPython", 1)[1]
    code = re.search(r"```python\n(.*?)```", sect, re.S).group(1)
    N = 13
    records = C.gen_constellations(N, 5)
    monkeypatch.chdir(ROOT)
    env = {"records": records, "N": N}
    exec(compile(code, "INTEGRATION.md", "exec"), env)
    want, _, _ = oracle.count_records(N, records)
    assert env["rc"] == 0
    assert np.array_equal(env["w"], want) and env["total"].value == 73712


@pytest.mark.parametrize("use_gpu", [True, False])
def test_exec_solutions_either_flag(p5_full, use_gpu):
    rec, sol = p5_full["rec12"], p5_full["sol12"]
    cs = [{"ld": int(r[0]), "rd": int(r[1]), "col": int(r[2]), "startijkl": int(r[3]), "solutions": 0} for r in rec]
    nqueens.exec_solutions(12, cs, use_gpu)
    assert [c["solutions"] for c in cs] == sol.tolist()
    assert sum(c["solutions"] for c in cs) == 14200


def test_small_boards_and_arguments(ctx):
    """N = 1 counts on the small-board kernel (one solution, one orbit); the
    small-board branch takes no chunks and reports its one run."""
    assert nqueens.count(1, ctx=ctx) == (1, 1)
    seen = []
    assert nqueens.count(5, ctx=ctx, progress=lambda i, k, t: seen.append((i, k, t))) == (10, 2)
    assert seen == [(0, 1, 10)]


@pytest.mark.parametrize("N", range(12, 20))
def test_count_through_reference_caches(ctx, digests, tmp_path, monkeypatch, N):
    """count(N, cache_dir=...) writes the reference's cache files: the list
    validates (.done = record count), the solutions v2 file read back holds
    the reference fixture's per-constellation counts (SHA-256 of the sorted
    stream), a second count is served from the files without a device count,
    and a damaged solutions file is counted again and rewritten."""
    from nqueens import cache, solver
    from tests.helpers import solutions_sha256
    known_total = digests[str(N)]["solutions_sum"]
    msgs = []
    total, unique = nqueens.count(N, preset=5, ctx=ctx, cache_dir=tmp_path, log=msgs.append)
    assert total == known_total and msgs == []
    rec = cache.load_constellations(tmp_path, N, 5)
    assert rec is not None and rec.shape[0] == digests[str(N)]["constellations"]
    frec, fsol = nqueens.constellations.read_solutions_bin(cache.solutions_path(tmp_path, N, 5))
    assert solutions_sha256(frec, fsol) == digests[str(N)]["solutions_sha256"]

    def no_device_count(*a, **k):
        raise AssertionError("the cache should have answered")
    with monkeypatch.context() as mp:
        mp.setattr(solver, "_count_list", no_device_count)
        assert nqueens.count(N, preset=5, ctx=ctx, cache_dir=tmp_path) == (total, unique)
    p = cache.solutions_path(tmp_path, N, 5)
    p.write_bytes(p.read_bytes()[:-40])
    assert nqueens.count(N, preset=5, ctx=ctx, cache_dir=tmp_path, log=msgs.append) == (total, unique)
    assert len(msgs) == 1 and "counting again" in msgs[0]
    frec, fsol = nqueens.constellations.read_solutions_bin(p)
    assert solutions_sha256(frec, fsol) == digests[str(N)]["solutions_sha256"]
