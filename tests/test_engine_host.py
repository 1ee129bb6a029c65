"""The device DFS engine (csrc/nq_task.h + nq_dfs.h: decode, schedule words,
expansion, walk_step with both stacks) built for the host and compared with
the oracle record by record: weighted counts and node counts, every split
depth, LDS and private stacks, and the stack-slot bound the kernels allocate."""
import ctypes
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from nqueens import constellations as C

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "host_engine" / "engine_host.cpp"


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    return None


@pytest.fixture(scope="module")
def engine(tmp_path_factory):
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("engine") / "engine_host.so"
    cmd = [hipcc, "--cuda-host-only", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", str(out), str(SRC)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    lib = ctypes.CDLL(str(out))
    u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
    u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
    lib.engine_count.argtypes = [ctypes.c_int, u32p, ctypes.c_long, u64p, u64p, ctypes.c_int, ctypes.c_int]
    lib.engine_count.restype = ctypes.c_int
    lib.engine_count_splits.argtypes = [ctypes.c_int, u32p, ctypes.c_long, u64p, u64p,
                                        np.ctypeslib.ndpointer(np.uint8, flags="C")]
    lib.engine_count_splits.restype = ctypes.c_int
    lib.engine_count_gate.argtypes = [ctypes.c_int, u32p, ctypes.c_long, u64p, u64p, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int]
    lib.engine_count_gate.restype = ctypes.c_int
    lib.engine_p5_records.argtypes = [ctypes.c_int, u32p, ctypes.c_long, ctypes.c_int]
    lib.engine_p5_records.restype = ctypes.c_long
    lib.engine_expand.argtypes = [ctypes.c_int, u32p, ctypes.c_long, u32p, ctypes.c_long]
    lib.engine_expand.restype = ctypes.c_long
    return lib


def _run(engine, N, rec, split, priv):
    rec = np.ascontiguousarray(rec, dtype=np.uint32)
    w = np.zeros(len(rec), np.uint64)
    n = np.zeros(len(rec), np.uint64)
    rc = engine.engine_count(N, rec, len(rec), w, n, split, priv)
    return rc, w, n


@pytest.mark.parametrize("N,preset", [(N, p) for N in range(7, 13) for p in (4, 5, 6)] + [(13, 5), (14, 4)])
def test_engine_matches_oracle(engine, oracle, N, preset):
    rec = C.gen_constellations(N, preset)
    w, _, nodes = oracle.count_records(N, rec)
    for split in (0, 1, 2, 3):
        for priv in (0, 1):
            rc, gw, gn = _run(engine, N, rec, split, priv)
            assert rc == 0, (split, priv, rc)
            assert np.array_equal(gw, w), (split, priv)
            assert np.array_equal(gn, nodes), (split, priv)


@pytest.mark.parametrize("N", [20, 21])
def test_engine_largest_boards_sample(engine, N):
    # records of the N=20/21 oracle fixture (terminal depth 14-15, long ld
    # histories: the lookahead's occupancy word fills all 32 bits)
    from tests.conftest import GOLDEN
    with np.load(GOLDEN / "oracle_sample_n20_n21.npz") as z:
        k = 12 if N == 20 else 40
        rec, w, nodes = z[f"rec{N}"][::k], z[f"w{N}"][::k], z[f"nodes{N}"][::k]
    rc, gw, gn = _run(engine, N, rec, 3, 0)
    assert rc == 0
    assert np.array_equal(gw, w) and np.array_equal(gn, nodes)


def test_engine_n24_sample(engine):
    # N=24 at preset 7 (terminal depth 16, 22-bit candidate sets, col's
    # board bits reach bit 23 under the parameter byte)
    from tests.conftest import GOLDEN
    with np.load(GOLDEN / "oracle_sample_n24.npz") as z:
        rec, w, nodes = z["rec24"][::28], z["w24"][::28], z["nodes24"][::28]
    rc, gw, gn = _run(engine, 24, rec, 1, 0)
    assert rc == 0
    assert np.array_equal(gw, w) and np.array_equal(gn, nodes)


@pytest.mark.parametrize("N", [22, 23])
def test_engine_n22_n23_sample(engine, N):
    # N=22 (preset 5, terminal depth 16, 20-bit candidate sets) and N=23
    # (preset 6, terminal depth 16, 21-bit candidate sets): the widest boards
    # the u32 frame layout holds
    from tests.conftest import GOLDEN
    with np.load(GOLDEN / "oracle_sample_n22_n23.npz") as z:
        k = 32 if N == 22 else 52
        rec, w, nodes = z[f"rec{N}"][::k], z[f"w{N}"][::k], z[f"nodes{N}"][::k]
    split, priv = (3, 0) if N == 22 else (5, 1)
    rc, gw, gn = _run(engine, N, rec, split, priv)
    assert rc == 0
    assert np.array_equal(gw, w) and np.array_equal(gn, nodes)


@pytest.mark.parametrize("N,lo,hi", [(12, 1, 8), (14, 1, 8), (14, 2, 12), (13, 0, 8)])
def test_engine_per_record_splits_set_layout(engine, oracle, N, lo, hi):
    """Per-record split depths (some beyond T-2, so those records walk from
    their roots) with every LDS stack sized for the whole set the way the
    library sizes it: no push goes past the allocated slots, and counts and
    node counts equal the oracle's."""
    rec = C.gen_constellations(N, 5)
    w, _, nodes = oracle.count_records(N, rec)
    depth = np.random.default_rng(N + lo).integers(lo, hi, rec.shape[0]).astype(np.uint8)
    T = N - 1 - 5
    assert (depth > T - 2).any() and (depth <= T - 2).any()
    gw = np.zeros(len(rec), np.uint64)
    gn = np.zeros(len(rec), np.uint64)
    rc = engine.engine_count_splits(N, np.ascontiguousarray(rec, np.uint32), len(rec), gw, gn, depth)
    assert rc == 0, rc
    assert np.array_equal(gw, w) and np.array_equal(gn, nodes)


@pytest.mark.parametrize("N", [9, 11, 13, 15, 16])
def test_engine_p5_records_under_narrow_gate(engine, oracle, N):
    """With build_soa's j-diagonal gate narrowed from 32 bits to W (N <= W),
    thousands of canonical records take the P5 functions (8..11) and insert
    that diagonal at its switch row: the device engine's weighted counts and
    node counts equal the oracle's under the same gate, at every split depth
    and with both stacks, and the weighted counts equal the 32-bit ones."""
    rec = np.ascontiguousarray(C.gen_constellations(N, 5), np.uint32)
    w32, _, _ = oracle.count_records(N, rec)
    for W in (N, N + 1, N + 3):
        n5 = engine.engine_p5_records(N, rec, len(rec), W)
        assert n5 > 0 or W > N
        with oracle.gate_bits(W):
            want, _, nodes = oracle.count_records(N, rec)
        assert np.array_equal(want, w32)
        for split, priv in ((0, 0), (2, 0), (3, 1)):
            gw = np.zeros(len(rec), np.uint64)
            gn = np.zeros(len(rec), np.uint64)
            rc = engine.engine_count_gate(N, rec, len(rec), gw, gn, split, priv, W)
            assert rc == 0, (W, split, rc)
            assert np.array_equal(gw, want), (W, split)
            assert np.array_equal(gn, nodes), (W, split)


def _expand(engine, N, rec):
    rec = np.ascontiguousarray(rec, dtype=np.uint32)
    out = np.zeros((rec.shape[0] * N + 1, 4), np.uint32)
    n = engine.engine_expand(N, rec, rec.shape[0], out, out.shape[0])
    assert n >= 0, n
    return out[:n]


@pytest.mark.parametrize("N,p", [(N, 4) for N in (8, 11, 14)] + [(N, 5) for N in range(7, 18)] + [(12, 6), (15, 6)])
def test_preset_expansion_matches_generator(engine, N, p):
    # one level of the device expansion (nqt::preset_children) turns the
    # preset-p records into exactly the preset-(p+1) record multiset
    got = _expand(engine, N, C.gen_constellations(N, p))
    want = C.gen_constellations(N, p + 1)
    assert np.array_equal(C.records_sorted(got), C.records_sorted(want))


def test_preset_expansion_reference_counts(engine, known):
    # record counts of the reference's own logs: N=18 preset 6 (571,705) and
    # preset 7 (3,653,940), N=21 preset 6 (2,025,282)
    p6 = _expand(engine, 18, C.gen_constellations(18, 5))
    assert p6.shape[0] == 571705
    assert _expand(engine, 18, p6).shape[0] == 3653940
    assert _expand(engine, 21, C.gen_constellations(21, 5)).shape[0] == 2025282
