"""ctypes wrapper of the CPU oracle (liboracle.so, built from nq_oracle.c).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker: tests, the package
smoke check and the CPU-baseline leg of bench.py load it to compute the
reference answers.  The product (n-queens_amd/) never imports this module.

Functions follow 13Bit_codon/18Py_constellations_GPU_cuda_codon.py (see the
header of nq_oracle.c for the per-function citations and parity pins).
The optional reference build in _ref/ (oracle/ref.mk) wraps the reference's
own C solver 10Bit_CUDA/004GCC_constellation.c (preset 4).
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref004.so"

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_lib = None
_ref = None


def build(ref: bool = False) -> None:
    """make -C oracle (and the reference build when ref=True)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", str(HERE), "-f", "ref.mk"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        L.nqo_gen_constellations_scoped.restype = ctypes.c_long
        L.nqo_gen_constellations_scoped.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                    ctypes.POINTER(ctypes.POINTER(ctypes.c_uint32)),
                                                    ctypes.POINTER(ctypes.c_long)]
        L.nqo_free.argtypes = [ctypes.c_void_p]
        L.nqo_count_records.argtypes = [ctypes.c_int, _u32p, ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.nqo_count_records_recursive.argtypes = [ctypes.c_int, _u32p, ctypes.c_long, _u64p, _u64p]
        L.nqo_count_records_13py.argtypes = [ctypes.c_int, _u32p, ctypes.c_long, _u64p, _u64p, ctypes.c_int]
        L.nqo_build_soa.argtypes = [ctypes.c_int, _u32p, ctypes.c_long, _i32p]
        L.nqo_symmetric_counts.argtypes = [ctypes.c_int, _u64p, _u64p]
        L.nqo_bruteforce.argtypes = [ctypes.c_int, _u64p]
        L.nqo_expand_records.restype = ctypes.c_long
        L.nqo_expand_records.argtypes = [ctypes.c_int, _u32p, ctypes.c_long, ctypes.c_int, _u32p, ctypes.c_long]
        L.nqo_set_gate_bits.argtypes = [ctypes.c_int]
        L.nqo_jasmin.argtypes = [ctypes.c_int, ctypes.c_int]
        L.nqo_symmetry.argtypes = [ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


SCOPE_CALL, SCOPE_START, SCOPE_NONE = 0, 1, 2


def gen_constellations(N: int, preset: int = 5, scope: int = SCOPE_NONE):
    """(records (m,4) u32, stats [starts, memo hits, cross-start hits, signature hits])."""
    L = lib()
    p = ctypes.POINTER(ctypes.c_uint32)()
    st = (ctypes.c_long * 4)()
    n = L.nqo_gen_constellations_scoped(N, preset, scope, ctypes.byref(p), st)
    if n < 0:
        L.nqo_free(p)
        raise OverflowError("oracle: record field exceeds 32 bits")
    a = np.ctypeslib.as_array(p, shape=(max(n, 1) * 4,)).copy().reshape(-1, 4)[:n]
    L.nqo_free(p)
    return np.ascontiguousarray(a), list(st)


def expand_records(N: int, rec, levels: int = 1) -> np.ndarray:
    """Records of preset p + levels reached from each record of preset p
    (set_pre_queens continued from the record; input order, columns ascending)."""
    rec = np.ascontiguousarray(np.asarray(rec, dtype=np.uint32))
    cap = max(1, rec.shape[0]) * N ** levels
    out = np.zeros(cap * 4, np.uint32)
    n = lib().nqo_expand_records(N, rec, rec.shape[0], levels, out, cap)
    if n < 0:
        raise OverflowError(f"oracle expand_records failed ({n})")
    return out[: 4 * n].reshape(-1, 4).copy()


def count_records(N: int, rec, threads: int = 0, schedule: int = 1, nodes: bool = True):
    """(weighted u64[m], raw counts u64[m], nodes u64[m] or None) via dfs_iter."""
    rec = np.ascontiguousarray(np.asarray(rec, dtype=np.uint32))
    m = rec.shape[0]
    w = np.zeros(m, np.uint64)
    raw = np.zeros(m, np.uint64)
    nn = np.zeros(m, np.uint64) if nodes else None
    rc = lib().nqo_count_records(N, rec, m, w.ctypes.data, raw.ctypes.data,
                                 nn.ctypes.data if nodes else None, threads, schedule)
    if rc:
        raise RuntimeError(f"oracle count failed: {rc}")
    return w, raw, nn


def count_records_recursive(N: int, rec):
    rec = np.ascontiguousarray(np.asarray(rec, dtype=np.uint32))
    m = rec.shape[0]
    w = np.zeros(m, np.uint64)
    nn = np.zeros(m, np.uint64)
    rc = lib().nqo_count_records_recursive(N, rec, m, w, nn)
    if rc:
        raise RuntimeError(f"oracle recursive count failed: {rc}")
    return w, nn


PLANT_BASE, PLANT_JMARK = 2, 4


def count_records_13py(N: int, rec, plant: int = 0):
    """(weighted, nodes) under the pruning of the reference's recursive SQ*
    functions (13Bit_codon/13Py_constellations_codon.py:183-678).  plant
    (tests only): PLANT_BASE counts every base frame as a node, PLANT_JMARK
    counts column 0 before the jmark mask."""
    rec = np.ascontiguousarray(np.asarray(rec, dtype=np.uint32))
    m = rec.shape[0]
    w = np.zeros(m, np.uint64)
    nn = np.zeros(m, np.uint64)
    rc = lib().nqo_count_records_13py(N, rec, m, w, nn, int(plant))
    if rc:
        raise RuntimeError(f"oracle 13Py-pruning count failed: {rc}")
    return w, nn


class gate_bits:
    """with oracle.gate_bits(w): build_soa's rd-word width for the j-diagonal
    gate (32 in the reference); records past it take the P5 functions."""

    def __init__(self, w: int):
        self.w = int(w)

    def __enter__(self):
        lib().nqo_set_gate_bits(self.w)
        return self

    def __exit__(self, *exc):
        lib().nqo_set_gate_bits(32)


def build_soa(N: int, rec) -> np.ndarray:
    """(m, 8) int32: fid, start, jmark, endmark, mark1, mark2, w, free."""
    rec = np.ascontiguousarray(np.asarray(rec, dtype=np.uint32))
    out = np.zeros((rec.shape[0], 8), np.int32)
    if lib().nqo_build_soa(N, rec, rec.shape[0], out.reshape(-1)):
        raise RuntimeError("oracle build_soa: shift out of range")
    return out


def symmetric_counts(N: int) -> tuple[int, int]:
    """(F90, F180): solutions fixed by the quarter turn and by the half turn."""
    a, b = np.zeros(1, np.uint64), np.zeros(1, np.uint64)
    if lib().nqo_symmetric_counts(N, a, b):
        raise ValueError(N)
    return int(a[0]), int(b[0])


def bruteforce(N: int) -> tuple[int, int, int, int]:
    """(total, unique, F90, F180) by enumerating every solution (N <= 14)."""
    o = np.zeros(4, np.uint64)
    if lib().nqo_bruteforce(N, o):
        raise ValueError(N)
    return tuple(int(x) for x in o)


def jasmin(ijkl: int, N: int) -> int:
    return lib().nqo_jasmin(ijkl, N)


def symmetry(ijkl: int, N: int) -> int:
    return lib().nqo_symmetry(ijkl, N)


def count(N: int, preset: int = 5, threads: int = 0):
    """Oracle (total, unique, nodes) for one N."""
    rec, _ = gen_constellations(N, preset)
    w, _, nn = count_records(N, rec, threads=threads)
    f90, f180 = symmetric_counts(N)
    total = int(w.sum())
    return total, (total + f180 + 2 * f90) // 8, int(nn.sum())


def ref_available() -> bool:
    return REF_LIB.exists()


def ref004_count(N: int):
    """Reference C solver 10Bit_CUDA/004GCC_constellation.c (preset 4) through
    oracle/_ref: (records (m,4) u32, weighted solutions i64[m])."""
    global _ref
    if _ref is None:
        R = ctypes.CDLL(str(REF_LIB))
        R.ref004_count.restype = ctypes.c_long
        R.ref004_count.argtypes = [ctypes.c_int, _u32p, _i64p, ctypes.c_long]
        _ref = R
    cap = 1 << 18
    rec = np.zeros(cap * 4, np.uint32)
    sol = np.zeros(cap, np.int64)
    m = _ref.ref004_count(N, rec, sol, cap)
    if m > cap:
        raise OverflowError("reference record list larger than the buffer")
    return rec[: 4 * m].reshape(-1, 4).copy(), sol[:m].copy()
