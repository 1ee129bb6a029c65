"""ctypes binding of libnq.so (the C ABI in include/nq.h).

This is the drop-in boundary: where the reference calls its Codon
``@gpu.kernel`` with ``gpu.raw(list)`` pointers
(13Bit_codon/18Py_constellations_GPU_cuda_codon.py:1200-1211), the host here
calls ``extern "C"`` entry points with plain pointers and sizes.  There is no
fallback: if the library is missing or the device cannot be opened the calls
raise.
"""
from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libnq.so"

NQ_ABI_VERSION = 1
NQ_MLP_IN = 24
NQ_MLP_H = 64
NQ_MLP_NPARAM = 64 * 24 + 64 + 64 * 64 + 64 + 64 + 1

STATUS = {0: "NQ_OK", -1: "NQ_EINVAL", -2: "NQ_EUNSUPPORTED", -3: "NQ_ECAPACITY",
          -4: "NQ_EHIP", -5: "NQ_ENOMEM", -6: "NQ_ESTATE"}


class NQError(RuntimeError):
    """A libnq entry point returned a negative nq_status."""

    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: {STATUS.get(code, code)}: {msg}")
        self.code = code


class Sched(ctypes.Structure):
    """nq_sched: batched-kernel schedule."""
    _fields_ = [("split_depth", ctypes.c_int32), ("block", ctypes.c_int32), ("lds_stack", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("chunk", ctypes.c_int32), ("waves_per_eu", ctypes.c_int32)]


class Timing(ctypes.Structure):
    """nq_timing: device times of one run."""
    _fields_ = [("kernel_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("items", ctypes.c_uint64), ("launches", ctypes.c_uint64)]


_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p

# name -> (restype, argtypes); every symbol declared in include/nq.h
SIGNATURES = {
    "nq_abi_version": (ctypes.c_int, []),
    "nq_last_error": (ctypes.c_char_p, []),
    "nq_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "nq_ctx_destroy": (ctypes.c_int, [_vp]),
    "nq_device_info": (ctypes.c_int, [_vp, _i32p]),
    "nq_issue_probe": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint32)]),
    "nq_ctx_load": (ctypes.c_int, [_vp, ctypes.c_int, _u32p, ctypes.c_uint64, _vp]),
    "nq_ctx_load_mixed": (ctypes.c_int, [_vp, _i32p, _u32p, ctypes.c_uint64, _vp]),
    "nq_expand_records": (ctypes.c_int, [_vp, ctypes.c_int, _u32p, ctypes.c_uint64, ctypes.c_int, _vp,
                                         ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    "nq_ctx_load_expanded": (ctypes.c_int, [_vp, ctypes.c_int, _u32p, ctypes.c_uint64, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_uint64)]),
    "nq_ctx_set_splits": (ctypes.c_int, [_vp, np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS"),
                                         ctypes.c_uint64]),
    "nq_ctx_run_plain": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(Timing)]),
    "nq_ctx_run_batched": (ctypes.c_int, [_vp, ctypes.POINTER(Sched), ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                          ctypes.POINTER(Timing)]),
    "nq_ctx_results": (ctypes.c_int, [_vp, _vp, _vp]),
    "nq_count": (ctypes.c_int, [_vp, ctypes.c_int, _u32p, ctypes.c_uint64, ctypes.POINTER(Sched), _vp, _vp,
                                ctypes.POINTER(ctypes.c_uint64)]),
    "nq_count_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, _u32p, ctypes.c_uint64, ctypes.POINTER(Sched), _vp, _vp,
                                        ctypes.POINTER(ctypes.c_uint64)]),
    "nq_symmetric_counts": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                           ctypes.POINTER(ctypes.c_uint64)]),
    "nq_small_total": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]),
    "nq_ctx_features": (ctypes.c_int, [_vp, _f32p]),
    "nq_mlp_forward": (ctypes.c_int, [_vp, _f32p, _f32p, ctypes.c_uint64, _f32p]),
    "nq_ctx_order_by_cost": (ctypes.c_int, [_vp, _f32p, _vp]),
    "nq_mlp_loss_grad": (ctypes.c_int, [_vp, _f32p, _f32p, _f32p, ctypes.c_uint64, _f32p, _f32p]),
    "nq_mlp_train": (ctypes.c_int, [_vp, _f32p, _f32p, _f32p, ctypes.c_uint64, _u32p, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_float, _f32p]),
    "nq_decode_host": (ctypes.c_int, [ctypes.c_int, _u32p, ctypes.c_uint64, _i32p]),
}

_lib = None
_lock = threading.Lock()


def load_library(path: Path | str | None = None) -> ctypes.CDLL:
    """Load libnq.so and bind every entry point; raises if it is not built."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path else LIB_PATH
        if not p.exists():
            raise RuntimeError(f"{p} is not built; run __graft_entry__.build() or python -m nqueens.build")
        lib = ctypes.CDLL(str(p))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.nq_abi_version() != NQ_ABI_VERSION:
            raise RuntimeError(f"libnq ABI {lib.nq_abi_version()} != {NQ_ABI_VERSION}")
        if path is None:
            _lib = lib
        return lib


def check(code: int, where: str) -> None:
    if code != 0:
        lib = load_library()
        raise NQError(code, where, (lib.nq_last_error() or b"").decode(errors="replace"))


def as_records(rec) -> np.ndarray:
    rec = np.ascontiguousarray(np.asarray(rec, dtype=np.uint32))
    if rec.ndim != 2 or rec.shape[1] != 4:
        raise ValueError(f"records must be (m, 4) uint32, got shape {rec.shape}")
    return rec


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One device context (nq_ctx): a HIP stream plus the resident task set."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = _vp()
        check(self.lib.nq_ctx_create(device, ctypes.byref(h)), "nq_ctx_create")
        self.h = h
        self.device = device
        self.N = None
        self.m = 0

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.nq_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self) -> dict:
        out = np.zeros(4, np.int32)
        check(self.lib.nq_device_info(self.h, out), "nq_device_info")
        return {"compute_units": int(out[0]), "lds_bytes": int(out[1]), "wavefront": int(out[2]),
                "clock_khz": int(out[3])}

    def issue_probe(self, waves_per_simd: int = 4, iters: int = 40000) -> dict:
        """VALU issue rate of independent v_or/v_and chains (nq_issue_probe):
        device ms, clocks per wave instruction per SIMD, lane-ops per second
        over the whole device, and the XOR of the lanes' results."""
        ms, clk, x = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_uint32(0)
        check(self.lib.nq_issue_probe(self.h, int(waves_per_simd), int(iters), ctypes.byref(ms),
                                      ctypes.byref(clk), ctypes.byref(x)), "nq_issue_probe")
        d = self.device_info()
        simds = 4 * d["compute_units"]
        lane_ops = 64.0 * simds * d["clock_khz"] * 1e3 / clk.value if clk.value > 0 else 0.0
        return {"waves_per_simd": int(waves_per_simd), "iters": int(iters), "ms": ms.value,
                "clk_per_wave_instr_per_simd": clk.value, "lane_ops_per_s": lane_ops, "xor": int(x.value)}

    def load(self, N: int, rec, order=None) -> None:
        rec = as_records(rec)
        ordp = None
        if order is not None:
            order = np.ascontiguousarray(np.asarray(order, dtype=np.uint32))
            if order.shape != (rec.shape[0],):
                raise ValueError("order must have one entry per record")
            ordp = _ptr(order)
        check(self.lib.nq_ctx_load(self.h, int(N), rec, rec.shape[0], ordp), "nq_ctx_load")
        self.N, self.m = int(N), rec.shape[0]

    def expand_records(self, N: int, rec, levels: int = 1) -> np.ndarray:
        """The records of preset p + levels from records of preset p,
        expanded on the device (record order kept, columns ascending)."""
        rec = as_records(rec)
        n = ctypes.c_uint64(0)
        check(self.lib.nq_expand_records(self.h, int(N), rec, rec.shape[0], int(levels), None, 0,
                                         ctypes.byref(n)), "nq_expand_records")
        out = np.zeros((int(n.value), 4), np.uint32)
        check(self.lib.nq_expand_records(self.h, int(N), rec, rec.shape[0], int(levels), _ptr(out), out.shape[0],
                                         ctypes.byref(n)), "nq_expand_records")
        return out

    def load_expanded(self, N: int, rec, levels: int) -> int:
        """load() of the records expand_records() would give, expanded and
        decoded on the device; returns the number of records loaded."""
        rec = as_records(rec)
        n = ctypes.c_uint64(0)
        check(self.lib.nq_ctx_load_expanded(self.h, int(N), rec, rec.shape[0], int(levels), ctypes.byref(n)),
              "nq_ctx_load_expanded")
        self.N, self.m = int(N), int(n.value)
        return self.m

    def load_mixed(self, Ns, rec, order=None) -> None:
        """A set mixing board sizes: Ns[i] is the N of record i."""
        rec = as_records(rec)
        Ns = np.ascontiguousarray(np.asarray(Ns, dtype=np.int32))
        if Ns.shape != (rec.shape[0],):
            raise ValueError("Ns must have one entry per record")
        ordp = None
        if order is not None:
            order = np.ascontiguousarray(np.asarray(order, dtype=np.uint32))
            if order.shape != (rec.shape[0],):
                raise ValueError("order must have one entry per record")
            ordp = _ptr(order)
        check(self.lib.nq_ctx_load_mixed(self.h, Ns, rec, rec.shape[0], ordp), "nq_ctx_load_mixed")
        self.N, self.m = 0, rec.shape[0]

    def set_splits(self, depth) -> None:
        """Per-record split depths, used by run_batched with split_depth = -1."""
        depth = np.ascontiguousarray(np.asarray(depth, dtype=np.uint8))
        check(self.lib.nq_ctx_set_splits(self.h, depth, depth.size), "nq_ctx_set_splits")

    def run_plain(self, count_nodes: bool = False) -> tuple[int, Timing]:
        tot = ctypes.c_uint64(0)
        t = Timing()
        check(self.lib.nq_ctx_run_plain(self.h, int(count_nodes), ctypes.byref(tot), ctypes.byref(t)),
              "nq_ctx_run_plain")
        return int(tot.value), t

    def run_batched(self, sched: Sched, count_nodes: bool = False) -> tuple[int, Timing]:
        tot = ctypes.c_uint64(0)
        t = Timing()
        check(self.lib.nq_ctx_run_batched(self.h, ctypes.byref(sched), int(count_nodes), ctypes.byref(tot),
                                          ctypes.byref(t)), "nq_ctx_run_batched")
        return int(tot.value), t

    def results(self, weighted: bool = True, nodes: bool = False):
        w = np.zeros(self.m, np.uint64) if weighted else None
        n = np.zeros(self.m, np.uint64) if nodes else None
        check(self.lib.nq_ctx_results(self.h, _ptr(w), _ptr(n)), "nq_ctx_results")
        return w, n

    def symmetric_counts(self, N: int) -> tuple[int, int]:
        f90, f180 = ctypes.c_uint64(0), ctypes.c_uint64(0)
        check(self.lib.nq_symmetric_counts(self.h, int(N), ctypes.byref(f90), ctypes.byref(f180)),
              "nq_symmetric_counts")
        return int(f90.value), int(f180.value)

    def small_total(self, N: int) -> int:
        """Every solution of the N x N board by a plain DFS on the device (1 <= N <= 12)."""
        t = ctypes.c_uint64(0)
        check(self.lib.nq_small_total(self.h, int(N), ctypes.byref(t)), "nq_small_total")
        return int(t.value)

    def features(self) -> np.ndarray:
        out = np.zeros((self.m, NQ_MLP_IN), np.float32)
        check(self.lib.nq_ctx_features(self.h, out.reshape(-1)), "nq_ctx_features")
        return out

    def order_by_cost(self, params: np.ndarray, want: bool = False):
        """Orders the resident set by the model's predicted cost on the device
        (heaviest first); returns the order when `want`."""
        params = np.ascontiguousarray(params, np.float32)
        if params.shape != (NQ_MLP_NPARAM,):
            raise ValueError(f"expected {NQ_MLP_NPARAM} parameters")
        out = np.zeros(self.m, np.uint32) if want else None
        check(self.lib.nq_ctx_order_by_cost(self.h, params, _ptr(out)), "nq_ctx_order_by_cost")
        return out

    def mlp_forward(self, params: np.ndarray, x: np.ndarray) -> np.ndarray:
        params = np.ascontiguousarray(params, np.float32)
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(x.shape[0], np.float32)
        check(self.lib.nq_mlp_forward(self.h, params, x.reshape(-1), x.shape[0], y), "nq_mlp_forward")
        return y

    def mlp_loss_grad(self, params: np.ndarray, x: np.ndarray, y: np.ndarray) -> tuple[float, np.ndarray]:
        params = np.ascontiguousarray(params, np.float32)
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = np.zeros(1, np.float32)
        grad = np.zeros(NQ_MLP_NPARAM, np.float32)
        check(self.lib.nq_mlp_loss_grad(self.h, params, x.reshape(-1), y, x.shape[0], loss, grad),
              "nq_mlp_loss_grad")
        return float(loss[0]), grad

    def mlp_train(self, params: np.ndarray, x: np.ndarray, y: np.ndarray, perm: np.ndarray, epochs: int,
                  batch: int, lr: float) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        perm = np.ascontiguousarray(perm, np.uint32)
        if perm.size != epochs * x.shape[0]:
            raise ValueError("perm must hold epochs * m row indices")
        losses = np.zeros(epochs, np.float32)
        check(self.lib.nq_mlp_train(self.h, params, x.reshape(-1), y, x.shape[0], perm, int(epochs), int(batch),
                                    float(lr), losses), "nq_mlp_train")
        return losses


def decode_host(N: int, rec) -> np.ndarray:
    """Host run of the device record decode: (m, 10) int32 rows
    {fid, start, jmark, endmark, mark1, mark2, w, free, T, status}."""
    lib = load_library()
    rec = as_records(rec)
    out = np.zeros((rec.shape[0], 10), np.int32)
    check(lib.nq_decode_host(int(N), rec, rec.shape[0], out.reshape(-1)), "nq_decode_host")
    return out
