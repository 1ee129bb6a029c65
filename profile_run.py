#!/usr/bin/env python3
"""Fixed workload for profiler runs (rocprofv3 kernel traces and counters).

    python3 profile_run.py --n 19 --preset 5 --mode batched --reps 2

Loads the constellations of (N, preset) once and runs the chosen kernel
`reps` times outside a graph, so that every dispatch is traced on its own.
With --calib-mib M it first copies M MiB from one device buffer to another
(torch, one plain copy kernel: M MiB read, M MiB written), the reference
that the FETCH_SIZE / WRITE_SIZE counters of the other kernels are scaled by.
"""
from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT / "n-queens_amd", ROOT):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))


def main(argv=None) -> None:
    from nqueens import autotune
    from nqueens.constellations import gen_constellations
    from nqueens.solver import default_context
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=19)
    ap.add_argument("--preset", type=int, default=5)
    ap.add_argument("--mode", choices=["batched", "plain"], default="batched")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--split", type=int, default=None)
    ap.add_argument("--block", type=int, default=None)
    ap.add_argument("--lds", type=int, default=None)
    ap.add_argument("--calib-mib", type=int, default=0)
    ap.add_argument("--scheduler", choices=["fifo", "mlp"], default="fifo")
    a = ap.parse_args(argv)
    if a.calib_mib > 0:
        import torch
        src = torch.ones(a.calib_mib << 18, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": a.calib_mib << 20}), flush=True)
        del src, dst
    ctx = default_context()
    rec = gen_constellations(a.n, a.preset)
    ctx.load(a.n, rec)
    params = None
    if a.scheduler == "mlp":
        from nqueens.costmodel import default_model
        params = default_model().params
    s = autotune.best_sched(a.n, a.preset)
    for k, v in (("split_depth", a.split), ("block", a.block), ("lds_stack", a.lds)):
        if v is not None:
            setattr(s, k, v)
    for r in range(a.reps):
        if params is not None:
            ctx.order_by_cost(params)        # the cost-model order, made on the device as in bench.py
        if a.mode == "batched":
            tot, t = ctx.run_batched(s)
        else:
            tot, t = ctx.run_plain()
        print(json.dumps({"rep": r, "mode": a.mode, "n": a.n, "preset": a.preset, "total": tot,
                          "kernel_ms": t.kernel_ms, "run_ms": t.total_ms, "items": int(t.items),
                          "split": s.split_depth, "block": s.block, "lds": s.lds_stack}), flush=True)


if __name__ == "__main__":
    main()
