"""Autotuner for the batched kernel: split depth x workgroup size x stack
placement (LDS or private memory).

Each configuration is timed on the resident task set with device events
(the library brackets the persistent DFS kernel and the whole run on its
stream); totals are checked on every repetition.  The table `TUNED` holds the
best measured configuration per (N, preset) and is what `best_sched` returns;
`sweep` regenerates it.
"""
from __future__ import annotations

import itertools
import json
import statistics
from pathlib import Path

from ._lib import Context, Sched

__all__ = ["TUNED", "best_sched", "sweep", "sched_from"]

TUNED_PATH = Path(__file__).resolve().parent / "tuned.json"
# fallback when no measurement exists for (N, preset)
DEFAULT = {"split_depth": 3, "block": 128, "lds_stack": 1, "chunk": 64}


def _load() -> dict:
    if TUNED_PATH.exists():
        return json.loads(TUNED_PATH.read_text())
    return {}


TUNED = _load()


def sched_from(d: dict, use_graph: int = 0) -> Sched:
    return Sched(split_depth=int(d["split_depth"]), block=int(d["block"]), lds_stack=int(d["lds_stack"]),
                 use_graph=use_graph, chunk=int(d.get("chunk", 64)), waves_per_eu=int(d.get("waves_per_eu", 0)))


def best_sched(N: int, preset: int, use_graph: int = 0) -> Sched:
    """Best measured schedule for (N, preset), else the nearest N measured at
    the same preset, else the default."""
    key = f"{N},{preset}"
    if key in TUNED:
        return sched_from(TUNED[key], use_graph)
    same = [(abs(int(k.split(",")[0]) - N), k) for k in TUNED if k.endswith(f",{preset}")]
    if same:
        return sched_from(TUNED[min(same)[1]], use_graph)
    return sched_from(DEFAULT, use_graph)


def sweep(ctx: Context, N: int, expect_total: int, splits=(2, 3, 4, 5, 6), blocks=(64, 128, 256),
          lds=(1, 0), chunks=(64,), reps: int = 3, warmup: int = 1, log=None) -> list[dict]:
    """Time every configuration on the task set loaded in `ctx`.

    Returns one row per configuration: its parameters, the work-item count and
    the median/best DFS-kernel and whole-run device times (ms)."""
    rows = []
    for split, block, st, chunk in itertools.product(splits, blocks, lds, chunks):
        s = Sched(split_depth=split, block=block, lds_stack=st, use_graph=0, chunk=chunk, waves_per_eu=0)
        for _ in range(warmup):
            tot, _ = ctx.run_batched(s)
            if tot != expect_total:
                raise RuntimeError(f"autotune: wrong total {tot} at {split},{block},{st}")
        ks, rs = [], []
        items = 0
        for _ in range(reps):
            tot, t = ctx.run_batched(s)
            if tot != expect_total:
                raise RuntimeError(f"autotune: wrong total {tot} at {split},{block},{st}")
            ks.append(t.kernel_ms)
            rs.append(t.total_ms)
            items = int(t.items)
        row = {"split_depth": split, "block": block, "lds_stack": st, "chunk": chunk, "items": items,
               "kernel_ms_median": statistics.median(ks), "kernel_ms_best": min(ks),
               "run_ms_median": statistics.median(rs), "run_ms_best": min(rs)}
        rows.append(row)
        if log:
            log(row)
    return rows


def record_best(N: int, preset: int, rows: list[dict], path: Path = TUNED_PATH) -> dict:
    """Store the configuration with the lowest median whole-run time."""
    best = min(rows, key=lambda r: r["run_ms_median"])
    table = _load()
    table[f"{N},{preset}"] = {k: best[k] for k in ("split_depth", "block", "lds_stack", "chunk")} | {
        "run_ms_median": best["run_ms_median"]}
    path.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
    TUNED.clear()
    TUNED.update(table)
    return best


def main(argv=None) -> None:
    """python -m nqueens.autotune --n 19 --preset 5 [--out PATH]: sweep, print
    one JSON row per configuration, store the best in tuned.json (and PATH)."""
    import argparse
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=19)
    ap.add_argument("--preset", type=int, default=5)
    ap.add_argument("--splits", default="2,3,4,5,6")
    ap.add_argument("--blocks", default="64,128,256")
    ap.add_argument("--lds", default="1,0")
    ap.add_argument("--chunks", default="64")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--scheduler", choices=["fifo", "mlp"], default="fifo")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    from .constellations import gen_constellations
    from .solver import default_context
    known = json.loads((Path(__file__).resolve().parents[2] / "tests" / "golden" / "known.json").read_text())
    ctx = default_context()
    rec = gen_constellations(a.n, a.preset)
    order = None
    if a.scheduler == "mlp":
        from .costmodel import default_model
        order = default_model().order(a.n, rec, ctx=ctx)
    ctx.load(a.n, rec, order)
    ints = lambda s: tuple(int(x) for x in s.split(","))
    rows = sweep(ctx, a.n, known["total_a000170"][str(a.n)], ints(a.splits), ints(a.blocks), ints(a.lds),
                 ints(a.chunks), reps=a.reps, log=lambda r: print(json.dumps(r), flush=True))
    best = record_best(a.n, a.preset, rows)
    print(json.dumps({"best": best, "n": a.n, "preset": a.preset, "scheduler": a.scheduler}), flush=True)
    if a.out:
        Path(a.out).write_text(json.dumps({"rows": rows, "best": best, "n": a.n, "preset": a.preset,
                                           "scheduler": a.scheduler}, indent=1) + "\n")
    sys.stdout.flush()


if __name__ == "__main__":
    main()
