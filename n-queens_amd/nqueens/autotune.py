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
