"""Autotuner for the batched kernel: preset-queen depth, split depth,
workgroup size and stack placement (LDS or private memory).

Each configuration is timed on the resident task set with device events
(the library brackets the persistent DFS kernel and the whole run on its
stream); totals are checked on every repetition.  The table `TUNED` holds the
best measured configuration per (N, preset) and is what `best_sched` returns;
`sweep` regenerates it.  `sweep_presets` repeats the sweep for several preset
depths of one N, timing also the load of each preset's records (host
generation, device expansion beyond preset 5, decode), and `TUNED_PRESETS`
keeps the preset whose whole count (load + run) was fastest: what
`solver.preset_for` returns.
"""
from __future__ import annotations

import itertools
import json
import statistics
import time
from pathlib import Path

from ._lib import Context, NQError, Sched

NQ_ECAPACITY = -3   # nq_status of a schedule beyond the library's limits

__all__ = ["TUNED", "TUNED_PRESETS", "best_sched", "tuned_preset", "sweep", "sweep_presets", "sched_from"]

TUNED_PATH = Path(__file__).resolve().parent / "tuned.json"
PRESETS_PATH = Path(__file__).resolve().parent / "tuned_presets.json"
# fallback when no measurement exists for (N, preset)
DEFAULT = {"split_depth": 3, "block": 128, "lds_stack": 1, "chunk": 64}


def _load(path: Path | None = None) -> dict:
    path = TUNED_PATH if path is None else path
    if path.exists():
        return json.loads(path.read_text())
    return {}


TUNED = _load()


def _load_presets(path: Path | None = None) -> dict:
    path = PRESETS_PATH if path is None else path
    if path.exists():
        return json.loads(path.read_text())
    return {}


TUNED_PRESETS = _load_presets()


def tuned_preset(N: int) -> int | None:
    """The preset depth measured fastest for a whole count of N, if swept."""
    e = TUNED_PRESETS.get(str(N))
    return int(e["preset"]) if e else None


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
    the median/best DFS-kernel and whole-run device times (ms).  A
    configuration the library refuses for capacity (a split whose items
    overflow the u32 item indices, a stack larger than the LDS) gives a row
    with its error and no times."""
    rows = []
    for split, block, st, chunk in itertools.product(splits, blocks, lds, chunks):
        s = Sched(split_depth=split, block=block, lds_stack=st, use_graph=0, chunk=chunk, waves_per_eu=0)
        ks, rs = [], []
        items = 0
        try:
            for i in range(warmup + reps):
                tot, t = ctx.run_batched(s)
                if tot != expect_total:
                    raise RuntimeError(f"autotune: wrong total {tot} at {split},{block},{st}")
                if i >= warmup:
                    ks.append(t.kernel_ms)
                    rs.append(t.total_ms)
                    items = int(t.items)
        except NQError as e:
            if e.code != NQ_ECAPACITY:
                raise
            row = {"split_depth": split, "block": block, "lds_stack": st, "chunk": chunk, "error": str(e)}
            rows.append(row)
            if log:
                log(row)
            continue
        row = {"split_depth": split, "block": block, "lds_stack": st, "chunk": chunk, "items": items,
               "kernel_ms_median": statistics.median(ks), "kernel_ms_best": min(ks),
               "run_ms_median": statistics.median(rs), "run_ms_best": min(rs)}
        rows.append(row)
        if log:
            log(row)
    return rows


def load_preset(ctx: Context, N: int, preset: int) -> float:
    """Loads the records of (N, preset) into ctx -- generated on the host up
    to preset 5, the further queens placed on the device -- and returns the
    wall time of generation + load in ms."""
    from .constellations import gen_constellations
    t0 = time.perf_counter()
    if preset <= 5:
        ctx.load(N, gen_constellations(N, preset))
    else:
        ctx.load_expanded(N, gen_constellations(N, 5), preset - 5)
    return 1000.0 * (time.perf_counter() - t0)


def sweep_presets(ctx: Context, N: int, expect_total: int, presets=(4, 5, 6), splits=(2, 3, 4),
                  blocks=(128, 256), chunks=(64,), reps: int = 3, warmup: int = 1, order=None,
                  log=None) -> list[dict]:
    """`sweep` (LDS stack) for each preset depth of N.

    A preset whose records the walk cannot take (terminal row deeper than 16,
    or a P5 functionid) is reported as a row with its error and skipped.
    Each row carries the preset, its record count and its load time (best of
    two loads); order(ctx) is applied after every load (the cost-model order
    on the device, or None for generation order)."""
    rows = []
    for p in presets:
        try:
            load_ms = min(load_preset(ctx, N, p) for _ in range(2))
        except (RuntimeError, ValueError) as e:
            row = {"preset": p, "error": str(e)}
            rows.append(row)
            if log:
                log(row)
            continue
        if order is not None:
            order(ctx)
        for r in sweep(ctx, N, expect_total, splits, blocks, (1,), chunks, reps, warmup):
            r = {"preset": p, "records": ctx.m, "load_ms": round(load_ms, 3)} | r
            rows.append(r)
            if log:
                log(r)
    return rows


def record_presets(N: int, rows: list[dict], path: Path = PRESETS_PATH, tuned_path: Path = TUNED_PATH) -> dict:
    """Stores, per preset, the best schedule (tuned.json) and, for N, the
    preset with the lowest load + median run time (tuned_presets.json)."""
    ok = [r for r in rows if "error" not in r]
    if not ok:
        raise ValueError(f"N={N}: no preset could be counted")
    per = {}
    for p in sorted({r["preset"] for r in ok}):
        rp = [r for r in ok if r["preset"] == p]
        best = record_best(N, p, rp, tuned_path)
        per[p] = {"records": best["records"], "load_ms": best["load_ms"],
                  "run_ms_median": round(best["run_ms_median"], 3), "split_depth": best["split_depth"],
                  "block": best["block"]}
    pbest = min(per, key=lambda p: per[p]["load_ms"] + per[p]["run_ms_median"])
    table = _load_presets(path)
    table[str(N)] = {"preset": pbest, "per_preset": {str(p): v for p, v in per.items()}}
    path.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
    TUNED_PRESETS.clear()
    TUNED_PRESETS.update(table)
    return table[str(N)]


def record_best(N: int, preset: int, rows: list[dict], path: Path = TUNED_PATH) -> dict:
    """Store the configuration with the lowest median whole-run time (rows
    with an error are skipped)."""
    ok = [r for r in rows if "error" not in r]
    if not ok:
        raise ValueError(f"N={N} preset {preset}: no configuration could be counted")
    best = min(ok, key=lambda r: r["run_ms_median"])
    table = _load(path)
    table[f"{N},{preset}"] = {k: best[k] for k in ("split_depth", "block", "lds_stack", "chunk")} | {
        "run_ms_median": round(best["run_ms_median"], 3)}
    path.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
    TUNED.clear()
    TUNED.update(table)
    return best


def main(argv=None) -> None:
    """python -m nqueens.autotune --n 19 --preset 5 [--out PATH]: sweep, print
    one JSON row per configuration, store the best in tuned.json (and PATH).
    With --presets 4,5,6 the sweep runs for each preset depth (LDS stack) and
    the fastest whole count is stored in tuned_presets.json as well."""
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
    ap.add_argument("--presets", default=None, help="sweep these preset depths too (comma-separated)")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    from .constellations import gen_constellations
    from .solver import default_context
    from .expected import TOTAL
    ctx = default_context()
    ints = lambda s: tuple(int(x) for x in s.split(","))
    if a.presets:
        order = None
        if a.scheduler == "mlp":
            from .costmodel import default_model
            params = default_model().params
            order = lambda c: c.order_by_cost(params)
        rows = sweep_presets(ctx, a.n, TOTAL[a.n], ints(a.presets), ints(a.splits),
                             ints(a.blocks), ints(a.chunks), reps=a.reps, order=order,
                             log=lambda r: print(json.dumps(r), flush=True))
        best = record_presets(a.n, rows)
        print(json.dumps({"best": best, "n": a.n, "scheduler": a.scheduler}), flush=True)
        if a.out:
            Path(a.out).write_text(json.dumps({"rows": rows, "best": best, "n": a.n, "scheduler": a.scheduler},
                                              indent=1) + "\n")
        return
    rec = gen_constellations(a.n, a.preset)
    order = None
    if a.scheduler == "mlp":
        from .costmodel import default_model
        order = default_model().order(a.n, rec, ctx=ctx)
    ctx.load(a.n, rec, order)
    rows = sweep(ctx, a.n, TOTAL[a.n], ints(a.splits), ints(a.blocks), ints(a.lds),
                 ints(a.chunks), reps=a.reps, log=lambda r: print(json.dumps(r), flush=True))
    best = record_best(a.n, a.preset, rows)
    print(json.dumps({"best": best, "n": a.n, "preset": a.preset, "scheduler": a.scheduler}), flush=True)
    if a.out:
        Path(a.out).write_text(json.dumps({"rows": rows, "best": best, "n": a.n, "preset": a.preset,
                                           "scheduler": a.scheduler}, indent=1) + "\n")
    sys.stdout.flush()


if __name__ == "__main__":
    main()
