#!/usr/bin/env python3
"""Benchmark: exact N-Queens count on one MI355X, DFS nodes/s.

Workload (default): N=19, preset 5 -- all 100,706 canonical constellations
counted exactly in one batched run (work items split at depth `--split`,
one persistent launch pulling them from a device queue).  A "step" is one
complete exact count of N=19 with the task set already resident in HBM:
expansion, persistent DFS and the weighted reduction; every step's total is
checked against A000170.  The plain one-lane-per-constellation kernel is
timed on the same records beside it, and the CPU oracle (a C/OpenMP
restatement of the reference's dfs_iter) on a bounded sample of them.

One node = one candidate extraction of the reference dfs_iter
(13Bit_codon/18Py_constellations_GPU_cuda_codon.py:368-562); node counts come
from a node-counting run of the same kernel (checked against the oracle by
the parity tests).  Roofline: 12 algorithmic int32 ops per node against the
INT32 VALU peak of 256 CUs x 128 lanes x 2.4 GHz = 78.6 T ops/s (a CU issues
two wave64 32-bit VALU instructions per clock: the published 157.3 TFLOP/s
FP32 vector peak is the same 128 lanes counting an FMA as two flops).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for _p in (ROOT / "n-queens_amd", ROOT):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

METRIC = "N=19/21 exact-count wall time & DFS nodes/s: batched+cost-model vs plain kernel"
OPS_PER_NODE = 12
PEAK_INT32_TOPS = 256 * 128 * 2.4e9 / 1e12   # 78.64 T int32 VALU ops/s


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=19)
    ap.add_argument("--preset", type=int, default=5)
    ap.add_argument("--split", type=int, default=None)
    ap.add_argument("--block", type=int, default=None)
    ap.add_argument("--lds", type=int, default=None)
    ap.add_argument("--chunk", type=int, default=None)
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--scheduler", choices=["fifo", "mlp"], default="mlp")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--big-n", type=int, default=21, help="also time one exact count of this N (0 = skip)")
    ap.add_argument("--big-plain", type=int, default=1, help="time the plain kernel on the --big-n records too")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline duration (0 = skip)")
    ap.add_argument("--json-out", default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 (a count over "
                                   "several GPUs would give each a share of the constellations and add "
                                   "the totals)"}))
        return 2

    import nqueens
    from nqueens import autotune
    from nqueens.expected import TOTAL, UNIQUE
    N, preset = args.n, args.preset
    expect_total, expect_unique = TOTAL[N], UNIQUE[N]

    ctx = nqueens.Context(0)
    info = ctx.device_info()
    log(f"[bench] device {info}")
    t0 = time.perf_counter()
    rec = nqueens.gen_constellations(N, preset)
    gen_s = time.perf_counter() - t0
    log(f"[bench] N={N} preset={preset} constellations={rec.shape[0]} gen {gen_s:.2f}s")

    sched = autotune.best_sched(N, preset)
    for k, v in (("split_depth", args.split), ("block", args.block), ("lds_stack", args.lds), ("chunk", args.chunk)):
        if v is not None:
            setattr(sched, k, v)
    sched.use_graph = args.graph
    mlp_info = None
    ctx.load(N, rec)
    params = None
    if args.scheduler == "mlp":
        # the cost model orders the records on the device inside every timed
        # count (features, forward pass, radix sort), before the batched run
        from nqueens import costmodel
        model = costmodel.default_model()
        params = model.params
        ctx.order_by_cost(params)
        t0 = time.perf_counter()
        for _ in range(5):
            ctx.order_by_cost(params)
        mlp_info = {"order_ms": round((time.perf_counter() - t0) / 5 * 1000, 3), "trained_on": model.trained_on,
                    "in_timed_step": True}

    def count_once(nodes_on: bool = False):
        if params is not None:
            ctx.order_by_cost(params)
        return ctx.run_batched(sched, count_nodes=nodes_on)

    # node count of the workload (untimed, same kernel with the counter on)
    total, _ = count_once(True)
    _, nodes = ctx.results(weighted=False, nodes=True)
    nodes_total = int(nodes.sum())
    assert total == expect_total, (total, expect_total)
    log(f"[bench] nodes={nodes_total} sched={dict((f, getattr(sched, f)) for f, _ in sched._fields_)}")

    for _ in range(args.warmup):
        tot, _ = count_once()
        assert tot == expect_total
    kernel_ms, run_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tot, t = count_once()
        if tot != expect_total:
            raise SystemExit(f"wrong total {tot} != {expect_total}")
        kernel_ms.append(t.kernel_ms)
        run_ms.append(t.total_ms)
        items = int(t.items)
    wall = time.perf_counter() - t0
    ms_per_step = 1000.0 * wall / args.steps
    value = nodes_total / (wall / args.steps)
    f90, f180 = ctx.symmetric_counts(N)
    unique = nqueens.unique_count(expect_total, f90, f180)
    assert unique == expect_unique
    log(f"[bench] batched {ms_per_step:.2f} ms/count, {value / 1e9:.2f} G nodes/s, "
        f"device run {np.mean(run_ms):.2f} ms")

    # the same schedule in generation (FIFO) order: what the cost model's
    # device ordering adds to the batched kernel
    fifo = None
    if params is not None:
        ctx.load(N, rec)
        tot, _ = ctx.run_batched(sched)
        assert tot == expect_total
        t0 = time.perf_counter()
        for _ in range(args.steps):
            tot, _ = ctx.run_batched(sched)
            assert tot == expect_total
        fwall = (time.perf_counter() - t0) / args.steps
        fifo = {"nodes_per_s": round(nodes_total / fwall, 1), "ms_per_count": round(1000 * fwall, 3),
                "cost_model_speedup": round(fwall / (wall / args.steps), 4)}
        log(f"[bench] fifo order {1000 * fwall:.2f} ms/count")

    # one whole count from scratch as a caller makes it: host generation,
    # load and decode, cost order, batched run, Unique (nqueens.count)
    t0 = time.perf_counter()
    ct, cu = nqueens.count(N, preset=preset, scheduler=args.scheduler, ctx=ctx)
    count_wall = time.perf_counter() - t0
    assert (ct, cu) == (expect_total, expect_unique)
    log(f"[bench] nqueens.count({N}) from scratch {count_wall * 1000:.1f} ms")

    # plain kernel on the same records (record order)
    ctx.load(N, rec)
    tot, _ = ctx.run_plain()
    assert tot == expect_total
    pk = []
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        tot, t = ctx.run_plain()
        assert tot == expect_total
        pk.append(t.kernel_ms)
    plain_wall = (time.perf_counter() - t0) / max(1, args.plain_steps)
    plain_value = nodes_total / plain_wall
    log(f"[bench] plain {1000 * plain_wall:.2f} ms/count, {plain_value / 1e9:.2f} G nodes/s")

    # DFS-kernel time for the roofline: the same runs outside the graph,
    # bracketed by events around the persistent kernel alone
    if sched.use_graph:
        eager = nqueens.Sched(**{f: getattr(sched, f) for f, _ in sched._fields_})
        eager.use_graph = 0
        ctx.load(N, rec)
        if params is not None:
            ctx.order_by_cost(params)
        kernel_ms = []
        for _ in range(max(2, args.steps)):
            tot, t = ctx.run_batched(eager)
            assert tot == expect_total
            kernel_ms.append(t.kernel_ms)
    kavg = float(np.mean(kernel_ms))
    achieved = nodes_total * OPS_PER_NODE / (kavg / 1000.0) / 1e12
    roofline = {"bound": "valu", "achieved": round(achieved, 4), "peak": round(PEAK_INT32_TOPS, 2),
                "unit": "TOP/s", "frac": round(achieved / PEAK_INT32_TOPS, 4),
                "traffic": profiled_figure(N, preset, "traffic_bytes_per_launch", int),
                "kernel": "k_batched", "kernel_ms": round(kavg, 3),
                "ops_per_node": OPS_PER_NODE, "nodes_per_launch": nodes_total,
                "issue_probe": issue_probe(ctx, N, preset, nodes_total, kavg)}

    big = None
    if args.big_n:
        big = big_count(nqueens, ctx, args.big_n, preset, args.scheduler, args.big_plain)
        log(f"[bench] N={args.big_n}: {big}")

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(N, rec, nodes, args.cpu_seconds)
        log(f"[bench] cpu {cpu}")

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "nodes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "u32", "data": "real",
        "config": {"workload": f"exact count N={N}, preset {preset}: {rec.shape[0]} constellations, "
                               f"batched persistent kernel, {args.scheduler} order",
                   "N": N, "preset": preset, "constellations": int(rec.shape[0]), "work_items": items,
                   "split_depth": sched.split_depth, "block": sched.block, "lds_stack": sched.lds_stack,
                   "chunk": sched.chunk, "graph": sched.use_graph, "scheduler": args.scheduler,
                   "parallelism": "single GPU"},
        "total": expect_total, "unique": unique, "nodes": nodes_total,
        "wall_s_per_count": round(ms_per_step / 1000, 5), "run_ms_device": round(float(np.mean(run_ms)), 3),
        "count_wall_s": round(count_wall, 4), "fifo": fifo,
        "plain": {"nodes_per_s": round(plain_value, 1), "ms_per_count": round(1000 * plain_wall, 3),
                  "kernel_ms": round(float(np.mean(pk)), 3)},
        "speedup_vs_plain": round(value / plain_value, 3),
        "roofline": roofline, "cpu_baseline": cpu, "cost_model": mlp_info, "big_n": big,
        "device": info,
    }
    line = json.dumps(out)
    print(line, flush=True)
    if args.json_out:
        Path(args.json_out).write_text(line + "\n")
    ctx.close()
    return 0


def profiled_figure(N: int, preset: int, key: str, conv=float):
    """A figure of the committed rocprofv3 summary of this workload
    (`key: value` line of profiles/k_batched_N{N}_p{preset}.md): the HBM bytes
    per k_batched launch (FETCH_SIZE + WRITE_SIZE), the VALU instructions per
    node (SQ_INSTS_VALU per wave step); None when no summary matches."""
    p = ROOT / "profiles" / f"k_batched_N{N}_p{preset}.md"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith(key + ":"):
            return conv(line.split(":")[1])
    return None


def issue_probe(ctx, N: int, preset: int, nodes: int, kernel_ms: float) -> dict:
    """The VALU issue ceiling measured live (nq_issue_probe: independent
    v_or_b32 / v_and_b32 chains at the DFS kernel's four waves per SIMD), and
    the DFS kernel's own VALU issue rate against it: the profiled VALU
    instructions per node times nodes per second of the event-timed kernel."""
    ctx.issue_probe(4, 2000)
    p = ctx.issue_probe(4, 40000)
    out = {"ops": "v_or_b32 / v_and_b32, 8 independent chains per lane", "waves_per_simd": 4,
           "clk_per_wave_instr_per_simd": round(p["clk_per_wave_instr_per_simd"], 3),
           "lane_ops_per_s": round(p["lane_ops_per_s"], 1),
           "frac_of_peak": round(p["lane_ops_per_s"] / (PEAK_INT32_TOPS * 1e12), 4)}
    vpn = profiled_figure(N, preset, "valu_per_node")
    if vpn:
        kernel_ops = vpn * nodes / (kernel_ms / 1000.0)
        out.update({"kernel_valu_per_node": vpn, "kernel_lane_ops_per_s": round(kernel_ops, 1),
                    "kernel_frac_of_probe": round(kernel_ops / p["lane_ops_per_s"], 4)})
    return out


def big_count(nqueens, ctx, N: int, preset: int, scheduler: str, plain: bool) -> dict:
    """One exact count of a larger N (wall time of the count call and its
    node rate; nodes from a node-counting run of the same set), and the plain
    one-lane-per-constellation kernel on the same records."""
    from nqueens import autotune
    from nqueens.expected import TOTAL
    rec = nqueens.gen_constellations(N, preset)
    ctx.load(N, rec)
    params = None
    if scheduler == "mlp":
        from nqueens import costmodel
        params = costmodel.default_model().params
        ctx.order_by_cost(params)
    s = autotune.best_sched(N, preset)
    s.use_graph = 1
    tot, _ = ctx.run_batched(s, count_nodes=True)
    _, nodes = ctx.results(weighted=False, nodes=True)
    t0 = time.perf_counter()
    if params is not None:
        ctx.order_by_cost(params)
    tot2, t = ctx.run_batched(s)
    wall = time.perf_counter() - t0
    want = TOTAL[N]
    if tot != want or tot2 != want:
        raise SystemExit(f"N={N}: wrong total {tot}/{tot2} != {want}")
    f90, f180 = ctx.symmetric_counts(N)
    out = {"N": N, "preset": preset, "total": tot2, "unique": nqueens.unique_count(tot2, f90, f180),
           "constellations": int(rec.shape[0]), "nodes": int(nodes.sum()), "wall_s": round(wall, 4),
           "nodes_per_s": round(float(nodes.sum()) / wall, 1), "split_depth": s.split_depth, "block": s.block}
    if plain:
        ctx.load(N, rec)
        t0 = time.perf_counter()
        tp, tt = ctx.run_plain()
        pwall = time.perf_counter() - t0
        if tp != want:
            raise SystemExit(f"N={N} plain kernel: wrong total {tp} != {want}")
        out["plain"] = {"wall_s": round(pwall, 3), "kernel_ms": round(tt.kernel_ms, 1),
                        "nodes_per_s": round(float(nodes.sum()) / pwall, 1)}
        out["speedup_vs_plain"] = round(pwall / wall, 3)
    return out


def cpu_baseline(N: int, rec: np.ndarray, nodes: np.ndarray, seconds: float) -> dict:
    """Oracle dfs_iter (C, OpenMP dynamic,64) on every k-th record, k chosen
    from the GPU node counts so that the sample is about `seconds` of work;
    the same sample is timed again with schedule(static), the Codon @par
    default (18Py:1050)."""
    from oracle import oracle
    cores = len(os.sched_getaffinity(0))
    # calibrate on a small stride sample
    probe = rec[::max(1, rec.shape[0] // 512)]
    t0 = time.perf_counter()
    _, _, pn = oracle.count_records(N, probe, threads=cores, schedule=1)
    dt = max(time.perf_counter() - t0, 1e-3)
    rate = float(pn.sum()) / dt
    want_nodes = rate * seconds * 3.0      # the small probe underestimates the rate (thread start-up dominates it)
    stride = max(1, int(np.ceil(float(nodes.sum()) / max(want_nodes, 1.0))))
    sample = np.ascontiguousarray(rec[::stride])
    t0 = time.perf_counter()
    w, _, sn = oracle.count_records(N, sample, threads=cores, schedule=1)
    dt = time.perf_counter() - t0
    if not np.array_equal(sn, nodes[::stride]):
        raise SystemExit("cpu baseline: node counts disagree with the GPU")
    # the same sample with the static schedule of the reference's Codon @par
    t0 = time.perf_counter()
    oracle.count_records(N, sample, threads=cores, schedule=0, nodes=False)
    dts = time.perf_counter() - t0
    return {"value": round(float(sn.sum()) / dt, 1), "unit": "nodes/s", "cores": cores, "kind": "port",
            "sample": f"every {stride}th of the {rec.shape[0]} N={N} constellations "
                      f"({sample.shape[0]} records, {int(sn.sum())} nodes, {dt:.1f} s), OpenMP dynamic,64",
            "seconds": round(dt, 2),
            "static_schedule": {"value": round(float(sn.sum()) / dts, 1), "seconds": round(dts, 2)}}


if __name__ == "__main__":
    sys.exit(main())
