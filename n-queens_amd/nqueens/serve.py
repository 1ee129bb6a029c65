"""Request loop: count(N) requests of mixed N sharing one device queue.

Requests arrive as a seeded Poisson process with N drawn uniformly from
[nmin, nmax].  Whenever the GPU is free, every request that has arrived is
admitted in arrival order while the batch stays within a record budget (HBM
for descriptors and work items); the admitted requests' constellations form
one mixed-N resident set (nq_ctx_load_mixed), ordered by the cost model, and
one persistent batched run counts them all.  Per-request totals are segment
sums of the per-record counts; Unique comes from the symmetric-solution
kernel.  Every result is checked against the published answers.

Time is virtual: arrivals follow the schedule, and each batch advances the
clock by its measured wall time (host preparation + GPU), so a run does not
sleep through idle gaps.  Latency = completion - arrival.

No reference counterpart: the reference counts one N per process run.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from collections import deque
from dataclasses import asdict, dataclass

import numpy as np

from . import autotune
from .constellations import gen_constellations
from .expected import TOTAL, UNIQUE
from .solver import default_context, unique_count

__all__ = ["Request", "Result", "poisson_requests", "RequestLoop", "latency_stats"]


@dataclass
class Request:
    rid: int
    N: int
    arrival: float


@dataclass
class Result:
    rid: int
    N: int
    arrival: float
    start: float
    done: float
    batch: int
    total: int
    unique: int
    ok: bool

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def poisson_requests(n: int, rate: float, nmin: int, nmax: int, seed: int = 20261203) -> list[Request]:
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n)
    Ns = rng.integers(nmin, nmax + 1, n)
    t = np.cumsum(gaps)
    return [Request(i, int(Ns[i]), float(t[i])) for i in range(n)]


def latency_stats(results: list[Result]) -> dict:
    lat = np.array([r.latency for r in results])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}


class RequestLoop:
    def __init__(self, ctx=None, preset: int = 5, max_records: int = 2_000_000, scheduler: str = "mlp",
                 sched=None, log=None):
        self.ctx = ctx or default_context()
        self.preset = preset
        self.max_records = max_records
        self.scheduler = scheduler
        self.sched = sched
        self.log = log or (lambda *_: None)
        self._rec: dict[int, np.ndarray] = {}
        self.nodes_per_n: dict[int, int] = {}

    def records(self, N: int) -> np.ndarray:
        """Constellations of N (host generation is a pure function of N and is kept)."""
        r = self._rec.get(N)
        if r is None:
            r = gen_constellations(N, self.preset)
            self._rec[N] = r
        return r

    def _sched(self, Ns):
        if self.sched is not None:
            return self.sched
        return autotune.best_sched(max(Ns), self.preset)

    def run_batch(self, batch: list[Request], count_nodes: bool = False):
        """Count one admitted batch; returns per-request (total, unique) and the run's node total."""
        recs = [self.records(r.N) for r in batch]
        rec = np.concatenate(recs)
        Ns = np.concatenate([np.full(x.shape[0], r.N, np.int32) for x, r in zip(recs, batch)])
        self.ctx.load_mixed(Ns, rec)
        if self.scheduler == "mlp":
            from .costmodel import default_model
            self.ctx.order_by_cost(default_model().params)   # on the device, heaviest first
        self.ctx.run_batched(self._sched([r.N for r in batch]), count_nodes=count_nodes)
        w, nodes = self.ctx.results(weighted=True, nodes=count_nodes)
        bounds = np.cumsum([0] + [x.shape[0] for x in recs])
        sym = {N: self.ctx.symmetric_counts(N) for N in sorted({r.N for r in batch})}
        out = []
        for i, r in enumerate(batch):
            total = int(w[bounds[i]:bounds[i + 1]].sum())
            f90, f180 = sym[r.N]
            out.append((total, unique_count(total, f90, f180)))
        return out, (int(nodes.sum()) if count_nodes else None)

    def calibrate(self, Ns) -> dict:
        """Service time and node count of each N alone (warm)."""
        times = {}
        for N in sorted(set(Ns)):
            req = [Request(-1, N, 0.0)]
            _, nodes = self.run_batch(req, count_nodes=True)
            self.nodes_per_n[N] = nodes
            t0 = time.perf_counter()
            self.run_batch(req)
            times[N] = time.perf_counter() - t0
            self.log(f"[serve] calibrate N={N}: {times[N] * 1e3:.1f} ms, {nodes} nodes")
        return times

    def run(self, requests: list[Request]) -> list[Result]:
        reqs = sorted(requests, key=lambda r: r.arrival)
        clock, i, nb = 0.0, 0, 0
        pending: deque[Request] = deque()
        results: list[Result] = []
        while i < len(reqs) or pending:
            if not pending and reqs[i].arrival > clock:
                clock = reqs[i].arrival
            while i < len(reqs) and reqs[i].arrival <= clock:
                pending.append(reqs[i])
                i += 1
            batch, nrec = [], 0
            while pending and (not batch or nrec + self.records(pending[0].N).shape[0] <= self.max_records):
                r = pending.popleft()
                batch.append(r)
                nrec += self.records(r.N).shape[0]
            t0 = time.perf_counter()
            answers, _ = self.run_batch(batch)
            dt = time.perf_counter() - t0
            start, clock = clock, clock + dt
            for r, (tot, uni) in zip(batch, answers):
                ok = TOTAL.get(r.N) == tot and UNIQUE.get(r.N) == uni
                results.append(Result(r.rid, r.N, r.arrival, start, clock, nb, tot, uni, ok))
            self.log(f"[serve] batch {nb}: {len(batch)} requests (N={sorted(r.N for r in batch)}), "
                     f"{nrec} constellations, {dt * 1e3:.1f} ms, clock {clock:.2f} s, {len(pending)} waiting")
            nb += 1
        return results


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m nqueens.serve")
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--nmin", type=int, default=14)
    ap.add_argument("--nmax", type=int, default=21)
    ap.add_argument("--load", type=float, default=0.8, help="arrival rate as a fraction of measured capacity")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--max-records", type=int, default=2_000_000)
    ap.add_argument("--scheduler", choices=["fifo", "mlp"], default="mlp")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    loop = RequestLoop(max_records=a.max_records, scheduler=a.scheduler, log=log)
    Ns = list(range(a.nmin, a.nmax + 1))
    times = loop.calibrate(Ns)
    mean_service = float(np.mean([times[N] for N in Ns]))
    rate = a.load / mean_service
    reqs = poisson_requests(a.requests, rate, a.nmin, a.nmax, a.seed)
    t0 = time.perf_counter()
    res = loop.run(reqs)
    wall = time.perf_counter() - t0
    span = max(r.done for r in res) - min(r.arrival for r in res)
    nodes = sum(loop.nodes_per_n[r.N] for r in res)
    busy = sum({r.batch: r.done - r.start for r in res}.values())
    out = {"requests": len(res), "all_exact": all(r.ok for r in res), "rate_per_s": rate,
           "mean_service_s": mean_service, "batches": len({r.batch for r in res}),
           **latency_stats(res), "span_s": span, "gpu_busy_s": busy, "nodes": nodes,
           "nodes_per_s_busy": nodes / busy, "nodes_per_s_span": nodes / span, "wall_s": wall,
           "seed": a.seed, "nmin": a.nmin, "nmax": a.nmax, "load": a.load, "scheduler": a.scheduler}
    print(json.dumps(out), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump({"summary": out, "results": [asdict(r) for r in res]}, f)
    return 0 if out["all_exact"] else 1


if __name__ == "__main__":
    sys.exit(main())
