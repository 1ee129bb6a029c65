"""python -m nqueens [nmin] [nmax] [--mode plain|batched] [--scheduler fifo|mlp] [--preset P] [--chunks K]
                  [--chunk-range FIRST LAST] [--progress-file PATH] [--cache-dir DIR]

Prints the reference solver's table (13Bit_codon/18Py_constellations_GPU_cuda_codon.py:2115-2160)
-- N, Total, Unique, hh:mm:ss.ms and ok/ng against A000170 (18Py:2089) --
with Unique filled in rather than 0, from N=5 by default as the reference's
main (N <= 5 by a plain full-board DFS, 18Py:2118-2126).  Every count runs on
the GPU.  With
--chunks K each count runs as K device runs over interleaved slices of the
constellations and prints a progress line after each; --chunk-range runs
only some of them (a count split over several runs), --progress-file logs
each counted chunk and skips the chunks the log already holds (resume),
--cache-dir keeps the constellation list and per-constellation counts in the
reference's cache files (nqueens.cache).
"""
from __future__ import annotations

import argparse
import sys
import time

from . import solver
from .expected import TOTAL as EXPECTED


def fmt(sec: float) -> str:
    ms = int(round(sec * 1000))
    h, ms = divmod(ms, 3600_000)
    m, ms = divmod(ms, 60_000)
    s, ms = divmod(ms, 1000)
    return f"{h}:{m:02d}:{s:02d}.{ms:03d}"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m nqueens")
    ap.add_argument("nmin", type=int, nargs="?", default=5)
    ap.add_argument("nmax", type=int, nargs="?", default=None)
    ap.add_argument("--mode", choices=["plain", "batched"], default="batched")
    ap.add_argument("--scheduler", choices=["fifo", "mlp"], default="mlp")
    ap.add_argument("--preset", type=int, default=None)
    ap.add_argument("--chunks", type=int, default=1)
    ap.add_argument("--chunk-range", type=int, nargs=2, metavar=("FIRST", "LAST"), default=None,
                    help="count only chunks FIRST..LAST-1 of --chunks and print their totals "
                         "(the totals of all chunks, from any runs, add up to the N-Queens total)")
    ap.add_argument("--cache-dir", default=None,
                    help="keep constellations_N{N}_{preset}.bin (+ .done) and solutions_N{N}_{preset}_v2.bin "
                         "here, as the reference does: valid files are read, others rebuilt")
    ap.add_argument("--progress-file", default=None,
                    help="tab-separated per-chunk log (see nqueens.progress): chunks it already holds "
                         "are skipped, so a rerun resumes; the N-Queens row is printed once it is complete")
    a = ap.parse_args(argv)
    if a.chunk_range is not None or a.progress_file is not None:
        return chunked(a)
    nmax = a.nmax if a.nmax is not None else a.nmin
    ctx = solver.default_context()
    print(" N:             Total           Unique         hh:mm:ss.ms")
    bad = 0
    for N in range(a.nmin, nmax + 1):
        t0 = time.perf_counter()
        def progress(i, k, t, N=N, t0=t0, acc=[0]):
            acc[0] += t
            print(f"  N={N}: chunk {i + 1}/{k} counted, partial total {acc[0]}, "
                  f"{time.perf_counter() - t0:.1f} s", flush=True)

        total, unique = solver.count(N, preset=a.preset, mode=a.mode, scheduler=a.scheduler, ctx=ctx,
                                     chunks=a.chunks, progress=progress if a.chunks > 1 else None,
                                     cache_dir=a.cache_dir, log=lambda m: print(f"  [cache] {m}", flush=True))
        dt = time.perf_counter() - t0
        ok = "ok" if EXPECTED.get(N) == total else f"ng({total}!={EXPECTED.get(N)})"
        bad += ok != "ok"
        print(f"{N:2d}:{total:18d}{unique:17d}{fmt(dt):>20s}    {ok}", flush=True)
    return 1 if bad else 0


def chunked(a) -> int:
    """One board size counted chunk by chunk (--chunk-range / --progress-file)."""
    from .progress import ChunkLog
    N = a.nmin
    if a.nmax is not None and a.nmax != N:
        raise SystemExit("--chunk-range / --progress-file count one board size")
    first, last = a.chunk_range if a.chunk_range is not None else (0, a.chunks)
    preset = a.preset or solver.preset_for(N)
    log = ChunkLog(a.progress_file, N, a.chunks, preset) if a.progress_file else None
    if log and a.preset is None and log.other_presets and not log.totals:
        # the measured preset of N changed since the log was started: its
        # chunks partition another record set
        raise SystemExit(f"{a.progress_file} holds chunks of N={N} / {a.chunks} at preset "
                         f"{sorted(log.other_presets)}, not at {preset}: pass --preset to resume it")
    todo = log.missing(first, last) if log else list(range(first, last))
    t0 = time.perf_counter()
    ctx = solver.default_context()
    done = []
    for i in todo:
        t1 = time.perf_counter()
        t = solver.count_chunks(N, a.chunks, preset=preset, mode=a.mode, scheduler=a.scheduler, ctx=ctx, ids=[i])[0]
        if log:
            log.record(i, t, time.perf_counter() - t1)
        done.append(t)
        print(f"N={N} chunk {i}/{a.chunks} total {t} ({time.perf_counter() - t0:.1f} s)", flush=True)
    print(f"N={N}: {len(done)} chunks counted now, sum {sum(done)}", flush=True)
    if log and log.complete():
        total = log.total()
        f90, f180 = ctx.symmetric_counts(N)
        unique = solver.unique_count(total, f90, f180)
        ok = "ok" if EXPECTED.get(N) == total else f"ng({total}!={EXPECTED.get(N)})"
        print(" N:             Total           Unique")
        print(f"{N:2d}:{total:18d}{unique:17d}    {ok}", flush=True)
        return 0 if ok == "ok" else 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
