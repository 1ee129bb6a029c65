"""count(N) -> (total, unique): the constellation solver's public API.

Mirrors the reference's per-N driver (13Bit_codon/18Py_constellations_GPU_cuda_codon.py:2116-2160):
generate the constellations (gen_constellations, :2139), count each one
(exec_solutions, :2153 -> :1057-1275) and sum the weighted counts (:2156).
Counting runs on the GPU through libnq; there is no host fallback.

Unique: the reference's constellation solvers print 0 (:2160).  Unique here
is the number of solution orbits under the 8 board symmetries, the quantity
the reference's bit-symmetry solver reports
(13Bit_codon/07Py_bit_symmetry_codon.py:410-411, unique = c2 + c4 + c8).
With F90 and F180 the solutions fixed by the quarter and half turns, and no
solution fixed by a reflection, Burnside's lemma gives
unique = (total + F180 + 2 * F90) / 8; F90 and F180 come from a GPU
enumerator of rotation-symmetric solutions.
"""
from __future__ import annotations

import threading

import numpy as np

from . import constellations as _cons
from ._lib import Context, Sched, as_records

__all__ = ["count", "count_chunks", "count_tasks", "exec_solutions", "unique_count", "default_context", "default_sched",
           "preset_for", "constellations_for", "MIN_N", "MAX_N"]

MIN_N, MAX_N = 6, 27      # the constellation path (the reference's range, 18Py:2089)
SMALL_N = range(1, 6)     # plain full DFS, as 18Py's main does for N <= 5 (:2118-2126)

_ctx: dict[int, Context] = {}
_ctx_lock = threading.Lock()


def default_context(device: int = 0) -> Context:
    """A process-wide context per device (created on first use)."""
    with _ctx_lock:
        c = _ctx.get(device)
        if c is None:
            c = Context(device)
            _ctx[device] = c
        return c


def preset_for(N: int) -> int:
    """Preset-queen depth of a count of N: the one the autotuner measured
    fastest for a whole count (load + run, autotune.tuned_preset), else 5 as
    18Py always uses (:2139), up to N = 22.

    The walk keeps 16 parameter bytes per record, and a record of preset p
    has its terminal row at depth N-1-p, so N = 23 needs preset 6 or more,
    N = 24 preset 7 and N = 25 preset 8 (the reference's choice there,
    359Py:5776-5785), N = 26 and 27 presets 9 and 10; a measured preset
    outside 4..10 or below N-17 is not used."""
    from .autotune import tuned_preset
    lo = max(5, N - 17)
    p = tuned_preset(N)
    if p is not None and max(4, N - 17) <= p <= 10:
        return p
    return lo


def constellations_for(N: int, preset: int, ctx: Context | None = None):
    """The constellation records of (N, preset): generated on the host up to
    preset 5, the further preset queens placed on the device
    (Context.expand_records, the same records as the host generator gives)."""
    if preset <= 5:
        return _cons.gen_constellations(N, preset)
    ctx = ctx or default_context()
    return ctx.expand_records(N, _cons.gen_constellations(N, 5), preset - 5)


def default_sched(N: int | None = None) -> Sched:
    """Batched schedule used when none is given (split 3, 128-lane groups, LDS stack)."""
    return Sched(split_depth=3, block=128, lds_stack=1, use_graph=0, chunk=64, waves_per_eu=0)


def unique_count(total: int, f90: int, f180: int) -> int:
    num = int(total) + int(f180) + 2 * int(f90)
    if num % 8:
        raise ArithmeticError(f"orbit count is not integral: ({total} + {f180} + 2*{f90}) / 8")
    return num // 8


def _check_n(N: int) -> None:
    if not MIN_N <= int(N) <= MAX_N:
        raise ValueError(f"N={N} outside the GPU path ({MIN_N}..{MAX_N})")


def count_tasks(N: int, records, *, mode: str = "batched", sched: Sched | None = None, order=None,
                nodes: bool = False, ctx: Context | None = None):
    """Count a list of constellation records on the GPU.

    Returns (weighted per record u64 array, node counts or None, total).
    `mode` is "plain" (one lane per constellation) or "batched" (split into
    work items, persistent kernel).  `order` is the batched processing order."""
    _check_n(N)
    ctx = ctx or default_context()
    rec = as_records(records)
    ctx.load(N, rec, order)
    if mode == "plain":
        total, _ = ctx.run_plain(count_nodes=nodes)
    elif mode == "batched":
        total, _ = ctx.run_batched(sched or default_sched(N), count_nodes=nodes)
    else:
        raise ValueError(f"unknown mode {mode!r}")
    w, n = ctx.results(weighted=True, nodes=nodes)
    return w, n, total


def exec_solutions(N: int, constellations: list, use_gpu: bool = True) -> None:
    """exec_solutions(N, constellations, use_gpu) of 18Py:1057: writes each
    dict's "solutions" (weighted count) in place.

    The reference runs its CPU dfs_iter under @par when use_gpu is False
    (18Py:1237-1266) and its GPU kernel otherwise; both give the same
    counts.  This library has one counting path, the GPU kernels, and takes
    it for either value of use_gpu, so a caller that passes False gets the
    same numbers the reference's CPU path would write."""
    del use_gpu   # one path: the counts do not depend on where they are computed
    if not constellations:
        return
    rec = np.array([[c["ld"], c["rd"], c["col"], c["startijkl"]] for c in constellations], dtype=np.uint64)
    if (rec >> 32).any():
        raise OverflowError("a record field does not fit 32 bits")
    w, _, _ = count_tasks(N, rec.astype(np.uint32))
    for c, v in zip(constellations, w):
        c["solutions"] = int(v)


def count_chunks(N: int, chunks: int, first: int = 0, last: int | None = None, *, preset: int | None = None,
                 mode: str = "batched", scheduler: str = "fifo", sched: Sched | None = None,
                 ctx: Context | None = None, progress=None, ids=None) -> list[int]:
    """Weighted totals of chunks first..last-1 (or of the chunk indices `ids`)
    of a count split in `chunks`.

    Chunk i holds the constellations that descend from every chunks-th
    record of preset min(preset, 5), starting at record i; deeper presets are
    expanded on the device.  The chunks partition the constellation set, so
    the totals of all chunks add up to the N-Queens total, whichever runs
    (or processes) count them: a long count can be split over several runs
    and resumed.  progress(i, chunks, total_i) is called after each chunk."""
    _check_n(N)
    last = chunks if last is None else last
    if not (chunks >= 1 and 0 <= first <= last <= chunks):
        raise ValueError(f"chunk range {first}..{last} of {chunks}")
    ids = list(range(first, last)) if ids is None else [int(i) for i in ids]
    if any(not 0 <= i < chunks for i in ids):
        raise ValueError(f"chunk index outside 0..{chunks - 1}")
    if scheduler not in ("fifo", "mlp"):
        raise ValueError(f"unknown scheduler {scheduler!r}")
    ctx = ctx or default_context()
    preset = preset or preset_for(N)
    # One chunk is the whole set: the host generates preset 4 (a tenth of
    # preset 5's records at N=19, a fifth of its time) and the device places
    # the rest; the same multiset of records, in another order.
    host_preset = min(preset, 4 if ids == [0] and chunks == 1 else 5)
    base = _cons.gen_constellations(N, host_preset)
    levels = preset - host_preset
    if mode == "batched" and sched is None:
        from .autotune import best_sched
        sched = best_sched(N, preset)
    totals = []
    for i in ids:
        part = base[i::chunks] if chunks > 1 else base
        if levels:
            ctx.load_expanded(N, part, levels)
        else:
            ctx.load(N, part)
        if scheduler == "mlp" and ctx.m:
            from .costmodel import default_model
            ctx.order_by_cost(default_model().params)       # on the device, heaviest first
        if mode == "plain":
            t, _ = ctx.run_plain()
        elif mode == "batched":
            t, _ = ctx.run_batched(sched)
        else:
            raise ValueError(f"unknown mode {mode!r}")
        totals.append(int(t))
        if progress is not None:
            progress(i, chunks, int(t))
    return totals


def _count_list(ctx: Context, N: int, rec, mode: str, scheduler: str, sched: Sched | None, preset: int):
    """Weighted count of every record of one list (one device run)."""
    ctx.load(N, rec)
    if scheduler == "mlp" and ctx.m:
        from .costmodel import default_model
        ctx.order_by_cost(default_model().params)
    if mode == "plain":
        ctx.run_plain()
    else:
        if sched is None:
            from .autotune import best_sched
            sched = best_sched(N, preset)
        ctx.run_batched(sched)
    w, _ = ctx.results(weighted=True)
    return w


def count(N: int, *, preset: int | None = None, mode: str = "batched", scheduler: str = "fifo",
          sched: Sched | None = None, ctx: Context | None = None, chunks: int = 1,
          progress=None, cache_dir=None, log=None) -> tuple[int, int]:
    """Total and unique N-Queens solutions for 1 <= N <= 27, counted on the GPU.

    N = 1..5 are counted by a plain full DFS over the board, as 18Py's main
    does for N <= 5 (_bit_total, :2071-2084, :2118-2126); N >= 6 through
    constellations.  chunks > 1 counts the constellations in that many
    slices, one device run each (as the reference counts fixed-size chunks
    of its list, 359Py:5621-5626; see count_chunks), and calls
    progress(i, chunks, chunk_total) after each: a long count then reports
    as it goes and no single launch holds the GPU for the whole run.

    cache_dir: keep the constellation list and the per-constellation counts
    on disk in the reference's formats (nqueens.cache:
    constellations_N{N}_{preset}.bin + .done, solutions_N{N}_{preset}_v2.bin,
    18Py:1788-2068, 359Py:2960-3134): a valid list is read instead of
    generated, a complete solutions file replaces the device count, and
    whatever does not validate is built again and rewritten.  log(msg) hears
    about rebuilt caches.  The list is counted in one run (chunks must be 1;
    a chunked count resumes from a progress file instead)."""
    if chunks < 1:
        raise ValueError("chunks must be >= 1")
    if mode not in ("plain", "batched"):
        raise ValueError(f"unknown mode {mode!r}")
    if scheduler not in ("fifo", "mlp"):
        raise ValueError(f"unknown scheduler {scheduler!r}")
    if cache_dir is not None and chunks != 1:
        raise ValueError("cache_dir counts the whole list in one run (chunks must be 1)")
    if int(N) in SMALL_N:
        # one device launch; chunks, preset and schedule do not apply
        ctx = ctx or default_context()
        total = ctx.small_total(int(N))
        if progress is not None:
            progress(0, 1, total)
        if int(N) == 1:
            # the one-square board: its solution is fixed by every symmetry,
            # reflections included, so the rotation-only orbit formula does
            # not apply; one solution, one orbit
            return total, total
        f90, f180 = ctx.symmetric_counts(int(N))
        return total, unique_count(total, f90, f180)
    _check_n(N)
    ctx = ctx or default_context()
    if cache_dir is not None:
        from . import cache
        preset = preset or preset_for(N)
        rec, _ = cache.load_or_build_constellations(cache_dir, N, preset,
                                                    lambda: constellations_for(N, preset, ctx), log)
        w, _ = cache.load_or_build_solutions(cache_dir, N, preset, rec,
                                             lambda r: _count_list(ctx, N, r, mode, scheduler, sched, preset), log)
        total = int(w.sum(dtype=np.uint64))
        if progress is not None:
            progress(0, 1, total)
    else:
        total = sum(count_chunks(N, chunks, preset=preset, mode=mode, scheduler=scheduler, sched=sched, ctx=ctx,
                                 progress=progress))
    f90, f180 = ctx.symmetric_counts(N)
    return total, unique_count(total, f90, f180)
