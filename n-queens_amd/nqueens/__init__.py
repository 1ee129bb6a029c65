"""nqueens: MI355X (gfx950) N-Queens constellation counter.

``count(N) -> (total, unique)`` counts every solution of the N x N board
exactly on the GPU.  The host builds the canonical constellations
(:mod:`nqueens.constellations`); libnq.so (HIP, loaded through ctypes in
:mod:`nqueens._lib`) runs the bitboard DFS.
"""
from .constellations import gen_constellations, start_constellations, jasmin, symmetry
from .solver import count, count_tasks, exec_solutions, unique_count, default_context, default_sched
from ._lib import Context, Sched, Timing, NQError, load_library, decode_host

__all__ = ["count", "count_tasks", "exec_solutions", "unique_count", "default_context", "default_sched",
           "gen_constellations", "start_constellations", "jasmin", "symmetry", "Context", "Sched", "Timing",
           "NQError", "load_library", "decode_host"]
