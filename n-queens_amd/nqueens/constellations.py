"""Canonical start constellations and their preset-queen expansions.

Host side of the constellation counter: builds the 16-byte records
``{ld, rd, col, startijkl}`` that the GPU counts.  Mirrors
``gen_constellations`` of 13Bit_codon/18Py_constellations_GPU_cuda_codon.py
(:1622-1785) and its helpers (:1396-1613):

* the ijkl packing, rotations, mirror and Jasmin normal form (:1396-1453);
* the order-dependent rotation check while the start set is built (:1687-1716);
* the preset-queen expansion ``set_pre_queens`` (:1535-1613), here as a
  vectorised breadth-first walk over all start constellations at once.

Expansion semantics.  The reference memoises expansion states.  18Py keeps
one memo for the whole call (:1681), which drops records whenever two start
constellations reach the same state (preset >= 6; its own logs show N=18
preset 6 falling to 571,640 records and a wrong total).  The later solver
clears it per start constellation
(14Bit_codon_GPU/359Py_pushbranch_partial.py:2919).  This module enumerates
every path (no memo), which is exact; for every N <= 21 at presets 4..6 it
produces the same record multiset as the per-start memo, and at preset 7 it
reproduces the reference's preset-7 record count (N=18: 3,653,940).
"""
from __future__ import annotations

import os

import numpy as np

__all__ = [
    "to_ijkl", "geti", "getj", "getk", "getl", "rot90", "rot180", "mirvert",
    "symmetry", "symmetry90", "jasmin", "check_rotations", "start_constellations", "start_array",
    "gen_constellations", "read_constellations_bin", "write_constellations_bin",
    "read_solutions_bin", "write_solutions_bin", "records_sorted",
]


# --- A1: ijkl packing (18Py:1396-1402) ------------------------------------
def to_ijkl(i: int, j: int, k: int, l: int) -> int:
    return (i << 15) + (j << 10) + (k << 5) + l


def geti(x: int) -> int:
    return (x >> 15) & 0x1F


def getj(x: int) -> int:
    return (x >> 10) & 0x1F


def getk(x: int) -> int:
    return (x >> 5) & 0x1F


def getl(x: int) -> int:
    return x & 0x1F


def rot90(x: int, N: int) -> int:
    """Quarter turn (r, c) -> (c, N-1-r) of the four border queens (18Py:1405)."""
    return ((N - 1 - getk(x)) << 15) + ((N - 1 - getl(x)) << 10) + (getj(x) << 5) + geti(x)


def rot180(x: int, N: int) -> int:
    return ((N - 1 - getj(x)) << 15) + ((N - 1 - geti(x)) << 10) + ((N - 1 - getl(x)) << 5) + (N - 1 - getk(x))


def mirvert(x: int, N: int) -> int:
    """Vertical mirror, which also swaps the k and l columns (18Py:1397)."""
    return to_ijkl(N - 1 - geti(x), N - 1 - getj(x), getl(x), getk(x))


def symmetry90(x: int, N: int) -> bool:
    return x == rot90(x, N)


def symmetry(x: int, N: int) -> int:
    """Weight of a start constellation: 2 if fixed by the quarter turn, 4 if by
    the half turn, else 8 (18Py:1407)."""
    if symmetry90(x, N):
        return 2
    if geti(x) == N - 1 - getj(x) and getk(x) == N - 1 - getl(x):
        return 4
    return 8


def jasmin(x: int, N: int) -> int:
    """Jasmin normal form (18Py:1429-1453): turn the border queen nearest an
    edge onto the j side, then mirror so that j lies in the right half."""
    arg = 0
    mn = min(getj(x), N - 1 - getj(x))
    if min(geti(x), N - 1 - geti(x)) < mn:
        arg, mn = 2, min(geti(x), N - 1 - geti(x))
    if min(getk(x), N - 1 - getk(x)) < mn:
        arg, mn = 3, min(getk(x), N - 1 - getk(x))
    if min(getl(x), N - 1 - getl(x)) < mn:
        arg, mn = 1, min(getl(x), N - 1 - getl(x))
    for _ in range(arg):
        x = rot90(x, N)
    if getj(x) < N - 1 - getj(x):
        x = mirvert(x, N)
    return x


def check_rotations(seen: set, i: int, j: int, k: int, l: int, N: int) -> bool:
    """True when a 90/180/270-degree image of (i,j,k,l) is already in `seen`
    (18Py:1411-1412)."""
    return (((N - 1 - k) << 15) + ((N - 1 - l) << 10) + (j << 5) + i) in seen or \
        (((N - 1 - j) << 15) + ((N - 1 - i) << 10) + ((N - 1 - l) << 5) + (N - 1 - k)) in seen or \
        ((l << 15) + (k << 10) + ((N - 1 - i) << 5) + (N - 1 - j)) in seen


def _grid(*ranges) -> list[np.ndarray]:
    """Index arrays of nested loops over `ranges` (first range outermost), in
    loop order."""
    g = np.meshgrid(*[np.asarray(r, np.int64) for r in ranges], indexing="ij")
    return [a.ravel() for a in g]


def _jasmin_array(x: np.ndarray, N: int) -> np.ndarray:
    """`jasmin` of every element of x (same comparisons, same tie order)."""
    N1 = N - 1
    i, j, k, l = (x >> 15) & 31, (x >> 10) & 31, (x >> 5) & 31, x & 31
    arg = np.zeros_like(x)
    mn = np.minimum(j, N1 - j)
    for a, v in ((2, i), (3, k), (1, l)):
        d = np.minimum(v, N1 - v)
        t = d < mn
        arg = np.where(t, a, arg)
        mn = np.where(t, d, mn)
    for r in range(3):                                    # arg quarter turns
        t = arg > r
        i, j, k, l = (np.where(t, N1 - k, i), np.where(t, N1 - l, j), np.where(t, j, k), np.where(t, i, l))
    mv = j < N1 - j                                       # mirvert
    return (np.where(mv, N1 - i, i) << 15) | (np.where(mv, N1 - j, j) << 10) | (np.where(mv, l, k) << 5) | \
        np.where(mv, k, l)


def _first_occurrences(key: np.ndarray) -> np.ndarray:
    """Positions of the first occurrence of each distinct key, ascending."""
    _, first = np.unique(key, return_index=True)
    return np.sort(first)


def start_array(N: int) -> np.ndarray:
    """`start_constellations` as an int64 array."""
    N1, N2, half = N - 1, N - 2, (N + 1) // 2
    blocks = []
    centre = [N // 2] if N % 2 == 1 else []               # centre block, 18Py:1687-1699
    for k in centre + list(range(1, half)):               # no corner queen, 18Py:1703-1712
        l, i, j = _grid(range(k + 1, N1), range(k + 1, N1), range(N - k - 2, 0, -1))
        m = (i != N1 - l) & (j != i) & (j != l)
        blocks.append((i[m], j[m], np.full(int(m.sum()), k, np.int64), l[m]))
    if not blocks:
        return np.zeros(0, np.int64)
    i, j, k, l = (np.concatenate(a) for a in zip(*blocks))
    x = (i << 15) | (j << 10) | (k << 5) | l
    # a candidate is skipped when a quarter, half or three-quarter turn of it
    # was added before it; the turns form orbits, so exactly the first
    # candidate of each orbit (in loop order) is added
    r90 = ((N1 - k) << 15) | ((N1 - l) << 10) | (j << 5) | i
    r180 = ((N1 - j) << 15) | ((N1 - i) << 10) | ((N1 - l) << 5) | (N1 - k)
    r270 = (l << 15) | (k << 10) | ((N1 - i) << 5) | (N1 - j)
    orbit = np.minimum(np.minimum(x, r90), np.minimum(r180, r270))
    kept = x[_first_occurrences(orbit)]
    j, l = _grid(range(1, N2), range(0, N1))              # corner queen, 18Py:1716 (no check)
    m = l > j
    seen = np.concatenate([kept, (j[m] << 10) | l[m]])
    jas = _jasmin_array(seen, N)                          # 18Py:1719, first occurrence kept
    return jas[_first_occurrences(jas)]


def start_constellations(N: int) -> list[int]:
    """Canonical start constellations (Jasmin forms), in insertion order.

    Candidates are added one at a time, each skipped when one of its rotations
    was added before it -- the incremental ``set.update`` of 18Py:1687-1712 --
    then every survivor is mapped to its Jasmin form (18Py:1719).  Computed
    over arrays: within each rotation orbit the first candidate in loop order
    is the one added (`start_array`)."""
    return start_array(N).tolist()


def gen_constellations(N: int, preset: int = 5) -> np.ndarray:
    """All constellations of board size N with `preset` queens placed.

    Returns an (m, 4) uint32 array of records ``ld, rd, col, startijkl``
    (``startijkl = row << 20 | ijkl``), grouped by start constellation.
    The per-start setup is 18Py:1737-1750; the expansion is set_pre_queens
    (18Py:1535-1613): rows k and l are skipped, queens are placed on free
    columns ``~(ld|rd|col|LD>>(N-1-row)|RD<<(N-1-row))`` until `preset`
    queens stand (3 already stand when j == N-1, else 4).
    """
    if not 5 <= N <= 27:
        raise ValueError(f"N={N} outside 5..27")
    if not 3 <= preset <= 10:
        raise ValueError(f"preset={preset} outside 3..10")
    scs = start_array(N)
    if scs.size == 0:
        return np.zeros((0, 4), np.uint32)
    i, j, k, l = (scs >> 15) & 31, (scs >> 10) & 31, (scs >> 5) & 31, scs & 31
    L = np.int64(1) << (N - 1)
    one = np.int64(1)
    ld = np.where(i > 0, L >> np.maximum(i - 1, 0), 0) | (one << (N - k))
    rd = (L >> (i + 1)) | (one << (l - 1))
    col = 1 | L | (L >> i) | (L >> j)
    LD = (L >> j) | (L >> l)
    RD = (L >> j) | (one << k)
    mask = (one << N) - 1
    sc = np.arange(scs.size, dtype=np.int64)
    row = np.ones(scs.size, dtype=np.int64)
    queens = np.where(j == N - 1, 3, 4).astype(np.int64)
    done_ld, done_rd, done_col, done_row, done_sc = [], [], [], [], []
    while sc.size:
        kk, ll = k[sc], l[sc]
        skip = (row == kk) | (row == ll)
        leaf = ~skip & (queens == preset)
        if leaf.any():
            done_ld.append(ld[leaf]); done_rd.append(rd[leaf]); done_col.append(col[leaf])
            done_row.append(row[leaf]); done_sc.append(sc[leaf])
        # states passing a pre-placed row move down without a placement
        s_ld, s_rd, s_col = ld[skip] << 1, rd[skip] >> 1, col[skip]
        s_row, s_q, s_sc = row[skip] + 1, queens[skip], sc[skip]
        grow = ~skip & ~leaf
        g_ld, g_rd, g_col, g_row, g_q, g_sc = ld[grow], rd[grow], col[grow], row[grow], queens[grow], sc[grow]
        sh = N - 1 - g_row
        if (sh < 0).any():
            raise ValueError("expansion passed the last row before placing the preset queens")
        free = ~(g_ld | g_rd | g_col | (LD[g_sc] >> sh) | (RD[g_sc] << sh)) & mask
        parts = [(s_ld, s_rd, s_col, s_row, s_q, s_sc)]
        for b in range(N):
            hit = ((free >> b) & 1).astype(bool)
            if not hit.any():
                continue
            bit = one << b
            parts.append(((g_ld[hit] | bit) << 1, (g_rd[hit] | bit) >> 1, g_col[hit] | bit,
                          g_row[hit] + 1, g_q[hit] + 1, g_sc[hit]))
        ld, rd, col, row, queens, sc = (np.concatenate([p[a] for p in parts]) for a in range(6))
    if not done_sc:
        return np.zeros((0, 4), np.uint32)
    ld, rd, col = np.concatenate(done_ld), np.concatenate(done_rd), np.concatenate(done_col)
    row, sc = np.concatenate(done_row), np.concatenate(done_sc)
    order = np.argsort(sc, kind="stable")                 # group by start constellation
    ld, rd, col, row, sc = ld[order], rd[order], col[order], row[order], sc[order]
    startijkl = (row << 20) | scs[sc]                     # 18Py:1777-1781
    out = np.stack([ld, rd, col, startijkl], axis=1)
    if (out >> 32).any():
        raise OverflowError("a record field does not fit 32 bits")
    return np.ascontiguousarray(out.astype(np.uint32))


def records_sorted(rec: np.ndarray) -> np.ndarray:
    """Records in lexicographic order of (ld, rd, col, startijkl): a canonical
    order for comparing record multisets."""
    rec = np.asarray(rec)
    return rec[np.lexsort(rec.T[::-1])]


# --- on-disk formats (18Py:1788-2068) --------------------------------------
def read_constellations_bin(path) -> np.ndarray:
    """16-byte records, u32 little endian ld, rd, col, startijkl (18Py:2009-2032)."""
    if os.path.getsize(path) % 16:
        raise ValueError(f"{path}: size is not a multiple of 16 bytes")
    raw = np.fromfile(path, dtype="<u4")
    return np.ascontiguousarray(raw.reshape(-1, 4).astype(np.uint32))


def write_constellations_bin(path, rec: np.ndarray) -> None:
    np.ascontiguousarray(np.asarray(rec, dtype="<u4").reshape(-1, 4)).tofile(path)


def read_solutions_bin(path) -> tuple[np.ndarray, np.ndarray]:
    """40-byte records, u64 little endian startijkl, ld, rd, col, solutions
    (18Py:1841-1927).  Returns (records (m,4) u32, weighted solutions (m,) u64)."""
    if os.path.getsize(path) % 40:
        raise ValueError(f"{path}: size is not a multiple of 40 bytes")
    raw = np.fromfile(path, dtype="<u8")
    raw = raw.reshape(-1, 5)
    rec = np.stack([raw[:, 1], raw[:, 2], raw[:, 3], raw[:, 0]], axis=1).astype(np.uint32)
    return np.ascontiguousarray(rec), np.ascontiguousarray(raw[:, 4].astype(np.uint64))


def write_solutions_bin(path, rec: np.ndarray, solutions: np.ndarray) -> None:
    rec = np.asarray(rec, dtype=np.uint64).reshape(-1, 4)
    out = np.stack([rec[:, 3], rec[:, 0], rec[:, 1], rec[:, 2],
                    np.asarray(solutions, dtype=np.uint64)], axis=1)
    np.ascontiguousarray(out.astype("<u8")).tofile(path)
