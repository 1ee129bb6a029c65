"""Helpers shared by the tests (digests in the fixture format)."""
import hashlib

import numpy as np


def sorted_rows(a: np.ndarray) -> np.ndarray:
    return a[np.lexsort(a.T[::-1])]


def records_sha256(rec: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(sorted_rows(np.asarray(rec)).astype("<u4")).tobytes()).hexdigest()


def solutions_sha256(rec: np.ndarray, sol: np.ndarray) -> str:
    a = np.concatenate([np.asarray(rec).astype(np.uint64), np.asarray(sol).astype(np.uint64)[:, None]], axis=1)
    return hashlib.sha256(np.ascontiguousarray(sorted_rows(a).astype("<u8")).tobytes()).hexdigest()


def p5_records_n28(oracle, parents=2, stride=40000):
    """P5 records with a terminal depth the walk holds: N=28 at preset 11,
    j = 14 and the start row 8 (the j diagonal enters the board on row 12),
    from the preset-5 records with j = 14, k, l >= 8 (oracle generator and
    expansion)."""
    N = 28
    p5, _ = oracle.gen_constellations(N, 5)
    ijkl = p5[:, 3] & 0xFFFFF
    j, k, l, st = (ijkl >> 10) & 31, (ijkl >> 5) & 31, ijkl & 31, p5[:, 3] >> 20
    par = p5[np.nonzero((j == 14) & (k >= 8) & (l >= 8) & (st == 2))[0][:parents]]
    kids = oracle.expand_records(N, par, 6)
    return par, kids, np.ascontiguousarray(kids[::stride])
