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
