"""On-disk caches of a count, in the reference's formats and with its rules.

constellations_N{N}_{preset}.bin + .done
    The constellation list: 16-byte records, u32 little endian ld, rd, col,
    startijkl (18Py:2009-2032), and beside it a text file holding the record
    count, written after the records (359Py:2993-3012, 3020-3116).  The cache
    is used only when the file is not empty, its size is a multiple of 16
    (validate_bin_file, 18Py:1808-1815) and the .done count equals its record
    count (ensure_constellations_bin_stream, 359Py:3118-3134); otherwise the
    list is generated again and both files are rewritten.  A truncated,
    corrupt or half-written file is never read as a list.

solutions_N{N}_{preset}_v2.bin
    The weighted count of each constellation: 40-byte records, u64 little
    endian startijkl, ld, rd, col, solutions (save_solutions_bin_v2,
    18Py:1841-1855).  Loading fills the constellation list by the key
    (startijkl, ld, rd, col) (load_solutions_bin_into_v2, 18Py:1896-1927); a
    file whose size is not a non-zero multiple of 40, or any constellation
    without its key, is a miss: every constellation is counted again and the
    file rewritten (load_or_build_solutions_bin, 18Py:1946-1959).

Files are written to a temporary name and renamed, so that a reader never
sees part of a new file under the cache's name.
"""
from __future__ import annotations

import os
from pathlib import Path

import numpy as np

from .constellations import read_constellations_bin, write_constellations_bin, write_solutions_bin

__all__ = ["constellations_path", "solutions_path", "load_constellations", "save_constellations",
           "load_or_build_constellations", "load_solutions", "load_or_build_solutions"]


def constellations_path(cache_dir, N: int, preset: int) -> Path:
    return Path(cache_dir) / f"constellations_N{N}_{preset}.bin"


def solutions_path(cache_dir, N: int, preset: int) -> Path:
    return Path(cache_dir) / f"solutions_N{N}_{preset}_v2.bin"


def _done_count(path: Path) -> int:
    """The record count in the .done file, or -1 (359Py:2993-3003)."""
    try:
        text = Path(str(path) + ".done").read_text().strip()
        return int(text) if text else -1
    except (OSError, ValueError):
        return -1


def _replace(path: Path, write) -> None:
    tmp = path.with_name(path.name + f".tmp{os.getpid()}")
    write(tmp)
    os.replace(tmp, path)


def load_constellations(cache_dir, N: int, preset: int):
    """The cached list of (N, preset), or None when the cache is absent or
    does not validate (empty, size not a multiple of 16, .done count absent or
    different from the record count)."""
    path = constellations_path(cache_dir, N, preset)
    try:
        size = path.stat().st_size
    except OSError:
        return None
    if size == 0 or size % 16 or _done_count(path) != size // 16:
        return None
    return read_constellations_bin(path)


def save_constellations(cache_dir, N: int, preset: int, rec: np.ndarray) -> Path:
    """Writes the list, then the .done count."""
    path = constellations_path(cache_dir, N, preset)
    path.parent.mkdir(parents=True, exist_ok=True)
    rec = np.ascontiguousarray(np.asarray(rec, dtype=np.uint32).reshape(-1, 4))
    _replace(path, lambda p: write_constellations_bin(p, rec))
    _replace(Path(str(path) + ".done"), lambda p: p.write_text(f"{rec.shape[0]}\n"))
    return path


def load_or_build_constellations(cache_dir, N: int, preset: int, build, log=None):
    """(records, True) from a valid cache, else (build(), False) after saving
    it (load_or_build_constellations_bin, 18Py:2034-2053)."""
    rec = load_constellations(cache_dir, N, preset)
    if rec is not None:
        return rec, True
    path = constellations_path(cache_dir, N, preset)
    if log is not None and path.exists():
        log(f"constellation cache {path} does not validate; rebuilding")
    rec = np.ascontiguousarray(np.asarray(build(), dtype=np.uint32))
    save_constellations(cache_dir, N, preset, rec)
    return rec, False


def load_solutions(path, rec: np.ndarray):
    """Weighted counts of the records from a v2 file, in record order, or None
    on a miss (unreadable file, size not a non-zero multiple of 40, or a
    record whose key (startijkl, ld, rd, col) the file lacks)."""
    path = Path(path)
    try:
        size = path.stat().st_size
    except OSError:
        return None
    if size == 0 or size % 40:
        return None
    raw = np.fromfile(path, dtype="<u8").reshape(-1, 5)
    rec = np.asarray(rec, dtype=np.uint64).reshape(-1, 4)
    if rec.shape[0] == 0:
        return np.zeros(0, np.uint64)
    # keys (startijkl, ld, rd, col) as 32-byte big-endian strings: byte order
    # is numeric order, so one sort and a binary search match them
    kf = np.ascontiguousarray(raw[:, :4].astype(">u8")).view("V32").ravel()
    kr = np.ascontiguousarray(np.stack([rec[:, 3], rec[:, 0], rec[:, 1], rec[:, 2]], axis=1)
                              .astype(">u8")).view("V32").ravel()
    # a key the file holds twice keeps its last record (the reference fills a
    # dict, 18Py:1913-1920): sort the file backwards, keep the first of each key
    rev = np.arange(kf.shape[0] - 1, -1, -1)
    order = rev[np.argsort(kf[rev], kind="stable")]
    ks = kf[order]
    pos = np.minimum(np.searchsorted(ks, kr, side="left"), len(ks) - 1)
    if not (ks[pos] == kr).all():
        return None
    return np.ascontiguousarray(raw[order[pos], 4].astype(np.uint64))


def load_or_build_solutions(cache_dir, N: int, preset: int, rec: np.ndarray, count, log=None):
    """(weighted counts, True) from the v2 cache, else (count(rec), False)
    after writing them to it."""
    path = solutions_path(cache_dir, N, preset)
    w = load_solutions(path, rec)
    if w is not None:
        return w, True
    if log is not None and path.exists():
        log(f"solutions cache {path} misses or does not validate; counting again")
    w = np.ascontiguousarray(np.asarray(count(rec), dtype=np.uint64))
    path.parent.mkdir(parents=True, exist_ok=True)
    _replace(path, lambda p: write_solutions_bin(p, rec, w))
    return w, False
