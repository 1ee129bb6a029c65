"""The on-disk caches (nqueens.cache) with the reference's rules, on the CPU
(the oracle stands in for the device count; GPU counts through the caches
are in test_gpu_parity.py): valid files are read, anything that does not
validate is rebuilt and rewritten, never read in part."""
import struct

import numpy as np
import pytest

from nqueens import cache
from nqueens import constellations as C
from tests.helpers import solutions_sha256


class Calls:
    def __init__(self, fn):
        self.fn, self.n = fn, 0

    def __call__(self, *a):
        self.n += 1
        return self.fn(*a)


def test_constellation_cache_hit_and_rebuilds(tmp_path, digests):
    N, p = 12, 5
    build = Calls(lambda: C.gen_constellations(N, p))
    rec, hit = cache.load_or_build_constellations(tmp_path, N, p, build)
    assert not hit and build.n == 1
    path = cache.constellations_path(tmp_path, N, p)
    assert path.stat().st_size == 16 * rec.shape[0] == 16 * 4300
    assert (tmp_path / (path.name + ".done")).read_text().strip() == "4300"
    rec2, hit = cache.load_or_build_constellations(tmp_path, N, p, build)
    assert hit and build.n == 1 and np.array_equal(rec2, rec)
    done = tmp_path / (path.name + ".done")
    damages = [
        lambda: path.write_bytes(path.read_bytes()[:-5]),        # size not a multiple of 16
        lambda: path.write_bytes(path.read_bytes()[:-16]),       # whole records missing: .done disagrees
        lambda: done.unlink(),                                   # no .done: the write never finished
        lambda: done.write_text("oops\n"),                       # unreadable count
        lambda: done.write_text(""),
        lambda: path.write_bytes(b""),                           # empty list
    ]
    for k, damage in enumerate(damages, start=2):
        damage()
        rec3, hit = cache.load_or_build_constellations(tmp_path, N, p, build)
        assert not hit and build.n == k and np.array_equal(rec3, rec), k
        assert cache.load_constellations(tmp_path, N, p) is not None
    assert C.records_sorted(rec).tobytes() == C.records_sorted(C.gen_constellations(N, p)).tobytes()


def test_constellation_cache_format_is_the_references(tmp_path):
    """16-byte records, u32 little endian ld, rd, col, startijkl (18Py:2009-2016)."""
    rec = C.gen_constellations(9, 5)
    cache.save_constellations(tmp_path, 9, 5, rec)
    raw = cache.constellations_path(tmp_path, 9, 5).read_bytes()
    assert raw[:16] == struct.pack("<4I", *map(int, rec[0]))
    assert len(raw) == 16 * len(rec)


def test_solutions_cache_hit_miss_and_rewrite(tmp_path, oracle, digests):
    N, p = 13, 5
    rec = C.gen_constellations(N, p)
    count = Calls(lambda r: oracle.count_records(N, r)[0])
    w, hit = cache.load_or_build_solutions(tmp_path, N, p, rec, count)
    assert not hit and count.n == 1
    assert solutions_sha256(rec, w) == digests[str(N)]["solutions_sha256"]
    path = cache.solutions_path(tmp_path, N, p)
    assert path.stat().st_size == 40 * len(rec)
    # the reference's layout: u64 LE startijkl, ld, rd, col, solutions (18Py:1841-1855)
    assert path.read_bytes()[:40] == struct.pack("<5Q", int(rec[0, 3]), int(rec[0, 0]), int(rec[0, 1]),
                                                 int(rec[0, 2]), int(w[0]))
    # a hit fills the list by key, in any record order, without counting
    perm = np.random.default_rng(1).permutation(len(rec))
    w2, hit = cache.load_or_build_solutions(tmp_path, N, p, rec[perm], count)
    assert hit and count.n == 1 and np.array_equal(w2, w[perm])
    raw = path.read_bytes()
    damages = [
        raw[:-40],            # one key missing: a miss, everything counted again
        raw[:-3],             # size not a multiple of 40
        b"",
        raw[:40 * 7] + struct.pack("<5Q", 1, 2, 3, 4, 5) + raw[40 * 8:],   # one record replaced by a foreign key
    ]
    for k, bad in enumerate(damages, start=2):
        path.write_bytes(bad)
        w3, hit = cache.load_or_build_solutions(tmp_path, N, p, rec, count)
        assert not hit and count.n == k and np.array_equal(w3, w), k
        assert cache.load_solutions(path, rec) is not None          # rewritten whole
    # a key held twice: the later record wins, as the reference's dict fill
    dup = rec[:2]
    C.write_solutions_bin(path, np.concatenate([dup, dup[:1]]), np.array([7, 8, 9], np.uint64))
    assert cache.load_solutions(path, dup).tolist() == [9, 8]


def test_reference_fixture_round_trip(tmp_path, p5_full, digests):
    """The reference's per-constellation counts (solutions_N11_5_v2.bin, kept
    as arrays) written in the v2 format and read back by key."""
    rec, sol = p5_full["rec11"], p5_full["sol11"]
    path = tmp_path / "solutions_N11_5_v2.bin"
    C.write_solutions_bin(path, rec, sol)
    got = cache.load_solutions(path, C.gen_constellations(11, 5))
    assert got is not None
    assert solutions_sha256(C.gen_constellations(11, 5), got) == digests["11"]["solutions_sha256"]


def test_count_cache_arguments():
    import nqueens
    with pytest.raises(ValueError):
        nqueens.count(12, cache_dir="/tmp/unused", chunks=2)
