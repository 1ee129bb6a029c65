"""The CPU oracle against everything the reference fixes (CPU only).

Pins (see oracle/nq_oracle.c):
  * constellation records: constellations_N{6..21}_5.bin multisets
  * per-record weighted counts: solutions_N{6..19}_5_v2.bin
  * totals: A000170 (18Py:2089); unique: README.en:12-33 through Burnside
  * F90/F180 and unique: brute-force enumeration for N <= 12
  * preset-6/7 record counts from the reference's logs
  * node counts: iterative and recursive restatements agree per record
  * the reference's own C solver (oracle/_ref) at preset 4
"""
import numpy as np
import pytest

from tests.helpers import records_sha256, solutions_sha256


@pytest.mark.parametrize("N", range(6, 22))
def test_records_match_fixture(oracle, digests, p5_full, N):
    rec, stats = oracle.gen_constellations(N, 5, scope=oracle.SCOPE_CALL)
    d = digests[str(N)]
    assert rec.shape[0] == d["constellations"]
    assert records_sha256(rec) == d["records_sha256"]
    assert stats[1] == 0 and stats[3] == 0   # no memo or signature hit at preset 5
    if N <= 13:
        assert np.array_equal(np.sort(rec.view([("", "<u4")] * 4), axis=0),
                              np.sort(p5_full[f"rec{N}"].view([("", "<u4")] * 4), axis=0))


@pytest.mark.parametrize("N", range(6, 18))
def test_counts_match_fixture(oracle, digests, p5_full, N):
    d = digests[str(N)]
    if N <= 13:
        rec, want = p5_full[f"rec{N}"], p5_full[f"sol{N}"]
        w, raw, nodes = oracle.count_records(N, rec)
        assert np.array_equal(w, want)
    else:
        rec, _ = oracle.gen_constellations(N, 5)
        w, raw, nodes = oracle.count_records(N, rec)
    assert int(w.sum()) == d["solutions_sum"]
    assert solutions_sha256(rec, w) == d["solutions_sha256"]
    # w = raw * weight with weight in {2,4,8}
    ratio = w[raw > 0] // raw[raw > 0]
    assert set(np.unique(ratio).tolist()) <= {2, 4, 8}


@pytest.mark.parametrize("N", range(6, 14))
def test_recursive_restatement_agrees(oracle, N):
    rec, _ = oracle.gen_constellations(N, 5)
    w, _, nodes = oracle.count_records(N, rec)
    w2, nodes2 = oracle.count_records_recursive(N, rec)
    assert np.array_equal(w, w2)
    assert np.array_equal(nodes, nodes2)
    assert nodes.sum() > 0 or N < 7


@pytest.mark.parametrize("N", range(2, 13))
def test_bruteforce_pins_symmetric_counts(oracle, known, N):
    total, unique, f90, f180 = oracle.bruteforce(N)
    assert oracle.symmetric_counts(N) == (f90, f180)
    assert 8 * unique == total + f180 + 2 * f90
    if str(N) in known["readme_total_unique"]:
        assert [total, unique] == known["readme_total_unique"][str(N)]


@pytest.mark.parametrize("N", range(4, 25))
def test_unique_table_through_burnside(oracle, known, N):
    total, unique = known["readme_total_unique"][str(N)]
    f90, f180 = oracle.symmetric_counts(N)
    assert total + f180 + 2 * f90 == 8 * unique
    if N >= 6:
        assert known["total_a000170"][str(N)] == total


@pytest.mark.parametrize("preset", [4, 5, 6])
@pytest.mark.parametrize("N", range(6, 15))
def test_totals_every_preset(oracle, known, N, preset):
    if preset >= N:
        pytest.skip("every queen preset: no DFS left")
    rec, _ = oracle.gen_constellations(N, preset)
    if rec.shape[0] == 0:
        pytest.skip("no constellation at this preset")
    w, _, _ = oracle.count_records(N, rec)
    assert int(w.sum()) == known["total_a000170"][str(N)]


def test_preset_record_counts(oracle, known):
    pc = known["preset_record_counts"]
    rec, st = oracle.gen_constellations(18, 6, scope=oracle.SCOPE_START)
    assert rec.shape[0] == pc["18,6"] and st[1] == 0
    rec, st = oracle.gen_constellations(18, 6, scope=oracle.SCOPE_CALL)
    assert rec.shape[0] == known["preset_record_counts_18py_memo"]["18,6"]
    assert st[2] == pc["18,6"] - rec.shape[0]           # every lost record is a cross-start memo hit
    rec, _ = oracle.gen_constellations(21, 6, scope=oracle.SCOPE_NONE)
    assert rec.shape[0] == pc["21,6"]
    rec, _ = oracle.gen_constellations(18, 7, scope=oracle.SCOPE_NONE)
    assert rec.shape[0] == pc["18,7"]


def test_18py_memo_total_at_preset6(oracle, known):
    """The whole-call memo of 18Py reproduces the reference log's wrong total."""
    rec, _ = oracle.gen_constellations(18, 6, scope=oracle.SCOPE_CALL)
    w, _, _ = oracle.count_records(18, rec)
    assert int(w.sum()) == known["preset_totals_18py_memo"]["18,6"]


@pytest.mark.skipif(not __import__("oracle.oracle", fromlist=["x"]).ref_available(),
                    reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("N", range(6, 15))
def test_reference_c_solver_preset4(oracle, N):
    """10Bit_CUDA/004GCC_constellation.c, compiled from its source: same records
    and weighted counts per record for even N; odd N differ in the Jasmin
    representative of the centre-row starts (that file has no odd-N centre
    block, 18Py:1687-1699), so only totals and record counts are compared."""
    rrec, rsol = oracle.ref004_count(N)
    rec, _ = oracle.gen_constellations(N, 4)
    w, _, _ = oracle.count_records(N, rec)
    assert rrec.shape[0] == rec.shape[0]
    assert int(rsol.sum()) == int(w.sum())
    if N % 2 == 0:
        a = sorted(zip(map(tuple, rrec.tolist()), rsol.tolist()))
        b = sorted(zip(map(tuple, rec.tolist()), w.tolist()))
        assert a == b


def test_pins_detect_planted_mistakes(oracle, p5_full, digests):
    """A wrong weight or a dropped lookahead term changes the pinned digests."""
    N = 12
    rec = p5_full[f"rec{N}"]
    w, raw, _ = oracle.count_records(N, rec)
    wrong_w = raw * np.where(w == raw * 2, 4, w // np.maximum(raw, 1)).astype(np.uint64)
    assert solutions_sha256(rec, wrong_w) != digests[str(N)]["solutions_sha256"]
    dropped = w.copy()
    dropped[np.argmax(w)] -= 1
    assert int(dropped.sum()) != digests[str(N)]["solutions_sum"]
