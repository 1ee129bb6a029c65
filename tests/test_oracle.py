"""The CPU oracle against everything the reference fixes (CPU only).

Pins (see oracle/nq_oracle.c):
  * constellation records: constellations_N{6..21}_5.bin multisets
  * per-record weighted counts: solutions_N{6..19}_5_v2.bin
  * totals: A000170 (18Py:2089); unique: README.en:12-33 through Burnside
  * F90/F180 and unique: brute-force enumeration for N <= 12
  * preset-6/7 record counts from the reference's logs
  * node counts: iterative and recursive restatements agree per record;
    under the pruning of the reference's recursive SQ* functions the oracle
    reproduces that solver's own traced node counts record by record
    (nodes_13py.npz, N=6..14), and planted accounting mistakes break that
  * P5 functionids: with build_soa's j-diagonal gate narrowed to W >= N
    bits, every record past it takes the P5 branch and the per-record counts
    are unchanged; the 18 P5 records of N=24 preset 5 count what their
    non-P5 descendants two queens deeper count (p5_n24.npz)
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


@pytest.fixture(scope="module")
def census():
    from tests.conftest import GOLDEN
    with np.load(GOLDEN / "nodes_13py.npz") as z:
        return {k: z[k] for k in z.files}


def _census_cases(census):
    return sorted(k[3:] for k in census if k.startswith("rec"))


def test_node_counts_match_reference_census(oracle, census):
    """The reference's recursive solver (13Py, traced line by line) against
    the oracle's node accounting with the same pruning: equal node counts for
    every record of every case, and equal weighted counts except where 13Py
    counts a root on the terminal row before testing its free set (N=6,
    preset 5: the 18Py fixture has those as 0)."""
    assert len(_census_cases(census)) == 17
    for tag in _census_cases(census):
        N = int(tag.split("_")[0])
        rec = census["rec" + tag]
        w, nodes = oracle.count_records_13py(N, rec)
        assert np.array_equal(nodes, census["nodes" + tag]), tag
        bad = np.nonzero(w != census["sol" + tag])[0]
        if tag == "6_5":
            soa = oracle.build_soa(N, rec[bad])
            assert len(bad) and (soa[:, 7] == 0).all() and (soa[:, 1] == soa[:, 3]).all()
        else:
            assert len(bad) == 0, tag
            w18, _, _ = oracle.count_records(N, rec)
            assert np.array_equal(w18, w), tag


@pytest.mark.parametrize("plant", ["base", "jmark"])
def test_planted_node_mistakes_fail_the_census(oracle, census, plant):
    """Counting a base frame as a node, or column 0 before the jmark mask,
    changes node counts the reference's census fixes."""
    flag = {"base": oracle.PLANT_BASE, "jmark": oracle.PLANT_JMARK}[plant]
    wrong = 0
    for tag in _census_cases(census):
        N = int(tag.split("_")[0])
        rec = census["rec" + tag]
        w, nodes = oracle.count_records_13py(N, rec, plant=flag)
        wrong += int((nodes != census["nodes" + tag]).sum())
        if tag != "6_5":
            assert np.array_equal(w, census["sol" + tag])   # the count is untouched
    assert wrong > 100


def test_node_definitions_differ_only_on_pruned_functionids(oracle, census):
    """The product's node definition (18Py's availptn table) and 13Py's
    differ only through SQBjrB (fid 2, reached from targets 0..7) and the
    SQd1BlB mark step (fid 20, from targets 19 and 20): records of the other
    targets have equal node counts under both."""
    for tag in _census_cases(census):
        N = int(tag.split("_")[0])
        rec = census["rec" + tag]
        fid = oracle.build_soa(N, rec)[:, 0]
        same = ~np.isin(fid, list(range(0, 8)) + [19, 20])
        _, _, nodes = oracle.count_records(N, rec)
        assert np.array_equal(nodes[same], census["nodes" + tag][same]), tag


@pytest.mark.parametrize("N", range(9, 17))
def test_p5_branch_under_narrow_gate(oracle, digests, N):
    """build_soa's gate sets the j queen's rd bit at the start row only when
    it fits the rd word (j >= 2N-33-start at 32 bits); other records take the
    P5 functions, which set it on the row it enters the board.  Narrowing the
    word to W bits (W >= N) moves records onto the P5 branch without changing
    any count: the per-record weighted counts keep the reference fixture's
    digest, the iterative and recursive restatements agree on counts and
    nodes, and all four P5 functionids occur.  At W = N-1 the gate also cuts
    the j >= N-3 families, which have no P5 branch, and the total is wrong."""
    from nqueens import constellations as C
    from tests.helpers import solutions_sha256
    rec = C.gen_constellations(N, 5)
    d = digests[str(N)]
    fids = set()
    for W in (N, N + 1, N + 3):
        with oracle.gate_bits(W):
            fid = oracle.build_soa(N, rec)[:, 0]
            w, _, nodes = oracle.count_records(N, rec)
            w2, nodes2 = oracle.count_records_recursive(N, rec)
        fids |= set(fid[(fid >= 8) & (fid <= 11)].tolist())
        assert int(w.sum()) == d["solutions_sum"], W
        assert solutions_sha256(rec, w) == d["solutions_sha256"], W
        assert np.array_equal(w, w2) and np.array_equal(nodes, nodes2), W
    assert fids == {8, 9, 10, 11} or N < 12
    with oracle.gate_bits(N - 1):
        assert int(oracle.count_records(N, rec)[0].sum()) != d["solutions_sum"]


def test_p5_records_of_n24_partition(oracle):
    """The 18 P5 records of N=24 preset 5 (cached oracle counts): each counts
    what its descendants two preset queens deeper (no P5 among them) count,
    and its node count is theirs plus one per extraction on the two expanded
    rows (placed without the lookahead); the records, their P5 targets and
    their expansions are recomputed here."""
    from nqueens import constellations as C
    from tests.conftest import GOLDEN
    with np.load(GOLDEN / "p5_n24.npz") as z:
        f = {k: z[k] for k in z.files}
    rec = C.gen_constellations(24, 5)
    fid = oracle.build_soa(24, rec)[:, 0]
    idx = np.nonzero((fid >= 8) & (fid <= 11))[0]
    assert np.array_equal(idx, f["index"]) and np.array_equal(rec[idx], f["rec"]) and len(idx) == 18
    assert np.array_equal(f["w"], f["kids_w"])
    assert np.array_equal(f["nodes"], f["kids_nodes"] + f["kids6"] + f["kids"])
    for a in range(len(idx)):
        assert oracle.expand_records(24, f["rec"][a:a + 1], 1).shape[0] == f["kids6"][a]
        k7 = oracle.expand_records(24, f["rec"][a:a + 1], 2)
        assert k7.shape[0] == f["kids"][a]
        kf = oracle.build_soa(24, k7)[:, 0]
        assert not ((kf >= 8) & (kf <= 11)).any()


def test_p5_records_of_n28_partition(oracle):
    """P5 records at a depth the u32 walk holds exist first at N=28 (preset
    11, j = 14, start row 8; for N <= 27 every P5 record ends 17 or more rows
    below its start): each counts, by the oracle's P5 branch, what its
    descendants two preset queens deeper (none of them P5) count."""
    from tests.helpers import p5_records_n28
    N = 28
    _, _, rec = p5_records_n28(oracle, parents=1, stride=80000)
    fid = oracle.build_soa(N, rec)[:, 0]
    assert np.isin(fid, [8, 9, 10, 11]).all() and len(rec) >= 40
    w, _, nodes = oracle.count_records(N, rec)
    for a in range(rec.shape[0]):
        kids = oracle.expand_records(N, rec[a:a + 1], 2)
        assert not np.isin(oracle.build_soa(N, kids)[:, 0], [8, 9, 10, 11]).any()
        kw, _, kn = oracle.count_records(N, kids)
        n6 = oracle.expand_records(N, rec[a:a + 1], 1).shape[0]
        assert int(kw.sum()) == int(w[a])
        assert int(kn.sum()) + n6 + kids.shape[0] == int(nodes[a])
