"""GPU parity: libnq (through its C ABI) against the oracle and the
reference's fixtures.  Bit-exact everywhere: weighted counts, node counts,
totals and unique counts."""
import numpy as np
import pytest

import nqueens
from nqueens import constellations as C
from tests.helpers import solutions_sha256

pytestmark = pytest.mark.gpu


def sched(split=3, block=128, lds=1, graph=0, chunk=64):
    return nqueens.Sched(split_depth=split, block=block, lds_stack=lds, use_graph=graph, chunk=chunk,
                         waves_per_eu=0)


def test_device_info(ctx):
    info = ctx.device_info()
    print(info, flush=True)
    assert info["wavefront"] == 64 and info["compute_units"] > 0


@pytest.mark.parametrize("N", range(6, 18))
def test_plain_preset4_per_record(ctx, oracle, N):
    rec = C.gen_constellations(N, 4)
    w, _, nodes = oracle.count_records(N, rec)
    ctx.load(N, rec)
    total, _ = ctx.run_plain(count_nodes=True)
    gw, gn = ctx.results(weighted=True, nodes=True)
    assert np.array_equal(gw, w)
    assert np.array_equal(gn, nodes)
    assert total == int(w.sum())


@pytest.mark.parametrize("N", range(6, 20))
def test_plain_preset5_fixture_digest(ctx, digests, N):
    rec = C.gen_constellations(N, 5)
    ctx.load(N, rec)
    total, _ = ctx.run_plain()
    gw, _ = ctx.results()
    d = digests[str(N)]
    assert total == d["solutions_sum"]
    assert solutions_sha256(rec, gw) == d["solutions_sha256"]


@pytest.mark.parametrize("split", [0, 1, 2, 3, 5])
@pytest.mark.parametrize("block,lds", [(64, 1), (128, 1), (256, 1), (64, 0), (256, 0)])
def test_batched_per_record(ctx, oracle, split, block, lds):
    N = 13
    rec = C.gen_constellations(N, 5)
    w, _, nodes = oracle.count_records(N, rec)
    ctx.load(N, rec)
    total, t = ctx.run_batched(sched(split, block, lds), count_nodes=True)
    gw, gn = ctx.results(weighted=True, nodes=True)
    assert np.array_equal(gw, w)
    assert np.array_equal(gn, nodes)
    assert total == int(w.sum())
    assert t.items >= 1


@pytest.mark.parametrize("N,preset", [(14, 4), (15, 6), (16, 5), (17, 4)])
def test_batched_other_presets(ctx, oracle, N, preset):
    rec = C.gen_constellations(N, preset)
    w, _, nodes = oracle.count_records(N, rec)
    ctx.load(N, rec)
    total, _ = ctx.run_batched(sched(3, 128, 1), count_nodes=True)
    gw, gn = ctx.results(weighted=True, nodes=True)
    assert np.array_equal(gw, w) and np.array_equal(gn, nodes)


def test_batched_order_is_schedule_only(ctx, oracle):
    N = 14
    rec = C.gen_constellations(N, 5)
    w, _, _ = oracle.count_records(N, rec)
    perm = np.random.default_rng(7).permutation(rec.shape[0]).astype(np.uint32)
    ctx.load(N, rec, order=perm)
    total, _ = ctx.run_batched(sched(2, 64, 1))
    gw, _ = ctx.results()
    assert np.array_equal(gw, w)


def test_batched_graph_replays(ctx, oracle):
    """The captured sequence replayed several times, with eager runs between."""
    N = 14
    rec = C.gen_constellations(N, 5)
    w, _, nodes = oracle.count_records(N, rec)
    ctx.load(N, rec)
    for rep in range(4):
        total, _ = ctx.run_batched(sched(3, 128, 1, graph=1), count_nodes=True)
        gw, gn = ctx.results(weighted=True, nodes=True)
        assert np.array_equal(gw, w) and np.array_equal(gn, nodes), rep
        t2, _ = ctx.run_plain()
        assert t2 == int(w.sum())


@pytest.mark.parametrize("N", [17, 18, 19])
def test_batched_fixture_digest(ctx, digests, N):
    rec = C.gen_constellations(N, 5)
    ctx.load(N, rec)
    total, _ = ctx.run_batched(sched(3, 128, 1))
    gw, _ = ctx.results()
    assert total == digests[str(N)]["solutions_sum"]
    assert solutions_sha256(rec, gw) == digests[str(N)]["solutions_sha256"]


@pytest.mark.parametrize("N", range(4, 24))
def test_symmetric_counts(ctx, oracle, N):
    assert ctx.symmetric_counts(N) == oracle.symmetric_counts(N)


def test_symmetric_counts_n24(ctx, known):
    # the oracle's N=24 values (43 s on one core, recomputed by
    # tests/test_oracle.py::test_unique_table_through_burnside[24])
    f90, f180 = ctx.symmetric_counts(24)
    assert (f90, f180) == (3328, 11675080)
    total, unique = known["readme_total_unique"]["24"]
    assert total + f180 + 2 * f90 == 8 * unique


@pytest.mark.parametrize("N", range(6, 19))
@pytest.mark.parametrize("mode", ["plain", "batched"])
def test_count_api_matches_published_table(known, N, mode):
    assert list(nqueens.count(N, mode=mode)) == known["readme_total_unique"][str(N)]


def test_exec_solutions_fills_dicts(p5_full):
    rec, sol = p5_full["rec11"], p5_full["sol11"]
    cs = [{"ld": int(r[0]), "rd": int(r[1]), "col": int(r[2]), "startijkl": int(r[3]), "solutions": 0} for r in rec]
    nqueens.exec_solutions(11, cs)
    assert [c["solutions"] for c in cs] == sol.tolist()


def test_empty_set_and_errors(ctx):
    ctx.load(12, np.zeros((0, 4), np.uint32))
    assert ctx.run_plain()[0] == 0
    assert ctx.run_batched(sched())[0] == 0
    with pytest.raises(nqueens.NQError):
        ctx.load(29, C.gen_constellations(12, 5))
    # a record the decode cannot place (its row walk never reaches a base
    # frame) is refused, never counted
    bad = np.array([[0, 0, 0, (2 << 20) | C.to_ijkl(5, 3, 7, 9)]], np.uint32)
    assert nqueens.decode_host(21, bad)[0, 9] == 3
    with pytest.raises(nqueens.NQError) as e:
        ctx.load(21, bad)
    assert e.value.code == -2
    with pytest.raises(nqueens.NQError):
        ctx.load(12, C.gen_constellations(12, 5), order=np.zeros(4300, np.uint32))


@pytest.mark.parametrize("mode", ["plain", "batched"])
def test_mixed_board_sizes_in_one_set(ctx, oracle, mode):
    """Records of several N in one resident set (one persistent queue)."""
    parts, Ns, want_w, want_n = [], [], [], []
    for N, preset in [(9, 5), (12, 4), (13, 5), (15, 5), (16, 6), (11, 5)]:
        rec = C.gen_constellations(N, preset)
        w, _, nodes = oracle.count_records(N, rec)
        parts.append(rec); Ns.append(np.full(rec.shape[0], N, np.int32)); want_w.append(w); want_n.append(nodes)
    rec, Ns = np.concatenate(parts), np.concatenate(Ns)
    perm = np.random.default_rng(3).permutation(rec.shape[0]).astype(np.uint32)
    ctx.load_mixed(Ns, rec, order=perm)
    if mode == "plain":
        total, _ = ctx.run_plain(count_nodes=True)
    else:
        total, _ = ctx.run_batched(sched(3, 128, 1), count_nodes=True)
    gw, gn = ctx.results(weighted=True, nodes=True)
    assert np.array_equal(gw, np.concatenate(want_w))
    assert np.array_equal(gn, np.concatenate(want_n))
    assert total == int(np.concatenate(want_w).sum())


def test_per_record_split_depths(ctx, oracle):
    N = 14
    rec = C.gen_constellations(N, 5)
    w, _, nodes = oracle.count_records(N, rec)
    ctx.load(N, rec)
    depth = np.random.default_rng(5).integers(0, 8, rec.shape[0]).astype(np.uint8)   # some beyond T-2
    ctx.set_splits(depth)
    for graph in (0, 1, 1):
        total, t = ctx.run_batched(sched(-1, 128, 1, graph=graph), count_nodes=True)
        gw, gn = ctx.results(weighted=True, nodes=True)
        assert np.array_equal(gw, w) and np.array_equal(gn, nodes)
    with pytest.raises(nqueens.NQError):
        ctx.set_splits(np.zeros(3, np.uint8))
    ctx.load(N, rec)
    with pytest.raises(nqueens.NQError):
        ctx.run_batched(sched(-1, 128, 1))


@pytest.mark.parametrize("N,lo,hi", [(14, 1, 8), (12, 1, 8), (15, 2, 12)])
@pytest.mark.parametrize("block", [64, 256])
def test_per_record_split_depths_without_root_depth(ctx, oracle, N, lo, hi, block):
    """Every depth >= 1 and some beyond T-2 (those records are walked from
    their roots): the LDS stacks must be sized for the root walks, not for
    the smallest requested depth."""
    rec = C.gen_constellations(N, 5)
    w, _, nodes = oracle.count_records(N, rec)
    depth = np.random.default_rng(N).integers(lo, hi, rec.shape[0]).astype(np.uint8)
    assert depth.min() >= 1 and (depth > N - 1 - 5 - 2).any()
    ctx.load(N, rec)
    ctx.set_splits(depth)
    for graph in (0, 1):
        total, _ = ctx.run_batched(sched(-1, block, 1, graph=graph), count_nodes=True)
        gw, gn = ctx.results(weighted=True, nodes=True)
        assert np.array_equal(gw, w) and np.array_equal(gn, nodes)
        assert total == int(w.sum())


@pytest.fixture(scope="module")
def oracle_samples():
    from tests.conftest import GOLDEN
    out = {}
    for name in ("oracle_sample_n20_n21.npz", "oracle_sample_n22_n23.npz", "oracle_sample_n24.npz"):
        with np.load(GOLDEN / name) as z:
            out.update({k: z[k] for k in z.files})
    return out


# The largest boards: every 400th record of N=20/21 (preset 5), every
# 2,000th of N=22 (preset 5), every 20,000th of N=23 (preset 6) and every
# 600,000th of N=24 (preset 7), with
# cached oracle counts.  Long ld histories fill all 32 bits of the
# lookahead's occupancy word here, which N <= 19 rarely does; N=22/23 use
# terminal depth 16 and 20/21-bit candidate sets.  The plain kernel (one lane
# per record) is left out at N=22/23, where one record is up to 4e8 nodes.
@pytest.mark.parametrize("N,mode", [(20, "plain"), (20, "batched"), (21, "plain"), (21, "batched"),
                                    (22, "batched"), (23, "batched"), (24, "batched")])
def test_largest_n_oracle_sample(ctx, oracle_samples, N, mode):
    rec = oracle_samples[f"rec{N}"]
    ctx.load(N, rec)
    if mode == "plain":
        total, _ = ctx.run_plain(count_nodes=True)
    else:
        total, _ = ctx.run_batched(sched(3, 256, 1), count_nodes=True)
    gw, gn = ctx.results(weighted=True, nodes=True)
    assert np.array_equal(gw, oracle_samples[f"w{N}"])
    assert np.array_equal(gn, oracle_samples[f"nodes{N}"])
    assert total == int(oracle_samples[f"w{N}"].sum())


@pytest.mark.parametrize("N", [20, 21, 22])
def test_largest_n_full_count(ctx, known, N):
    total, unique = nqueens.count(N, ctx=ctx)
    assert [total, unique] == list(known["readme_total_unique"][str(N)])


def test_count_in_chunks(ctx, known):
    """Interleaved slices counted one device run each add up to the whole."""
    seen = []
    total, unique = nqueens.count(16, ctx=ctx, chunks=3, progress=lambda i, k, t: seen.append((i, k, t)))
    assert [total, unique] == known["readme_total_unique"]["16"]
    assert [i for i, _, _ in seen] == [0, 1, 2] and sum(t for _, _, t in seen) == total
    # a count split over separate calls (chunk ranges) adds up to the same total,
    # also with the deeper presets expanded on the device
    from nqueens.solver import count_chunks
    parts = count_chunks(16, 5, 0, 2, ctx=ctx) + count_chunks(16, 5, 2, 5, ctx=ctx)
    assert sum(parts) == total
    assert sum(count_chunks(16, 4, preset=7, ctx=ctx)) == total


@pytest.mark.parametrize("N,p,levels", [(N, 5, 1) for N in (8, 12, 16, 19, 21, 23)] + [(N, 4, 1) for N in (6, 7, 14, 19, 22)]
                         + [(13, 5, 2), (16, 5, 3)])
def test_device_preset_expansion(ctx, N, p, levels):
    """nq_expand_records gives exactly the host generator's records of the
    deeper preset (as multisets; the device keeps record order, columns
    ascending)."""
    got = ctx.expand_records(N, C.gen_constellations(N, p), levels)
    want = C.gen_constellations(N, p + levels)
    assert np.array_equal(C.records_sorted(got), C.records_sorted(want))


def test_device_preset_expansion_reference_counts(ctx):
    # record counts of the reference's logs: N=18 presets 6 and 7, N=21 preset 6
    rec5 = C.gen_constellations(18, 5)
    assert ctx.expand_records(18, rec5, 1).shape[0] == 571705
    assert ctx.expand_records(18, rec5, 2).shape[0] == 3653940
    assert ctx.expand_records(21, C.gen_constellations(21, 5), 1).shape[0] == 2025282


@pytest.mark.parametrize("N,levels", [(12, 1), (15, 2), (17, 1)])
def test_load_expanded_counts(ctx, oracle, known, N, levels):
    """Records expanded and decoded on the device count the same as the
    host-generated records of that preset, record by record."""
    rec = C.gen_constellations(N, 5 + levels)
    w, _, _ = oracle.count_records(N, rec)
    m = ctx.load_expanded(N, C.gen_constellations(N, 5), levels)
    assert m == rec.shape[0]
    total, _ = ctx.run_batched(sched(2, 128, 1))
    assert total == known["total_a000170"][str(N)] == int(w.sum())
    gw, _ = ctx.results()
    got = ctx.expand_records(N, C.gen_constellations(N, 5), levels)
    key = lambda r, v: sorted(zip(map(tuple, r.tolist()), v.tolist()))
    assert key(got, gw) == key(rec, w)


def test_expansion_errors(ctx):
    bad = np.array([[0, 0, 0, (3 << 20) | C.to_ijkl(5, 3, 3, 9)]], np.uint32)   # start row is row k
    with pytest.raises(nqueens.NQError):
        ctx.expand_records(12, bad, 1)
    with pytest.raises(nqueens.NQError):
        ctx.expand_records(12, C.gen_constellations(12, 5), 0)


def test_resumed_count_from_progress_file(tmp_path, known, capsys):
    """A count interrupted after two of four chunks resumes from its progress
    log and ends with the published total and unique count."""
    from nqueens.__main__ import main
    from nqueens.progress import ChunkLog
    log = tmp_path / "n14.tsv"
    assert main(["14", "--chunks", "4", "--chunk-range", "0", "2", "--progress-file", str(log)]) == 0
    assert ChunkLog(log, 14, 4, 5).missing() == [2, 3]
    assert main(["14", "--chunks", "4", "--progress-file", str(log)]) == 0
    out = capsys.readouterr().out
    total, unique = known["readme_total_unique"]["14"]
    assert f"{total:18d}{unique:17d}    ok" in out
    assert ChunkLog(log, 14, 4, 5).total() == total


@pytest.mark.parametrize("N", [13, 16])
def test_device_cost_order(ctx, known, N):
    """The device ordering (features, MLP forward, stable radix sort) is the
    host's argsort of the same predictions, and leaves the count unchanged."""
    from nqueens import costmodel
    params = costmodel.default_model().params
    ctx.load(N, C.gen_constellations(N, 5))
    pred = ctx.mlp_forward(params, ctx.features())
    want = np.argsort(-pred, kind="stable").astype(np.uint32)
    got = ctx.order_by_cost(params, want=True)
    assert np.array_equal(got, want)
    total, _ = ctx.run_batched(sched(3, 128, 1, graph=1))
    assert total == known["total_a000170"][str(N)]


def test_trained_model_gains_over_fifo(ctx, oracle):
    """What the trained network (the shipped model, trained on N=16..19) gains
    over FIFO: N=17's constellations as whole work items on 16,384 lanes,
    list-scheduled with their measured node counts as costs.  FIFO ends 68 %
    above the work bound, heaviest-first by the true counts 14 %."""
    from nqueens import costmodel
    rec = C.gen_constellations(17, 5)
    _, _, nodes = oracle.count_records(17, rec)
    cost = nodes.astype(np.float64)
    ctx.load(17, rec)
    order = ctx.order_by_cost(costmodel.default_model().params, want=True)
    lanes = 16384
    lower = cost.sum() / lanes
    fifo = costmodel.simulated_makespan(cost, np.arange(cost.size), lanes) / lower
    model = costmodel.simulated_makespan(cost, order, lanes) / lower
    best = costmodel.simulated_makespan(cost, np.argsort(-cost, kind="stable"), lanes) / lower
    print(f"makespan / work bound: FIFO {fifo:.3f}, model order {model:.3f}, true order {best:.3f}")
    assert best <= model < fifo
    assert model < 1.25 and fifo > 1.6


def test_autotune_preset_sweep(ctx, known, tmp_path, monkeypatch):
    """The preset sweep counts N=13 at presets 4..7 (host generation up to 5,
    device expansion beyond), checks every total, and stores the fastest
    whole count; the records of each preset are the generator's."""
    from nqueens import autotune
    tp, pp = tmp_path / "tuned.json", tmp_path / "tuned_presets.json"
    monkeypatch.setattr(autotune, "TUNED", {})
    monkeypatch.setattr(autotune, "TUNED_PRESETS", {})
    N = 13
    rows = autotune.sweep_presets(ctx, N, known["total_a000170"][str(N)], presets=(4, 5, 6, 7), splits=(2, 3),
                                  blocks=(64,), reps=1, warmup=0)
    assert all("error" not in r for r in rows), rows
    for p in (4, 5, 6, 7):
        rp = [r for r in rows if r["preset"] == p]
        assert len(rp) == 2 and all(r["records"] == C.gen_constellations(N, p).shape[0] for r in rp)
    best = autotune.record_presets(N, rows, path=pp, tuned_path=tp)
    assert best["preset"] in (4, 5, 6, 7)
    assert nqueens.solver.preset_for(N) == best["preset"]
    total, unique = nqueens.count(N, ctx=ctx)
    assert (total, unique) == tuple(known["readme_total_unique"][str(N)])


@pytest.mark.parametrize("m", [1, 1023, 1024, 1025, 2049])
def test_batched_scan_tile_boundaries(ctx, oracle, m):
    """Record counts on either side of the item-offset scan's 1024-entry
    tiles: every record's weighted count and node count, and the item count
    the run reports, match the oracle and a second split depth."""
    N = 13
    rec = C.gen_constellations(N, 5)[:m]
    w, _, nodes = oracle.count_records(N, rec)
    ctx.load(N, rec)
    for split in (2, 3):
        total, t = ctx.run_batched(sched(split, 256, 1), count_nodes=True)
        gw, gn = ctx.results(weighted=True, nodes=True)
        assert np.array_equal(gw, w) and np.array_equal(gn, nodes)
        assert total == int(w.sum()) and t.items >= 1


def test_small_boards(ctx, known, oracle):
    """N <= 5 by the plain full-board DFS kernel, as the reference's main
    counts them (_bit_total): totals against A000170 and the brute force,
    Unique against the published table; N = 6..9 cross-check the kernel
    against the constellation path."""
    A000170 = {1: 1, 2: 0, 3: 0, 4: 2, 5: 10}
    for N, want in A000170.items():
        assert ctx.small_total(N) == want, N
    for N in (4, 5):
        assert oracle.bruteforce(N)[0] == A000170[N]
    assert nqueens.count(2, ctx=ctx) == (0, 0) and nqueens.count(3, ctx=ctx) == (0, 0)
    for N in (4, 5):
        assert list(nqueens.count(N, ctx=ctx)) == known["readme_total_unique"][str(N)]
    for N in (6, 7, 8, 9):
        assert ctx.small_total(N) == nqueens.count(N, ctx=ctx)[0] == known["total_a000170"][str(N)]
    with pytest.raises(nqueens.NQError):
        ctx.small_total(13)


@pytest.fixture(scope="module")
def p5_n24():
    from tests.conftest import GOLDEN
    with np.load(GOLDEN / "p5_n24.npz") as z:
        return {k: z[k] for k in z.files}


def test_p5_records_of_n24_through_descendants(ctx, p5_n24):
    """The 18 P5 records of N=24 preset 5 start on row 2, so their terminal
    row lies 18 rows down, past the walk's 16-depth table: they are counted
    through their descendants two preset queens deeper, expanded on the
    device (no P5 among them), and each record's sum equals the oracle's P5
    branch on the record itself (weighted count and node count less the
    extractions of the two expanded rows)."""
    rec = p5_n24["rec"]
    d = nqueens.decode_host(24, rec)
    assert set(d[:, 0].tolist()) <= {8, 9, 10, 11} and (d[:, 8] == 18).all()
    with pytest.raises(nqueens.NQError):
        ctx.load(24, rec)
    kids = [ctx.expand_records(24, rec[a:a + 1], 2) for a in range(rec.shape[0])]
    assert [k.shape[0] for k in kids] == p5_n24["kids"].tolist()
    allk = np.concatenate(kids)
    assert not np.isin(nqueens.decode_host(24, allk)[:, 0], [8, 9, 10, 11]).any()
    ctx.load(24, allk)
    ctx.run_batched(sched(3, 256, 1), count_nodes=True)
    kw, kn = ctx.results(weighted=True, nodes=True)
    b = np.cumsum([0] + [k.shape[0] for k in kids])
    assert [int(kw[b[a]:b[a + 1]].sum()) for a in range(len(kids))] == p5_n24["w"].tolist()
    assert [int(kn[b[a]:b[a + 1]].sum()) for a in range(len(kids))] == p5_n24["kids_nodes"].tolist()


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_p5_records_n28_preset11(ctx, oracle, mode):
    """Real P5 records that the walk takes directly (terminal depth 16): the
    device expansion gives the oracle's records, the P5 switch row sets the j
    diagonal (parameter bit 7), and weighted and node counts equal the
    oracle's P5 branch record by record; each record also counts what its
    descendants two queens deeper (no P5 among them) count on the device."""
    from tests.helpers import p5_records_n28
    N = 28
    par, kids, rec = p5_records_n28(oracle)
    dev = ctx.expand_records(N, par, 6)
    assert np.array_equal(dev, kids)
    d = nqueens.decode_host(N, rec)
    assert np.isin(d[:, 0], [8, 9, 10, 11]).all() and (d[:, 8] == 16).all() and (d[:, 9] == 0).all()
    w, _, nodes = oracle.count_records(N, rec)
    ctx.load(N, rec)
    if mode == "plain":
        total, _ = ctx.run_plain(count_nodes=True)
    else:
        total, _ = ctx.run_batched(sched(3, 256, 1), count_nodes=True)
    gw, gn = ctx.results(weighted=True, nodes=True)
    assert np.array_equal(gw, w) and np.array_equal(gn, nodes)
    sub = [ctx.expand_records(N, rec[a:a + 1], 2) for a in range(0, rec.shape[0], 8)]
    allk = np.concatenate(sub)
    assert not np.isin(nqueens.decode_host(N, allk)[:, 0], [8, 9, 10, 11]).any()
    ctx.load(N, allk)
    ctx.run_batched(sched(3, 256, 1))
    kw, _ = ctx.results(weighted=True)
    b = np.cumsum([0] + [x.shape[0] for x in sub])
    assert [int(kw[b[a]:b[a + 1]].sum()) for a in range(len(sub))] == gw[::8].tolist()


def test_n25_preset8_oracle_sample(ctx):
    """N=25 at preset 8, the reference's preset there (terminal depth 16, a
    25-bit board): the device expansion of every 20,000th preset-5 record by
    three preset queens gives the oracle's records (sampled every 448th), and
    their weighted and node counts equal the oracle's, batched and plain."""
    from tests.conftest import GOLDEN
    with np.load(GOLDEN / "oracle_sample_n25.npz") as z:
        f = {k: z[k] for k in z.files}
    parents = C.gen_constellations(25, 5)[::20000]
    assert np.array_equal(parents, f["parents"])
    kids = ctx.expand_records(25, parents, 3)
    assert np.array_equal(kids[::448], f["rec25"])
    ctx.load(25, f["rec25"])
    for mode in ("batched", "plain"):
        if mode == "plain":
            total, _ = ctx.run_plain(count_nodes=True)
        else:
            total, _ = ctx.run_batched(sched(3, 256, 1), count_nodes=True)
        gw, gn = ctx.results(weighted=True, nodes=True)
        assert np.array_equal(gw, f["w25"]) and np.array_equal(gn, f["nodes25"]), mode
        assert total == int(f["w25"].sum())


def test_symmetric_counts_n25_n27(ctx, known):
    """F90/F180 past N=24: at N=25 they give the published Unique through
    Burnside; at N=26 and 27 (no published Unique in the reference) the
    orbit count is integral (seconds each)."""
    from nqueens.expected import TOTAL, UNIQUE
    f90, f180 = ctx.symmetric_counts(25)
    assert TOTAL[25] + f180 + 2 * f90 == 8 * UNIQUE[25]
    for N in (26, 27):
        f90, f180 = ctx.symmetric_counts(N)
        assert (TOTAL[N] + f180 + 2 * f90) % 8 == 0 and f180 >= f90 >= 0
        print(f"N={N}: F90={f90} F180={f180} unique={(TOTAL[N] + f180 + 2 * f90) // 8}", flush=True)
