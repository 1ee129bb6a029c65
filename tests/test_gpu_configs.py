"""The benchmark configurations at their stated workloads, on the GPU:

  N=19, preset 5, persistent batched kernel replayed as a hipGraph (the
      bench's schedule) and at a capped occupancy;
  N=20, preset 5, learned cost model: the device cost order and the
      model-binned per-record split depths, every count exact, the
      prediction error against the measured node counts asserted;
  request loop: seeded Poisson arrivals of N drawn from 14..21
      (seed 20261203), every answer exact.
"""
import math

import numpy as np
import pytest

import nqueens
from nqueens import autotune, costmodel, serve
from nqueens import constellations as C
from nqueens.expected import TOTAL, UNIQUE
from tests.helpers import solutions_sha256

pytestmark = pytest.mark.gpu


def sched(split=3, block=256, lds=1, graph=0, chunk=64, wpe=0):
    return nqueens.Sched(split_depth=split, block=block, lds_stack=lds, use_graph=graph, chunk=chunk,
                         waves_per_eu=wpe)


@pytest.fixture(scope="module")
def sample20():
    from tests.conftest import GOLDEN
    with np.load(GOLDEN / "oracle_sample_n20_n21.npz") as z:
        return z["rec20"], z["w20"], z["nodes20"]


def test_n19_graph_replays_and_occupancy(ctx, digests):
    """The bench's N=19 schedule: graph replays (with eager runs and a
    node-counting run in between) and occupancy caps all give the fixture's
    per-record counts."""
    N = 19
    rec = C.gen_constellations(N, 5)
    d = digests[str(N)]
    ctx.load(N, rec)
    ctx.order_by_cost(costmodel.default_model().params)
    s = autotune.best_sched(N, 5)
    runs = [(1, 0, False), (1, 0, False), (0, 0, True), (1, 0, False), (0, 2, False), (0, 1, False), (1, 0, False)]
    for graph, wpe, nodes in runs:
        s.use_graph, s.waves_per_eu = graph, wpe
        total, t = ctx.run_batched(s, count_nodes=nodes)
        w, n = ctx.results(weighted=True, nodes=nodes)
        print(f"graph={graph} waves_per_eu={wpe} nodes={nodes}: {t.total_ms:.2f} ms", flush=True)
        assert total == d["solutions_sum"] == TOTAL[N]
        assert solutions_sha256(rec, w) == d["solutions_sha256"], (graph, wpe, nodes)
        if nodes:
            assert int(n.sum()) == 69277295929


def test_config4_cost_model_n20(ctx, sample20):
    """N=20, preset 5 (139,232 constellations) under the shipped model's device
    order, then with model-binned per-record split depths: the total, and the
    weighted and node counts of every oracle-sampled record, are unchanged;
    the model's prediction error against the measured node counts."""
    N, preset = 20, 5
    rec = C.gen_constellations(N, preset)
    assert rec.shape[0] == 139232
    srec, sw, sn = sample20
    idx = np.arange(0, rec.shape[0], 400)
    assert np.array_equal(rec[idx], srec)
    model = costmodel.default_model()
    ctx.load(N, rec)
    order = ctx.order_by_cost(model.params, want=True)
    assert np.array_equal(np.sort(order), np.arange(rec.shape[0]))
    total, t1 = ctx.run_batched(sched(3, 256, 1, graph=1), count_nodes=True)
    w, nodes = ctx.results(weighted=True, nodes=True)
    assert total == TOTAL[N] == 39029188884
    assert np.array_equal(w[idx], sw) and np.array_equal(nodes[idx], sn)
    # prediction error of log(1 + nodes), as the model is trained on
    pred = model.predict_features(ctx, ctx.features()).astype(np.float64)
    truth = np.log1p(nodes.astype(np.float64))
    rho = costmodel.spearman(pred, truth)
    mae = float(np.mean(np.abs(pred - truth)))
    print(f"N=20 prediction: Spearman {rho:.4f}, MAE {mae:.4f} (log nodes), {int(nodes.sum())} nodes", flush=True)
    assert rho > 0.85 and mae < 0.6
    # model-binned per-record split depths, model order
    T = N - 1 - preset
    depths = costmodel.binned_splits(pred, T, quantiles=(0.5, 0.9), base=2)
    assert len(set(depths.tolist())) == 3
    ctx.load(N, rec, order)
    ctx.set_splits(depths)
    for graph, nodes_on in ((0, True), (1, False), (1, False)):
        total2, t2 = ctx.run_batched(sched(-1, 256, 1, graph=graph), count_nodes=nodes_on)
        w2, n2 = ctx.results(weighted=True, nodes=nodes_on)
        assert total2 == TOTAL[N]
        assert np.array_equal(w2, w)
        if nodes_on:
            assert np.array_equal(n2, nodes)
    print(f"N=20 model order split 3: {t1.total_ms:.1f} ms; binned splits: {t2.total_ms:.1f} ms", flush=True)


def test_config5_request_loop(ctx):
    """Seeded Poisson arrivals (seed 20261203) of N uniform in 14..21 on the
    shared persistent queue: every total and unique exact, finite latency
    statistics."""
    loop = serve.RequestLoop(ctx=ctx, log=lambda m: print(m, flush=True))
    reqs = serve.poisson_requests(20, 1.0, 14, 21, seed=20261203)
    assert {r.N for r in reqs} <= set(range(14, 22))
    res = loop.run(reqs)
    assert len(res) == 20 and sorted(r.rid for r in res) == list(range(20))
    assert all(r.ok for r in res)
    for r in res:
        assert (r.total, r.unique) == (TOTAL[r.N], UNIQUE[r.N])
        assert r.done >= r.start >= r.arrival - 1e-9
    st = serve.latency_stats(res)
    print(f"request loop: {st}, {len({r.batch for r in res})} batches", flush=True)
    assert all(math.isfinite(v) and v >= 0 for v in st.values())
    assert st["p50_s"] <= st["p99_s"] <= st["max_s"]
