"""Request loop: arrival process and admission logic (CPU, with a stand-in
counter), and the real loop on the GPU."""
import numpy as np
import pytest

from nqueens import serve
from nqueens.expected import TOTAL, UNIQUE


def test_poisson_requests_seeded():
    a = serve.poisson_requests(50, 2.0, 14, 21, seed=7)
    b = serve.poisson_requests(50, 2.0, 14, 21, seed=7)
    assert [(r.N, r.arrival) for r in a] == [(r.N, r.arrival) for r in b]
    assert all(14 <= r.N <= 21 for r in a)
    assert all(x.arrival < y.arrival for x, y in zip(a, a[1:]))


class _FakeLoop(serve.RequestLoop):
    """Counts by table lookup: exercises batching, admission and the clock."""

    def __init__(self, sizes, max_records):
        self.sizes = sizes
        self.max_records = max_records
        self.log = lambda *_: None
        self.batches = []

    def records(self, N):
        return np.zeros((self.sizes[N], 4), np.uint32)

    def run_batch(self, batch, count_nodes=False):
        self.batches.append([r.N for r in batch])
        return [(TOTAL[r.N], UNIQUE[r.N]) for r in batch], None


def test_admission_respects_budget_and_order():
    loop = _FakeLoop({10: 100, 11: 300, 12: 700}, max_records=800)
    reqs = [serve.Request(i, N, 0.0) for i, N in enumerate([10, 11, 12, 10, 12])]
    res = loop.run(reqs)
    assert loop.batches == [[10, 11], [12, 10], [12]]
    assert all(r.ok for r in res) and sorted(r.rid for r in res) == list(range(5))
    assert all(r.done >= r.arrival for r in res)
    st = serve.latency_stats(res)
    assert st["p50_s"] <= st["p99_s"] <= st["max_s"]


def test_idle_gaps_advance_the_clock():
    loop = _FakeLoop({10: 1}, max_records=10)
    res = loop.run([serve.Request(0, 10, 5.0), serve.Request(1, 10, 9.0)])
    assert [r.start for r in res] == [5.0, 9.0]


@pytest.mark.gpu
def test_request_loop_on_gpu():
    loop = serve.RequestLoop(max_records=50_000)
    reqs = serve.poisson_requests(12, 5000.0, 10, 15, seed=3)   # arrivals far faster than service
    res = loop.run(reqs)
    assert len(res) == 12 and all(r.ok for r in res)
    assert len({r.batch for r in res}) < 12      # requests shared launches
