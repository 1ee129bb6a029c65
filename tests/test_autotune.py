"""Autotuner tables: the preset depth chosen per N and the schedule per
(N, preset), written and read back without a GPU."""
import json

import pytest

from nqueens import autotune, solver


def _rows():
    # two presets of N=20 as sweep_presets reports them (times in ms)
    return [
        {"preset": 4, "error": "nq_ctx_load: P5 functionid"},
        {"preset": 5, "records": 139232, "load_ms": 40.0, "split_depth": 3, "block": 128, "lds_stack": 1,
         "chunk": 64, "items": 1, "kernel_ms_median": 500.0, "kernel_ms_best": 499.0,
         "run_ms_median": 520.0, "run_ms_best": 519.0},
        {"preset": 5, "records": 139232, "load_ms": 40.0, "split_depth": 4, "block": 256, "lds_stack": 1,
         "chunk": 64, "items": 1, "kernel_ms_median": 530.0, "kernel_ms_best": 529.0,
         "run_ms_median": 550.0, "run_ms_best": 549.0},
        {"preset": 6, "records": 1500000, "load_ms": 60.0, "split_depth": 2, "block": 256, "lds_stack": 1,
         "chunk": 64, "items": 1, "kernel_ms_median": 480.0, "kernel_ms_best": 478.0,
         "run_ms_median": 490.0, "run_ms_best": 488.0},
    ]


@pytest.fixture
def tables(tmp_path, monkeypatch):
    """The autotuner's tables redirected to empty files under tmp_path."""
    tp, pp = tmp_path / "tuned.json", tmp_path / "tuned_presets.json"
    monkeypatch.setattr(autotune, "TUNED_PATH", tp)
    monkeypatch.setattr(autotune, "PRESETS_PATH", pp)
    monkeypatch.setattr(autotune, "TUNED", {})
    monkeypatch.setattr(autotune, "TUNED_PRESETS", {})
    return tp, pp


def test_record_presets_picks_fastest_whole_count(tables):
    tp, pp = tables
    best = autotune.record_presets(20, _rows(), path=pp, tuned_path=tp)
    # preset 6: 60 + 490 < preset 5: 40 + 520
    assert best["preset"] == 6
    assert best["per_preset"]["5"]["split_depth"] == 3 and best["per_preset"]["5"]["block"] == 128
    assert json.loads(pp.read_text())["20"]["preset"] == 6
    t = json.loads(tp.read_text())
    assert t["20,5"]["split_depth"] == 3 and t["20,6"]["split_depth"] == 2
    assert autotune.tuned_preset(20) == 6 and autotune.tuned_preset(19) is None
    assert solver.preset_for(20) == 6
    s = autotune.best_sched(20, 6)
    assert (s.split_depth, s.block, s.lds_stack) == (2, 256, 1)


def test_record_presets_load_time_counts(tables):
    tp, pp = tables
    rows = _rows()
    rows[3]["load_ms"] = 100.0      # preset 6 now loses: 100 + 490 > 40 + 520
    assert autotune.record_presets(20, rows, path=pp, tuned_path=tp)["preset"] == 5


def test_record_presets_all_failed(tables):
    tp, pp = tables
    with pytest.raises(ValueError):
        autotune.record_presets(20, [_rows()[0]], path=pp, tuned_path=tp)


def test_preset_for_ignores_unusable_measurements(monkeypatch):
    # a table entry the walk cannot take (terminal row deeper than 16, or
    # outside 4..10) falls back to the rule
    monkeypatch.setattr(autotune, "TUNED_PRESETS", {"23": {"preset": 5}, "12": {"preset": 3},
                                                    "13": {"preset": 11}, "14": {"preset": 4}})
    assert solver.preset_for(23) == 6
    assert solver.preset_for(12) == 5
    assert solver.preset_for(13) == 5
    assert solver.preset_for(14) == 4
    assert solver.preset_for(24) == 7


def test_shipped_preset_table_is_usable():
    """Every preset the shipped table holds keeps the terminal row within the
    16 parameter bytes, and its schedule is in the schedule table."""
    for n, e in autotune.TUNED_PRESETS.items():
        N, p = int(n), int(e["preset"])
        assert max(4, N - 17) <= p <= 7
        assert solver.preset_for(N) == p
        assert f"{N},{p}" in autotune.TUNED


def test_sweep_presets_control_flow(monkeypatch):
    """sweep_presets with a stand-in context: a preset whose load fails is
    reported as an error row and skipped, every other preset is swept with
    the LDS stack only, checked against the expected total, ordered after
    each load, and its rows carry the preset, record count and load time."""
    from nqueens import _lib

    class FakeCtx:
        def __init__(self):
            self.m, self.loads, self.ordered, self.runs = 0, [], 0, []

        def load(self, N, rec, order=None):
            self.m = rec.shape[0]
            self.loads.append(("host", N, self.m))

        def load_expanded(self, N, rec, levels):
            if levels == 2:
                raise _lib.NQError(-2, "nq_ctx_load_expanded", "P5 target")
            self.m = rec.shape[0] * 3
            self.loads.append(("device", N, levels))
            return self.m

        def run_batched(self, s):
            self.runs.append((s.split_depth, s.block, s.lds_stack))
            return 73712, _lib.Timing(kernel_ms=1.0, total_ms=2.0, items=self.m, launches=1)

    ctx = FakeCtx()
    order = lambda c: setattr(c, "ordered", c.ordered + 1)
    rows = autotune.sweep_presets(ctx, 13, 73712, presets=(4, 6, 7), splits=(2, 3), blocks=(64,), reps=1,
                                  warmup=0, order=order)
    errs = [r for r in rows if "error" in r]
    assert [r["preset"] for r in errs] == [7] and "P5 target" in errs[0]["error"]
    ok = [r for r in rows if "error" not in r]
    assert [(r["preset"], r["split_depth"]) for r in ok] == [(4, 2), (4, 3), (6, 2), (6, 3)]
    assert all(r["lds_stack"] == 1 and r["load_ms"] >= 0 and r["records"] > 0 for r in ok)
    assert ctx.ordered == 2 and all(lds == 1 for _, _, lds in ctx.runs)
    with pytest.raises(RuntimeError):    # a wrong total is never recorded
        autotune.sweep_presets(ctx, 13, 1, presets=(4,), splits=(2,), blocks=(64,), reps=1, warmup=0)


def test_sweep_skips_capacity_refusals(tmp_path, monkeypatch):
    """A split whose items overflow the library's u32 indices is refused with
    NQ_ECAPACITY: the sweep reports that point and goes on; any other error,
    and a wrong total, still stop it; the refused point is never stored."""
    from nqueens import _lib

    class FakeCtx:
        m = 1

        def __init__(self, code=-3):
            self.code = code

        def run_batched(self, s):
            if s.split_depth == 5:
                raise _lib.NQError(self.code, "nq_ctx_run_batched", "split depth 5 gives 9348244174 work items")
            return 314666222712, _lib.Timing(kernel_ms=float(s.split_depth), total_ms=float(s.split_depth) + 1,
                                             items=10 ** s.split_depth, launches=1)

    monkeypatch.setattr(autotune, "TUNED", {})   # record_best refreshes the module table
    rows = autotune.sweep(FakeCtx(), 21, 314666222712, splits=(3, 4, 5), blocks=(256,), lds=(1,), reps=1)
    assert [r["split_depth"] for r in rows] == [3, 4, 5]
    assert "error" in rows[2] and "NQ_ECAPACITY" in rows[2]["error"] and "kernel_ms_median" not in rows[2]
    best = autotune.record_best(21, 5, rows, tmp_path / "tuned.json")
    assert best["split_depth"] == 3
    with pytest.raises(_lib.NQError):
        autotune.sweep(FakeCtx(code=-4), 21, 314666222712, splits=(5,), blocks=(256,), lds=(1,), reps=1)
    with pytest.raises(ValueError):
        autotune.record_best(21, 5, rows[2:], tmp_path / "tuned.json")
