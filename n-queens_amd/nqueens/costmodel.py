"""Learned cost model: predicts each constellation's DFS node count so that
the batched kernel can take heavy subtrees first.

The model is a 3-layer FP32 MLP (24 -> 64 -> 64 -> 1, ReLU) over 24
features of a constellation (border queens i,j,k,l, start row, function id,
rows to the mark/jmark/end rows, weight, popcounts of the start diagonals and
columns, free squares of the next five rows, ...).  Features, forward pass,
backward pass and Adam updates all run in libnq's own kernels; the labels are
log(1 + nodes) measured per constellation by the node-counting variant of
the batched kernel.

The reference orders its work by a hand-made key
(14Bit_codon_GPU/359Py_pushbranch_partial.py:4342-4404) whose rank
correlation with measured per-task work is -0.36
(14Bit_codon_GPU/README.md:4209); its notes list regression on measured work
as the untried next step (:4259).  The prediction only changes the order in
which work items are queued; counts never depend on it, and plain record
order (FIFO) is the fallback.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from ._lib import NQ_MLP_H, NQ_MLP_IN, NQ_MLP_NPARAM, Context

__all__ = ["CostModel", "init_params", "collect", "train", "default_model", "spearman", "binned_splits",
           "simulated_makespan", "MODEL_PATH"]

MODEL_PATH = Path(__file__).resolve().parent / "costmodel.npz"


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases, in the packed layout W1|b1|W2|b2|W3|b3."""
    rng = np.random.default_rng(seed)
    H, I = NQ_MLP_H, NQ_MLP_IN
    w1 = rng.uniform(-1, 1, (H, I)) * np.sqrt(6.0 / I)
    w2 = rng.uniform(-1, 1, (H, H)) * np.sqrt(6.0 / H)
    w3 = rng.uniform(-1, 1, H) * np.sqrt(6.0 / H) * 0.1
    p = np.concatenate([w1.ravel(), np.zeros(H), w2.ravel(), np.zeros(H), w3, np.zeros(1)]).astype(np.float32)
    assert p.size == NQ_MLP_NPARAM
    return p


def spearman(a: np.ndarray, b: np.ndarray) -> float:
    """Spearman rank correlation (ties broken by position)."""
    ra = np.empty(len(a)); ra[np.argsort(a, kind="stable")] = np.arange(len(a))
    rb = np.empty(len(b)); rb[np.argsort(b, kind="stable")] = np.arange(len(b))
    return float(np.corrcoef(ra, rb)[0, 1])


@dataclass
class CostModel:
    params: np.ndarray
    y_mean: float = 0.0
    trained_on: list = field(default_factory=list)
    history: list = field(default_factory=list)

    def predict_features(self, ctx: Context, feats: np.ndarray) -> np.ndarray:
        """Predicted log(1 + nodes) for feature rows (GPU forward kernel)."""
        return ctx.mlp_forward(self.params, feats) + np.float32(self.y_mean)

    def predict(self, N: int, rec, ctx: Context) -> np.ndarray:
        ctx.load(N, rec)
        return self.predict_features(ctx, ctx.features())

    def order(self, N: int, rec, ctx: Context) -> np.ndarray:
        """Processing order: predicted heaviest constellations first."""
        pred = self.predict(N, rec, ctx)
        return np.argsort(-pred, kind="stable").astype(np.uint32)

    def save(self, path: Path = MODEL_PATH) -> None:
        np.savez(path, params=self.params, y_mean=np.float64(self.y_mean),
                 meta=np.frombuffer(json.dumps({"trained_on": self.trained_on,
                                                "history": self.history}).encode(), dtype=np.uint8))

    @classmethod
    def load(cls, path: Path = MODEL_PATH) -> "CostModel":
        with np.load(path, allow_pickle=False) as z:
            meta = json.loads(bytes(z["meta"]).decode())
            return cls(params=z["params"].astype(np.float32), y_mean=float(z["y_mean"]),
                       trained_on=meta.get("trained_on", []), history=meta.get("history", []))


def collect(ctx: Context, Ns=(16, 17, 18, 19), preset: int = 5, sched=None):
    """Features and measured log(1 + nodes) of every constellation of each N
    (node-counting batched runs on the GPU)."""
    from .constellations import gen_constellations
    from .solver import default_sched
    xs, ys, ns = [], [], []
    for N in Ns:
        rec = gen_constellations(N, preset)
        ctx.load(N, rec)
        ctx.run_batched(sched or default_sched(N), count_nodes=True)
        _, nodes = ctx.results(weighted=False, nodes=True)
        xs.append(ctx.features())
        ys.append(np.log1p(nodes.astype(np.float64)).astype(np.float32))
        ns.append(np.full(rec.shape[0], N, np.int32))
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(ns)


def train(ctx: Context, X: np.ndarray, y: np.ndarray, epochs: int = 40, batch: int = 512, lr: float = 2e-3,
          seed: int = 0, shuffle_seed: int = 1, holdout: float = 0.1, trained_on=None) -> tuple[CostModel, dict]:
    """Train on the GPU (Adam, minibatches); returns the model and a report
    with the held-out MAE and Spearman correlation of log(1 + nodes)."""
    rng = np.random.default_rng(shuffle_seed)
    idx = rng.permutation(X.shape[0])
    nh = int(round(holdout * X.shape[0]))
    te, tr = idx[:nh], idx[nh:]
    Xtr, ytr = np.ascontiguousarray(X[tr]), np.ascontiguousarray(y[tr])
    y_mean = float(ytr.mean())
    params = init_params(seed)
    perm = np.concatenate([rng.permutation(Xtr.shape[0]) for _ in range(epochs)]).astype(np.uint32)
    losses = ctx.mlp_train(params, Xtr, ytr - np.float32(y_mean), perm, epochs, batch, lr)
    model = CostModel(params=params, y_mean=y_mean, trained_on=list(trained_on or []),
                      history=[float(v) for v in losses])
    report = {"train_loss_first": float(losses[0]), "train_loss_last": float(losses[-1]),
              "train_rows": int(tr.size), "holdout_rows": int(nh)}
    if nh:
        pred = model.predict_features(ctx, np.ascontiguousarray(X[te]))
        report["holdout_mae"] = float(np.mean(np.abs(pred - y[te])))
        report["holdout_spearman"] = spearman(pred, y[te])
    return model, report


def binned_splits(pred_log_nodes: np.ndarray, T: int, quantiles=(0.5, 0.8, 0.95), base: int = 1) -> np.ndarray:
    """Per-record split depths from predicted work: constellations are binned
    by predicted node count at the given quantiles and each heavier bin is
    split one level deeper than the last (lightest bin at `base`), never
    deeper than T-2."""
    pred = np.asarray(pred_log_nodes, dtype=np.float64)
    cuts = np.quantile(pred, quantiles) if pred.size else np.zeros(len(quantiles))
    d = base + np.searchsorted(cuts, pred, side="right")
    return np.clip(d, 0, max(T - 2, 0)).astype(np.uint8)


def simulated_makespan(costs: np.ndarray, order: np.ndarray, lanes: int) -> float:
    """Makespan of list scheduling: items taken in `order` by the first free of
    `lanes` identical workers (the persistent queue's policy), in cost units."""
    import heapq
    heap = [0.0] * min(lanes, len(order))
    heapq.heapify(heap)
    for i in order:
        t = heapq.heappop(heap)
        heapq.heappush(heap, t + float(costs[i]))
    return max(heap) if heap else 0.0


_default = None


def default_model() -> CostModel:
    """The committed model (trained by `python -m nqueens.costmodel`)."""
    global _default
    if _default is None:
        if not MODEL_PATH.exists():
            raise RuntimeError(f"{MODEL_PATH} missing: train it with python -m nqueens.costmodel")
        _default = CostModel.load(MODEL_PATH)
    return _default


def study(ctx: Context, N: int, preset: int = 5, reps: int = 2, block: int = 256, log=print) -> dict:
    """What the model buys at one N: prediction error against measured node
    counts, and the batched run time for FIFO, model and true-work orders at
    uniform split depths, and for model-binned per-record depths."""
    import statistics
    from .constellations import gen_constellations
    from .expected import TOTAL
    from ._lib import Sched
    rec = gen_constellations(N, preset)
    ctx.load(N, rec)
    ctx.run_batched(Sched(split_depth=3, block=block, lds_stack=1, use_graph=0, chunk=64, waves_per_eu=0),
                    count_nodes=True)
    _, nodes = ctx.results(weighted=False, nodes=True)
    truth = np.log1p(nodes.astype(np.float64))
    model = default_model()
    pred = model.predict_features(ctx, ctx.features()).astype(np.float64)
    out = {"N": N, "preset": preset, "constellations": int(rec.shape[0]), "nodes": int(nodes.sum()),
           "pred_mae_log": float(np.mean(np.abs(pred - truth))), "pred_spearman": spearman(pred, truth),
           "runs": []}
    log(json.dumps({k: v for k, v in out.items() if k != "runs"}))
    T = N - 1 - preset
    orders = {"fifo": None, "mlp": np.argsort(-pred, kind="stable").astype(np.uint32),
              "true": np.argsort(-truth, kind="stable").astype(np.uint32)}

    def timed(label, order, split, depths=None):
        ctx.load(N, rec, order)
        if depths is not None:
            ctx.set_splits(depths)
        s = Sched(split_depth=split, block=block, lds_stack=1, use_graph=0, chunk=64, waves_per_eu=0)
        ms, items = [], 0
        for _ in range(reps + 1):
            tot, t = ctx.run_batched(s)
            if tot != TOTAL[N]:
                raise RuntimeError(f"{label}: wrong total {tot}")
            ms.append(t.total_ms)
            items = int(t.items)
        row = {"config": label, "split": split, "items": items, "run_ms": statistics.median(ms[1:])}
        out["runs"].append(row)
        log(json.dumps(row))

    for split in (1, 2, 3):
        for name, order in orders.items():
            timed(f"uniform split {split}, {name} order", order, split)
    for q, base in (((0.5, 0.8, 0.95), 1), ((0.5, 0.9), 1), ((0.7,), 2), ((0.5, 0.9), 2)):
        d = binned_splits(pred, T, quantiles=q, base=base)
        timed(f"mlp bins at quantiles {list(q)} from depth {base}, mlp order", orders["mlp"], -1, d)
    return out


def main(argv=None) -> None:
    """Collect labels at N=16..19 (preset 5), train, report, save
    (python -m nqueens.costmodel [--out PATH] to also write a copy);
    --study N: compare schedules at N with the saved model."""
    import argparse
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--study", type=int, default=0)
    a = ap.parse_args(argv)
    ctx = Context(0)
    if a.study:
        res = study(ctx, a.study, log=lambda m: print(m, flush=True))
        if a.out:
            Path(a.out).write_text(json.dumps(res, indent=1) + "\n")
        return
    t0 = time.perf_counter()
    X, y, ns = collect(ctx)
    t1 = time.perf_counter()
    model, rep = train(ctx, X, y, epochs=a.epochs, trained_on=[16, 17, 18, 19])
    t2 = time.perf_counter()
    rep.update({"collect_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2), "rows": int(X.shape[0])})
    model.save()
    if a.out:
        model.save(Path(a.out))
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
