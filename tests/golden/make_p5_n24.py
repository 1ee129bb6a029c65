"""Writes p5_n24.npz: the N=24 preset-5 constellations whose target is a P5
functionid (8..11: the j diagonal enters the board below the start row,
13Bit_codon/13Py_constellations_codon.py:620-678), with

  rec      (18, 4) u32   the records (nqueens.gen_constellations order)
  index    (18,)         their positions in that list
  w        (18,)  u64    weighted count of each, by the oracle's P5 branch
  nodes    (18,)  u64    its node count
  kids     (18,)         records it expands to two preset queens deeper (preset 7)
  kids_w   (18,)  u64    their weighted counts summed (no P5 target among them)
  kids_nodes (18,) u64   their node counts summed

The partition of a record's completions by the next two queens gives
w == kids_w exactly; the node counts differ by the extractions of the two
expanded rows, which the P5 rows place without the one-row lookahead:
nodes == kids_nodes + (records of preset 6 it expands to) + kids.

    PYTHONPATH=n-queens_amd python tests/golden/make_p5_n24.py
"""
from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parents[1] / "n-queens_amd"))


def main() -> None:
    from nqueens import constellations
    from oracle import oracle
    N = 24
    rec = constellations.gen_constellations(N, 5)
    soa = oracle.build_soa(N, rec)
    idx = np.where((soa[:, 0] >= 8) & (soa[:, 0] <= 11))[0]
    sel = np.ascontiguousarray(rec[idx])
    t0 = time.time()
    w, _, nodes = oracle.count_records(N, sel, schedule=0)   # one record per thread
    print(f"{len(idx)} P5 records counted in {time.time() - t0:.0f} s", flush=True)
    kids6 = [oracle.expand_records(N, sel[a:a + 1], 1) for a in range(len(idx))]
    kids7 = [oracle.expand_records(N, sel[a:a + 1], 2) for a in range(len(idx))]
    allk = np.concatenate(kids7)
    ks = oracle.build_soa(N, allk)
    assert not ((ks[:, 0] >= 8) & (ks[:, 0] <= 11)).any()
    t0 = time.time()
    kw, _, kn = oracle.count_records(N, allk)
    print(f"{allk.shape[0]} descendants counted in {time.time() - t0:.0f} s", flush=True)
    bounds = np.cumsum([0] + [k.shape[0] for k in kids7])
    kids_w = np.array([kw[bounds[a]:bounds[a + 1]].sum() for a in range(len(idx))], np.uint64)
    kids_n = np.array([kn[bounds[a]:bounds[a + 1]].sum() for a in range(len(idx))], np.uint64)
    n6 = np.array([k.shape[0] for k in kids6], np.int64)
    n7 = np.array([k.shape[0] for k in kids7], np.int64)
    for a in range(len(idx)):
        print(int(idx[a]), int(soa[idx[a], 0]), int(w[a]), int(kids_w[a]), int(nodes[a]), int(kids_n[a]),
              int(n6[a]), int(n7[a]), flush=True)
    np.savez_compressed(HERE / "p5_n24.npz", rec=sel, index=idx.astype(np.int64), w=w, nodes=nodes,
                        kids6=n6, kids=n7, kids_w=kids_w, kids_nodes=kids_n)


if __name__ == "__main__":
    main()
