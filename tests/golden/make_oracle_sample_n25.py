"""Writes oracle_sample_n25.npz: per-record oracle counts of N=25 at preset 8,
the preset the reference uses for N >= 25 (14Bit_codon_GPU/359Py_pushbranch_partial.py:5776-5785;
terminal depth 25-1-8 = 16).

Every 20,000th preset-5 record of N=25 (nqueens.gen_constellations) is
expanded three preset queens deeper by the oracle (set_pre_queens continued
from the record), and every 448th of those children is kept:

  parents  (p, 4) u32   the preset-5 records
  rec25    (m, 4) u32   the preset-8 records, in expansion order
  parent25 (m,)         index of each one's parent in `parents`
  w25, nodes25 (m,) u64 the oracle's weighted and node counts

    PYTHONPATH=n-queens_amd python tests/golden/make_oracle_sample_n25.py
"""
from __future__ import annotations

import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parents[1] / "n-queens_amd"))

N, PARENT_STRIDE, CHILD_STRIDE = 25, 20000, 448


def main() -> None:
    from nqueens import constellations
    from oracle import oracle
    parents = np.ascontiguousarray(constellations.gen_constellations(N, 5)[::PARENT_STRIDE])
    kids, owner = [], []
    for a in range(parents.shape[0]):
        k = oracle.expand_records(N, parents[a:a + 1], 3)
        kids.append(k)
        owner.append(np.full(k.shape[0], a, np.int64))
    allk, own = np.concatenate(kids), np.concatenate(owner)
    rec, par = np.ascontiguousarray(allk[::CHILD_STRIDE]), own[::CHILD_STRIDE]
    fid = oracle.build_soa(N, rec)[:, 0]
    t0 = time.time()
    w, _, nodes = oracle.count_records(N, rec, threads=len(os.sched_getaffinity(0)))
    print(f"N={N} preset 8: {parents.shape[0]} parents, {allk.shape[0]} children, {rec.shape[0]} sampled; "
          f"P5 among them: {int(((fid >= 8) & (fid <= 11)).sum())}; total {int(w.sum())}, nodes {int(nodes.sum())}, "
          f"{time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(HERE / "oracle_sample_n25.npz", parents=parents, rec25=rec, parent25=par, w25=w,
                        nodes25=nodes)


if __name__ == "__main__":
    main()
