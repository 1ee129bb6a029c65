"""Writes tests/golden/oracle_sample_n20_n21.npz: every 400th constellation
record of N=20 and N=21 (preset 5, generated by nqueens.gen_constellations)
with its weighted count and node count from the oracle's dfs_iter
restatement (oracle/nq_oracle.c, OpenMP).  The reference ships no per-record
solutions for N >= 20 (its solutions_N*_5_v2.bin stop at N=19), so these
cached oracle results let the GPU tests check per-record exactness at the
largest sizes in seconds; the totals they sum to are pinned by the
reference's N=20/21 totals in known.json.

    PYTHONPATH=n-queens_amd python tests/golden/make_oracle_sample.py
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parents[1] / "n-queens_amd"))

STRIDE = 400


def main() -> None:
    from nqueens import constellations
    from oracle import oracle
    out = {}
    for N in (20, 21):
        rec = constellations.gen_constellations(N, 5)[::STRIDE].copy()
        w, _, nodes = oracle.count_records(N, rec, threads=len(os.sched_getaffinity(0)))
        out[f"rec{N}"] = rec
        out[f"w{N}"] = w
        out[f"nodes{N}"] = nodes
        print(N, rec.shape[0], int(w.sum()), int(nodes.sum()), flush=True)
    np.savez_compressed(HERE / "oracle_sample_n20_n21.npz", **out)


if __name__ == "__main__":
    main()
