"""Writes the oracle sample fixtures:

  oracle_sample_n20_n21.npz  every 400th constellation record of N=20 and
                             N=21 (preset 5)
  oracle_sample_n22_n23.npz  every 2,000th record of N=22 (preset 5) and
                             every 20,000th of N=23 (preset 6)
  oracle_sample_n24.npz      every 600,000th record of N=24 (preset 7)

each with its weighted count and node count from the oracle's dfs_iter
restatement (oracle/nq_oracle.c, OpenMP), records generated by
nqueens.gen_constellations.  The reference ships no per-record solutions for
N >= 20 (its solutions_N*_5_v2.bin stop at N=19), so these cached oracle
results let the GPU tests check per-record exactness at the largest sizes in
seconds; the N=20..24 totals of whole runs are pinned by A000170 in
known.json.

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

SAMPLES = {
    "oracle_sample_n20_n21.npz": [(20, 5, 400), (21, 5, 400)],
    "oracle_sample_n22_n23.npz": [(22, 5, 2000), (23, 6, 20000)],
    "oracle_sample_n24.npz": [(24, 7, 600000)],
}


def main() -> None:
    from nqueens import constellations
    from oracle import oracle
    threads = len(os.sched_getaffinity(0))
    for name, plan in SAMPLES.items():
        if (HERE / name).exists() and "--all" not in sys.argv:
            continue
        out = {}
        for N, preset, stride in plan:
            rec = constellations.gen_constellations(N, preset)[::stride].copy()
            w, _, nodes = oracle.count_records(N, rec, threads=threads)
            out[f"rec{N}"] = rec
            out[f"w{N}"] = w
            out[f"nodes{N}"] = nodes
            out[f"preset{N}"] = np.array(preset)
            print(name, N, preset, rec.shape[0], int(w.sum()), int(nodes.sum()), flush=True)
        np.savez_compressed(HERE / name, **out)


if __name__ == "__main__":
    main()
