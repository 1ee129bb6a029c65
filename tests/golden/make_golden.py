"""Writes the golden fixtures of tests/golden from the reference's own data.

Run in a checkout that has the reference mounted (default /root/reference):

    python tests/golden/make_golden.py [REFERENCE_ROOT]

Sources (data files and tables of the reference, read as data):
  13Bit_codon/Archive/bin/constellations_N{6..21}_5.bin   16-byte records
      ld, rd, col, startijkl (u32 LE), written by
      13Bit_codon/18Py_constellations_GPU_cuda_codon.py:2009-2016
  13Bit_codon/Archive/bin/solutions_N{6..19}_5_v2.bin     40-byte records
      startijkl, ld, rd, col, solutions (u64 LE), written by 18Py:1841-1855
  18Py:2089  expected totals (A000170)
  README.en:12-33  Total/Unique table of the bit-symmetry GPU solver
  14Bit_codon_GPU/359Py_pushbranch_partial_validate_N21_full_once.sh:108
      EXPECTED_TASKS=2025282 (N=21, preset 6)
  13Bit_codon/Archive/bc_p6_N18_boundary_start.log:3   N=18 preset 6: 571705
  13Bit_codon/Archive/bc_p7_N18_boundary_start_80.log:3 N=18 preset 7: 3653940
  13Bit_codon/Archive/bc_sol_N18_N20_p6_78.log:3-4  18Py-semantics N=18 preset 6:
      571640 records (the whole-call memo drops 65), total 666053384

Outputs:
  p5_full.npz     every record + weighted count for N = 6..13 (preset 5)
  p5_digests.json per N: record counts, sum of solutions, SHA-256 of the
                  sorted record stream (16-byte records) and of the sorted
                  (ld, rd, col, startijkl, solutions) stream (u64 LE)
  known.json      totals, unique counts and preset-6/7 record counts
"""
from __future__ import annotations

import hashlib
import json
import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def sorted_records(rec: np.ndarray) -> np.ndarray:
    return rec[np.lexsort(rec.T[::-1])]


def digest_records(rec: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(sorted_records(rec).astype("<u4")).tobytes()).hexdigest()


def digest_solutions(rec: np.ndarray, sol: np.ndarray) -> str:
    a = np.concatenate([rec.astype(np.uint64), sol.astype(np.uint64)[:, None]], axis=1)
    a = a[np.lexsort(a.T[::-1])]
    return hashlib.sha256(np.ascontiguousarray(a.astype("<u8")).tobytes()).hexdigest()


def main(ref: Path) -> None:
    binp = ref / "13Bit_codon/Archive/bin"
    full = {}
    dig = {}
    for N in range(6, 22):
        c = np.fromfile(binp / f"constellations_N{N}_5.bin", dtype="<u4").reshape(-1, 4).astype(np.uint32)
        d = {"constellations": int(c.shape[0]), "records_sha256": digest_records(c)}
        sp = binp / f"solutions_N{N}_5_v2.bin"
        if sp.exists():
            s = np.fromfile(sp, dtype="<u8").reshape(-1, 5)
            rec = np.stack([s[:, 1], s[:, 2], s[:, 3], s[:, 0]], axis=1).astype(np.uint32)
            sol = s[:, 4].astype(np.uint64)
            assert rec.shape[0] == c.shape[0]
            assert digest_records(rec) == d["records_sha256"]
            d["solutions_sum"] = int(sol.sum())
            d["solutions_sha256"] = digest_solutions(rec, sol)
            w = np.bincount(np.minimum(sol, 2**62).astype(np.int64) == 0, minlength=2)
            d["zero_records"] = int(w[1])
            if N <= 13:
                full[f"rec{N}"] = rec
                full[f"sol{N}"] = sol
        dig[str(N)] = d
    np.savez_compressed(HERE / "p5_full.npz", **full)
    (HERE / "p5_digests.json").write_text(json.dumps(dig, indent=1, sort_keys=True) + "\n")

    src = (ref / "13Bit_codon/18Py_constellations_GPU_cuda_codon.py").read_text()
    m = re.search(r"expected:List\[int\]=\[([0-9,]+)\]", src)
    expected = [int(x) for x in m.group(1).split(",")]
    totals = {str(n): v for n, v in enumerate(expected) if v}
    table = {}
    for line in (ref / "README.en").read_text().splitlines()[8:34]:
        mm = re.match(r"\s*(\d+):\s+(\d+)\s+(\d+)\s", line)
        if mm:
            table[mm.group(1)] = [int(mm.group(2)), int(mm.group(3))]
    known = {
        "total_a000170": totals,
        "readme_total_unique": table,
        "preset_record_counts": {
            "18,6": 571705, "21,6": 2025282, "18,7": 3653940,
        },
        "preset_record_counts_18py_memo": {"18,6": 571640},
        "preset_totals_18py_memo": {"18,6": 666053384},
        "sources": {
            "total_a000170": "13Bit_codon/18Py_constellations_GPU_cuda_codon.py:2089",
            "readme_total_unique": "README.en:12-33",
            "18,6": "13Bit_codon/Archive/bc_p6_N18_boundary_start.log:3",
            "21,6": "14Bit_codon_GPU/359Py_pushbranch_partial_validate_N21_full_once.sh:108",
            "18,7": "13Bit_codon/Archive/bc_p7_N18_boundary_start_80.log:3",
            "18py_memo 18,6": "13Bit_codon/Archive/bc_sol_N18_N20_p6_78.log",
            "18py_memo total 18,6": "13Bit_codon/Archive/bc_sol_N18_N20_p6_78.log:4",
        },
    }
    (HERE / "known.json").write_text(json.dumps(known, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference"))
