"""Writes nodes_13py.npz: per-constellation DFS node counts of the
reference's own recursive solver, the plain-CPython NQueens13 class of
13Bit_codon/13Py_constellations_codon.py (28 SQ* functions, :183-678;
dispatch exec_solutions, :775-930).

A node is one candidate extraction, `bit=free&-free` in the `while free:`
loop of an SQ* function.  The script imports that module unchanged (in this
container only: the reference is not present where the tests run), runs its
exec_solutions on one constellation at a time and counts, with sys.settrace,
how often the extraction lines of its SQ* methods execute.  It stores, for
each (N, preset), the records (from nqueens.gen_constellations), the
weighted counts the reference writes into constellation["solutions"] and
the node counts:

  rec{N}_{p}   (m, 4) u32      sol{N}_{p}   (m,) u64      nodes{N}_{p}   (m,) u64

    PYTHONPATH=n-queens_amd python tests/golden/make_13py_nodes.py
"""
from __future__ import annotations

import importlib.util
import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1] / "n-queens_amd"))

REF = Path("/root/reference/13Bit_codon/13Py_constellations_codon.py")
CASES = [(N, 4) for N in range(6, 14)] + [(N, 5) for N in range(6, 15)]
EXTRACT = re.compile(r"^\s*bit\s*(:\s*int)?\s*=\s*(free\s*&\s*-\s*free|-\s*free\s*&\s*free)\s*$")


def load_reference():
    spec = importlib.util.spec_from_file_location("ref13py", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main() -> None:
    from nqueens import constellations
    mod = load_reference()
    lines = REF.read_text().splitlines()
    extract = {i + 1 for i, s in enumerate(lines) if EXTRACT.match(s)}
    solver = mod.NQueens13()
    counter = [0]

    def local(frame, event, arg):
        if event == "line" and frame.f_lineno in extract:
            counter[0] += 1
        return local

    def tracer(frame, event, arg):
        if frame.f_code.co_filename == str(REF) and frame.f_code.co_name.startswith("SQ"):
            return local
        return None

    out = {}
    for N, p in CASES:
        rec = constellations.gen_constellations(N, p)
        sol = np.zeros(rec.shape[0], np.uint64)
        nodes = np.zeros(rec.shape[0], np.uint64)
        for a, r in enumerate(rec):
            c = [{"ld": int(r[0]), "rd": int(r[1]), "col": int(r[2]), "startijkl": int(r[3]), "solutions": 0}]
            counter[0] = 0
            sys.settrace(tracer)
            try:
                solver.exec_solutions(c, N)
            finally:
                sys.settrace(None)
            sol[a] = c[0]["solutions"]
            nodes[a] = counter[0]
        out[f"rec{N}_{p}"] = rec
        out[f"sol{N}_{p}"] = sol
        out[f"nodes{N}_{p}"] = nodes
        print(N, p, rec.shape[0], int(sol.sum()), int(nodes.sum()), flush=True)
    np.savez_compressed(HERE / "nodes_13py.npz", **out)


if __name__ == "__main__":
    main()
