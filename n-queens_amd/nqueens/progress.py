"""Per-chunk progress log of a chunked count: audit and resume.

A long count (``solver.count_chunks``) writes one tab-separated line per
counted chunk -- board size, chunk count, preset, chunk index, the chunk's
weighted total, seconds -- so that a run can be audited, or resumed after an
interruption by counting only the chunks the log lacks (the reference's
per-chunk progress TSV with a running total,
14Bit_codon_GPU/359Py_pushbranch_partial.py:3362-3370, 5560-5642).  Chunk
totals are exact integers; the N-Queens total is their sum once every chunk
is present.  A line that does not parse, or that contradicts an earlier line
for the same chunk, is an error: an exact count never guesses.
"""
from __future__ import annotations

from pathlib import Path

HEADER = "#N\tchunks\tpreset\tchunk\ttotal\tseconds"


class ChunkLog:
    """The chunks of one (N, chunks, preset) count recorded in `path`."""

    def __init__(self, path, N: int, chunks: int, preset: int):
        self.path = Path(path)
        self.key = (int(N), int(chunks), int(preset))
        self.totals: dict[int, int] = {}
        self.other_presets: set[int] = set()   # presets of lines for the same N and chunk count
        if self.path.exists():
            self._read()

    def _read(self) -> None:
        for no, line in enumerate(self.path.read_text().splitlines(), 1):
            if not line.strip() or line.startswith("#"):
                continue
            f = line.split("\t")
            try:
                if len(f) != 6:
                    raise ValueError
                n, k, p, i, t = (int(x) for x in f[:5])
                float(f[5])
            except ValueError:
                raise ValueError(f"{self.path}:{no}: not a progress line: {line!r}") from None
            if (n, k, p) != self.key:
                if (n, k) == self.key[:2]:
                    self.other_presets.add(p)
                continue
            if not 0 <= i < k or t < 0:
                raise ValueError(f"{self.path}:{no}: chunk {i} of {k}, total {t}")
            if i in self.totals and self.totals[i] != t:
                raise ValueError(f"{self.path}:{no}: chunk {i} recorded twice with different totals")
            self.totals[i] = t

    def missing(self, first: int = 0, last: int | None = None) -> list[int]:
        last = self.key[1] if last is None else last
        return [i for i in range(first, last) if i not in self.totals]

    def record(self, chunk: int, total: int, seconds: float) -> None:
        if chunk in self.totals and self.totals[chunk] != total:
            raise ValueError(f"chunk {chunk}: total {total} differs from the recorded {self.totals[chunk]}")
        new = not self.path.exists() or self.path.stat().st_size == 0
        with self.path.open("a") as fh:
            if new:
                fh.write(HEADER + "\n")
            n, k, p = self.key
            fh.write(f"{n}\t{k}\t{p}\t{chunk}\t{total}\t{seconds:.3f}\n")
            fh.flush()
        self.totals[chunk] = int(total)

    def complete(self) -> bool:
        return len(self.totals) == self.key[1]

    def total(self) -> int:
        if not self.complete():
            raise ValueError(f"{len(self.totals)} of {self.key[1]} chunks recorded")
        return sum(self.totals.values())
