"""The per-chunk progress log (nqueens.progress): resume bookkeeping and its
refusal of logs it cannot trust.  Runs on the CPU."""
import pytest

from nqueens.progress import ChunkLog


def test_log_records_and_resumes(tmp_path):
    p = tmp_path / "run.tsv"
    log = ChunkLog(p, 16, 4, 5)
    assert log.missing() == [0, 1, 2, 3] and not log.complete()
    log.record(0, 100, 1.5)
    log.record(2, 300, 1.0)
    again = ChunkLog(p, 16, 4, 5)                 # a later run reads what the first wrote
    assert again.missing() == [1, 3] and again.missing(2, 4) == [3]
    again.record(1, 200, 0.5)
    again.record(3, 400, 0.5)
    assert again.complete() and again.total() == 1000
    other = ChunkLog(p, 16, 8, 5)                 # another chunking of the same N shares nothing
    assert other.missing() == list(range(8))


def test_log_refuses_bad_lines(tmp_path):
    p = tmp_path / "run.tsv"
    p.write_text("#N\tchunks\tpreset\tchunk\ttotal\tseconds\n16\t4\t5\t0\t100\t1.0\n16\t4\t5\t0\t101\t1.0\n")
    with pytest.raises(ValueError):
        ChunkLog(p, 16, 4, 5)                     # one chunk, two totals
    p.write_text("16\t4\t5\tzero\t100\t1.0\n")
    with pytest.raises(ValueError):
        ChunkLog(p, 16, 4, 5)
    p.write_text("16\t4\t5\t7\t100\t1.0\n")
    with pytest.raises(ValueError):
        ChunkLog(p, 16, 4, 5)                     # chunk index beyond the chunk count
    log = ChunkLog(tmp_path / "new.tsv", 16, 4, 5)
    log.record(0, 5, 0.1)
    with pytest.raises(ValueError):
        log.record(0, 6, 0.1)
    with pytest.raises(ValueError):
        log.total()


def test_log_notes_other_presets(tmp_path):
    """Lines of the same N and chunk count at another preset are not this
    count's chunks, but the log says they are there (the CLI then refuses to
    start over silently)."""
    p = tmp_path / "run.tsv"
    ChunkLog(p, 20, 4, 5).record(0, 100, 1.0)
    log = ChunkLog(p, 20, 4, 4)
    assert log.missing() == [0, 1, 2, 3] and log.other_presets == {5}
    assert ChunkLog(p, 20, 8, 4).other_presets == set()


def test_cli_refuses_log_of_another_preset(tmp_path):
    from nqueens.__main__ import main
    p = tmp_path / "run.tsv"
    ChunkLog(p, 20, 4, 7).record(0, 100, 1.0)     # a preset the table never picks for N=20
    with pytest.raises(SystemExit) as e:
        main(["20", "--chunks", "4", "--progress-file", str(p)])
    assert "pass --preset" in str(e.value)
