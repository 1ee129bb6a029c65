"""The VALU issue probe behind the bench's roofline (nq_issue_probe)."""
import numpy as np
import pytest

from nqueens import _lib

pytestmark = pytest.mark.gpu

K_OR, K_AND = 0x01234567, 0xFEDCBA98


def expected_xor(lanes: int) -> int:
    """Every lane g runs chains a_i = (a_i | k_i) & m_i from a_i = g + i,
    k_i = K_OR << i, m_i = K_AND >> i; the map is idempotent, so any number of
    iterations leaves (g + i | k_i) & m_i.  The XOR over chains and lanes."""
    g = np.arange(lanes, dtype=np.uint64)
    x = np.zeros(lanes, dtype=np.uint64)
    for i in range(8):
        k = (K_OR << i) & 0xFFFFFFFF
        m = K_AND >> i
        x ^= ((g + i) | k) & m
    return int(np.bitwise_xor.reduce(x & 0xFFFFFFFF))


@pytest.mark.parametrize("waves", [1, 4])
def test_issue_probe_runs_and_times(ctx, waves):
    d = ctx.device_info()
    p = ctx.issue_probe(waves, 3000)
    print(f"issue probe: {waves} waves/SIMD, {p['clk_per_wave_instr_per_simd']:.3f} clk per wave instruction per "
          f"SIMD, {p['lane_ops_per_s'] / 1e12:.2f} T lane-ops/s")
    assert p["xor"] == expected_xor(d["compute_units"] * 4 * waves * 64)
    # a wave64 instruction on a 16-lane SIMD takes 4 clocks; gfx950 issues
    # simple ops of several waves faster, never below 1 clock
    assert 1.0 < p["clk_per_wave_instr_per_simd"] < 40.0
    assert p["ms"] > 0


def test_issue_probe_rejects_bad_arguments(ctx):
    with pytest.raises(_lib.NQError):
        ctx.issue_probe(0, 100)
    with pytest.raises(_lib.NQError):
        ctx.issue_probe(9, 100)
    with pytest.raises(_lib.NQError):
        ctx.issue_probe(4, 0)
