"""The 32-point line transform over eight lanes of four values (the per-axis
chain's line transforms, csrc/fft.hpp fft32_l8_lds) restated in NumPy: the
in-lane 4-point DFTs of x[r + 8 m], the twiddles w32^(r k), and per lane
r = 2 k + h the parity half of the 8-point DFT across the lanes,
X[k + 4 (h + 2 j)] = sum_s w8^(s h) w4^(s j) (y_s + (-1)^h y_{s+4}),
give numpy.fft.fft in both directions, every output position written once."""
import numpy as np
import pytest


def _line32(x, sign):
    w = lambda n, e: np.exp(sign * 2j * np.pi * e / n)
    # stage 1: lane r holds x[r + 8 m], m < 4
    y = np.empty((8, 4), dtype=complex)
    for r in range(8):
        v = x[r + 8 * np.arange(4)]
        y[r] = [np.sum(v * w(4, np.arange(4) * k)) * w(32, r * k) for k in range(4)]
    out = np.full(32, np.nan, dtype=complex)
    for r in range(8):
        k, h = r >> 1, r & 1
        z = np.array([(y[s, k] + (-1) ** h * y[s + 4, k]) * w(8, s * h) for s in range(4)])
        for j in range(4):
            pos = k + 4 * (h + 2 * j)
            assert np.isnan(out[pos])
            out[pos] = np.sum(z * w(4, np.arange(4) * j))
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_line32_equals_fft(seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(32) + 1j * rng.standard_normal(32)
    assert np.allclose(_line32(x, -1), np.fft.fft(x), rtol=0, atol=1e-12)
    assert np.allclose(_line32(x, +1), np.fft.ifft(x) * 32, rtol=0, atol=1e-12)
