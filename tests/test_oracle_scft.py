"""Pins the SCFT oracle (oracle/scft.py) by closed forms and by the full-grid
pseudo-spectral computation (CPU only).

- homogeneous fields: Q = exp(-(f wA + (1-f) wB)), phiA = f, phiB = 1 - f;
- disordered melt: wA = chiN (1-f), wB = chiN f is a fixed point of the field
  update with F = chiN f (1-f);
- Simpson weights integrate cubics exactly (even n) and sum to n;
- the subgrid sweep (KRFFT diffusion on the stride-L subgrid) equals the
  full-grid sweep (numpy FFT on N^3) restricted to the subgrid, and the full-grid
  mean of q(1) equals its subgrid mean;
- Anderson mixing solves a linear fixed-point problem in at most
  dim + 1 iterations and beats simple mixing.
"""
import numpy as np
import pytest

import oracle as O
from conftest import group_setup


def test_simpson_weights():
    for n in (2, 4, 10, 70):
        w = O.contour_weights(n)
        x = np.linspace(0.0, 1.0, n + 1)
        for p in range(4):
            assert abs(np.sum(w * x ** p) / n - 1.0 / (p + 1)) < 1e-13
    for n in (1, 3, 7, 101):
        w = O.contour_weights(n)
        assert abs(w.sum() - n) < 1e-12
        x = np.linspace(0.0, 1.0, n + 1)
        assert abs(np.sum(w * x) / n - 0.5) < 1e-13


@pytest.mark.parametrize("f,ns", [(0.3, 20), (0.35, 40), (0.5, 10)])
def test_homogeneous_closed_form(f, ns):
    shape = (4, 4, 4)
    wA, wB = 1.7 * np.ones(shape), -0.4 * np.ones(shape)
    identity = lambda x: x  # D acts trivially on constants
    out = O.sweep(wA, wB, f, ns, identity)
    nA = int(round(f * ns))
    fr = nA / ns
    assert abs(out["Q"] - O.homogeneous_Q(1.7, -0.4, fr)) < 1e-13
    assert np.max(np.abs(out["phiA"] - fr)) < 1e-12
    assert np.max(np.abs(out["phiB"] - (1 - fr))) < 1e-12


def test_disordered_fixed_point():
    f, chiN = 0.3, 14.0
    shape = (4, 4, 4)
    wA, wB = chiN * (1 - f) * np.ones(shape), chiN * f * np.ones(shape)
    out = O.sweep(wA, wB, f, 10, lambda x: x)
    tA, tB, res = O.field_update(wA, wB, out["phiA"], out["phiB"], chiN)
    assert res < 1e-12
    F = O.free_energy(wA, wB, out["phiA"], out["phiB"], out["Q"], chiN)
    assert abs(F - O.disordered_free_energy(f, chiN)) < 1e-12


def _full_diffuse(N, ds, cell):
    k = [np.fft.fftfreq(n, d=1.0 / n) * 2 * np.pi / cell for n in N]
    K2 = k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2 + k[2][None, None, :] ** 2
    Kf = np.exp(-K2 * ds)
    return lambda u: np.real(np.fft.ifftn(Kf * np.fft.fftn(u)))


@pytest.mark.parametrize("sg", [229, 47])
def test_subgrid_sweep_equals_full_grid(sg):
    N = (16, 16, 16)
    ns, f, cell = 10, 0.3, 4.0
    _, _, _, R, t = group_setup(sg, N)
    Lv = O.auto_L(t)
    WA = 2.0 * O.make_symmetric(R, t, N, seed=1)
    WB = -1.0 * O.make_symmetric(R, t, N, seed=2)
    full = O.sweep(WA, WB, f, ns, _full_diffuse(N, 1.0 / ns, cell))
    spec = O.calc_spectral_asu(R, t, N)
    table = O.m2_backward_table(spec, R, t, N)
    sub = O.sweep(O.extract_subgrid(WA, Lv), O.extract_subgrid(WB, Lv), f, ns,
                  lambda x: O.m2_scft(x, spec, R, t, N, 1.0 / ns, cell * np.eye(3), table))
    assert abs(sub["Q"] - full["Q"]) < 1e-12 * full["Q"]
    for key in ("phiA", "phiB"):
        assert np.max(np.abs(sub[key] - O.extract_subgrid(full[key], Lv))) < 1e-12
    F_full = O.free_energy(WA, WB, full["phiA"], full["phiB"], full["Q"], 20.0)
    F_sub = O.free_energy(O.extract_subgrid(WA, Lv), O.extract_subgrid(WB, Lv), sub["phiA"],
                          sub["phiB"], sub["Q"], 20.0)
    assert abs(F_full - F_sub) < 1e-12


def test_anderson_linear_fixed_point():
    """x = A x + b with a contraction A (n = 6): Anderson converges in <= n + 2
    steps to 1e-10, simple mixing does not."""
    rng = np.random.default_rng(0)
    n = 6
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ np.diag(np.linspace(0.5, 0.95, n)) @ Q.T
    b = rng.standard_normal(n)
    xs = np.linalg.solve(np.eye(n) - A, b)
    mix = O.AndersonMixer(m=10, lam=1.0, lam_simple=1.0, simple_above=np.inf)
    x = np.zeros(n)
    x_simple = np.zeros(n)
    mix.simple_above = 1e9
    for it in range(n + 2):
        d = A @ x + b - x
        x = mix.step(x, d, 0.0)
        x_simple = A @ x_simple + b
    assert np.max(np.abs(x - xs)) < 1e-10
    assert np.max(np.abs(x_simple - xs)) > 1e-3
