"""The dense real-ASU diffusion operator (plan_diffusion's default
method=:matrix, src/diffusion_solver.jl:20-71): host planning of M, Q and M^+
(CPU), pinned by the reference's own checks (test/test_diffusion_solver.jl:
6-49: M^+ M = I on the spectral ASU, M M^+ keeps real fields real) and by the
full-grid FFT . K . IFFT of a symmetric field at the real-ASU points; the
device operator (GPU) against the host product and the Strang step."""
import numpy as np
import pytest

import cfft_hip as cf
import oracle as O
from cfft_hip.asu import calc_asu
from cfft_hip.diffusion_solver import calc_transform_matrix, matrix_kernel
from cfft_hip.plans import calc_spectral_asu


def _plan(sg, dim, N):
    ops = cf.get_ops(sg, dim, N)
    shift, ops_s = cf.find_optimal_shift(ops, N)
    spec = calc_spectral_asu(ops_s, dim, N)
    pts = calc_asu(N, ops_s)
    M = calc_transform_matrix(pts, spec, ops_s, N, shift)
    return ops_s, spec, pts, M


def _fullgrid_diffusion(u, N, ds, lat):
    B = 2 * np.pi * np.linalg.inv(lat).T
    hs = [np.where(np.arange(n) >= n // 2, np.arange(n) - n, np.arange(n)) for n in N]
    H = np.stack(np.meshgrid(*hs, indexing="ij"), -1) @ B.T
    return np.real(np.fft.ifftn(np.fft.fftn(u) * np.exp(-(H * H).sum(-1) * ds)))


def test_transform_matrix_p2mm_reference_checks():
    """test/test_diffusion_solver.jl:6-49 (P2mm, 8 x 8): M^+ M = I, and a real
    vector of ASU values comes back real and unchanged through M M^+."""
    _, spec, pts, M = _plan(6, 2, (8, 8))
    Mi = np.linalg.pinv(M)
    assert np.allclose(Mi @ M, np.eye(len(spec)), atol=1e-10)
    u = np.random.default_rng(0).random(len(pts))
    v = M @ (Mi @ u)
    assert np.max(np.abs(v.imag)) < 1e-10
    assert np.allclose(v.real, u, atol=1e-10)


@pytest.mark.parametrize("sg", [47, 221, 229, 230])
def test_matrix_operator_equals_fullgrid_diffusion(sg):
    """A = Re(M diag(K) M^+) on the real-ASU values of a symmetric field equals
    the full-grid FFT . exp(-|k|^2 ds) . IFFT at the real-ASU points."""
    N = (8, 8, 8)
    ops_s, spec, pts, M = _plan(sg, 3, N)
    lat = 4.0 * np.eye(3)
    ds = 0.05
    R, t = cf.ops_arrays(ops_s)
    u = O.make_symmetric(R, t, N, seed=sg)
    A = np.real(M @ (matrix_kernel(spec, N, ds, lat)[:, None] * np.linalg.pinv(M)))
    idx = np.array([p.idx for p in pts])
    Du = _fullgrid_diffusion(u, N, ds, lat)
    got = A @ u[idx[:, 0], idx[:, 1], idx[:, 2]]
    assert np.max(np.abs(got - Du[idx[:, 0], idx[:, 1], idx[:, 2]])) < 1e-12 * np.max(np.abs(Du))


@pytest.mark.gpu
@pytest.mark.parametrize("sg,N", [(6, (8, 8)), (47, (8, 8, 8)), (229, (16, 16, 16))])
def test_matrix_solver_on_device(sg, N):
    """plan_diffusion(...; method=:matrix) (the reference's default): the
    device operator against the host product A u, and step_diffusion! against
    e^{-w dt/2} A e^{-w dt/2} u, 1e-13 relative."""
    dim = len(N)
    lat = 2.0 * np.eye(dim)
    ds, dt = 0.05, 0.1
    solver = cf.plan_diffusion(N, lat, sg, dim, ds)
    assert isinstance(solver, cf.MatrixDiffusionSolver)
    n = len(solver.real_asu)
    A = np.real(solver.M @ (solver.Q[:, None] * solver.M_inv))
    rng = np.random.default_rng(sg)
    u = rng.random(n)
    w = rng.random(n)
    v = u.copy()
    cf.apply_diffusion_operator(solver, v)
    ref = A @ u
    assert np.max(np.abs(v - ref)) < 1e-13 * np.max(np.abs(ref))
    s = u.copy()
    cf.step_diffusion(solver, s, w, dt)
    e = np.exp(-w * dt / 2)
    ref = e * (A @ (e * u))
    assert np.max(np.abs(s - ref)) < 1e-13 * np.max(np.abs(ref))
    assert not np.array_equal(s, u)
