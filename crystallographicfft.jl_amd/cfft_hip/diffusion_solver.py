"""Diffusion solver interface (src/diffusion_solver.jl).

``plan_diffusion(N, lattice, sg_num, dim, ds; method)`` returns a solver whose
``apply_diffusion_operator`` applies D = IFFT . exp(-|k|^2 ds) . FFT to a
stride-L subgrid field and whose ``step_diffusion`` applies the Strang step
u <- e^{-w dt/2} D e^{-w dt/2} u, both on the GPU.

Methods: ``"matrix"`` (the reference's default: the dense real-ASU operator
A = Re(M Q M^+) of src/diffusion_solver.jl:20-71, planned on the host, applied
on the GPU by ``cfft_dense_*``; fields are vectors over the real-ASU points),
``"qfused"`` (the reference's :qfused, M2 Q-fused multiply,
src/diffusion_solver.jl:245-248), ``"m2"`` (fwd + K + bwd, plan_m2_scft),
``"centered"`` (centring-fold fwd + K + bwd, plan_centered_scft) and
``"auto"`` (centered when the lattice is centred, else m2).  The legacy
``:krfft`` block solver (src/diffusion_solver.jl:93-160) is not built;
requesting it raises the reference's "Unknown method" error.

Quirk kept from the reference: fields are subgrid arrays (M = N / L); the
reference's ``interaction_step!`` has no 3-D method (src/diffusion_solver.jl:
270-282), so 3-D callers pass flattened vectors — here 3-D arrays are accepted
directly.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .plans import DevicePlan, find_optimal_shift, get_k_vector
from .symmetry_ops import get_ops

_METHODS = {"qfused": L.PATH_M2Q, "m2": L.PATH_M2, "centered": L.PATH_M7, "auto": L.PATH_AUTO}


def calc_transform_matrix(real_pts, spec, direct_ops, N, shift):
    """The real-ASU x spectral-ASU transform matrix of src/diffusion_solver.jl:
    29-60: M[i, j] = (1/|G|) sum_g exp(2 pi i (k'.r_i - k'.t_g)/N) times the
    column phase exp(2 pi i k_j.s), with k' = R*_g k_j (the spectral ASU's
    reciprocal op g on the signed frequency of point j), r_i the real-ASU grid
    point and t_g the translation of the magic-shifted op g (grid units), s the
    magic shift (fraction of the cell).  The shifted ops already carry the
    shift (t' = t + (R - 1) s N), so s enters only through the column phase,
    which places the coefficients at the physical positions (r + s N)/N and
    cancels in M diag(K) M^+; the reference's formula adds k'.s inside the sum
    as well, which counts the shift twice: for a group with a nonzero magic
    shift its A = Re(M diag(K) M^+) then differs from the full-grid operator
    restricted to the real ASU, while this form matches it to round-off."""
    Nv = np.asarray(N, dtype=np.float64)
    D = len(N)
    k = np.array([get_k_vector(spec, j) for j in range(len(spec))], dtype=np.float64)  # (n_spec, D)
    r = np.array([np.asarray(p.idx, dtype=np.float64)[:D] for p in real_pts])          # (n_real, D)
    s = np.asarray(shift, dtype=np.float64)[:D]
    M = np.zeros((r.shape[0], k.shape[0]), dtype=np.complex128)
    for rop, op in zip(spec.ops, direct_ops):
        kp = k @ np.asarray(rop.Rm, dtype=np.float64).T                                 # R* k, (n_spec, D)
        ph_t = -(kp * (np.asarray(op.tv, dtype=np.float64) / Nv)).sum(axis=1)           # (n_spec,)
        M += np.exp(2j * np.pi * ((r / Nv) @ kp.T + ph_t[None, :]))
    return M * (np.exp(2j * np.pi * (k @ s)) / len(direct_ops))[None, :]


def matrix_kernel(spec, N, ds, lattice):
    """src/matrix_q.jl:18-79 (calc_gradient_term, calc_matrix_q): the diagonal
    of Q, K_j = exp(-|2 pi A^-T k_j|^2 ds) at the signed frequencies."""
    B = 2.0 * np.pi * np.linalg.inv(np.asarray(lattice, dtype=np.float64)).T
    k = np.array([get_k_vector(spec, j) for j in range(len(spec))], dtype=np.float64)
    kv = k @ B.T
    return np.exp(-(kv * kv).sum(axis=1) * ds)


class MatrixDiffusionSolver:
    """src/diffusion_solver.jl:20-28: the real-ASU points, the spectral ASU,
    the diagonal of Q, the transform matrix M and its pseudo-inverse, and the
    device operator A = Re(M Q M^+) (fields are real vectors over the
    real-ASU points, in the order of ``real_asu``)."""

    def __init__(self, real_asu, spec_asu, Q, M, M_inv):
        import ctypes as C
        self.real_asu = real_asu
        self.spec_asu = spec_asu
        self.Q = Q
        self.M = M
        self.M_inv = M_inv
        A = np.ascontiguousarray(np.real(M @ (Q[:, None] * M_inv)))
        self.n = A.shape[0]
        h = C.c_void_p()
        L.check(L.lib().cfft_dense_create(self.n, A.ctypes.data_as(L.PD), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            L.lib().cfft_dense_destroy(h)
            self._h = None

    def _vec(self, u):
        a = np.ascontiguousarray(np.asarray(u, dtype=np.float64).ravel())
        if a.size != self.n:
            raise ValueError(f"field has {a.size} entries, the real ASU has {self.n}")
        return a


def _plan_matrix(N, lattice, sg_num, dim, ds):
    """src/diffusion_solver.jl:165-180 (method = :matrix)."""
    from .asu import calc_asu
    from .plans import calc_spectral_asu
    ops = get_ops(sg_num, dim, N)
    shift, ops_s = find_optimal_shift(ops, N)
    spec = calc_spectral_asu(ops_s, dim, N)
    Q = matrix_kernel(spec, N, ds, lattice)
    pts = calc_asu(tuple(N), ops_s)
    M = calc_transform_matrix(pts, spec, ops_s, N, shift)
    return MatrixDiffusionSolver(pts, spec, Q, M, np.linalg.pinv(M))


class DiffusionSolver:
    """A device plan plus the subgrid shape it acts on."""

    def __init__(self, dev: DevicePlan, method: str, shift):
        self.plan = dev
        self.method = method
        self.shift = shift
        self.M = dev.M
        self.L = dev.L

    def _as_sub(self, u):
        a = np.asarray(u, dtype=np.float64)
        if a.size != int(np.prod(self.M)):
            raise ValueError(f"field has {a.size} entries, subgrid has {int(np.prod(self.M))}")
        return a.reshape(self.M, order="F") if a.ndim == 1 else a


QFusedDiffusionSolver = DiffusionSolver


def plan_diffusion(N, lattice, sg_num, dim, ds, method="matrix", precision=L.FP64, batch=1):
    """src/diffusion_solver.jl:165-252 (default method :matrix, as there)."""
    method = str(method).lstrip(":")
    if method == "matrix":
        return _plan_matrix(tuple(N), lattice, sg_num, dim, ds)
    if method not in _METHODS:
        raise L.CFFTError(L.CFFT_ERR_UNKNOWN_METHOD,
                          f"Unknown method: {method}. Supported: :matrix, :qfused, :m2, :centered, :auto")
    ops = get_ops(sg_num, dim, N)
    shift, ops_s = find_optimal_shift(ops, N)
    dev = DevicePlan(N, ops_s, ds, lattice, _METHODS[method], precision, batch)
    return DiffusionSolver(dev, method, shift)


def apply_diffusion_operator(solver, u):
    """src/diffusion_solver.jl:87-91 (in place on the subgrid field), or
    :66-71 for the matrix solver (in place on the real-ASU vector)."""
    if isinstance(solver, MatrixDiffusionSolver):
        a = solver._vec(u)
        L.check(L.lib().cfft_dense_apply(solver._h, a.ctypes.data_as(L.PD)))
        u[...] = a.reshape(np.shape(u))
        return u
    a = np.asarray(u)
    out = solver.plan.diffuse(solver._as_sub(a))
    a[...] = out.reshape(a.shape, order="F") if a.ndim == 1 else out
    return u


def step_diffusion(solver, u, w, dt):
    """src/diffusion_solver.jl:254-267 (in place)."""
    if isinstance(solver, MatrixDiffusionSolver):
        a = solver._vec(u)
        wv = solver._vec(w)
        L.check(L.lib().cfft_dense_step(solver._h, a.ctypes.data_as(L.PD), wv.ctypes.data_as(L.PD), float(dt)))
        u[...] = a.reshape(np.shape(u))
        return u
    a = np.asarray(u)
    out = solver.plan.step(solver._as_sub(a), solver._as_sub(w), dt, 1)
    a[...] = out.reshape(a.shape, order="F") if a.ndim == 1 else out
    return u


def interaction_step(u, w, factor):
    """src/diffusion_solver.jl:270-282: u *= exp(-w factor) (host arrays)."""
    u *= np.exp(-np.asarray(w) * factor)
    return u
