"""Diffusion solver interface (src/diffusion_solver.jl).

``plan_diffusion(N, lattice, sg_num, dim, ds; method)`` returns a solver whose
``apply_diffusion_operator`` applies D = IFFT . exp(-|k|^2 ds) . FFT to a
stride-L subgrid field and whose ``step_diffusion`` applies the Strang step
u <- e^{-w dt/2} D e^{-w dt/2} u, both on the GPU.

Methods: ``"qfused"`` (the reference's :qfused, M2 Q-fused multiply,
src/diffusion_solver.jl:245-248), ``"m2"`` (fwd + K + bwd, plan_m2_scft),
``"centered"`` (centring-fold fwd + K + bwd, plan_centered_scft) and
``"auto"`` (centered when the lattice is centred, else m2).  The reference's
dense ``:matrix`` oracle and its legacy ``:krfft`` block solver
(src/diffusion_solver.jl:20-160) are not part of the GPU path; requesting them
raises the reference's "Unknown method" error.

Quirk kept from the reference: fields are subgrid arrays (M = N / L); the
reference's ``interaction_step!`` has no 3-D method (src/diffusion_solver.jl:
270-282), so 3-D callers pass flattened vectors — here 3-D arrays are accepted
directly.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .plans import DevicePlan, find_optimal_shift
from .symmetry_ops import get_ops

_METHODS = {"qfused": L.PATH_M2Q, "m2": L.PATH_M2, "centered": L.PATH_M7, "auto": L.PATH_AUTO}


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


def plan_diffusion(N, lattice, sg_num, dim, ds, method="qfused", precision=L.FP64, batch=1):
    """src/diffusion_solver.jl:165-252."""
    method = str(method).lstrip(":")
    if method not in _METHODS:
        raise L.CFFTError(L.CFFT_ERR_UNKNOWN_METHOD,
                          f"Unknown method: {method}. Supported: :qfused, :m2, :centered, :auto")
    ops = get_ops(sg_num, dim, N)
    shift, ops_s = find_optimal_shift(ops, N)
    dev = DevicePlan(N, ops_s, ds, lattice, _METHODS[method], precision, batch)
    return DiffusionSolver(dev, method, shift)


def apply_diffusion_operator(solver: DiffusionSolver, u):
    """src/diffusion_solver.jl:87-91 (in place on the subgrid field)."""
    a = np.asarray(u)
    out = solver.plan.diffuse(solver._as_sub(a))
    a[...] = out.reshape(a.shape, order="F") if a.ndim == 1 else out
    return u


def step_diffusion(solver: DiffusionSolver, u, w, dt):
    """src/diffusion_solver.jl:254-267 (in place)."""
    a = np.asarray(u)
    out = solver.plan.step(solver._as_sub(a), solver._as_sub(w), dt, 1)
    a[...] = out.reshape(a.shape, order="F") if a.ndim == 1 else out
    return u


def interaction_step(u, w, factor):
    """src/diffusion_solver.jl:270-282: u *= exp(-w factor) (host arrays)."""
    u *= np.exp(-np.asarray(w) * factor)
    return u
