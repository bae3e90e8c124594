"""Device-resident SCFT propagator sweeps on the stride-L subgrid.

No counterpart in the reference (it stops at the diffusion step,
src/diffusion_solver.jl:254-282).  An AB diblock propagator sweep advances
q (from the A end) and q+ (from the B end) together as a batch of two fields:

  q(s + ds) = e^{-w(s) ds/2} D e^{-w(s) ds/2} q(s),   ds = 1/Ns,

with w = w_A on the A block (the first round(f Ns) steps of q, the last of q+)
and w_B elsewhere.  Q = <q(1)> = <q+(1)> (grid mean; on the subgrid because the
full grid is the union of prod(L) images of it).

Device memory comes from torch (plumbing only); every kernel is the library's.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .plans import DevicePlan


def first_star_bcc(N, shift_grid, cell):
    """sum of cos 2 pi (x +- y)/a, (y +- z)/a, (z +- x)/a at x = (i + b) a/N."""
    x = [(np.arange(n) + b) / n for n, b in zip(N, shift_grid)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    tp = 2.0 * np.pi
    return (np.cos(tp * (X + Y)) + np.cos(tp * (X - Y)) + np.cos(tp * (Y + Z)) +
            np.cos(tp * (Y - Z)) + np.cos(tp * (Z + X)) + np.cos(tp * (Z - X)))


def gyroid_first_harmonic(N, shift_grid, cell):
    """sin x cos y + sin y cos z + sin z cos x, x in units of 2 pi / a."""
    x = [(np.arange(n) + b) / n for n, b in zip(N, shift_grid)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    tp = 2.0 * np.pi
    return (np.sin(tp * X) * np.cos(tp * Y) + np.sin(tp * Y) * np.cos(tp * Z) +
            np.sin(tp * Z) * np.cos(tp * X))


def symmetrize_subgrid(u, ops_shifted, Lv):
    """Group average of a full-grid field, evaluated on the stride-L subgrid:
    u_sym(x) = (1/|G|) sum_g u(R_g x + t_g) at x = L n (gather form of the
    scatter average of test/test_helpers.jl:11-29)."""
    N = u.shape
    M = tuple(n // l for n, l in zip(N, Lv))
    g = np.meshgrid(*[np.arange(m) * l for m, l in zip(M, Lv)], indexing="ij")
    X = np.stack([a.ravel(order="F") for a in g], axis=1)
    acc = np.zeros(X.shape[0])
    for op in ops_shifted:
        y = np.mod(X @ op.Rm.T + op.tv, np.asarray(N))
        acc += u[y[:, 0], y[:, 1], y[:, 2]]
    return (acc / len(ops_shifted)).reshape(M, order="F")


class PropagatorSweep:
    """q and q+ of one AB diblock melt, batched on one plan."""

    def __init__(self, N, ops_shifted, lattice, f, ns, precision=L.FP64, path=L.PATH_AUTO):
        import torch
        self.torch = torch
        self.ns = int(ns)
        self.f = float(f)
        self.nA = int(round(f * ns))
        self.ds = 1.0 / ns
        self.plan = DevicePlan(N, ops_shifted, self.ds, lattice, path, precision, batch=2)
        self.M = self.plan.M
        self.M3 = int(np.prod(self.M))
        self.dtype = torch.float64 if precision == L.FP64 else torch.float32
        self.dev = torch.device("cuda")
        self.q = torch.ones(2, self.M3, dtype=self.dtype, device=self.dev)
        self.EA = torch.empty(self.M3, dtype=self.dtype, device=self.dev)
        self.EB = torch.empty(self.M3, dtype=self.dtype, device=self.dev)
        self.w = torch.empty(2, self.M3, dtype=self.dtype, device=self.dev)

    def stream(self):
        return self.torch.cuda.current_stream().cuda_stream

    def set_fields(self, wA, wB):
        """wA, wB: subgrid arrays (M0, M1, M2) in x-fastest order."""
        t = self.torch
        self.w[0] = t.from_numpy(np.asarray(wA, dtype=np.float64).ravel(order="F")).to(self.dev, self.dtype)
        self.w[1] = t.from_numpy(np.asarray(wB, dtype=np.float64).ravel(order="F")).to(self.dev, self.dtype)
        s = self.stream()
        self.plan.exp_dev(self.w[0].data_ptr(), self.EA.data_ptr(), -self.ds / 2.0, self.M3, s)
        self.plan.exp_dev(self.w[1].data_ptr(), self.EB.data_ptr(), -self.ds / 2.0, self.M3, s)
        # q starts on the A block (E_A for steps < nA), q+ on the B block
        # (E_B for steps < ns - nA); cfft_sweep_dev switches per entry.
        pa, pb = self.EA.data_ptr(), self.EB.data_ptr()
        self._E1 = (C.c_void_p * 2)(pa, pb)
        self._E2 = (C.c_void_p * 2)(pb, pa)
        self._sw = (C.c_int32 * 2)(self.nA, self.ns - self.nA)

    def reset(self):
        self.q.fill_(1.0)

    def run(self, steps=None, start=0):
        """Advance q and q+ through `steps` contour steps (default Ns) starting
        at contour step `start`, in one library call."""
        n = self.ns - start if steps is None else steps
        L.check(L.lib().cfft_sweep_dev(self.plan.handle, C.c_void_p(self.q.data_ptr()), self._E1,
                                       self._E2, self._sw, int(start), int(n),
                                       C.c_void_p(self.stream())))

    def Q(self):
        """(Q from q, Q from q+), reduced in FP64."""
        qq = self.q.to(self.torch.float64).mean(dim=1)
        return float(qq[0]), float(qq[1])

    def fields(self):
        out = self.q.to(self.torch.float64).cpu().numpy()
        return [out[b].reshape(self.M, order="F") for b in range(2)]
