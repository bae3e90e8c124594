"""Device-resident SCFT propagator sweeps on the stride-L subgrid.

No counterpart in the reference (it stops at the diffusion step,
src/diffusion_solver.jl:254-282).  An AB diblock propagator sweep advances
q (from the A end) and q+ (from the B end) together as a batch of two fields:

  q(s + ds) = e^{-w(s) ds/2} D e^{-w(s) ds/2} q(s),   ds = 1/Ns,

with w = w_A on the A block (the first round(f Ns) steps of q, the last of q+)
and w_B elsewhere.  Q = <q(1)> = <q+(1)> (grid mean; on the subgrid because the
full grid is the union of prod(L) images of it).

Device memory comes from torch (plumbing only); every kernel is the library's.
``SCFTSolver`` keeps the whole iteration (sweep, densities, free energy, field
update, Anderson mixing) inside the library's ``cfft_scft`` context.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .plans import DevicePlan, qs_eligible


def first_star_bcc(N, shift_grid, cell):
    """sum of cos 2 pi (x +- y)/a, (y +- z)/a, (z +- x)/a at x = (i + b) a/N."""
    x = [(np.arange(n) + b) / n for n, b in zip(N, shift_grid)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    tp = 2.0 * np.pi
    return (np.cos(tp * (X + Y)) + np.cos(tp * (X - Y)) + np.cos(tp * (Y + Z)) +
            np.cos(tp * (Y - Z)) + np.cos(tp * (Z + X)) + np.cos(tp * (Z - X)))


def double_gyroid_211(N, shift_grid, cell):
    """sin 2x cos y sin z + sin 2y cos z sin x + sin 2z cos x sin y (x in units
    of 2 pi / a): the {211} star, the first harmonic of the double gyroid that
    Ia-3d does not extinguish (the {110} terms of the single-gyroid nodal
    function sin x cos y + ... are extinct in Ia-3d and symmetrise to zero).
    It is the {211} part of the square of that nodal function, so it is large
    on both networks."""
    x = [(np.arange(n) + b) / n for n, b in zip(N, shift_grid)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    tp = 2.0 * np.pi
    return (np.sin(2 * tp * X) * np.cos(tp * Y) * np.sin(tp * Z) +
            np.sin(2 * tp * Y) * np.cos(tp * Z) * np.sin(tp * X) +
            np.sin(2 * tp * Z) * np.cos(tp * X) * np.sin(tp * Y))


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
        # interaction factors: FP64 in both precisions
        self.EA = torch.empty(self.M3, dtype=torch.float64, device=self.dev)
        self.EB = torch.empty(self.M3, dtype=torch.float64, device=self.dev)
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


# Residual the plane-pair chain reaches in mixed precision with margin (its
# floor is ~4e-6 at Ia-3d 128^3, Ns 200)
QS_MIXED_TOL = 1e-5


def default_path(N, ops_shifted, lattice, precision, tol=None):
    """FP64: the plane-pair chain where it applies.  Mixed precision keeps the
    default path unless the solve stops at a residual tol >= QS_MIXED_TOL: on
    the default path every step passes through the spectral ASU, which
    projects FP32 rounding back onto the symmetric fields, while on the
    plane-pair chain the asymmetric part of the rounding accumulates over a
    sweep (residual floor ~4e-6 instead of ~7e-7 at Ia-3d 128^3, Ns 200)."""
    if not qs_eligible(N, ops_shifted, lattice):
        return L.PATH_AUTO
    if precision == L.FP64 or (tol is not None and tol >= QS_MIXED_TOL):
        return L.PATH_QS
    return L.PATH_AUTO


class SCFTSolver:
    """Device SCFT of an AB diblock melt on the stride-L subgrid (C context
    ``cfft_scft``): sweep (q, q+ batched, every contour point kept), Simpson
    densities, free energy, field update and Anderson mixing, FP64 except the
    propagator arithmetic in mixed precision.

    ``ops_shifted`` are the magic-shifted ops; fields are subgrid arrays of
    shape M (x fastest).  ``iterate()`` returns a dict with Q, Q+ (Qd), F, the
    residual, <phiA + phiB>, <phiA>, and whether Anderson mixing was used.
    """

    KEYS = ("Q", "Qd", "F", "res", "phi_total", "phiA_mean", "anderson", "hist")

    def __init__(self, N, ops_shifted, lattice, f, chiN, ns, precision=L.FP64, path=None,
                 m_hist=10, axis_split=False):
        self.ns = int(ns)
        self.f = float(f)
        self.chiN = float(chiN)
        path = default_path(N, ops_shifted, lattice, precision) if path is None else path
        self.plan = DevicePlan(N, ops_shifted, 1.0 / self.ns, lattice, path, precision, batch=2)
        if axis_split:  # the per-axis chain of path QS (CFFT_OPT_AXIS_SPLIT)
            self.plan.set_option(L.OPT_AXIS_SPLIT, 1)
        self.M = self.plan.M
        self.M3 = int(np.prod(self.M))
        h = C.c_void_p()
        L.check(L.lib().cfft_scft_create(self.plan.handle, self.f, self.chiN, self.ns, int(m_hist),
                                         C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.lib().cfft_scft_destroy(h)
            except Exception:
                pass
            self._h = None

    def update(self, lattice=None, f=None, chiN=None):
        """Reuse the plan and context for another cell (kernel update,
        src/centering_fold.jl:939-953) and/or another f, chiN."""
        if lattice is not None:
            self.plan.update_kernel(1.0 / self.ns, lattice)
        if f is not None or chiN is not None:
            self.f = self.f if f is None else float(f)
            self.chiN = self.chiN if chiN is None else float(chiN)
            L.check(L.lib().cfft_scft_set_params(self._h, self.f, self.chiN))

    def set_mixing(self, lam=1.0, lam_simple=0.1, simple_above=1e-2):
        L.check(L.lib().cfft_scft_set_mixing(self._h, float(lam), float(lam_simple), float(simple_above)))

    def continue_from(self, other):
        """Take over another solver's fields and Anderson history (same
        subgrid, instance count and history length; cfft_scft_copy_state):
        a mixed-precision stage handed to an FP64 finish keeps mixing
        instead of restarting from simple mixing."""
        L.check(L.lib().cfft_scft_copy_state(self._h, other._h))

    def set_fields(self, wA, wB):
        a = np.ascontiguousarray(np.asarray(wA, dtype=np.float64).ravel(order="F"))
        b = np.ascontiguousarray(np.asarray(wB, dtype=np.float64).ravel(order="F"))
        if a.size != self.M3 or b.size != self.M3:
            raise L.CFFTError(L.CFFT_ERR_INVALID, "fields must have the subgrid size")
        L.check(L.lib().cfft_scft_set_fields(self._h, L.f64p(a), L.f64p(b)))

    def _out(self, o):
        return {k: float(v) for k, v in zip(self.KEYS, o)}

    def iterate(self):
        o = np.zeros(8)
        L.check(L.lib().cfft_scft_iterate(self._h, L.f64p(o)))
        return self._out(o)

    def solve(self, tol=1e-8, max_iter=1000):
        o = np.zeros(8)
        it = C.c_int32()
        L.check(L.lib().cfft_scft_solve(self._h, float(tol), int(max_iter), C.byref(it), L.f64p(o)))
        return int(it.value), self._out(o)

    def _get(self, which):
        out = np.zeros(2 * self.M3)
        L.check(L.lib().cfft_scft_get(self._h, which, L.f64p(out)))
        return [out[i * self.M3:(i + 1) * self.M3].reshape(self.M, order="F") for i in range(2)]

    def fields(self):
        """[wA, wB] (the fields the next iteration will use)."""
        return self._get(L.SCFT_W)

    def densities(self):
        """[phiA, phiB] of the last iteration."""
        return self._get(L.SCFT_PHI)

    def deviations(self):
        return self._get(L.SCFT_DEV)


class SCFTBatch:
    """Several melts that share grid, space group, cell and Ns in one context
    (``cfft_scft_create_multi``): one plan with 2 n fields, one sweep per
    iteration advances q, q+ of every instance, and each instance keeps its
    own f, chiN, fields and Anderson history.  Per instance the iterations are
    those of an ``SCFTSolver`` bit for bit; ``solve`` stops each instance at
    its own convergence.  Outputs are lists of ``SCFTSolver``-style dicts."""

    KEYS = SCFTSolver.KEYS

    def __init__(self, N, ops_shifted, lattice, fs, chiNs, ns, precision=L.FP64, path=None, m_hist=10):
        self.ns = int(ns)
        self.fs = [float(v) for v in fs]
        self.chiNs = [float(v) for v in chiNs]
        if len(self.fs) != len(self.chiNs) or not self.fs:
            raise L.CFFTError(L.CFFT_ERR_INVALID, "need one f and one chiN per instance")
        self.n = len(self.fs)
        path = default_path(N, ops_shifted, lattice, precision) if path is None else path
        self.plan = DevicePlan(N, ops_shifted, 1.0 / self.ns, lattice, path, precision, batch=2 * self.n)
        self.M = self.plan.M
        self.M3 = int(np.prod(self.M))
        f = np.array(self.fs, dtype=np.float64)
        x = np.array(self.chiNs, dtype=np.float64)
        h = C.c_void_p()
        L.check(L.lib().cfft_scft_create_multi(self.plan.handle, self.n, L.f64p(f), L.f64p(x), self.ns,
                                               int(m_hist), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.lib().cfft_scft_destroy(h)
            except Exception:
                pass
            self._h = None

    def set_schedule(self, options):
        """Plan options (e.g. the best of ``tune.tune_schedule``); results do
        not depend on them."""
        for k, v in options.items():
            self.plan.set_option(int(k), int(v))

    def set_params(self, i, f, chiN):
        L.check(L.lib().cfft_scft_set_params_inst(self._h, int(i), float(f), float(chiN)))
        self.fs[i], self.chiNs[i] = float(f), float(chiN)

    def set_mixing(self, lam=1.0, lam_simple=0.1, simple_above=1e-2):
        L.check(L.lib().cfft_scft_set_mixing(self._h, float(lam), float(lam_simple), float(simple_above)))

    def continue_from(self, other):
        """Take over another solver's fields and Anderson history (same
        subgrid, instance count and history length; cfft_scft_copy_state):
        a mixed-precision stage handed to an FP64 finish keeps mixing
        instead of restarting from simple mixing."""
        L.check(L.lib().cfft_scft_copy_state(self._h, other._h))

    def set_fields(self, i, wA, wB):
        a = np.ascontiguousarray(np.asarray(wA, dtype=np.float64).ravel(order="F"))
        b = np.ascontiguousarray(np.asarray(wB, dtype=np.float64).ravel(order="F"))
        if a.size != self.M3 or b.size != self.M3:
            raise L.CFFTError(L.CFFT_ERR_INVALID, "fields must have the subgrid size")
        L.check(L.lib().cfft_scft_set_fields_inst(self._h, int(i), L.f64p(a), L.f64p(b)))

    def _outs(self, o):
        return [{k: float(v) for k, v in zip(self.KEYS, o[8 * i:8 * i + 8])} for i in range(self.n)]

    def iterate(self):
        o = np.zeros(8 * self.n)
        L.check(L.lib().cfft_scft_iterate_multi(self._h, L.f64p(o)))
        return self._outs(o)

    def solve(self, tol=1e-8, max_iter=1000):
        """(iterations per instance, result dict per instance)."""
        o = np.zeros(8 * self.n)
        it = np.zeros(self.n, dtype=np.int32)
        L.check(L.lib().cfft_scft_solve_multi(self._h, float(tol), int(max_iter), L.i32p(it), L.f64p(o)))
        return [int(v) for v in it], self._outs(o)

    def _get(self, i, which):
        out = np.zeros(2 * self.M3)
        L.check(L.lib().cfft_scft_get_inst(self._h, int(i), which, L.f64p(out)))
        return [out[k * self.M3:(k + 1) * self.M3].reshape(self.M, order="F") for k in range(2)]

    def fields(self, i):
        return self._get(i, L.SCFT_W)

    def densities(self, i):
        return self._get(i, L.SCFT_PHI)


class SCFTEnsemble:
    """Several ``SCFTBatch`` contexts (any grids, space groups and instance
    counts; one precision and Ns; path QS with channel grids H in {8, 12, 16,
    24, 32}) advanced together through ``cfft_scft_ensemble_*``: per
    iteration one launch per chain stage and contour step covers the plane
    pairs (or tiles) of every context, from a work list, and the sweep is
    replayed from one captured hipGraph.  Per instance the results are those
    of its own context bit for bit.  The batches must outlive the ensemble."""

    def __init__(self, batches):
        self.batches = list(batches)
        if not self.batches:
            raise L.CFFTError(L.CFFT_ERR_INVALID, "need at least one context")
        arr = (C.c_void_p * len(self.batches))(*[b._h.value for b in self.batches])
        h = C.c_void_p()
        L.check(L.lib().cfft_scft_ensemble_create(arr, len(self.batches), C.byref(h)))
        self._h = h
        self.n = sum(b.n for b in self.batches)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.lib().cfft_scft_ensemble_destroy(h)
            except Exception:
                pass
            self._h = None

    def _split(self, o, it=None):
        out, k = [], 0
        for b in self.batches:
            outs = b._outs(o[8 * k:8 * (k + b.n)])
            out.append(outs if it is None else ([int(v) for v in it[k:k + b.n]], outs))
            k += b.n
        return out

    def iterate(self):
        """One iteration of every instance: per context, a list of result dicts."""
        o = np.zeros(8 * self.n)
        L.check(L.lib().cfft_scft_ensemble_iterate(self._h, L.f64p(o)))
        return self._split(o)

    def solve(self, tol=1e-8, max_iter=1000):
        """Per context, (iterations per instance, result dict per instance)."""
        o = np.zeros(8 * self.n)
        it = np.zeros(self.n, dtype=np.int32)
        L.check(L.lib().cfft_scft_ensemble_solve(self._h, float(tol), int(max_iter), L.i32p(it), L.f64p(o)))
        return self._split(o, it)
