"""Many independent SCFT instances on one GPU.

Each instance owns a plan (its grid and space group), a ``cfft_scft`` context
and a stream; its contour sweeps are replayed from captured hipGraphs.  A pool
of host threads drives the instances (every library call releases the GIL), so
the instances' kernels run concurrently on their streams.  No counterpart in
the reference.
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .plans import auto_L, find_optimal_shift
from .scft import SCFTSolver, symmetrize_subgrid
from .symmetry_ops import get_ops


@dataclass
class Instance:
    """One AB diblock melt: grid, space group, cell, f, chiN, Ns, start fields."""
    N: tuple
    sg: int
    cell: float
    f: float
    chiN: float
    ns: int
    seed: int


def draw_instances(n, seed, sizes=(32, 48, 64, 96), groups=(47, 70, 223, 229, 230), f_range=(0.25, 0.45),
                   chiN_range=(12.0, 30.0), ns=100, cell=4.0):
    """n instances drawn from a seed: grid size and group uniformly from the
    lists, f and chiN uniform in their ranges."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        m = int(rng.choice(sizes))
        out.append(Instance((m, m, m), int(rng.choice(groups)), cell, float(rng.uniform(*f_range)),
                            float(rng.uniform(*chiN_range)), ns, int(rng.integers(1 << 30))))
    return out


# candidate first harmonics, by |h|^2: the start uses the first whose
# symmetrised cosine (or sine) does not vanish under the group
_STARS = ((1, 0, 0), (1, 1, 0), (1, 1, 1), (2, 0, 0), (2, 1, 0), (2, 1, 1), (2, 2, 0), (2, 2, 1), (3, 1, 0))


def lowest_star(N, ops_s, shift, Lv):
    """Symmetrised cos 2 pi h.x (or sin) for the smallest |h| whose star is not
    extinct in the group (x in cell units at (i + b)/N), normalised to unit RMS,
    on the subgrid, and h.  {100} is extinct in Pm-3n, Fddd, Im-3m and Ia-3d, so
    a fixed cos x + cos y + cos z start would symmetrise to zero there."""
    x = [(np.arange(n) + s * n) / n for n, s in zip(N, shift)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    for h in _STARS:
        arg = 2 * np.pi * (h[0] * X + h[1] * Y + h[2] * Z)
        for fn in (np.cos, np.sin):
            u = symmetrize_subgrid(fn(arg), ops_s, Lv)
            rms = float(np.sqrt(np.mean(u * u)))
            if rms > 1e-3:
                return u / rms, h
    raise ValueError("no harmonic up to |h|^2 = 10 survives the group average")


def start_fields(inst: Instance, ops_s, shift, amp=2.0, noise=0.1):
    """Seeded cosine-plus-noise start: w_A = +amp c + n_A, w_B = -amp c + n_B,
    c = the group's lowest non-extinct star (unit RMS), n_A, n_B symmetrised
    seeded N(0, noise) fields; on the subgrid."""
    N = inst.N
    Lv = auto_L(ops_s)
    rng = np.random.default_rng(inst.seed)
    c, _ = lowest_star(N, ops_s, shift, Lv)
    wA = amp * c + symmetrize_subgrid(noise * rng.standard_normal(N), ops_s, Lv)
    wB = -amp * c + symmetrize_subgrid(noise * rng.standard_normal(N), ops_s, Lv)
    return wA, wB


def _prepare(inst: Instance, precision, m_hist, finish):
    ops = get_ops(inst.sg, 3, inst.N)
    shift, ops_s = find_optimal_shift(ops, inst.N)
    mk = lambda prec: SCFTSolver(inst.N, ops_s, inst.cell * np.eye(3), inst.f, inst.chiN, inst.ns, prec,
                                 path=None, m_hist=m_hist)
    solver = mk(precision)
    fin = mk(L.FP64) if finish else None
    return solver, fin, start_fields(inst, ops_s, shift)


def solve_batch(instances, precision=L.MIXED, tol=1e-6, max_iter=400, workers=8, m_hist=10,
                finish_tol=None):
    """Plan every instance, then solve all of them concurrently.  With
    finish_tol, each instance continues in FP64 from its `precision` solution
    until the residual is below finish_tol (its FP64 plan is made in the
    planning phase).  Returns (results, plan_seconds, solve_seconds);
    results[i] = dict(iters, F, res, Q[, iters_finish])."""
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        prepared = list(ex.map(lambda inst: _prepare(inst, precision, m_hist, finish_tol is not None),
                               instances))
    t1 = time.perf_counter()

    def run(item):
        solver, fin, (wA, wB) = item
        solver.set_fields(wA, wB)
        it, out = solver.solve(tol=tol, max_iter=max_iter)
        r = dict(iters=it, F=out["F"], res=out["res"], Q=out["Q"])
        if fin is not None:
            fin.set_fields(*solver.fields())
            it2, o2 = fin.solve(tol=finish_tol, max_iter=max_iter)
            r.update(iters_finish=it2, F=o2["F"], res=o2["res"], Q=o2["Q"], F_first=out["F"],
                     res_first=out["res"])
        return r

    with ThreadPoolExecutor(max_workers=workers) as ex:
        results = list(ex.map(run, prepared))
    t2 = time.perf_counter()
    del prepared
    return results, t1 - t0, t2 - t1
