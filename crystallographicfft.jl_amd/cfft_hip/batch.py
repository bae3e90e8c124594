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


def start_fields(inst: Instance, ops_s, shift, amp=2.0, noise=0.1):
    """Symmetrised cosine-plus-noise start: w_A = +amp c + n_A, w_B = -amp c + n_B
    with c = cos 2 pi x + cos 2 pi y + cos 2 pi z (x in cell units), on the subgrid."""
    N = inst.N
    Lv = auto_L(ops_s)
    rng = np.random.default_rng(inst.seed)
    x = [(np.arange(n) + s * n) / n for n, s in zip(N, shift)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    c = np.cos(2 * np.pi * X) + np.cos(2 * np.pi * Y) + np.cos(2 * np.pi * Z)
    wA = symmetrize_subgrid(amp * c + noise * rng.standard_normal(N), ops_s, Lv)
    wB = symmetrize_subgrid(-amp * c + noise * rng.standard_normal(N), ops_s, Lv)
    return wA, wB


def _prepare(inst: Instance, precision, m_hist):
    ops = get_ops(inst.sg, 3, inst.N)
    shift, ops_s = find_optimal_shift(ops, inst.N)
    solver = SCFTSolver(inst.N, ops_s, inst.cell * np.eye(3), inst.f, inst.chiN, inst.ns, precision,
                        path=L.PATH_AUTO, m_hist=m_hist)
    wA, wB = start_fields(inst, ops_s, shift)
    return solver, (wA, wB)


def solve_batch(instances, precision=L.MIXED, tol=1e-6, max_iter=400, workers=8, m_hist=10):
    """Plan every instance, then solve all of them concurrently.  Returns
    (results, plan_seconds, solve_seconds); results[i] = dict(iters, F, res, Q)."""
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        prepared = list(ex.map(lambda inst: _prepare(inst, precision, m_hist), instances))
    t1 = time.perf_counter()

    def run(item):
        solver, (wA, wB) = item
        solver.set_fields(wA, wB)
        it, out = solver.solve(tol=tol, max_iter=max_iter)
        return dict(iters=it, F=out["F"], res=out["res"], Q=out["Q"])

    with ThreadPoolExecutor(max_workers=workers) as ex:
        results = list(ex.map(run, prepared))
    t2 = time.perf_counter()
    del prepared
    return results, t1 - t0, t2 - t1
