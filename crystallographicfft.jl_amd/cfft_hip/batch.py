"""Many independent SCFT instances on one GPU.

Instances that share grid, space group, cell and Ns form a group: one plan
with 2 n fields and one ``cfft_scft`` context (``SCFTBatch``), so one sweep of
two kernels per contour step advances all of them, and per instance the
iterations are those of a one-instance solver bit for bit.  Each group has its
own stream, its contour sweeps are replayed from captured hipGraphs, and a
pool of host threads drives the groups (every library call releases the GIL),
so the groups' kernels run concurrently.  ``group=False`` gives every instance
its own plan, context and stream instead.  ``ensemble=True`` advances the
groups together instead (``SCFTEnsemble``): one launch per chain stage and
contour step covers every group's plane pairs or tiles, the sweep replayed
from one hipGraph, each instance's results unchanged bit for bit.  No
counterpart in the reference.
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .plans import auto_L, find_optimal_shift
from .scft import SCFTBatch, SCFTEnsemble, SCFTSolver, default_path
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


def star_wave(h, fn, N, ops_s, shift, Lv):
    """Group average of fn(2 pi h.(X + b)/N) (fn = np.cos or np.sin, b the
    magic shift in grid units) at the stride-L subgrid points X = L x:
    (1/|G|) sum_g fn(2 pi (kappa_g . X + c_g)) with kappa_g = R_g^T (h / N)
    and c_g = h.(t_g + b)/N, summed over the distinct kappa_g (the star of h)
    as products of 1-D exponentials.  Equals symmetrize_subgrid of the full-grid
    wave (the gather form of test/test_helpers.jl:11-29) without building it."""
    N = np.asarray(N, dtype=np.int64)
    Lv = np.asarray(Lv, dtype=np.int64)
    M = N // Lv
    b = np.asarray(shift, dtype=np.float64) * N
    hv = np.asarray(h, dtype=np.float64)
    terms = {}
    for op in ops_s:
        kap = (hv / N) @ np.asarray(op.Rm, dtype=np.float64)
        c = float(np.sum(hv * (np.asarray(op.tv, dtype=np.float64) + b) / N))
        key = tuple(np.round(kap * N, 9))
        a = np.exp(2j * np.pi * (c - np.floor(c)))
        k0, a0 = terms.get(key, (kap, 0.0))
        terms[key] = (k0, a0 + a)
    acc = np.zeros(tuple(M), dtype=np.complex128)
    for kap, a in terms.values():
        if abs(a) < 1e-13 * len(ops_s):
            continue
        e = [np.exp(2j * np.pi * kap[d] * Lv[d] * np.arange(M[d])) for d in range(3)]
        acc += a * e[0][:, None, None] * e[1][None, :, None] * e[2][None, None, :]
    acc /= len(ops_s)
    return acc.real if fn is np.cos else acc.imag


def lowest_star(N, ops_s, shift, Lv):
    """Symmetrised cos 2 pi h.x (or sin) for the smallest |h| whose star is not
    extinct in the group (x in cell units at (i + b)/N), normalised to unit RMS,
    on the subgrid, and h.  {100} is extinct in Pm-3n, Fddd, Im-3m and Ia-3d, so
    a fixed cos x + cos y + cos z start would symmetrise to zero there."""
    for h in _STARS:
        for fn in (np.cos, np.sin):
            u = star_wave(h, fn, N, ops_s, shift, Lv)
            rms = float(np.sqrt(np.mean(u * u)))
            if rms > 1e-3:
                return u / rms, h
    raise ValueError("no harmonic up to |h|^2 = 10 survives the group average")


def subgrid_maps(N, ops_s, Lv):
    """Index maps of the stride-L subgrid (x fastest) under the ops that map
    it onto itself, one per distinct map: x' = R (L x) + t over L, mod M."""
    N = np.asarray(N, dtype=np.int64)
    Lv = np.asarray(Lv, dtype=np.int64)
    M = N // Lv
    ax = [np.arange(m, dtype=np.int64) * l for m, l in zip(M, Lv)]
    grid = (ax[0][:, None, None], ax[1][None, :, None], ax[2][None, None, :])
    maps = []
    seen = set()
    for op in ops_s:
        R = np.asarray(op.Rm, dtype=np.int64)
        t = np.asarray(op.tv, dtype=np.int64)
        # the op keeps the subgrid iff t and every column L_j R_:j are
        # multiples of L (row-wise)
        if np.any(t % Lv != 0) or np.any((R * Lv[None, :]) % Lv[:, None] != 0):
            continue
        y = [np.mod(R[i, 0] * grid[0] + R[i, 1] * grid[1] + R[i, 2] * grid[2] + t[i], N[i]) // Lv[i]
             for i in range(3)]
        idx = (y[0] + M[0] * (y[1] + M[1] * y[2])).ravel(order="F")
        key = idx.tobytes()
        if key not in seen:
            seen.add(key)
            maps.append(idx)
    return maps


def project_subgrid(f, maps):
    """Mean of a subgrid field over the maps: the field symmetric under the
    group closest to f (a field invariant under the ops that keep the subgrid
    is the subgrid of exactly one symmetric full-grid field)."""
    v = np.asarray(f, dtype=np.float64).ravel(order="F")
    acc = np.zeros_like(v)
    for idx in maps:
        acc += v[idx]
    return (acc / len(maps)).reshape(np.shape(f), order="F")


def start_fields(inst: Instance, ops_s, shift, amp=2.0, noise=0.1):
    """Seeded cosine-plus-noise start: w_A = +amp c + n_A, w_B = -amp c + n_B,
    c = the group's lowest non-extinct star (unit RMS), n_A, n_B seeded
    N(0, noise sqrt(n_maps / |G|)) subgrid fields projected onto the symmetric
    ones (the scale gives the point variance of a full-grid N(0, noise) field
    averaged over the group); on the subgrid."""
    N = inst.N
    Lv = auto_L(ops_s)
    M = tuple(n // l for n, l in zip(N, Lv))
    rng = np.random.default_rng(inst.seed)
    c, _ = lowest_star(N, ops_s, shift, Lv)
    maps = subgrid_maps(N, ops_s, Lv)
    sig = noise * np.sqrt(len(maps) / len(ops_s))
    wA = amp * c + project_subgrid(sig * rng.standard_normal(M), maps)
    wB = -amp * c + project_subgrid(sig * rng.standard_normal(M), maps)
    return wA, wB


def _prepare(inst: Instance, precision, m_hist, finish, tol):
    ops = get_ops(inst.sg, 3, inst.N)
    shift, ops_s = find_optimal_shift(ops, inst.N)
    lat = inst.cell * np.eye(3)
    mk = lambda prec, path: SCFTSolver(inst.N, ops_s, lat, inst.f, inst.chiN, inst.ns, prec, path=path,
                                       m_hist=m_hist)
    solver = mk(precision, default_path(inst.N, ops_s, lat, precision, tol))
    fin = mk(L.FP64, None) if finish else None
    return solver, fin, start_fields(inst, ops_s, shift)


def _prepare_group(insts, precision, m_hist, finish, tol):
    i0 = insts[0]
    ops = get_ops(i0.sg, 3, i0.N)
    shift, ops_s = find_optimal_shift(ops, i0.N)
    lat = i0.cell * np.eye(3)
    mk = lambda prec, path: SCFTBatch(i0.N, ops_s, lat, [i.f for i in insts], [i.chiN for i in insts], i0.ns,
                                      prec, path=path, m_hist=m_hist)
    solver = mk(precision, default_path(i0.N, ops_s, lat, precision, tol))
    fin = mk(L.FP64, None) if finish else None
    return solver, fin, [start_fields(i, ops_s, shift) for i in insts]


MAX_GROUP = 32  # instances per context: the plan takes per-field factors for up to 64 fields


def groups_of(instances, max_group=MAX_GROUP):
    """Indices of the instances grouped by (grid, group, cell, Ns), each group
    at most max_group long, largest work (instances x grid points) first."""
    by = {}
    for k, inst in enumerate(instances):
        by.setdefault((tuple(inst.N), inst.sg, inst.cell, inst.ns), []).append(k)
    out = []
    for key, idx in by.items():
        for a in range(0, len(idx), max_group):
            out.append(idx[a:a + max_group])
    out.sort(key=lambda g: -len(g) * int(np.prod(instances[g[0]].N)))
    return out


def _result(it, out):
    return dict(iters=it, F=out["F"], res=out["res"], Q=out["Q"])


def _finish(r, it2, o2, out):
    r.update(iters_finish=it2, F=o2["F"], res=o2["res"], Q=o2["Q"], F_first=out["F"], res_first=out["res"])


ENSEMBLE_LENS = (8, 12, 16, 24, 32)  # channel transform lengths of the ensemble kernels


def ensemble_ok(plan):
    """Whether a plan can join an ensemble: path QS, a cubic channel grid of a
    supported length, the plane-pair kernels (the ensemble runs the
    FP64-twiddle plane-pair kernels, not the per-axis chain)."""
    return (plan.path == L.PATH_QS and plan.H[0] == plan.H[1] == plan.H[2]
            and plan.H[0] in ENSEMBLE_LENS and not plan.options.get(L.OPT_AXIS_SPLIT))


# Mixing of the batch's contexts: Anderson steps from residual 0.1 down
# (simple mixing, lam 0.1, above).  On the 64-instance batch of sizes
# 32-96^3 and groups 47/70/223/229/230 every instance converges to 1e-8 in
# 84.5 iterations on average (max 138), against 153 (one instance, Fddd 96^3
# near its order-disorder transition, creeping under simple mixing and
# unconverged after 400) with Anderson from 1e-2, the single solver's
# default; the converged solutions are the same.
BATCH_MIXING = dict(lam=1.0, lam_simple=0.1, simple_above=0.1)


def solve_batch(instances, precision=L.MIXED, tol=1e-6, max_iter=400, workers=8, m_hist=10,
                finish_tol=None, group=True, tune=False, ensemble=False, mixing=BATCH_MIXING):
    """Plan every instance (or group of instances), then solve all of them
    concurrently.  With finish_tol, each instance continues in FP64 from its
    `precision` solution, fields and Anderson history handed over
    (``continue_from``), until the residual is below finish_tol (its FP64
    plan is made in the planning phase).  The path follows
    ``scft.default_path`` for the stage's tolerance.  With tune (grouped
    only), each group's launch schedule is timed on its own workload
    (``tune.tune_schedule``, one group after another, in the planning phase)
    and the fastest applied; results[i]["schedule"] names it.  With ensemble
    (grouped only), the groups whose plans qualify (``ensemble_ok``) run as
    one ``SCFTEnsemble`` per stage (and one for the FP64 finish), the others
    as before, concurrently with it; tile-height schedules from the tuner
    apply to both.  mixing: keyword arguments of ``set_mixing`` for every
    context (None: the contexts' own default).  Returns (results, plan_seconds, solve_seconds);
    results[i] = dict(iters, F, res, Q[, iters_finish]) in the order of
    `instances`."""
    finish = finish_tol is not None
    t0 = time.perf_counter()
    if group:
        groups = groups_of(instances)
        with ThreadPoolExecutor(max_workers=workers) as ex:
            prepared = list(ex.map(lambda g: _prepare_group([instances[k] for k in g], precision, m_hist,
                                                            finish, tol), groups))
        schedules = [None] * len(groups)
        if tune:
            from .tune import tune_schedule
            for j, (solver, fin, _) in enumerate(prepared):
                for sv in (solver, fin):
                    if sv is None:
                        continue
                    p = sv.plan
                    best, times = tune_schedule(p.N, p.ops, p.lattice, sv.ns, p.precision, p.batch, p.path)
                    sv.set_schedule(best)
                    if sv is solver:
                        schedules[j] = min(times, key=times.get)
    else:
        with ThreadPoolExecutor(max_workers=workers) as ex:
            prepared = list(ex.map(lambda inst: _prepare(inst, precision, m_hist, finish, tol), instances))
    if mixing:
        for sv, fin, _ in prepared:
            for x in (sv, fin):
                if x is not None:
                    x.set_mixing(**mixing)
    t1 = time.perf_counter()

    def run(item):
        solver, fin, (wA, wB) = item
        solver.set_fields(wA, wB)
        it, out = solver.solve(tol=tol, max_iter=max_iter)
        r = _result(it, out)
        if fin is not None:
            fin.continue_from(solver)
            it2, o2 = fin.solve(tol=finish_tol, max_iter=max_iter)
            _finish(r, it2, o2, out)
        return r

    def run_group(item):
        solver, fin, starts = item
        for i, (wA, wB) in enumerate(starts):
            solver.set_fields(i, wA, wB)
        its, outs = solver.solve(tol=tol, max_iter=max_iter)
        rs = [_result(it, out) for it, out in zip(its, outs)]
        if fin is not None:
            fin.continue_from(solver)
            its2, outs2 = fin.solve(tol=finish_tol, max_iter=max_iter)
            for r, it2, o2, out in zip(rs, its2, outs2, outs):
                _finish(r, it2, o2, out)
        return rs

    def run_ensemble(items):
        solvers = [sv for sv, _, _ in items]
        ens = SCFTEnsemble(solvers)
        for sv, _, starts in items:
            for i, (wA, wB) in enumerate(starts):
                sv.set_fields(i, wA, wB)
        res = ens.solve(tol=tol, max_iter=max_iter)
        rss = [[_result(it, out) for it, out in zip(its, outs)] for its, outs in res]
        if finish:
            fins = [fin for _, fin, _ in items]
            ens2 = SCFTEnsemble(fins)
            for sv, fin, _ in items:
                fin.continue_from(sv)
            res2 = ens2.solve(tol=finish_tol, max_iter=max_iter)
            for rs, (its2, outs2), (_, outs) in zip(rss, res2, res):
                for r, it2, o2, out in zip(rs, its2, outs2, outs):
                    _finish(r, it2, o2, out)
        return rss

    with ThreadPoolExecutor(max_workers=workers) as ex:
        if group:
            results = [None] * len(instances)
            ens_idx = []
            if ensemble:
                ens_idx = [j for j, (sv, fin, _) in enumerate(prepared)
                           if ensemble_ok(sv.plan) and (fin is None or ensemble_ok(fin.plan))]
            rest = [j for j in range(len(prepared)) if j not in set(ens_idx)]
            fut = ex.submit(run_ensemble, [prepared[j] for j in ens_idx]) if ens_idx else None
            for j, rs in zip(rest, ex.map(run_group, [prepared[j] for j in rest])):
                for k, r in zip(groups[j], rs):
                    if schedules[j] is not None:
                        r["schedule"] = schedules[j]
                    results[k] = r
            if fut is not None:
                for j, rs in zip(ens_idx, fut.result()):
                    for k, r in zip(groups[j], rs):
                        if schedules[j] is not None:
                            r["schedule"] = schedules[j]
                        r["ensemble"] = True
                        results[k] = r
        else:
            results = list(ex.map(run, prepared))
    t2 = time.perf_counter()
    del prepared
    return results, t1 - t0, t2 - t1
