"""Launch-schedule tuning for a workload: the fold/unfold tile height,
hipGraph replay and, on path QS, the plane-pair workgroup size and the
radix schedule (lane split) of the transforms, timed on the plan's own grid,
group, precision and number of fields per launch.  No counterpart in the
reference.

Tile height, graph replay and workgroup size do not change results (every
schedule computes each field with the same arithmetic), only the
decomposition of the phase kernels and how the step sequence is launched.
The lane split R1 x R2 of a length-H transform (in-lane R2-point passes,
then a cross-lane radix-R1 pass) reorders the butterflies, so results move
by rounding only; it is searched only when `lanes` is set.  So does the
per-axis chain of path QS (CFFT_OPT_AXIS_SPLIT), searched when `axis` is set.  A timing is one
contour sweep of `steps` Strang steps over `batch` fields with all
interaction factors 1 (a plain multi-step sweep, the kernels of the stored
sweeps apart from the q store), after one untimed sweep that captures the
graph; the best of `reps` sweeps counts.  The grid of tile heights x graph
replay is timed in full, then each further dimension from the best so far.
Results are cached per workload in this process.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib as L
from .plans import DevicePlan

_CACHE: dict = {}


def schedule_candidates(H1):
    """Tile heights 8, 4, 2 that divide H1, each with and without graph replay."""
    rows = [r for r in (8, 4, 2) if H1 % r == 0] or [1]
    return [{L.OPT_TILE_ROWS: r, L.OPT_GRAPH: g} for r in rows for g in (1, 0)]


def extra_dimensions(plan, lanes, axis=False):
    """Further schedule dimensions of a plan on path QS: with `axis`, the
    per-axis chain (H = 8, 12, 16, 24, 32; plans without the separable
    multiply refuse it and keep the plane pairs); the plane-pair workgroup size (512 threads
    for H = 16, 24, 32) and, with `lanes`, the lane split R1 = 2 of the
    transforms (H = 16, 32)."""
    out = []
    if plan.path == L.PATH_QS:
        if axis and plan.H[0] in (8, 12, 16, 24, 32):
            out.append((L.OPT_AXIS_SPLIT, 1))
        if plan.H[1] in (16, 24, 32):
            out.append((L.OPT_PAIR_THREADS, 512))
        if lanes and plan.H[0] in (16, 32):
            out.append((L.OPT_FFT_LANES, 2))
    return out


def _label(opt):
    return (f"rows{opt[L.OPT_TILE_ROWS]}_graph{opt[L.OPT_GRAPH]}"
            + ("_axis" if opt.get(L.OPT_AXIS_SPLIT) else "")
            + (f"_pt{opt[L.OPT_PAIR_THREADS]}" if opt.get(L.OPT_PAIR_THREADS) else "")
            + (f"_lanes{opt[L.OPT_FFT_LANES]}" if opt.get(L.OPT_FFT_LANES) else ""))


def _key(N, ops_shifted, lattice, ns, precision, batch, path):
    ops = tuple((tuple(np.asarray(o.Rm).ravel().tolist()), tuple(np.asarray(o.tv).ravel().tolist()))
                for o in ops_shifted)
    return (tuple(N), hash(ops), tuple(np.asarray(lattice, dtype=float).ravel().tolist()), int(ns),
            int(precision), int(batch), int(path))


def tune_schedule(N, ops_shifted, lattice, ns, precision, batch, path, steps=None, reps=3, lanes=False,
                  axis=False):
    """(best options, {label: microseconds per step}) for a plan of `batch`
    fields with ds = 1/ns; options are {OPT_TILE_ROWS: r, OPT_GRAPH: g,
    OPT_PAIR_THREADS: t, OPT_FFT_LANES: l} (the last two 0 = default), plus
    OPT_AXIS_SPLIT with `axis`.  A candidate the plan refuses (the per-axis
    chain without the separable multiply, or with the R1 = 2 lane split) is
    skipped."""
    import torch
    key = _key(N, ops_shifted, lattice, ns, precision, batch, path) + (bool(lanes), bool(axis))
    if key in _CACHE:
        return _CACHE[key]
    steps = int(ns if steps is None else steps)
    plan = DevicePlan(N, ops_shifted, 1.0 / ns, lattice, path, precision, batch=batch)
    M3 = int(np.prod(plan.M))
    dt = torch.float64 if precision == L.FP64 else torch.float32
    q = torch.ones(batch * M3, dtype=dt, device="cuda")
    E = torch.ones(batch * M3, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    times = {}

    def timed(opt):
        # the per-axis chain first: it excludes the R1 = 2 lane split
        plan.set_option(L.OPT_AXIS_SPLIT, opt.get(L.OPT_AXIS_SPLIT, 0))
        for k, v in opt.items():
            plan.set_option(k, v)
        plan.step_dev(q.data_ptr(), E.data_ptr(), steps, s)  # captures the graph when enabled
        torch.cuda.synchronize()
        t = None
        for _ in range(reps):
            t0 = time.perf_counter()
            plan.step_dev(q.data_ptr(), E.data_ptr(), steps, s)
            torch.cuda.synchronize()
            el = (time.perf_counter() - t0) / steps * 1e6
            t = el if t is None else min(t, el)
        times[_label(opt)] = t
        return t

    best, best_t = None, None
    for opt in schedule_candidates(plan.H[1]):
        opt = dict(opt)
        opt.update({L.OPT_PAIR_THREADS: 0, L.OPT_FFT_LANES: 0})
        if axis:
            opt[L.OPT_AXIS_SPLIT] = 0
        t = timed(opt)
        if best_t is None or t < best_t:
            best, best_t = opt, t

    def trial(opt):
        nonlocal best, best_t
        try:
            t = timed(opt)
        except L.CFFTError as e:
            if e.code != L.CFFT_ERR_UNSUPPORTED:
                raise
            return False
        if t < best_t:
            best, best_t = opt, t
        return True

    grid_best = dict(best)
    for k, v in extra_dimensions(plan, lanes, axis):
        opt = dict(best)
        opt[k] = v
        if trial(opt) and k == L.OPT_AXIS_SPLIT:
            # the per-axis chain's kernels cut the tiles differently: every
            # tile height with the best graph setting
            for r in sorted({c[L.OPT_TILE_ROWS] for c in schedule_candidates(plan.H[1])}):
                if r != grid_best[L.OPT_TILE_ROWS]:
                    trial({**grid_best, L.OPT_TILE_ROWS: r, L.OPT_AXIS_SPLIT: 1})
    del plan
    _CACHE[key] = (best, times)
    return best, times
