"""The per-axis chain of path QS (CFFT_OPT_AXIS_SPLIT, GPU): the separable Q
multiply applied as one 2 x 2 mixing per axis, the x mixing with the x
transforms in the unfold/fold tile kernel and the y, z mixings in a kernel
over single x-planes (32-point lines over eight lanes).  The same operator as
the plane-pair chain, which the oracle pins at these sizes, up to rounding:
sweeps agree with it to 1e-12 relative (FP64) or 2e-6 (mixed), an SCFT
iteration's Q, F and densities to 1e-12; the option is refused where it does
not apply."""
import numpy as np
import pytest

import cfft_hip as cf
import oracle as O
from conftest import group_setup

pytestmark = pytest.mark.gpu


def _fields(sg, N):
    ops, ops_s, shift, R, t = group_setup(sg, N)
    Lv = O.auto_L(t)
    wA = 2.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=41), Lv)
    wB = -1.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=42), Lv)
    return ops_s, wA, wB


def _sweep(sg, N, prec, axis, ns=20, graph=1):
    import torch
    from cfft_hip.scft import PropagatorSweep
    ops_s, wA, wB = _fields(sg, N)
    sw = PropagatorSweep(N, ops_s, 4.0 * np.eye(3), 0.3, ns, prec, path=cf.PATH_QS)
    sw.plan.set_option(cf._lib.OPT_AXIS_SPLIT, axis)
    sw.plan.set_option(cf._lib.OPT_GRAPH, graph)
    sw.set_fields(wA, wB)
    sw.reset()
    sw.run()
    torch.cuda.synchronize()
    return [np.asarray(x, dtype=np.float64).copy() for x in sw.fields()]


@pytest.mark.parametrize("sg,n", [(229, 64), (229, 128), (47, 64), (47, 128), (223, 64), (229, 32), (229, 48),
                                  (229, 96), (47, 48), (47, 32), (47, 96), (223, 96)])
@pytest.mark.parametrize("prec,tol", [(cf.FP64, 1e-12), (cf.MIXED, 2e-6)])
def test_axis_split_sweep_matches_plane_pairs(sg, n, prec, tol):
    """20-step q and q+ sweep, per-axis chain against the plane-pair chain:
    channel length H = N / 4 (8 at 32^3 to 32 at 128^3; lengths other than
    32 on the default lane split); four channels (Im-3m, Pm-3n) or eight
    (Pmmm: 64 tile rows, two passes of the row transforms; at 32^3 eight
    channels of 8-row tiles fill the tile's 64 LDS rows)."""
    print(f"axis split sg={sg} N={n} prec={prec}", flush=True)
    N = (n, n, n)
    a = _sweep(sg, N, prec, 0)
    b = _sweep(sg, N, prec, 1)
    for x, y in zip(a, b):
        err = np.max(np.abs(x - y)) / np.max(np.abs(x))
        print(f"  rel err {err:.2e}", flush=True)
        assert err < tol


def test_axis_split_unfused_and_eager_equal_graph():
    """The unfused chain (unfold, then fold) and eager launches give the
    graph-replayed fused chain's result bit for bit (Im-3m 64^3)."""
    import torch
    from cfft_hip.scft import PropagatorSweep
    N = (64, 64, 64)
    ref = _sweep(229, N, cf.FP64, 1, graph=1)
    assert all(np.array_equal(x, y) for x, y in zip(ref, _sweep(229, N, cf.FP64, 1, graph=0)))
    ops_s, wA, wB = _fields(229, N)
    sw = PropagatorSweep(N, ops_s, 4.0 * np.eye(3), 0.3, 20, cf.FP64, path=cf.PATH_QS)
    sw.plan.set_option(cf._lib.OPT_AXIS_SPLIT, 1)
    sw.plan.set_option(cf._lib.OPT_FUSE, 0)
    sw.set_fields(wA, wB)
    sw.reset()
    sw.run()
    torch.cuda.synchronize()
    out = [np.asarray(x, dtype=np.float64) for x in sw.fields()]
    for x, y in zip(ref, out):
        assert np.max(np.abs(x - y)) / np.max(np.abs(x)) < 1e-13


@pytest.mark.parametrize("prec,tol", [(cf.FP64, 1e-12), (cf.MIXED, 5e-6)])
def test_axis_split_scft_iteration(prec, tol):
    """Three SCFT iterations of Im-3m 64^3 (the stored history sweep, whose
    tile kernel writes every contour slice): Q, F and the densities with the
    per-axis chain against the plane-pair chain."""
    from cfft_hip.scft import SCFTSolver
    N = (64, 64, 64)
    ops_s, wA, wB = _fields(229, N)
    res = []
    for axis in (0, 1):
        s = SCFTSolver(N, ops_s, 4.0 * np.eye(3), 0.3, 20.0, 50, prec, path=cf.PATH_QS)
        s.plan.set_option(cf._lib.OPT_AXIS_SPLIT, axis)
        s.set_fields(wA, wB)
        outs = [s.iterate() for _ in range(3)]
        res.append((outs, s.densities()))
    (o0, d0), (o1, d1) = res
    for a, b in zip(o0, o1):
        assert abs(a["Q"] - b["Q"]) / a["Q"] < tol
        assert abs(a["F"] - b["F"]) < tol * max(1.0, abs(a["F"]))
    for x, y in zip(d0, d1):
        assert np.max(np.abs(np.asarray(x) - np.asarray(y))) < tol


def test_axis_split_refused():
    """Off path QS, without the separable multiply (Ia-3d), with the R1 = 2
    lane split or FP32 twiddle products set, and for an ensemble of such a
    plan: CFFT_ERR_UNSUPPORTED."""
    from cfft_hip.scft import SCFTBatch, SCFTEnsemble
    lat = 4.0 * np.eye(3)
    E = cf._lib.CFFT_ERR_UNSUPPORTED

    def plan(n, path=cf.PATH_QS, prec=cf.FP64, sg=229):
        N = (n, n, n)
        _, ops_s = cf.find_optimal_shift(cf.get_ops(sg, 3, N), N)
        return cf.DevicePlan(N, ops_s, 0.01, lat, path, prec)

    for p in (plan(64, cf.PATH_M7), plan(64, sg=230)):
        with pytest.raises(cf.CFFTError) as e:
            p.set_option(cf._lib.OPT_AXIS_SPLIT, 1)
        assert e.value.code == E
    p = plan(64)
    p.set_option(cf._lib.OPT_FFT_LANES, 2)
    with pytest.raises(cf.CFFTError) as e:
        p.set_option(cf._lib.OPT_AXIS_SPLIT, 1)
    assert e.value.code == E
    p = plan(64, prec=cf.MIXED)
    p.set_option(cf._lib.OPT_AXIS_SPLIT, 1)
    for key, val in ((cf._lib.OPT_FP32_TWIDDLES, 1), (cf._lib.OPT_FFT_LANES, 2)):
        with pytest.raises(cf.CFFTError) as e:
            p.set_option(key, val)
        assert e.value.code == E
    N = (64, 64, 64)
    _, ops_s = cf.find_optimal_shift(cf.get_ops(229, 3, N), N)
    ctx = SCFTBatch(N, ops_s, lat, [0.3], [20.0], 20, cf.FP64, path=cf.PATH_QS)
    ctx.set_schedule({cf._lib.OPT_AXIS_SPLIT: 1})
    with pytest.raises(cf.CFFTError) as e:
        SCFTEnsemble([ctx])
    assert e.value.code == E


def test_tuner_searches_axis_split():
    """tune_schedule(axis=True) times the per-axis chain from the best tile
    and graph schedule where the plan takes it (Im-3m 64^3) and skips it
    where the plan refuses it (Ia-3d 64^3, no separable multiply)."""
    from cfft_hip.tune import tune_schedule
    lat = 4.0 * np.eye(3)
    for sg, takes in ((229, True), (230, False)):
        N = (64, 64, 64)
        _, ops_s = cf.find_optimal_shift(cf.get_ops(sg, 3, N), N)
        best, times = tune_schedule(N, ops_s, lat, 20, cf.FP64, 2, cf.PATH_QS, reps=2, axis=True)
        print(f"  sg={sg}", best, {k: round(v, 2) for k, v in times.items()}, flush=True)
        assert cf._lib.OPT_AXIS_SPLIT in best
        assert any("_axis" in k for k in times) == takes
        if not takes:
            assert best[cf._lib.OPT_AXIS_SPLIT] == 0
