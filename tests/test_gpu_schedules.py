"""Launch and radix schedules of the plane-pair chain (GPU): the plane-pair
workgroup size (CFFT_OPT_PAIR_THREADS) leaves every result bit for bit; the
lane split of the transforms (CFFT_OPT_FFT_LANES: R1 = 2 instead of 4, so
16- or 8-point in-lane passes and a cross-lane radix-2 pass) reorders the
butterflies and agrees with the default schedule, which the oracle pins at
these sizes, to rounding; unavailable schedules are refused; the tuner
searches both dimensions."""
import numpy as np
import pytest

import cfft_hip as cf
import oracle as O
from conftest import group_setup

pytestmark = pytest.mark.gpu


def _sweep(N, prec, options, ns=20):
    import torch
    from cfft_hip.scft import PropagatorSweep
    ops, ops_s, shift, R, t = group_setup(229, N)
    Lv = O.auto_L(t)
    wA = 2.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=21), Lv)
    wB = -1.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=22), Lv)
    sw = PropagatorSweep(N, ops_s, 4.0 * np.eye(3), 0.3, ns, prec, path=cf.PATH_QS)
    for k, v in options.items():
        sw.plan.set_option(k, v)
    sw.set_fields(wA, wB)
    sw.reset()
    sw.run()
    torch.cuda.synchronize()
    return [np.asarray(x, dtype=np.float64).copy() for x in sw.fields()]


@pytest.mark.parametrize("n", [64, 96, 128])
@pytest.mark.parametrize("prec", [cf.FP64, cf.MIXED])
def test_pair_threads_bit_identical(n, prec):
    """512-thread plane-pair workgroups (H = 16, 24, 32) give the 256-thread
    sweep bit for bit."""
    print(f"pair threads N={n} prec={prec}", flush=True)
    N = (n, n, n)
    a = _sweep(N, prec, {})
    b = _sweep(N, prec, {cf._lib.OPT_PAIR_THREADS: 512})
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("n", [64, 128])
@pytest.mark.parametrize("prec,tol", [(cf.FP64, 1e-12), (cf.MIXED, 2e-6)])
def test_lane_split_matches_default(n, prec, tol):
    """Lane split R1 = 2 (H = 16, 32), alone and with 512-thread plane pairs:
    the 20-step sweep agrees with the default schedule to 1e-12 relative
    (FP64) or 2e-6 (mixed: both schedules round to FP32 on every pass)."""
    print(f"lane split N={n} prec={prec}", flush=True)
    N = (n, n, n)
    a = _sweep(N, prec, {})
    for opts in ({cf._lib.OPT_FFT_LANES: 2},
                 {cf._lib.OPT_FFT_LANES: 2, cf._lib.OPT_PAIR_THREADS: 512}):
        b = _sweep(N, prec, opts)
        for x, y in zip(a, b):
            err = np.max(np.abs(x - y)) / np.max(np.abs(x))
            print(f"  {opts}: rel err {err:.2e}", flush=True)
            assert err < tol
            if prec == cf.FP64:
                assert not np.array_equal(x, y)  # the option reached the kernels


def test_schedule_options_refused():
    """Lane splits other than the default or R1 = 2 on H = 16, 32 (H = 24
    splits as 2 x 12 by default; 4 x 6 is no split), 512-thread pairs on H
    other than 16, 24, 32, either off path QS, and an ensemble of a context
    with a non-default lane split are refused."""
    from cfft_hip.scft import SCFTBatch, SCFTEnsemble
    lat = 4.0 * np.eye(3)
    E = cf._lib.CFFT_ERR_UNSUPPORTED

    def plan(n, path=cf.PATH_QS):
        N = (n, n, n)
        _, ops_s = cf.find_optimal_shift(cf.get_ops(229, 3, N), N)
        return cf.DevicePlan(N, ops_s, 0.01, lat, path, cf.FP64)

    for p, key, val in ((plan(96), cf._lib.OPT_FFT_LANES, 4), (plan(64), cf._lib.OPT_FFT_LANES, 3),
                        (plan(32), cf._lib.OPT_PAIR_THREADS, 512), (plan(64), cf._lib.OPT_PAIR_THREADS, 384),
                        (plan(64, cf.PATH_M7), cf._lib.OPT_FFT_LANES, 2),
                        (plan(64, cf.PATH_M7), cf._lib.OPT_PAIR_THREADS, 512)):
        with pytest.raises(cf.CFFTError) as e:
            p.set_option(key, val)
        assert e.value.code == E
    p = plan(64)
    for key, val in ((cf._lib.OPT_FFT_LANES, 0), (cf._lib.OPT_FFT_LANES, 4), (cf._lib.OPT_PAIR_THREADS, 256)):
        p.set_option(key, val)  # the defaults are always accepted
    plan(96).set_option(cf._lib.OPT_FFT_LANES, 2)  # H = 24's own default split
    N = (64, 64, 64)
    _, ops_s = cf.find_optimal_shift(cf.get_ops(229, 3, N), N)
    ctx = SCFTBatch(N, ops_s, lat, [0.3], [20.0], 20, cf.FP64, path=cf.PATH_QS)
    ctx.set_schedule({cf._lib.OPT_FFT_LANES: 2})
    with pytest.raises(cf.CFFTError) as e:
        SCFTEnsemble([ctx])
    assert e.value.code == E


def test_tuner_searches_radix_schedule():
    """tune_schedule(lanes=True) on Im-3m 64^3 (H = 16) times the tile/graph
    grid, then 512-thread pairs and the R1 = 2 lane split from the best; the
    returned options name all four settings."""
    from cfft_hip.tune import tune_schedule
    N = (64, 64, 64)
    _, ops_s = cf.find_optimal_shift(cf.get_ops(229, 3, N), N)
    best, times = tune_schedule(N, ops_s, 4.0 * np.eye(3), 20, cf.FP64, 2, cf.PATH_QS, reps=2, lanes=True)
    print("  ", best, {k: round(v, 2) for k, v in times.items()}, flush=True)
    assert set(best) == {cf._lib.OPT_TILE_ROWS, cf._lib.OPT_GRAPH, cf._lib.OPT_PAIR_THREADS, cf._lib.OPT_FFT_LANES}
    assert any("_pt512" in k for k in times) and any("_lanes2" in k for k in times)
