"""Ensembles of SCFT contexts (GPU): several contexts of different grids and
space groups advanced by one multi-plan sweep per iteration
(cfft_scft_ensemble_*).  Each instance's iterations, fields and densities
must equal those of its own context bit for bit: the ensemble's work items
run the contexts' phase functions with the same parameters and buffers."""
import numpy as np
import pytest

import cfft_hip as cf

pytestmark = pytest.mark.gpu

# (space group, grid, cell, [(f, chiN, seed), ...]): channel lengths 8, 12, 16
_CONTEXTS = [
    (229, (32, 32, 32), 4.0, [(0.30, 16.0, 11), (0.42, 22.0, 12)]),
    (230, (48, 48, 48), 3.8, [(0.35, 18.0, 13)]),
    (47, (32, 32, 32), 4.0, [(0.28, 14.0, 14), (0.40, 25.0, 15), (0.33, 19.0, 16)]),
    (70, (64, 64, 64), 4.0, [(0.38, 21.0, 17)]),
    (223, (48, 48, 48), 4.0, [(0.31, 24.0, 18)]),
]


def _setup(prec, ns=40, contexts=_CONTEXTS):
    from cfft_hip import batch
    from cfft_hip.scft import SCFTBatch
    made = []
    for sg, N, cell, params in contexts:
        shift, ops_s = cf.find_optimal_shift(cf.get_ops(sg, 3, N), N)
        insts = [batch.Instance(N, sg, cell, f, x, ns, seed) for f, x, seed in params]
        starts = [batch.start_fields(i, ops_s, shift) for i in insts]
        lat = cell * np.eye(3)
        mk = lambda: SCFTBatch(N, ops_s, lat, [p[0] for p in params], [p[1] for p in params], ns, prec,
                               path=cf.PATH_QS)
        made.append((mk(), mk(), starts))
    return made


def _set(ctx, starts):
    for i, (wA, wB) in enumerate(starts):
        ctx.set_fields(i, wA, wB)


@pytest.mark.parametrize("prec", [cf.FP64, cf.MIXED])
def test_ensemble_iterations_equal_own_contexts(prec):
    """Four iterations of five contexts (Im-3m 32^3, Ia-3d 48^3, Pmmm 32^3,
    Fddd 64^3, Pm-3n 48^3; 1-3 instances each) through one ensemble against
    the same contexts iterated on their own: results, fields and densities
    equal bit for bit."""
    from cfft_hip.scft import SCFTEnsemble
    made = _setup(prec)
    print(f"ensemble prec={prec}: {[m[0].plan.H for m in made]}", flush=True)
    ens = SCFTEnsemble([a for a, _, _ in made])
    for a, b, starts in made:
        _set(a, starts)
        _set(b, starts)
    for step in range(4):
        got = ens.iterate()
        want = [b.iterate() for _, b, _ in made]
        assert got == want, step
    for a, b, _ in made:
        for i in range(a.n):
            for x, y in zip(a.fields(i) + a.densities(i), b.fields(i) + b.densities(i)):
                assert np.array_equal(x, y)


def test_ensemble_solve_stops_each_instance():
    """solve() through the ensemble: every instance stops at its own
    convergence with the iteration count and results of its own context's
    solve (FP64, residual 1e-6)."""
    from cfft_hip.scft import SCFTEnsemble
    made = _setup(cf.FP64, contexts=_CONTEXTS[:3])
    ens = SCFTEnsemble([a for a, _, _ in made])
    for a, b, starts in made:
        _set(a, starts)
        _set(b, starts)
    got = ens.solve(tol=1e-6, max_iter=150)
    want = [b.solve(tol=1e-6, max_iter=150) for _, b, _ in made]
    print(f"  iterations {[g[0] for g in got]}", flush=True)
    assert got == want
    assert len({it for its, _ in got for it in its}) > 1  # the instances stopped at different iterations
    for a, b, _ in made:
        for i in range(a.n):
            assert np.array_equal(a.fields(i)[0], b.fields(i)[0])


def test_ensemble_batch_equals_grouped():
    """solve_batch(ensemble=True) gives every instance the iterations and
    results of the grouped contexts run on their own streams."""
    from cfft_hip import batch
    insts = batch.draw_instances(6, seed=31, sizes=(32, 48), groups=(229, 47, 230), ns=40)
    e, _, _ = batch.solve_batch(insts, precision=cf.FP64, tol=1e-7, max_iter=200, workers=4, ensemble=True)
    g, _, _ = batch.solve_batch(insts, precision=cf.FP64, tol=1e-7, max_iter=200, workers=4)
    assert all(r.pop("ensemble", False) for r in e)
    assert e == g


def test_ensemble_rejects_unsupported_plans():
    """A context off path QS, a channel length without an ensemble kernel, or
    contexts of different contour lengths or precisions are refused."""
    from cfft_hip.scft import SCFTBatch, SCFTEnsemble
    N = (32, 32, 32)
    _, ops_s = cf.find_optimal_shift(cf.get_ops(229, 3, N), N)
    lat = 4.0 * np.eye(3)
    good = SCFTBatch(N, ops_s, lat, [0.3], [20.0], 20, cf.FP64, path=cf.PATH_QS)
    centred = SCFTBatch(N, ops_s, lat, [0.3], [20.0], 20, cf.FP64, path=cf.PATH_M7)
    with pytest.raises(cf.CFFTError) as e:
        SCFTEnsemble([good, centred])
    assert e.value.code == cf._lib.CFFT_ERR_UNSUPPORTED
    N16 = (16, 16, 16)
    _, ops16 = cf.find_optimal_shift(cf.get_ops(229, 3, N16), N16)
    small = SCFTBatch(N16, ops16, lat, [0.3], [20.0], 20, cf.FP64, path=cf.PATH_QS)
    with pytest.raises(cf.CFFTError) as e:
        SCFTEnsemble([good, small])
    assert e.value.code == cf._lib.CFFT_ERR_UNSUPPORTED
    other_ns = SCFTBatch(N, ops_s, lat, [0.3], [20.0], 30, cf.FP64, path=cf.PATH_QS)
    with pytest.raises(cf.CFFTError):
        SCFTEnsemble([good, other_ns])
    mixed = SCFTBatch(N, ops_s, lat, [0.3], [20.0], 20, cf.MIXED, path=cf.PATH_QS)
    with pytest.raises(cf.CFFTError):
        SCFTEnsemble([good, mixed])
    with pytest.raises(cf.CFFTError):
        SCFTEnsemble([good, good])
