"""GPU contour sweeps (cfft_sweep_dev, persistent cooperative kernel on cubic
grids, one launch per phase otherwise) vs the oracle's Strang propagation."""
import numpy as np
import pytest

import cfft_hip as cf
import oracle as O
from conftest import group_setup

pytestmark = pytest.mark.gpu


def sweep_case(sg, N, f, ns, prec, seed=3, persistent=0):
    import torch
    from cfft_hip.scft import PropagatorSweep
    ops, ops_s, shift, R, t = group_setup(sg, N)
    Lv = O.auto_L(t)
    lat = 4.0 * np.eye(3)
    wA = 2.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=seed), Lv)
    wB = -1.5 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=seed + 1), Lv)
    sw = PropagatorSweep(N, ops_s, lat, f, ns, prec)
    sw.plan.set_option(cf._lib.OPT_PERSISTENT, persistent)
    sw.set_fields(wA, wB)
    sw.reset()
    sw.run()
    torch.cuda.synchronize()
    q, qd = sw.fields()
    spec = O.calc_spectral_asu(R, t, N)
    table = O.m2_backward_table(spec, R, t, N)
    ds = 1.0 / ns
    diffuse = lambda x: O.m2_scft(x, spec, R, t, N, ds, lat, table)
    nA = int(round(f * ns))
    one = np.ones_like(wA)
    ref_q = O.propagate(one, lambda j: wA if j < nA else wB, ns, ds, diffuse)[-1]
    ref_qd = O.propagate(one, lambda j: wB if j < ns - nA else wA, ns, ds, diffuse)[-1]
    return q, qd, ref_q, ref_qd, sw.Q()


@pytest.mark.parametrize("persistent", [0, 1])
@pytest.mark.parametrize("sg,N", [(229, (32, 32, 32)), (230, (32, 32, 32)), (47, (32, 32, 32)),
                                  (70, (32, 32, 32)), (10, (16, 16, 16))])
def test_sweep_fp64_vs_oracle(sg, N, persistent):
    print(f"sweep sg={sg} N={N} persistent={persistent}", flush=True)
    q, qd, rq, rqd, (Q1, Q2) = sweep_case(sg, N, 0.3, 20, cf.FP64, persistent=persistent)
    assert np.max(np.abs(q - rq)) / np.max(np.abs(rq)) < 1e-12
    assert np.max(np.abs(qd - rqd)) / np.max(np.abs(rqd)) < 1e-12
    # Q from q(1) and from q+(1) agree (symmetry of the chain)
    assert abs(Q1 - Q2) / abs(Q1) < 1e-12
    assert abs(Q1 - float(np.mean(rq))) / abs(Q1) < 1e-12


def test_sweep_mixed_vs_fp64():
    q, qd, rq, rqd, (Q1, Q2) = sweep_case(229, (32, 32, 32), 0.3, 20, cf.MIXED)
    assert np.max(np.abs(q - rq)) / np.max(np.abs(rq)) < 2e-5
    assert abs(Q1 - float(np.mean(rq))) / abs(Q1) < 1e-6
