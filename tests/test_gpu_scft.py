"""GPU contour sweeps (cfft_sweep_dev, persistent cooperative kernel on cubic
grids, one launch per phase otherwise) vs the oracle's Strang propagation."""
import numpy as np
import pytest

import cfft_hip as cf
import oracle as O
from conftest import group_setup

pytestmark = pytest.mark.gpu


def sweep_case(sg, N, f, ns, prec, seed=3, persistent=0, fuse=1):
    import torch
    from cfft_hip.scft import PropagatorSweep
    ops, ops_s, shift, R, t = group_setup(sg, N)
    Lv = O.auto_L(t)
    lat = 4.0 * np.eye(3)
    wA = 2.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=seed), Lv)
    wB = -1.5 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=seed + 1), Lv)
    sw = PropagatorSweep(N, ops_s, lat, f, ns, prec)
    sw.plan.set_option(cf._lib.OPT_PERSISTENT, persistent)
    sw.plan.set_option(cf._lib.OPT_FUSE, fuse)
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


@pytest.mark.parametrize("persistent,fuse", [(0, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("sg,N", [(229, (32, 32, 32)), (230, (32, 32, 32)), (47, (32, 32, 32)),
                                  (70, (32, 32, 32)), (10, (16, 16, 16)), (71, (24, 32, 48))])
def test_sweep_fp64_vs_oracle(sg, N, persistent, fuse):
    print(f"sweep sg={sg} N={N} persistent={persistent} fuse={fuse}", flush=True)
    q, qd, rq, rqd, (Q1, Q2) = sweep_case(sg, N, 0.3, 20, cf.FP64, persistent=persistent, fuse=fuse)
    assert np.max(np.abs(q - rq)) / np.max(np.abs(rq)) < 1e-12
    assert np.max(np.abs(qd - rqd)) / np.max(np.abs(rqd)) < 1e-12
    # Q from q(1) and from q+(1) agree (symmetry of the chain)
    assert abs(Q1 - Q2) / abs(Q1) < 1e-12
    assert abs(Q1 - float(np.mean(rq))) / abs(Q1) < 1e-12


def test_sweep_mixed_vs_fp64():
    q, qd, rq, rqd, (Q1, Q2) = sweep_case(229, (32, 32, 32), 0.3, 20, cf.MIXED)
    assert np.max(np.abs(q - rq)) / np.max(np.abs(rq)) < 2e-5
    assert abs(Q1 - float(np.mean(rq))) / abs(Q1) < 1e-6


# ---------------------------------------------------------------------------
# SCFT driver (cfft_scft context) vs the oracle
# ---------------------------------------------------------------------------

def _bcc_fields(sg, N, amp=3.0, noise=0.05, seed=7):
    from cfft_hip.scft import first_star_bcc, symmetrize_subgrid
    ops, ops_s, shift, R, t = group_setup(sg, N)
    Lv = O.auto_L(t)
    rng = np.random.default_rng(seed)
    sgrid = [s * n for s, n in zip(shift, N)]
    star = first_star_bcc(N, sgrid, 4.0)
    wA = symmetrize_subgrid(amp * star + noise * rng.standard_normal(N), ops_s, Lv)
    wB = symmetrize_subgrid(-amp * star + noise * rng.standard_normal(N), ops_s, Lv)
    return ops_s, R, t, Lv, wA, wB


@pytest.mark.parametrize("sg,N", [(229, (16, 16, 16)), (47, (16, 16, 16)), (230, (24, 24, 24))])
def test_scft_iterations_vs_oracle(sg, N):
    """Q, F, residual and the mixed fields of 6 iterations (simple mixing, then
    Anderson) against the oracle; FP64 bar 1e-11 relative."""
    from cfft_hip.scft import SCFTSolver
    print(f"scft iterations sg={sg} N={N}", flush=True)
    f, chiN, ns = 0.3, 20.0, 20
    lat = 4.0 * np.eye(3)
    ops_s, R, t, Lv, wA, wB = _bcc_fields(sg, N)
    solver = SCFTSolver(N, ops_s, lat, f, chiN, ns, cf.FP64, m_hist=4)
    solver.set_mixing(lam=1.0, lam_simple=0.2, simple_above=1e9)
    solver.set_fields(wA, wB)
    spec = O.calc_spectral_asu(R, t, N)
    table = O.m2_backward_table(spec, R, t, N)
    diffuse = lambda x: O.m2_scft(x, spec, R, t, N, 1.0 / ns, lat, table)
    mix = O.AndersonMixer(m=4, lam=1.0, lam_simple=0.2, simple_above=1e9)
    w = np.concatenate([wA.ravel(order="F"), wB.ravel(order="F")])
    M3 = wA.size
    for it in range(6):
        g = solver.iterate()
        a = w[:M3].reshape(wA.shape, order="F")
        b = w[M3:].reshape(wA.shape, order="F")
        out = O.sweep(a, b, f, ns, diffuse)
        tA, tB, res = O.field_update(a, b, out["phiA"], out["phiB"], chiN)
        F = O.free_energy(a, b, out["phiA"], out["phiB"], out["Q"], chiN)
        assert abs(g["Q"] - out["Q"]) < 1e-11 * out["Q"]
        assert abs(g["Qd"] - out["Q"]) < 1e-11 * out["Q"]
        assert abs(g["F"] - F) < 1e-11 * max(1.0, abs(F))
        assert abs(g["res"] - res) < 1e-9 * res
        assert g["anderson"] == (1.0 if it > 0 else 0.0)
        phiA, phiB = solver.densities()
        assert np.max(np.abs(phiA - out["phiA"])) < 1e-11
        assert np.max(np.abs(phiB - out["phiB"])) < 1e-11
        d = np.concatenate([(tA - a).ravel(order="F"), (tB - b).ravel(order="F")])
        w = mix.step(w, d, res)
        gw = np.concatenate([x.ravel(order="F") for x in solver.fields()])
        assert np.max(np.abs(gw - w)) < 1e-9 * np.max(np.abs(w))


def test_scft_homogeneous_closed_form():
    from cfft_hip.scft import SCFTSolver
    N = (16, 16, 16)
    ops_s, R, t, Lv, wA, wB = _bcc_fields(229, N)
    solver = SCFTSolver(N, ops_s, 4.0 * np.eye(3), 0.3, 20.0, 20, cf.FP64)
    solver.set_fields(1.5 * np.ones_like(wA), -0.5 * np.ones_like(wA))
    g = solver.iterate()
    assert abs(g["Q"] - O.homogeneous_Q(1.5, -0.5, 6 / 20)) < 1e-13
    phiA, phiB = solver.densities()
    assert np.max(np.abs(phiA - 0.3)) < 1e-13 and np.max(np.abs(phiB - 0.7)) < 1e-13


@pytest.mark.parametrize("prec", [cf.FP64, cf.MIXED])
def test_scft_converges_bcc(prec):
    """Im-3m 32^3, f = 0.3, chiN = 20, Ns = 50 from BCC first-star fields:
    FP64 converges to residual 1e-8; mixed precision reaches 1e-6 with F
    within 1e-6 of the FP64 solution."""
    from cfft_hip.scft import SCFTSolver
    N = (32, 32, 32)
    print(f"scft converge prec={prec}", flush=True)
    ops_s, R, t, Lv, wA, wB = _bcc_fields(229, N)
    res = {}
    for p in (cf.FP64, prec):
        solver = SCFTSolver(N, ops_s, 4.0 * np.eye(3), 0.3, 20.0, 50, p, m_hist=10)
        solver.set_fields(wA, wB)
        it, out = solver.solve(tol=1e-8 if p == cf.FP64 else 1e-6, max_iter=400)
        print(f"  prec={p} iters={it} F={out['F']:.12f} res={out['res']:.3e}", flush=True)
        res[p] = (it, out, solver.densities())
    it64, o64, ph64 = res[cf.FP64]
    assert o64["res"] < 1e-8
    assert abs(o64["phi_total"] - 1.0) < 1e-8
    assert o64["F"] < O.disordered_free_energy(0.3, 20.0)
    it, o, ph = res[prec]
    assert o["res"] < (1e-8 if prec == cf.FP64 else 1e-6)
    assert abs(o["F"] - o64["F"]) < 1e-6
    assert np.max(np.abs(ph[0] - ph64[0])) < 1e-4
