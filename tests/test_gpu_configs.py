"""Parity at the sizes of the benchmark workloads (GPU).

* Im-3m 128^3, f 0.3, chiN 20, Ns 100 (the timed workload): the plane-pair
  chain's q / q+ sweep and the first SCFT iteration (Q, F, densities) in FP64
  against the C restatement of the Strang step with the oracle's Simpson
  densities (1e-12), and mixed precision against FP64.
* Ia-3d 128^3, cell 3.8, f 0.35, chiN 18, Ns 200: mixed precision against
  FP64 over the first iterations (Q, F, densities), and mixed SCFT
  iterations against the oracle at 24^3.
* 64-instance shape: every grid size {32, 48, 64, 96}^3 with every group
  {47, 70, 223, 229, 230}, mixed precision against each instance's FP64
  solve (F and densities; Q depends on the gauge of w), and each group's
  first mixed iteration at 32^3 against the oracle.
"""
import numpy as np
import pytest

import cfft_hip as cf
import oracle as O
from oracle import cstep
from conftest import group_setup

pytestmark = pytest.mark.gpu

THREADS = 16  # the C oracle's threads (the host share of one GPU)


def _bcc128():
    """The timed workload's fields: w_A = +3 S + n_A, w_B = -3 S + n_B on the
    subgrid, S the BCC {110} star, n seeded N(0, 0.05) fields symmetrised."""
    from cfft_hip.scft import first_star_bcc, symmetrize_subgrid
    N = (128, 128, 128)
    ops, ops_s, shift, R, t = group_setup(229, N)
    Lv = O.auto_L(t)
    rng = np.random.default_rng(2024)
    S = first_star_bcc(N, [s * n for s, n in zip(shift, N)], 4.0)
    S0 = np.ascontiguousarray(S[tuple(slice(0, None, int(l)) for l in Lv)])
    nA = symmetrize_subgrid(rng.normal(0.0, 0.05, N), ops_s, Lv)
    nB = symmetrize_subgrid(rng.normal(0.0, 0.05, N), ops_s, Lv)
    return N, ops_s, R, t, 3.0 * S0 + nA, -3.0 * S0 + nB


@pytest.fixture(scope="module")
def im3m128():
    N, ops_s, R, t, wA, wB = _bcc128()
    cs = cstep.CStep(R, t, N, 1.0 / 100, 4.0 * np.eye(3), nthreads=THREADS)
    ref = cstep.first_iteration(cs, wA, wB, 0.3, 20.0, 100)
    return dict(N=N, ops_s=ops_s, wA=wA, wB=wB, ref=ref)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _sweep(case, prec, path, axis=0):
    import torch
    from cfft_hip.scft import PropagatorSweep
    sw = PropagatorSweep(case["N"], case["ops_s"], 4.0 * np.eye(3), 0.3, 100, prec, path=path)
    if axis:
        sw.plan.set_option(cf._lib.OPT_AXIS_SPLIT, 1)
    sw.set_fields(case["wA"], case["wB"])
    sw.reset()
    sw.run()
    torch.cuda.synchronize()
    return sw.fields()


@pytest.mark.parametrize("axis", [0, 1])
def test_im3m128_plane_pair_sweep_fp64_vs_c_oracle(im3m128, axis):
    """100-step q and q+ sweep on path QS (the timed chain, H = 32; plane
    pairs, and the per-axis chain the tuned bench runs) vs the C restatement
    of the step: 1e-12 relative."""
    print(f"Im-3m 128^3 QS sweep FP64 axis={axis}", flush=True)
    q, qd = _sweep(im3m128, cf.FP64, cf.PATH_QS, axis)
    ref = im3m128["ref"]
    assert _rel(q, ref["q"]) < 1e-12
    assert _rel(qd, ref["qd"]) < 1e-12


@pytest.mark.parametrize("axis", [0, 1])
def test_im3m128_plane_pair_sweep_mixed_vs_fp64(im3m128, axis):
    print(f"Im-3m 128^3 QS sweep mixed axis={axis}", flush=True)
    q, qd = _sweep(im3m128, cf.MIXED, cf.PATH_QS, axis)
    ref = im3m128["ref"]
    assert _rel(q, ref["q"]) < 2e-5
    assert _rel(qd, ref["qd"]) < 2e-5


@pytest.mark.parametrize("prec,path,axis", [(cf.FP64, cf.PATH_QS, False), (cf.FP64, cf.PATH_QS, True),
                                            (cf.MIXED, cf.PATH_QS, False), (cf.MIXED, cf.PATH_AUTO, False)])
def test_im3m128_first_scft_iteration(im3m128, prec, path, axis):
    """First SCFT iteration at the timed workload: Q, F, phi_A, phi_B against
    the C oracle, 1e-12 (FP64) / 2e-5 relative (mixed); path QS also as the
    per-axis chain (the bench's SCFT solve)."""
    from cfft_hip.scft import SCFTSolver
    print(f"Im-3m 128^3 first SCFT iteration prec={prec} path={path} axis={axis}", flush=True)
    tol = 1e-12 if prec == cf.FP64 else 2e-5
    s = SCFTSolver(im3m128["N"], im3m128["ops_s"], 4.0 * np.eye(3), 0.3, 20.0, 100, prec, path=path,
                   axis_split=axis)
    s.set_fields(im3m128["wA"], im3m128["wB"])
    g = s.iterate()
    ref = im3m128["ref"]
    assert abs(g["Q"] - ref["Q"]) < tol * ref["Q"]
    assert abs(g["Qd"] - ref["Q"]) < tol * ref["Q"]
    assert abs(g["F"] - ref["F"]) < tol * max(1.0, abs(ref["F"]))
    phiA, phiB = s.densities()
    assert _rel(phiA, ref["phiA"]) < tol
    assert _rel(phiB, ref["phiB"]) < tol


# ---------------------------------------------------------------------------
# Ia-3d
# ---------------------------------------------------------------------------

def _gyroid(N, cell=3.8, amp=4.0, noise=0.05, seed=7):
    """Double-gyroid {211} start (the harmonic Ia-3d keeps), w_A low on the
    networks, seeded symmetrised noise."""
    from cfft_hip.scft import double_gyroid_211, symmetrize_subgrid
    ops, ops_s, shift, R, t = group_setup(230, N)
    Lv = O.auto_L(t)
    rng = np.random.default_rng(seed)
    g = double_gyroid_211(N, [s * n for s, n in zip(shift, N)], cell)
    g0 = np.ascontiguousarray(g[tuple(slice(0, None, int(l)) for l in Lv)])
    nA = symmetrize_subgrid(rng.normal(0.0, noise, N), ops_s, Lv)
    nB = symmetrize_subgrid(rng.normal(0.0, noise, N), ops_s, Lv)
    return ops_s, R, t, -amp * g0 + nA, amp * g0 + nB


@pytest.mark.parametrize("path", [cf.PATH_AUTO, cf.PATH_QS])
def test_ia3d128_mixed_vs_fp64_iterations(path):
    """Ia-3d 128^3, cell 3.8, f 0.35, chiN 18, Ns 200: four iterations (simple
    mixing) from the gyroid start in mixed precision and in FP64; Q, F and the
    densities agree to 2e-5 relative at every iteration."""
    from cfft_hip.scft import SCFTSolver
    print(f"Ia-3d 128^3 mixed vs FP64 path={path}", flush=True)
    N = (128, 128, 128)
    ops_s, R, t, wA, wB = _gyroid(N)
    lat = 3.8 * np.eye(3)
    s64 = SCFTSolver(N, ops_s, lat, 0.35, 18.0, 200, cf.FP64)
    s32 = SCFTSolver(N, ops_s, lat, 0.35, 18.0, 200, cf.MIXED, path=path)
    for s in (s64, s32):
        s.set_mixing(lam=1.0, lam_simple=0.2, simple_above=-1.0)  # simple mixing only
        s.set_fields(wA, wB)
    for it in range(4):
        a, b = s64.iterate(), s32.iterate()
        print(f"  it {it}: F64 {a['F']:.12f} dF {abs(a['F'] - b['F']):.2e} dQ/Q {abs(a['Q'] - b['Q']) / a['Q']:.2e}",
              flush=True)
        assert abs(a["Q"] - b["Q"]) < 2e-5 * a["Q"]
        assert abs(a["F"] - b["F"]) < 2e-5 * max(1.0, abs(a["F"]))
        for x, y in zip(s32.densities(), s64.densities()):
            assert _rel(x, y) < 2e-5


@pytest.mark.parametrize("path", [cf.PATH_AUTO, cf.PATH_QS])
def test_ia3d_mixed_scft_iterations_vs_oracle(path):
    """Mixed-precision SCFT iterations (Ia-3d 24^3, gyroid start, simple then
    Anderson mixing) against the FP64 oracle: before each of 5 iterations the
    solver's current fields are read back and the oracle sweeps them; Q, F
    and the densities of the mixed iteration agree within 2e-5 relative.  (The
    oracle evaluates the solver's own fields, so the check measures the
    precision of each iteration; two trajectories that mix independently
    drift apart through the Anderson least squares, which is not a precision
    error.  Mixing itself is checked in FP64.)"""
    from cfft_hip.scft import SCFTSolver
    N = (24, 24, 24)
    f, chiN, ns = 0.35, 18.0, 40
    lat = 3.8 * np.eye(3)
    ops_s, R, t, wA, wB = _gyroid(N)
    s = SCFTSolver(N, ops_s, lat, f, chiN, ns, cf.MIXED, path=path, m_hist=4)
    s.set_mixing(lam=1.0, lam_simple=0.2, simple_above=1e9)
    s.set_fields(wA, wB)
    spec = O.calc_spectral_asu(R, t, N)
    table = O.m2_backward_table(spec, R, t, N)
    diffuse = lambda x: O.m2_scft(x, spec, R, t, N, 1.0 / ns, lat, table)
    for it in range(5):
        a, b = (np.asarray(x).reshape(wA.shape, order="F") for x in s.fields())
        g = s.iterate()
        out = O.sweep(a, b, f, ns, diffuse)
        F = O.free_energy(a, b, out["phiA"], out["phiB"], out["Q"], chiN)
        phiA, phiB = s.densities()
        print(f"  it {it}: dQ/Q {abs(g['Q'] - out['Q']) / out['Q']:.2e} dF {abs(g['F'] - F):.2e} "
              f"dphiA {_rel(phiA, out['phiA']):.2e}", flush=True)
        assert abs(g["Q"] - out["Q"]) < 2e-5 * out["Q"]
        assert abs(g["F"] - F) < 2e-5 * max(1.0, abs(F))
        assert _rel(phiA, out["phiA"]) < 2e-5
        assert _rel(phiB, out["phiB"]) < 2e-5


def test_ia3d_ordered_gyroid_mixed_vs_fp64_solve():
    """The Ia-3d melt (f 0.35, chiN 18, Ns 200) in a cell near the gyroid's
    own period (9.0; at 3.8 it relaxes to the disordered state), 64^3: the
    FP64 solve reaches an ordered gyroid (F well below chiN f (1 - f), phi_A
    from < 0.1 to > 0.9); the mixed-precision solve to residual 1e-6 lands on
    the same solution (|dF| < 2e-5, max |dphi| < 1e-4; measured 8e-7 and
    2e-6), and a mixed stage to 1e-5 followed by an FP64 finish to 1e-8
    matches the FP64 solve to 2e-9 (two FP64 solves stopped at residual 1e-8
    from different fields: measured 1.8e-10)."""
    from cfft_hip.scft import QS_MIXED_TOL, SCFTSolver, default_path
    N = (64, 64, 64)
    cell = 9.0
    ops_s, R, t, wA, wB = _gyroid(N, cell=cell)
    lat = cell * np.eye(3)
    mk = lambda prec, path=None: SCFTSolver(N, ops_s, lat, 0.35, 18.0, 200, prec, path=path)
    s64 = mk(cf.FP64)
    s64.set_fields(wA, wB)
    i64, o64 = s64.solve(tol=1e-8, max_iter=600)
    p64 = s64.densities()
    assert o64["res"] < 1e-8
    assert o64["F"] < 18.0 * 0.35 * 0.65 - 0.2
    assert p64[0].min() < 0.1 and p64[0].max() > 0.9
    s32 = mk(cf.MIXED)
    s32.set_fields(wA, wB)
    i32, o32 = s32.solve(tol=1e-6, max_iter=600)
    dphi = max(float(np.max(np.abs(a - b))) for a, b in zip(s32.densities(), p64))
    print(f"  FP64 {i64} it F {o64['F']:.10f}; mixed {i32} it res {o32['res']:.1e} "
          f"dF {abs(o32['F'] - o64['F']):.2e} dphi {dphi:.2e}", flush=True)
    assert o32["res"] < 1e-6
    assert abs(o32["F"] - o64["F"]) < 2e-5
    assert dphi < 1e-4
    hand = mk(cf.MIXED, default_path(N, ops_s, lat, cf.MIXED, QS_MIXED_TOL))
    hand.set_fields(wA, wB)
    hand.solve(tol=QS_MIXED_TOL, max_iter=600)
    fin = mk(cf.FP64)
    fin.continue_from(hand)
    _, of = fin.solve(tol=1e-8, max_iter=600)
    print(f"  mixed then FP64: dF {abs(of['F'] - o64['F']):.2e}", flush=True)
    assert of["res"] < 1e-8 and abs(of["F"] - o64["F"]) < 2e-9


# ---------------------------------------------------------------------------
# the 64-instance shape: every grid size with every group
# ---------------------------------------------------------------------------

SIZES = (32, 48, 64, 96)
GROUPS = (47, 70, 223, 229, 230)


def _instances():
    from cfft_hip.batch import Instance
    rng = np.random.default_rng(64)
    return [Instance((m, m, m), sg, 4.0, float(rng.uniform(0.28, 0.4)), float(rng.uniform(18.0, 26.0)), 100,
                     int(rng.integers(1 << 30)))
            for m in SIZES for sg in GROUPS]


def test_batch_shape_mixed_vs_fp64_iterations():
    """Each of the 20 (size, group) instances: three iterations from its start
    in mixed precision and in FP64 on one grouped batch each; per instance F
    and the densities agree to 2e-5 relative at every iteration."""
    from cfft_hip.batch import groups_of, start_fields
    from cfft_hip.plans import find_optimal_shift
    from cfft_hip.scft import SCFTBatch, default_path
    from cfft_hip.symmetry_ops import get_ops
    insts = _instances()
    for g in groups_of(insts):
        i0 = insts[g[0]]
        print(f"batch shape: N={i0.N} sg={i0.sg} instances={len(g)}", flush=True)
        ops = get_ops(i0.sg, 3, i0.N)
        shift, ops_s = find_optimal_shift(ops, i0.N)
        lat = i0.cell * np.eye(3)
        starts = [start_fields(insts[k], ops_s, shift) for k in g]
        runs = {}
        for prec in (cf.FP64, cf.MIXED):
            b = SCFTBatch(i0.N, ops_s, lat, [insts[k].f for k in g], [insts[k].chiN for k in g], i0.ns, prec,
                          path=default_path(i0.N, ops_s, lat, prec, 1e-5))
            b.set_mixing(lam=1.0, lam_simple=0.2, simple_above=-1.0)  # simple mixing only
            for i, (wA, wB) in enumerate(starts):
                b.set_fields(i, wA, wB)
            its = []
            for _ in range(3):
                its.append((b.iterate(), [b.densities(i) for i in range(b.n)]))
            runs[prec] = its
            del b
        for (o64, p64), (o32, p32) in zip(runs[cf.FP64], runs[cf.MIXED]):
            for i in range(len(g)):
                assert abs(o64[i]["F"] - o32[i]["F"]) < 2e-5 * max(1.0, abs(o64[i]["F"]))
                for x, y in zip(p32[i], p64[i]):
                    assert _rel(x, y) < 2e-5


def test_batch_shape_mixed_solve_vs_fp64_solve():
    """The 20 instances solved concurrently (grouped contexts, hipGraph sweeps,
    host threads): mixed precision to residual 1e-6 and FP64 to 1e-8 from the
    same starts.  All but at most two instances converge in both precisions
    (an instance near its order-disorder transition may not within 400
    iterations); for every instance that both solves bring below their
    tolerance, |dF| < 2e-5 (F is stationary, so its error is second order in
    the residual, plus the precision error)."""
    from cfft_hip import batch
    insts = _instances()
    r64, _, _ = batch.solve_batch(insts, precision=cf.FP64, tol=1e-8, max_iter=400, workers=8)
    r32, _, _ = batch.solve_batch(insts, precision=cf.MIXED, tol=1e-6, max_iter=400, workers=8)
    n_both = 0
    for inst, a, b in zip(insts, r64, r32):
        print(f"  N={inst.N[0]} sg={inst.sg} f={inst.f:.3f} chiN={inst.chiN:.2f}: fp64 {a['iters']} it "
              f"res {a['res']:.1e}, mixed {b['iters']} it res {b['res']:.1e}, dF {abs(a['F'] - b['F']):.2e}",
              flush=True)
        if a["res"] < 1e-8 and b["res"] < 1e-6:
            n_both += 1
            assert abs(a["F"] - b["F"]) < 2e-5 * max(1.0, abs(a["F"]))
    assert n_both >= len(insts) - 2


@pytest.mark.parametrize("sg", GROUPS)
def test_batch_groups_first_mixed_iteration_vs_oracle(sg):
    """One instance per group at 32^3, Ns 100, mixed precision: the first
    iteration's Q, F and densities against the C oracle (FP64), 2e-5."""
    from cfft_hip.batch import Instance, start_fields
    from cfft_hip.scft import SCFTSolver, default_path
    N = (32, 32, 32)
    ops, ops_s, shift, R, t = group_setup(sg, N)
    inst = Instance(N, sg, 4.0, 0.33, 22.0, 100, 1234 + sg)
    wA, wB = start_fields(inst, ops_s, shift)
    lat = 4.0 * np.eye(3)
    s = SCFTSolver(N, ops_s, lat, inst.f, inst.chiN, inst.ns, cf.MIXED,
                   path=default_path(N, ops_s, lat, cf.MIXED, 1e-5))
    s.set_fields(wA, wB)
    g = s.iterate()
    cs = cstep.CStep(R, t, N, 1.0 / inst.ns, lat, nthreads=THREADS)
    ref = cstep.first_iteration(cs, wA, wB, inst.f, inst.chiN, inst.ns)
    assert abs(g["Q"] - ref["Q"]) < 2e-5 * ref["Q"]
    assert abs(g["F"] - ref["F"]) < 2e-5 * max(1.0, abs(ref["F"]))
    phiA, phiB = s.densities()
    assert _rel(phiA, ref["phiA"]) < 2e-5
    assert _rel(phiB, ref["phiB"]) < 2e-5
