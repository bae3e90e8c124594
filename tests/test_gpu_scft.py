"""GPU contour sweeps (cfft_sweep_dev, persistent cooperative kernel on cubic
grids, one launch per phase otherwise) vs the oracle's Strang propagation."""
import numpy as np
import pytest

import cfft_hip as cf
import oracle as O
from conftest import group_setup

pytestmark = pytest.mark.gpu


def sweep_case(sg, N, f, ns, prec, seed=3, persistent=0, fuse=1, graph=1, path=cf.PATH_AUTO):
    import torch
    from cfft_hip.scft import PropagatorSweep
    ops, ops_s, shift, R, t = group_setup(sg, N)
    Lv = O.auto_L(t)
    lat = 4.0 * np.eye(3)
    wA = 2.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=seed), Lv)
    wB = -1.5 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=seed + 1), Lv)
    sw = PropagatorSweep(N, ops_s, lat, f, ns, prec, path=path)
    sw.plan.set_option(cf._lib.OPT_PERSISTENT, persistent)
    sw.plan.set_option(cf._lib.OPT_FUSE, fuse)
    sw.plan.set_option(cf._lib.OPT_GRAPH, graph)
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


@pytest.mark.parametrize("persistent,fuse,graph", [(0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 1, 0)])
@pytest.mark.parametrize("sg,N", [(229, (32, 32, 32)), (230, (32, 32, 32)), (47, (32, 32, 32)),
                                  (70, (32, 32, 32)), (10, (16, 16, 16)), (71, (24, 32, 48))])
def test_sweep_fp64_vs_oracle(sg, N, persistent, fuse, graph):
    print(f"sweep sg={sg} N={N} persistent={persistent} fuse={fuse} graph={graph}", flush=True)
    q, qd, rq, rqd, (Q1, Q2) = sweep_case(sg, N, 0.3, 20, cf.FP64, persistent=persistent, fuse=fuse,
                                          graph=graph)
    assert np.max(np.abs(q - rq)) / np.max(np.abs(rq)) < 1e-12
    assert np.max(np.abs(qd - rqd)) / np.max(np.abs(rqd)) < 1e-12
    # Q from q(1) and from q+(1) agree (symmetry of the chain)
    assert abs(Q1 - Q2) / abs(Q1) < 1e-12
    assert abs(Q1 - float(np.mean(rq))) / abs(Q1) < 1e-12


@pytest.mark.parametrize("path", [cf.PATH_AUTO, cf.PATH_QR, cf.PATH_QS])
def test_live_stage_timing(path):
    """cfft_time_stage_dev: positive per-launch times for P5+P1 and P2,
    other stages rejected, and the plan still sweeps to the oracle afterwards
    (the timed launches only touch the scratch field and the channel array)."""
    import ctypes as C
    import torch
    from cfft_hip.scft import PropagatorSweep
    N, sg = (32, 32, 32), 229
    ops, ops_s, shift, R, t = group_setup(sg, N)
    Lv = O.auto_L(t)
    wA = 2.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=3), Lv)
    wB = -1.5 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=4), Lv)
    sw = PropagatorSweep(N, ops_s, 4.0 * np.eye(3), 0.3, 20, cf.FP64, path=path)
    sw.set_fields(wA, wB)
    lib = cf._lib.lib()
    scratch = sw.q.clone()
    for name in ("p51_unfold_fold", "p2q_plane_pair" if path == cf.PATH_QS else "p2_fft_yz"):
        ms = C.c_double()
        cf._lib.check(lib.cfft_time_stage_dev(sw.plan.handle, cf._lib.STAGES.index(name),
                                              C.c_void_p(scratch.data_ptr()), sw._E1, 5,
                                              C.c_void_p(sw.stream()), C.byref(ms)))
        assert 0.0 < ms.value < 10.0
    with pytest.raises(cf.CFFTError):
        cf._lib.check(lib.cfft_time_stage_dev(sw.plan.handle, cf._lib.STAGES.index("p3_recon"),
                                              C.c_void_p(scratch.data_ptr()), sw._E1, 5,
                                              C.c_void_p(sw.stream()), C.byref(ms)))
    sw.reset()
    sw.run()
    torch.cuda.synchronize()
    q, _ = sw.fields()
    spec = O.calc_spectral_asu(R, t, N)
    table = O.m2_backward_table(spec, R, t, N)
    lat = 4.0 * np.eye(3)
    diffuse = lambda x: O.m2_scft(x, spec, R, t, N, 1.0 / 20, lat, table)
    ref = O.propagate(np.ones_like(wA), lambda j: wA if j < 6 else wB, 20, 1.0 / 20, diffuse)[-1]
    assert np.max(np.abs(q - ref)) / np.max(np.abs(ref)) < 1e-12


@pytest.mark.parametrize("fuse,graph", [(1, 1), (0, 0), (1, 0)])
@pytest.mark.parametrize("sg,N", [(229, (32, 32, 32)), (230, (32, 32, 32)), (47, (32, 32, 32)),
                                  (70, (32, 32, 32)), (223, (24, 24, 24))])
def test_sweep_plane_pair_q_vs_oracle(sg, N, fuse, graph):
    """Contour sweeps on path QS (plane-pair kernel + fused unfold/fold chain,
    eager or replayed from a hipGraph) vs the oracle's Strang propagation."""
    print(f"QS sweep sg={sg} N={N} fuse={fuse} graph={graph}", flush=True)
    q, qd, rq, rqd, (Q1, Q2) = sweep_case(sg, N, 0.3, 20, cf.FP64, fuse=fuse, graph=graph, path=cf.PATH_QS)
    assert np.max(np.abs(q - rq)) / np.max(np.abs(rq)) < 1e-12
    assert np.max(np.abs(qd - rqd)) / np.max(np.abs(rqd)) < 1e-12
    assert abs(Q1 - Q2) / abs(Q1) < 1e-12


@pytest.mark.parametrize("path", [cf.PATH_AUTO, cf.PATH_QS])
def test_sweep_mixed_vs_fp64(path):
    """Mixed precision (FP32 data; on path QS also FP32 twiddle, fold and Q
    arithmetic from the FP64 tables) against the FP64 oracle."""
    q, qd, rq, rqd, (Q1, Q2) = sweep_case(229, (32, 32, 32), 0.3, 20, cf.MIXED, path=path)
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
    Anderson) against the oracle.  FP64 bar: 1e-12 on the first iteration;
    afterwards the Anderson least-squares step amplifies the rounding
    differences between the two FFT implementations, so 1e-9."""
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
        tol = 1e-12 if it == 0 else 1e-9
        g = solver.iterate()
        a = w[:M3].reshape(wA.shape, order="F")
        b = w[M3:].reshape(wA.shape, order="F")
        out = O.sweep(a, b, f, ns, diffuse)
        tA, tB, res = O.field_update(a, b, out["phiA"], out["phiB"], chiN)
        F = O.free_energy(a, b, out["phiA"], out["phiB"], out["Q"], chiN)
        assert abs(g["Q"] - out["Q"]) < tol * out["Q"]
        assert abs(g["Qd"] - out["Q"]) < tol * out["Q"]
        assert abs(g["F"] - F) < tol * max(1.0, abs(F))
        # the residual is a small difference of O(1) quantities: absolute bar
        assert abs(g["res"] - res) < 1e-10 + 1e-9 * res
        assert g["anderson"] == (1.0 if it > 0 else 0.0)
        phiA, phiB = solver.densities()
        assert np.max(np.abs(phiA - out["phiA"])) < tol
        assert np.max(np.abs(phiB - out["phiB"])) < tol
        d = np.concatenate([(tA - a).ravel(order="F"), (tB - b).ravel(order="F")])
        w = mix.step(w, d, res)
        gw = np.concatenate([x.ravel(order="F") for x in solver.fields()])
        assert np.max(np.abs(gw - w)) < 10 * tol * np.max(np.abs(w))


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


def test_scft_converges_bcc_fp64_and_mixed():
    """Im-3m 32^3, f = 0.3, chiN = 20, Ns = 50 from BCC first-star fields.
    FP64 converges to residual 1e-8.  Mixed precision converges to 1e-6, where
    its F is within 1e-5 of the FP64 solution (the free-energy error tracks the
    residual); an FP64 finish from the mixed fields reaches 1e-8 in fewer FP64
    iterations than the cold start and matches the FP64 F to 1e-8."""
    from cfft_hip.scft import SCFTSolver
    N = (32, 32, 32)
    ops_s, R, t, Lv, wA, wB = _bcc_fields(229, N)
    lat = 4.0 * np.eye(3)
    s64 = SCFTSolver(N, ops_s, lat, 0.3, 20.0, 50, cf.FP64, m_hist=10)
    s64.set_fields(wA, wB)
    it64, o64 = s64.solve(tol=1e-8, max_iter=400)
    print(f"  fp64 iters={it64} F={o64['F']:.12f} res={o64['res']:.3e}", flush=True)
    assert o64["res"] < 1e-8
    assert abs(o64["phi_total"] - 1.0) < 1e-8
    assert o64["F"] < O.disordered_free_energy(0.3, 20.0)
    ph64 = s64.densities()
    s32 = SCFTSolver(N, ops_s, lat, 0.3, 20.0, 50, cf.MIXED, m_hist=10)
    s32.set_fields(wA, wB)
    it32, o32 = s32.solve(tol=1e-6, max_iter=400)
    print(f"  mixed iters={it32} F={o32['F']:.12f} res={o32['res']:.3e}", flush=True)
    assert o32["res"] < 1e-6
    assert abs(o32["F"] - o64["F"]) < 1e-5
    assert np.max(np.abs(s32.densities()[0] - ph64[0])) < 1e-4
    fin = SCFTSolver(N, ops_s, lat, 0.3, 20.0, 50, cf.FP64, m_hist=10)
    fin.set_fields(*s32.fields())
    itf, of = fin.solve(tol=1e-8, max_iter=400)
    print(f"  finish iters={itf} F={of['F']:.12f} res={of['res']:.3e}", flush=True)
    assert of["res"] < 1e-8
    assert itf < it64
    assert abs(of["F"] - o64["F"]) < 1e-8


@pytest.mark.parametrize("prec", [cf.FP64, cf.MIXED])
def test_graph_replays_match_eager(prec):
    """Chains replayed from captured hipGraphs equal the eager launches bit for
    bit, on every replay, with eager work and synchronisations in between."""
    import torch
    from cfft_hip.scft import PropagatorSweep
    N = (32, 32, 32)
    ops, ops_s, shift, R, t = group_setup(229, N)
    Lv = O.auto_L(t)
    wA = 2.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=11), Lv)
    wB = -1.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=12), Lv)
    sw = PropagatorSweep(N, ops_s, 4.0 * np.eye(3), 0.3, 30, prec)
    sw.set_fields(wA, wB)
    sw.plan.set_option(cf._lib.OPT_GRAPH, 0)
    sw.reset()
    sw.run()
    torch.cuda.synchronize()
    eager = [x.copy() for x in sw.fields()]
    sw.plan.set_option(cf._lib.OPT_GRAPH, 1)
    f0 = O.extract_subgrid(O.make_symmetric(R, t, N, seed=13), Lv)
    for rep in range(4):
        sw.reset()
        sw.run()
        torch.cuda.synchronize()
        got = sw.fields()
        for a, b in zip(got, eager):
            assert np.array_equal(a, b), f"replay {rep} differs"
        sw.plan.forward(np.stack([f0, f0]))  # eager work on the same plan between replays
    # split runs (two graphs, different step ranges) equal the whole sweep
    sw.reset()
    sw.run(12, start=0)
    sw.run(18, start=12)
    torch.cuda.synchronize()
    for a, b in zip(sw.fields(), eager):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("side_stream", [False, True])
@pytest.mark.parametrize("graph", [0, 1])
def test_sweep_ordered_with_caller_stream(side_stream, graph):
    """*_dev calls run on exactly the caller's stream (NULL = the null stream,
    PyTorch's default): set_fields -> reset -> run issued with no host
    synchronisation, each reset queued behind a long matrix product on the
    same stream, gives the eager result bit for bit on every repetition,
    on the default stream and on a side stream."""
    import torch
    from cfft_hip.scft import PropagatorSweep
    print(f"ordered sweep side_stream={side_stream} graph={graph}", flush=True)
    N = (32, 32, 32)
    ops, ops_s, shift, R, t = group_setup(229, N)
    Lv = O.auto_L(t)
    wA = 2.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=21), Lv)
    wB = -1.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=22), Lv)
    ref = PropagatorSweep(N, ops_s, 4.0 * np.eye(3), 0.3, 30, cf.FP64)
    ref.plan.set_option(cf._lib.OPT_GRAPH, 0)
    ref.set_fields(wA, wB)
    torch.cuda.synchronize()
    ref.reset()
    torch.cuda.synchronize()
    ref.run()
    torch.cuda.synchronize()
    eager = [x.copy() for x in ref.fields()]
    sw = PropagatorSweep(N, ops_s, 4.0 * np.eye(3), 0.3, 30, cf.FP64)
    sw.plan.set_option(cf._lib.OPT_GRAPH, graph)
    big = torch.randn(3072, 3072, device="cuda")
    torch.cuda.synchronize()
    stream = torch.cuda.Stream() if side_stream else torch.cuda.default_stream()
    outs = []
    with torch.cuda.stream(stream):
        sw.set_fields(wA, wB)
        for _ in range(3):
            big = (big @ big).clamp_(-1.0, 1.0)  # keeps the stream busy ahead of the reset
            sw.reset()
            sw.run()
            outs.append(sw.q.clone())
        stream.synchronize()
    for k, q in enumerate(outs):
        got = q.cpu().numpy()
        for b in range(2):
            assert np.array_equal(got[b].reshape(sw.M, order="F"), eager[b]), f"repetition {k} field {b}"


def test_batched_instances_match_sequential():
    """Concurrent instances (own plan, context and stream each, hipGraph sweeps)
    give the same solutions as solving them one at a time."""
    from cfft_hip import batch
    insts = batch.draw_instances(6, seed=5, sizes=(16, 24, 32), ns=40)
    res, t_plan, t_solve = batch.solve_batch(insts, precision=cf.FP64, tol=1e-8, max_iter=300, workers=6)
    print(f"  batch of {len(insts)}: plan {t_plan:.2f}s solve {t_solve:.2f}s", flush=True)
    for inst, r in zip(insts, res):
        assert r["res"] < 1e-8, (inst, r)
        seq, _, _ = batch.solve_batch([inst], precision=cf.FP64, tol=1e-8, max_iter=300, workers=1)
        assert seq[0]["iters"] == r["iters"]
        assert abs(seq[0]["F"] - r["F"]) < 1e-12 * max(1.0, abs(r["F"]))


def test_solver_update_equals_fresh_solver():
    """Re-pointing a solver at another cell, f and chiN (kernel update +
    context parameters) gives the fresh solver's iterations bit for bit."""
    from cfft_hip.scft import SCFTSolver
    N = (16, 16, 16)
    ops_s, R, t, Lv, wA, wB = _bcc_fields(229, N)
    s = SCFTSolver(N, ops_s, 4.0 * np.eye(3), 0.3, 20.0, 40, cf.FP64)
    s.set_fields(wA, wB)
    s.solve(tol=1e-6, max_iter=50)
    s.update(lattice=3.7 * np.eye(3), f=0.34, chiN=17.0)
    s.set_fields(wA, wB)
    a = [s.iterate() for _ in range(5)]
    fresh = SCFTSolver(N, ops_s, 3.7 * np.eye(3), 0.34, 17.0, 40, cf.FP64)
    fresh.set_fields(wA, wB)
    b = [fresh.iterate() for _ in range(5)]
    for x, y in zip(a, b):
        assert x == y


@pytest.mark.parametrize("sg,N,prec", [(229, (16, 16, 16), cf.FP64), (229, (16, 16, 16), cf.MIXED),
                                       (47, (16, 16, 16), cf.FP64), (70, (32, 32, 32), cf.FP64)])
def test_scft_batch_equals_single_instances(sg, N, prec):
    """Three melts (different f, chiN and start fields) in one multi-instance
    context: every iteration's results, the fields, the densities and the
    per-instance stop of solve() are those of one-instance solvers bit for
    bit (the sweep computes each field independently; every reduction has the
    same block decomposition)."""
    from cfft_hip import batch
    from cfft_hip.scft import SCFTBatch, SCFTSolver
    print(f"scft batch sg={sg} N={N} prec={prec}", flush=True)
    shift, ops_s = cf.find_optimal_shift(cf.get_ops(sg, 3, N), N)
    params = [(0.30, 16.0, 11), (0.42, 22.0, 12), (0.35, 28.0, 13)]
    insts = [batch.Instance(N, sg, 4.0, f, x, 40, seed) for f, x, seed in params]
    starts = [batch.start_fields(i, ops_s, shift) for i in insts]
    lat = 4.0 * np.eye(3)
    multi = SCFTBatch(N, ops_s, lat, [p[0] for p in params], [p[1] for p in params], 40, prec)
    singles = [SCFTSolver(N, ops_s, lat, f, x, 40, prec) for f, x, _ in params]
    assert multi.plan.path == singles[0].plan.path
    for i, (wA, wB) in enumerate(starts):
        multi.set_fields(i, wA, wB)
        singles[i].set_fields(wA, wB)
    for _ in range(4):
        a = multi.iterate()
        b = [s.iterate() for s in singles]
        assert a == b
    for i, s in enumerate(singles):
        for x, y in zip(multi.fields(i) + multi.densities(i), s.fields() + s.densities()):
            assert np.array_equal(x, y)
    # solve(): instances stop at their own convergence
    tol = 1e-6 if prec == cf.FP64 else 1e-4
    for i, (wA, wB) in enumerate(starts):
        multi.set_fields(i, wA, wB)
        singles[i].set_fields(wA, wB)
    its, outs = multi.solve(tol=tol, max_iter=120)
    ref = [s.solve(tol=tol, max_iter=120) for s in singles]
    print(f"  iterations {its}", flush=True)
    assert its == [r[0] for r in ref]
    assert outs == [r[1] for r in ref]
    for i, s in enumerate(singles):
        assert np.array_equal(multi.fields(i)[0], s.fields()[0])


def test_scft_batch_rejects_single_calls():
    """A context of several instances takes the indexed calls only; the plan's
    batch must be twice the instance count."""
    from cfft_hip.scft import SCFTBatch
    N = (16, 16, 16)
    _, ops_s = cf.find_optimal_shift(cf.get_ops(229, 3, N), N)
    m = SCFTBatch(N, ops_s, 4.0 * np.eye(3), [0.3, 0.4], [20.0, 20.0], 20, cf.FP64)
    z = np.zeros(m.M3)
    with pytest.raises(cf.CFFTError):
        cf._lib.check(cf._lib.lib().cfft_scft_set_fields(m._h, cf._lib.f64p(z), cf._lib.f64p(z)))
    with pytest.raises(cf.CFFTError):
        m.set_fields(2, z, z)
    import ctypes as C
    h = C.c_void_p()
    f = np.array([0.3, 0.4, 0.5])
    with pytest.raises(cf.CFFTError):
        cf._lib.check(cf._lib.lib().cfft_scft_create_multi(m.plan.handle, 3, cf._lib.f64p(f), cf._lib.f64p(f),
                                                           20, 4, C.byref(h)))


def test_batched_grouped_equals_ungrouped():
    """solve_batch with instances grouped by (grid, group) into shared
    contexts gives every instance the iterations and results of solving it
    in a context of its own."""
    from cfft_hip import batch
    insts = batch.draw_instances(8, seed=21, sizes=(16, 24), groups=(229, 47), ns=40)
    assert len(batch.groups_of(insts)) < len(insts)
    g, _, _ = batch.solve_batch(insts, precision=cf.FP64, tol=1e-7, max_iter=200, workers=4, group=True)
    u, _, _ = batch.solve_batch(insts, precision=cf.FP64, tol=1e-7, max_iter=200, workers=4, group=False)
    assert g == u


@pytest.mark.parametrize("path", [cf.PATH_QS, cf.PATH_AUTO])
def test_schedule_options_do_not_change_results(path):
    """Every launch schedule the tuner may pick (tile rows 8, 4, 2; graph
    replay on and off) gives the same sweep bit for bit; tile heights that are
    not 1, 2, 4, 8 dividing H1 are refused."""
    import torch
    from cfft_hip.scft import PropagatorSweep
    from cfft_hip.tune import schedule_candidates
    N = (32, 32, 32)
    ops, ops_s, shift, R, t = group_setup(229, N)
    Lv = O.auto_L(t)
    wA = 2.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=21), Lv)
    wB = -1.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=22), Lv)
    sw = PropagatorSweep(N, ops_s, 4.0 * np.eye(3), 0.3, 20, cf.FP64, path=path)
    sw.set_fields(wA, wB)
    ref = None
    for opt in schedule_candidates(sw.plan.H[1]):
        for k, v in opt.items():
            sw.plan.set_option(k, v)
        sw.reset()
        sw.run()
        torch.cuda.synchronize()
        got = [x.copy() for x in sw.fields()]
        if ref is None:
            ref = got
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), opt
    for bad in (3, 16, -1):
        with pytest.raises(cf.CFFTError):
            sw.plan.set_option(cf._lib.OPT_TILE_ROWS, bad)


def test_tuned_batch_equals_untuned():
    """solve_batch(tune=True) times each group's schedules, applies the
    fastest and gives the untuned iterations and results."""
    from cfft_hip import batch
    insts = batch.draw_instances(5, seed=23, sizes=(16, 24), groups=(229, 47), ns=30)
    t, _, _ = batch.solve_batch(insts, precision=cf.FP64, tol=1e-7, max_iter=150, workers=3, tune=True)
    u, _, _ = batch.solve_batch(insts, precision=cf.FP64, tol=1e-7, max_iter=150, workers=3)
    print("  schedules", sorted({r["schedule"] for r in t}), flush=True)
    for a, b in zip(t, u):
        assert a.pop("schedule").startswith("rows")
        assert a == b


def test_warm_start_pipeline_small():
    """End-to-end at 16^3 with a small network: generate converged solutions,
    train, predict, decode and solve from the prediction (fallback counted)."""
    from cfft_hip import warmstart as W
    ps = W.draw(24, seed=3)
    X, T, ok, its, els, groups = W.generate(ps, (16, 16, 16), workers=4, max_iter=400)
    assert ok.sum() >= 12
    ws = W.WarmStart(dims=(6, 64, 64, 64, 4 * W.N_STARS), seed=1)
    hist = ws.fit(X[ok], T[ok], epochs=200, batch=8, lr=2e-3)
    assert hist[-1] < hist[0]
    rows = W.evaluate(ws, W.draw(4, seed=4), groups, max_iter=400, workers=2)
    summ = W.summary(rows)
    print("  warm start (16^3, small):", summ, flush=True)
    assert summ["n"] == 4


def test_trained_warm_start_gains_over_cold_start():
    """The committed network (trained on 2000 converged 64^3 solutions of
    Im-3m, Ia-3d and Fddd) against the seeded cold start on held-out parameter
    sets, two per group: same solver, residual 1e-8.  The network's guesses
    reach the cold start's solutions (|dF| < 1e-6) with no fallback, in at
    most 0.6x its iterations and at most 0.5x its wall time (inference and
    decoding included)."""
    import os
    from cfft_hip import warmstart as W
    path = os.path.join(os.path.dirname(__file__), "..", "crystallographicfft.jl_amd", "models",
                        "warmstart.npz")
    ws = W.WarmStart.load(path)
    assert ws.meta["grid"] == 64 and ws.meta["n_train"] == 2000
    held_out = W.draw(40, seed=2)
    ps = [p for sg in W.GROUPS for p in [q for q in held_out if q.sg == sg][:2]]
    groups = {sg: W.Group(sg, (64, 64, 64)) for sg in W.GROUPS}
    rows = W.evaluate(ws, ps, groups, workers=3)
    summ = W.summary(rows)
    print("  trained warm start (64^3):", {k: summ[k] for k in ("cold_iters_mean", "warm_iters_mean",
                                                               "time_ratio", "max_abs_dF")}, flush=True)
    assert summ["fallbacks"] == 0 and summ["other_solution"] == 0
    assert summ["max_abs_dF"] < 1e-6
    assert summ["iter_ratio"] < 0.6
    assert summ["time_ratio"] <= 0.5


def test_continue_from_hands_over_the_mixing():
    """cfft_scft_copy_state: a context that takes over another's fields and
    Anderson history continues its iterations bit for bit (FP64 to FP64,
    one- and two-instance contexts); a mixed-precision stage handed to an
    FP64 finish this way needs fewer FP64 iterations than restarting the
    mixing from its fields, and reaches the same solution."""
    from cfft_hip import batch
    from cfft_hip.scft import SCFTBatch, SCFTSolver
    N = (32, 32, 32)
    shift, ops_s = cf.find_optimal_shift(cf.get_ops(229, 3, N), N)
    lat = 4.0 * np.eye(3)
    insts = [batch.Instance(N, 229, 4.0, 0.3, 20.0, 40, s) for s in (31, 32)]
    starts = [batch.start_fields(i, ops_s, shift) for i in insts]
    a = SCFTSolver(N, ops_s, lat, 0.3, 20.0, 40, cf.FP64)
    a.set_fields(*starts[0])
    for _ in range(12):
        a.iterate()
    b = SCFTSolver(N, ops_s, lat, 0.3, 20.0, 40, cf.FP64)
    b.continue_from(a)
    for _ in range(6):
        assert a.iterate() == b.iterate()
    m1 = SCFTBatch(N, ops_s, lat, [0.3, 0.3], [20.0, 20.0], 40, cf.FP64)
    for i, w in enumerate(starts):
        m1.set_fields(i, *w)
    for _ in range(12):
        m1.iterate()
    m2 = SCFTBatch(N, ops_s, lat, [0.3, 0.3], [20.0, 20.0], 40, cf.FP64)
    m2.continue_from(m1)
    for _ in range(6):
        assert m1.iterate() == m2.iterate()
    # mixed stage to 1e-5, then FP64 to 1e-9: hand-off against restart
    s32 = SCFTSolver(N, ops_s, lat, 0.3, 20.0, 40, cf.MIXED, path=cf.PATH_QS)
    s32.set_fields(*starts[0])
    s32.solve(tol=1e-5, max_iter=300)
    f1 = SCFTSolver(N, ops_s, lat, 0.3, 20.0, 40, cf.FP64)
    f1.continue_from(s32)
    i1, o1 = f1.solve(tol=1e-9, max_iter=300)
    f2 = SCFTSolver(N, ops_s, lat, 0.3, 20.0, 40, cf.FP64)
    f2.set_fields(*s32.fields())
    i2, o2 = f2.solve(tol=1e-9, max_iter=300)
    print(f"  FP64 finish: {i1} iterations with the history, {i2} restarting", flush=True)
    assert o1["res"] < 1e-9 and o2["res"] < 1e-9
    assert i1 < i2
    assert abs(o1["F"] - o2["F"]) < 1e-9
