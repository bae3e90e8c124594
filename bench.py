"""Benchmark: diffusion steps/s on the Im-3m 128^3 SCFT propagator sweep.

Workload (one GPU): space group Im-3m (#229), 128^3 grid, cubic cell 4.0, AB
diblock f = 0.3, chiN = 20, Ns = 100 contour steps (ds = 0.01) for q and q+
batched together; w_A = +3 S(r) + noise, w_B = -3 S(r) + noise with S the BCC
{110} first-star cosine sum and seeded N(0, 0.05) noise, symmetrised over the
96 operations.  One bench "step" (--steps, --warmup) = one full contour sweep
of the workload: q and q+ each take Ns = 100 Strang steps
q <- e^{-w ds/2} D e^{-w ds/2} q through the fastest launch strategy (the
plane-pair Q chain or the centred KRFFT pipeline); `value` counts diffusion
steps of one field (2 Ns per sweep) per second, inputs resident in HBM.

Prints one JSON line (rank 0).  --gpus above 1 exits with a message: this
build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "crystallographicfft.jl_amd"))

import numpy as np  # noqa: E402

N = (128, 128, 128)
SG = 229
CELL = 4.0
F_A = 0.3
CHIN = 20.0
NS = 100
HBM_PEAK_GBS = 8000.0


def make_fields(ops_s, shift, Lv, seed=2024):
    from cfft_hip.scft import first_star_bcc, symmetrize_subgrid
    b = [s * n for s, n in zip(shift, N)]
    rng = np.random.default_rng(seed)
    S = first_star_bcc(N, b, CELL)
    sub = tuple(slice(0, None, int(l)) for l in Lv)
    S0 = np.ascontiguousarray(S[sub])
    nA = symmetrize_subgrid(rng.normal(0.0, 0.05, N), ops_s, Lv)
    nB = symmetrize_subgrid(rng.normal(0.0, 0.05, N), ops_s, Lv)
    return 3.0 * S0 + nA, -3.0 * S0 + nB


def stage_bytes(plan, e):
    """Algorithmic HBM bytes per launch of each stage (tables excluded):
    each distinct input array read once, each output written once; fields of
    e bytes, interaction factors E of 8 bytes in both precisions."""
    B = plan.batch
    eE = 8
    M3 = int(np.prod(plan.M))
    chan = plan.n_ch * int(np.prod(plan.H)) * 2 * e
    spec = plan.n_spec * 2 * e
    rep = plan.n_rep * 2 * e
    qr = plan.path == 5  # orbit-reduced Q chain
    return {
        "p1_fold_fftx": B * (M3 * (e + eE) + chan),        # f0, E -> channels
        "fft_x": B * 2 * chan, "p2_fft_yz": B * 2 * chan, "fft_z": B * 2 * chan,
        "p3_recon": B * (chan + spec),                     # channels -> F (+K)
        # F -> channels, or (QR) representative values -> channels
        "p4_irecon_ifft_zy": B * ((rep if qr else spec) + chan),
        "ifft_x": B * 2 * chan, "ifft_y": B * 2 * chan, "ifft_z": B * 2 * chan,
        "p5_ifftx_unfold": B * (chan + M3 * (e + eE)),     # channels, E -> f0
        # (QR) channels gathered at the representatives -> representative values
        "qmul": B * ((min(chan, plan.n_rep * plan.d * 2 * e) + rep) if qr else 2 * chan),
        # between two steps of a plain sweep: channels and E in, channels
        # out (q(s + 1) stays in registers; a stored sweep adds M3 e per field)
        "p51_unfold_fold": B * (2 * chan + M3 * eE),
        # (QS) channels -> FFT y, z, Q multiply, IFFT z, y -> channels
        "p2q_plane_pair": B * 2 * chan,
    }


def launch_bytes(plan, e, stage, steps_per_launch):
    """Algorithmic bytes of one launch of `stage` (the persistent sweep moves
    one Strang step's bytes per field per contour step)."""
    if stage == "sweep":
        return steps_per_launch * plan.batch * step_bytes(plan, e)
    return stage_bytes(plan, e)[stage]


def step_bytes(plan, e):
    """Algorithmic bytes of one Strang step per field:
    4 M^3 e + 8 n_ch H^3 2e + 2 n_spec 2e."""
    M3 = int(np.prod(plan.M))
    return 4 * M3 * e + 8 * plan.n_ch * int(np.prod(plan.H)) * 2 * e + 2 * plan.n_spec * 2 * e


def sweep_calls(sw, sweeps):
    """The library calls of `sweeps` full contour sweeps of the batch (each
    restarts from q = 1 and runs Ns Strang steps of q and q+)."""
    for _ in range(sweeps):
        sw.reset()
        sw.run(NS, start=0)


def time_sweeps(sw, sweeps, warmup, torch):
    """Wall time of `sweeps` full contour sweeps back to back (one bench step
    = one sweep of the workload), after `warmup` untimed sweeps making exactly
    the same calls, so that every hipGraph the timed calls replay is captured
    before the timed region."""
    sweep_calls(sw, max(1, warmup))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sweep_calls(sw, sweeps)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def median_sweep_ms(sw, torch, trials=30, warmup=5):
    """The reference's timing method (median of 30 trials after 5 warm-ups)
    applied to one full contour sweep of Ns steps: median ms per step."""
    times = []
    for i in range(warmup + trials):
        sw.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sw.run()
        torch.cuda.synchronize()
        if i >= warmup:
            times.append((time.perf_counter() - t0) / NS * 1e3)
    return float(np.median(times))


def profile_stages(sw, steps, torch):
    import cfft_hip._lib as L
    import ctypes as C
    lib = L.lib()
    h = sw.plan.handle
    L.check(lib.cfft_profile(h, 1))
    sw.reset()
    sw.run(min(steps, NS))
    torch.cuda.synchronize()
    ms = np.zeros(len(L.STAGES))
    cnt = np.zeros(len(L.STAGES), dtype=np.int64)
    L.check(lib.cfft_profile_read(h, ms.ctypes.data_as(L.PD), cnt.ctypes.data_as(L.PI64)))
    L.check(lib.cfft_profile(h, 0))
    return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(L.STAGES) if cnt[i]}


STRATEGY_PATH = {"m7_chain": 0, "m7_persistent": 0, "qr_chain": 5, "qs_chain": 6, "qs_chain_eager": 6,
                 "qs_chain_axis": 6, "qs_chain_tuned": 6}
# the FETCH_SIZE / WRITE_SIZE record of each pipeline path's chain kernels
# (eager launches and graph replays run the same kernels)
TRAFFIC_RECORD = {0: "m7_chain", 5: "qr_chain", 6: "qs_chain"}


def profiled_traffic(strategy, stage_name, prec_name, axis=False):
    """HBM bytes per launch of a chain kernel at this workload (2 fields per
    launch), from the FETCH_SIZE / WRITE_SIZE profiler runs recorded under
    profiles/ for the pipeline path of that launch strategy (each counter
    scaled by what it reads for a plain 64 MiB copy), the per-axis chain's
    kernels under their own keys; None when no such record exists."""
    name = TRAFFIC_RECORD.get(STRATEGY_PATH.get(strategy))
    if name is None:
        return None
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_im3m128_{name}.json")
    try:
        with open(path) as fh:
            rec = json.load(fh)["kernels"].get(f"{stage_name}/{prec_name}" + ("/axis" if axis else ""))
    except (OSError, ValueError, KeyError):
        return None
    return None if rec is None else float(rec["bytes"])


# chain kernels that run once per contour step and that cfft_time_stage_dev times
TIMED_STAGES = ("p51_unfold_fold", "p2_fft_yz", "p2q_plane_pair")


def dominant_stage(stages):
    """The per-step chain kernel with the largest event-timed total."""
    cand = [k for k in stages if k in TIMED_STAGES] or list(stages)
    return max(cand, key=lambda k: stages[k][0])


def live_launch_ms(plan, stage_name, q, E_ptrs, stream, reps=100, trials=3):
    """Average duration of one launch of a chain kernel, timed live with two
    hipEvents on the stream it runs on: `reps` identical launches back to back
    (cfft_time_stage_dev) on q, which is overwritten; the smallest average of
    `trials` such loops (a stray host or driver pause inflates one loop)."""
    import ctypes as C
    import cfft_hip._lib as L
    stage = L.STAGES.index(stage_name)
    best = None
    for _ in range(trials):
        ms = C.c_double()
        L.check(L.lib().cfft_time_stage_dev(plan.handle, stage, C.c_void_p(q), E_ptrs, int(reps),
                                            C.c_void_p(stream), C.byref(ms)))
        best = ms.value if best is None else min(best, ms.value)
    return best


def cpu_threads():
    """Host threads for the multi-core CPU leg: every core this process may
    use, i.e. the affinity set, bounded by OMP_NUM_THREADS when it is set (the
    CPU share that the machine gives one GPU's job)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    try:
        n = min(n, int(os.environ["OMP_NUM_THREADS"]))
    except (KeyError, ValueError):
        pass
    return max(1, n)


def cpu_baseline(ops_s, wA, Lv, budget_s):
    """The C restatement of the oracle's Strang step (oracle/cstep.c, tables
    from the NumPy oracle, FP64) on the same workload: Strang steps of q along
    the A block, timed for `budget_s` seconds at 1 thread and at all cores."""
    import oracle as O
    from oracle import cstep
    import cfft_hip as cf
    R, t = cf.ops_arrays(ops_s)
    if not os.path.exists(cstep.LIB):
        cstep.build()
    ds = 1.0 / NS
    E = np.asfortranarray(np.exp(-wA * ds / 2.0))
    out = {}
    t0 = time.perf_counter()
    cs = cstep.CStep(R, t, N, ds, CELL * np.eye(3), nthreads=1)
    setup = time.perf_counter() - t0
    for nt in (1, cpu_threads()):
        cs.set_threads(nt)
        q = np.ones(cs.M, order="F")
        for _ in range(5):  # warm-ups, as in the reference's benches
            q = cs.step(q, E)
        n = 0
        dts = []
        t0 = time.perf_counter()
        while True:
            t1 = time.perf_counter()
            q = cs.step(q, E)
            dts.append(time.perf_counter() - t1)
            n += 1
            if time.perf_counter() - t0 > budget_s:
                break
        out[nt] = (n / (time.perf_counter() - t0), n, setup, 1.0 / float(np.median(dts)))
    return out


def gyroid_fields(cf, N, sg, cell, amp, noise, seed):
    """Ia-3d start on the subgrid: w_A = -amp * g + n_A, w_B = +amp * g + n_B with
    g the {211} star sin 2x cos y sin z + sin 2y cos z sin x + sin 2z cos x sin y
    (x in units of 2 pi / cell) and n seeded N(0, noise) fields, symmetrised.
    The single-gyroid nodal function sin x cos y + sin y cos z + sin z cos x is
    made of {110} waves, which Ia-3d extinguishes (hhl: 2h + l = 4n): averaged
    over the group it vanishes and the melt relaxes to the disordered state.
    The {211} star is the part of its square that Ia-3d keeps (large on both
    networks), so the start is a double gyroid with the A block on the networks."""
    from cfft_hip.scft import double_gyroid_211, symmetrize_subgrid
    ops = cf.get_ops(sg, 3, N)
    shift, ops_s = cf.find_optimal_shift(ops, N)
    Lv = cf.auto_L(ops_s)
    b = [s * n for s, n in zip(shift, N)]
    rng = np.random.default_rng(seed)
    g = double_gyroid_211(N, b, cell)
    sub = tuple(slice(0, None, int(l)) for l in Lv)
    g0 = np.ascontiguousarray(g[sub])
    nA = symmetrize_subgrid(rng.normal(0.0, noise, N), ops_s, Lv)
    nB = symmetrize_subgrid(rng.normal(0.0, noise, N), ops_s, Lv)
    return ops_s, Lv, -amp * g0 + nA, amp * g0 + nB


def scft_im3m(cf, ops_s, wA, wB, tol, max_iter):
    """Plain cold start of the Im-3m 128^3 melt (f 0.3, chiN 20, Ns 100) on the
    device, FP64: first-iteration Q, F, densities; wall time to `tol`."""
    lat = CELL * np.eye(3)
    solver = cf.scft.SCFTSolver(N, ops_s, lat, F_A, CHIN, NS, cf.FP64, m_hist=10, axis_split=True)
    solver.set_fields(wA, wB)
    first = solver.iterate()
    phi0 = solver.densities()
    solver.set_fields(wA, wB)
    t0 = time.perf_counter()
    it, out = solver.solve(tol=tol, max_iter=max_iter)
    el = time.perf_counter() - t0
    return first, phi0, it, out, el


def scft_ia3d(cf, tol64, tol32, max_iter, cell=3.8):
    """Ia-3d 128^3, cell 3.8 (or `cell`), f 0.35, chiN 18, Ns 200 from the
    double-gyroid {211} harmonic: solve in FP64 and in mixed precision from
    the same start; errors of Q, F and the densities of the mixed solution
    against the FP64 one; F of the disordered melt, chiN f (1 - f), and the
    range of phi_A (uniform 0.35 when disordered) beside them."""
    n3 = (128, 128, 128)
    ops_s, Lv, wA, wB = gyroid_fields(cf, n3, 230, cell, 4.0, 0.05, 7)
    lat = cell * np.eye(3)
    out = {}
    for name, prec, tol in (("fp64", cf.FP64, tol64), ("mixed", cf.MIXED, tol32)):
        solver = cf.scft.SCFTSolver(n3, ops_s, lat, 0.35, 18.0, 200, prec, m_hist=10)
        solver.set_fields(wA, wB)
        t0 = time.perf_counter()
        it, o = solver.solve(tol=tol, max_iter=max_iter)
        el = time.perf_counter() - t0
        out[name] = (it, o, el, solver.densities())
        del solver
    # mixed precision to the hand-off residual, then an FP64 finish to 1e-8
    # from those fields (the mixed stage on the plane-pair chain, whose
    # mixed-precision floor lies below the hand-off)
    hand = cf.scft.QS_MIXED_TOL
    s32 = cf.scft.SCFTSolver(n3, ops_s, lat, 0.35, 18.0, 200, cf.MIXED, m_hist=10,
                             path=cf.scft.default_path(n3, ops_s, lat, cf.MIXED, hand))
    s32.set_fields(wA, wB)
    t0 = time.perf_counter()
    ia, oa = s32.solve(tol=hand, max_iter=max_iter)
    fin = cf.scft.SCFTSolver(n3, ops_s, lat, 0.35, 18.0, 200, cf.FP64, m_hist=10)
    fin.continue_from(s32)  # fields and Anderson history
    ib, ob = fin.solve(tol=tol64, max_iter=max_iter)
    e_mf = time.perf_counter() - t0
    del s32, fin
    i64, o64, e64, p64 = out["fp64"]
    i32, o32, e32, p32 = out["mixed"]
    refine = {"hand_off": hand, "iters_mixed": ia, "iters_fp64": ib, "res": ob["res"], "F": ob["F"], "wall_s": e_mf,
              "abs_dF_vs_fp64": abs(ob["F"] - o64["F"]), "speedup_vs_fp64": e64 / e_mf}
    return {
        "cell": cell,
        "start": "double gyroid {211} star, amplitude 4, w_A low on the networks",
        "F_disordered": 18.0 * 0.35 * 0.65,
        "phiA_range_fp64": [float(p64[0].min()), float(p64[0].max())],
        "mixed_then_fp64": refine,
        "fp64": {"iters": i64, "res": o64["res"], "F": o64["F"], "Q": o64["Q"], "wall_s": e64},
        "mixed": {"iters": i32, "res": o32["res"], "F": o32["F"], "Q": o32["Q"], "wall_s": e32},
        "abs_dF": abs(o32["F"] - o64["F"]),
        "rel_dQ": abs(o32["Q"] - o64["Q"]) / abs(o64["Q"]),
        "max_abs_dphi": float(max(np.max(np.abs(a - b)) for a, b in zip(p32, p64))),
    }


def batched_fields(cf, torch, ops_s, wA, prec, e, batch, steps, path, options=None):
    """Field-steps/s and the dominant kernel's bandwidth when one plan carries
    `batch` Im-3m 128^3 fields (the same workload's q, replicated): the same
    kernels (and plan options) once there is enough work per launch to leave
    the latency floor."""
    import ctypes as C
    lat = CELL * np.eye(3)
    plan = cf.DevicePlan(N, ops_s, 1.0 / NS, lat, path, prec, batch=batch)
    for k, v in (options or {}).items():
        plan.set_option(k, v)
    M3 = int(np.prod(plan.M))
    dt = torch.float64 if prec == cf.FP64 else torch.float32
    w = torch.from_numpy(np.asarray(wA, dtype=np.float64).ravel(order="F")).to("cuda")
    E = torch.exp(-w * (0.5 / NS)).repeat(batch)  # FP64 in both precisions
    q = torch.ones(batch * M3, dtype=dt, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    # warm-up with the timed call's own arguments: it captures the hipGraph
    # that the timed call replays (a call with another step count would
    # capture a graph of its own inside the timed region)
    plan.step_dev(q.data_ptr(), E.data_ptr(), steps, s)
    torch.cuda.synchronize()
    q.fill_(1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.step_dev(q.data_ptr(), E.data_ptr(), steps, s)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    stages = {}
    import cfft_hip._lib as L
    lib = L.lib()
    L.check(lib.cfft_profile(plan.handle, 1))
    plan.step_dev(q.data_ptr(), E.data_ptr(), 20, s)
    torch.cuda.synchronize()
    ms = np.zeros(len(L.STAGES))
    cnt = np.zeros(len(L.STAGES), dtype=np.int64)
    L.check(lib.cfft_profile_read(plan.handle, ms.ctypes.data_as(L.PD), cnt.ctypes.data_as(L.PI64)))
    L.check(lib.cfft_profile(plan.handle, 0))
    stages = {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(L.STAGES) if cnt[i]}
    dom = dominant_stage(stages)
    E_ptrs = (C.c_void_p * batch)(*[E.data_ptr() + b * M3 * 8 for b in range(batch)])
    avg = live_launch_ms(plan, dom, q.data_ptr(), E_ptrs, s)
    nb = stage_bytes(plan, e)[dom]
    return {"batch": batch, "value": batch * steps / el, "unit": "diffusion steps/s",
            "ms_per_step": el / steps * 1e3, "dominant_kernel": dom, "avg_launch_ms": avg,
            "achieved_GBs": nb / (avg * 1e-3) / 1e9, "frac": nb / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS}


def batched_instances(cf, n, seed, workers):
    """64 (default) independent instances, sizes {32,48,64,96}^3 and groups
    {47,70,223,229,230} drawn from a seed, f ~ U(.25,.45), chiN ~ U(12,30),
    Ns 100, each started from its group's lowest non-extinct star: solved in
    mixed precision (FP32 propagators, FP64 reductions) to residual 1e-5,
    then finished in FP64 to 1e-8; plain FP64 to 1e-8 from the same starts
    for the per-instance errors and the like-for-like time.  Instances of one
    grid and space group share a context; the contexts run as one ensemble
    (per transform-length class one launch per chain stage and contour step
    over the active fields, cfft_hip.scft.SCFTEnsemble).  Beside it: the same
    contexts each on its own stream (grouped), a context per instance, and
    the tuned launch schedules."""
    from cfft_hip import batch
    insts = batch.draw_instances(n, seed)
    strip = lambda rs: [{k: v for k, v in r.items() if k not in ("ensemble", "schedule")} for r in rs]
    solve = lambda prec, tol, **kw: batch.solve_batch(insts, precision=prec, tol=tol, max_iter=400,
                                                      workers=workers, **kw)
    ru, pu, su = solve(cf.FP64, 1e-8, group=False)
    rg, pg, sg = solve(cf.FP64, 1e-8)
    r32g, p32g, s32g = solve(cf.MIXED, 1e-5)
    rt, pt, st = solve(cf.FP64, 1e-8, tune=True, ensemble=True)
    sched = {}
    for r in rt:
        sched[r["schedule"]] = sched.get(r["schedule"], 0) + 1
    r32, p32, s32 = solve(cf.MIXED, 1e-5, ensemble=True)
    rmf, pmf, smf = solve(cf.MIXED, 1e-5, finish_tol=1e-8, ensemble=True)
    r64, p64, s64 = solve(cf.FP64, 1e-8, ensemble=True)
    in_ens = sum(bool(r.get("ensemble")) for r in r64)
    r32, rmf, r64, rt = strip(r32), strip(rmf), strip(r64), strip(rt)
    dF = np.array([abs(a["F"] - b["F"]) for a, b in zip(r32, r64)])
    dFf = np.array([abs(a["F"] - b["F"]) for a, b in zip(rmf, r64)])
    return {
        "instances": n, "workers": workers, "groups": len(batch.groups_of(insts)),
        "instances_in_ensemble": in_ens, "mixing": dict(batch.BATCH_MIXING),
        "mixed": {"tol": 1e-5, "instances_per_s": n / s32, "solve_s": s32, "plan_s": p32,
                  "converged": int(sum(a["res"] < 1e-5 for a in r32)),
                  "iters_mean": float(np.mean([a["iters"] for a in r32])),
                  "abs_dF_vs_fp64": {"max": float(dF.max()), "median": float(np.median(dF))},
                  "other_solution": int(np.sum(dF > 1e-3)), "speedup_vs_fp64": s64 / s32},
        "mixed_then_fp64": {"tol": 1e-8, "instances_per_s": n / smf, "solve_s": smf, "plan_s": pmf,
                            "converged": int(sum(a["res"] < 1e-8 for a in rmf)),
                            "iters_mean": float(np.mean([a["iters"] + a["iters_finish"] for a in rmf])),
                            "abs_dF_vs_fp64": {"max": float(dFf.max()), "median": float(np.median(dFf))},
                            "speedup_vs_fp64": s64 / smf},
        "fp64": {"tol": 1e-8, "instances_per_s": n / s64, "solve_s": s64, "plan_s": p64,
                 "converged": int(sum(a["res"] < 1e-8 for a in r64)),
                 "iters_mean": float(np.mean([a["iters"] for a in r64])),
                 "iters_max": int(max(a["iters"] for a in r64))},
        # the same contexts each on its own stream, host threads driving them
        "fp64_grouped": {"instances_per_s": n / sg, "solve_s": sg, "plan_s": pg,
                         "same_results": strip(rg) == r64, "ensemble_speedup": sg / s64},
        "mixed_grouped": {"instances_per_s": n / s32g, "solve_s": s32g, "same_results": strip(r32g) == r32,
                          "ensemble_speedup": s32g / s32},
        "fp64_context_per_instance": {"instances_per_s": n / su, "solve_s": su, "plan_s": pu,
                                      "same_results": strip(ru) == r64, "ensemble_speedup": su / s64},
        # each group's launch schedule (tile rows, graph replay) timed on its
        # own workload and the fastest applied (cfft_hip.tune)
        "fp64_tuned": {"instances_per_s": n / st, "solve_s": st, "plan_and_tune_s": pt,
                       "instances_by_schedule": sched, "same_results": rt == r64,
                       "speedup_vs_untuned": s64 / st},
    }


def families_fm3m(cf, torch, n=128, reps=100):
    """The sparse-stage KRFFT families (P3c: recursive, sparse, selective-G0,
    G0-ASU; the fractal decimation tree) on Fm-3m (#225) n^3, device-resident
    forward u -> F at the spectral ASU (us per call, back to back after one
    warm-up), beside the device pipeline's forward of the stride-2 subgrid of
    the same field; plan time and table sizes per family."""
    from cfft_hip import families as FAM
    N3 = (n, n, n)
    _, ops_s = cf.find_optimal_shift(cf.get_ops(225, 3, N3), N3)
    spec = cf.calc_spectral_asu(ops_s, 3, N3)
    rng = np.random.default_rng(225)
    u = torch.from_numpy(rng.standard_normal(n ** 3)).to("cuda")  # timing only: any field
    F = torch.empty(2 * len(spec.points), dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream

    def timed(call):
        call()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e6

    out = {"grid": list(N3), "n_spec": len(spec.points)}
    # the fractal plan takes the group's own ops and spectral ASU (its field is
    # symmetric under the half-grid shifted ops)
    ops_raw = cf.get_ops(225, 3, N3)
    spec_raw = cf.calc_spectral_asu(ops_raw, 3, N3)
    for name, planner in (("recursive", FAM.plan_krfft_recursive), ("sparse", FAM.plan_krfft_sparse),
                          ("selective", FAM.plan_krfft_selective), ("g0asu", FAM.plan_krfft_g0asu),
                          ("fractal", FAM.plan_fractal_krfft)):
        t0 = time.perf_counter()
        plan = planner(spec_raw, ops_raw) if name == "fractal" else planner(spec, ops_s)
        plan_s = time.perf_counter() - t0
        eng = plan.p3c
        out[name] = {"us_per_forward": timed(lambda: eng.forward_dev(u.data_ptr(), F.data_ptr(), s)),
                     "plan_s": plan_s,
                     "tables": plan.fractal if name == "fractal" else plan.p3c_tables}
        del plan, eng
    dev = cf.DevicePlan(N3, ops_s, 0.0, None, cf.PATH_AUTO)
    Lv = tuple(int(x) for x in dev.L)
    f0 = u.view(n, n, n)[::Lv[2], ::Lv[1], ::Lv[0]].contiguous().view(-1)
    out["device_pipeline_subgrid_forward_us"] = timed(lambda: dev.forward_dev(f0.data_ptr(), F.data_ptr(), s))
    out["subgrid"] = [int(x) for x in dev.M]
    return out


def ensemble_iterations(cf, n, seed, iters):
    """The batched-instances workload built as one ensemble (contexts of one
    grid and group each, planned one after another on this thread) and
    advanced `iters` iterations in FP64, then in mixed precision: the launch
    sequence of every solve_batch(..., ensemble=True) iteration, for a
    kernel trace.  Returns the wall time per iteration."""
    from cfft_hip import batch
    from cfft_hip.scft import SCFTEnsemble
    insts = batch.draw_instances(n, seed)
    out = {}
    for name, prec, tol in (("fp64", cf.FP64, 1e-8), ("mixed", cf.MIXED, 1e-5)):
        prep = [batch._prepare_group([insts[k] for k in g], prec, 10, False, tol) for g in batch.groups_of(insts)]
        ens = SCFTEnsemble([p[0] for p in prep])
        for sv, _, starts in prep:
            sv.set_mixing(**batch.BATCH_MIXING)
            for i, (wA, wB) in enumerate(starts):
                sv.set_fields(i, wA, wB)
        t0 = time.perf_counter()
        for _ in range(iters):
            ens.iterate()
        out[name + "_ms_per_iteration"] = (time.perf_counter() - t0) / iters * 1e3
        del ens, prep
    return out


def config_pmmm_64(cf, torch):
    """Pmmm (#47) 64^3, cell 4.0, ds 0.01, one field (the reference's
    CPU-runnable configuration): device-resident forward and backward KRFFT
    (us per call, 200 calls) and 100-step sweeps (field-steps/s, best of 5) on
    each path that applies (path QS also as the per-axis chain), FP64;
    w = 5 cos cos cos symmetrised (no noise)."""
    from cfft_hip.scft import symmetrize_subgrid
    n3 = (64, 64, 64)
    ops = cf.get_ops(47, 3, n3)
    shift, ops_s = cf.find_optimal_shift(ops, n3)
    Lv = cf.auto_L(ops_s)
    lat = CELL * np.eye(3)
    a = [2 * np.pi * (np.arange(n) + sh * n) / n for n, sh in zip(n3, shift)]
    X, Y, Z = np.meshgrid(*a, indexing="ij")
    w = symmetrize_subgrid(5.0 * np.cos(X) * np.cos(Y) * np.cos(Z), ops_s, Lv)
    s = torch.cuda.current_stream().cuda_stream
    out = {}
    for name, path, axis in (("m2", cf.PATH_M2, 0), ("qs_chain", cf.PATH_QS, 0), ("qs_chain_axis", cf.PATH_QS, 1)):
        plan = cf.DevicePlan(n3, ops_s, 0.01, lat, path)
        if axis:  # the per-axis chain (H = 16 here)
            plan.set_option(cf._lib.OPT_AXIS_SPLIT, 1)
        M3 = int(np.prod(plan.M))
        wd = torch.from_numpy(np.asarray(w, dtype=np.float64).ravel(order="F")).cuda()
        E = torch.exp(-wd * 0.005)
        q = torch.ones(M3, dtype=torch.float64, device="cuda")
        r = {}
        F = torch.empty(2 * plan.n_spec, dtype=torch.float64, device="cuda")
        for what, call in (("forward", lambda: plan.forward_dev(q.data_ptr(), F.data_ptr(), s)),
                           ("backward", lambda: plan.backward_dev(F.data_ptr(), q.data_ptr(), s))):
            if what == "backward" and path == cf.PATH_QS:
                continue  # the plane-pair chain has no F_spec backward
            call()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(200):
                call()
            torch.cuda.synchronize()
            r[what + "_us"] = (time.perf_counter() - t0) / 200 * 1e6
        plan.step_dev(q.data_ptr(), E.data_ptr(), 100, s)  # warm-up: captures the graph
        torch.cuda.synchronize()
        best = None
        for _ in range(5):
            q.fill_(1.0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            plan.step_dev(q.data_ptr(), E.data_ptr(), 100, s)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            best = el if best is None else min(best, el)
        r["steps_per_s"] = 100 / best
        r["us_per_step"] = best / 100 * 1e6
        out[name] = r
        del plan
    return out


def learned_warm_start(n_eval, workers):
    """Cold start vs the trained network's start on held-out parameter sets
    (Im-3m, Ia-3d, Fddd at the model's grid), residual 1e-8, FP64."""
    from cfft_hip import warmstart as W
    path = os.path.join(ROOT, "crystallographicfft.jl_amd", "models", "warmstart.npz")
    if not os.path.exists(path):
        return None
    ws = W.WarmStart.load(path)
    grid = int(ws.meta.get("grid", 64))
    ps = W.draw(n_eval, seed=2)
    groups = {sg: W.Group(sg, (grid,) * 3) for sg in sorted({p.sg for p in ps})}
    rows = W.evaluate(ws, ps, groups, workers=workers)
    out = W.summary(rows)
    out["grid"] = grid
    out["trained_on"] = ws.meta.get("n_train")
    # the 200-set evaluation recorded in the model file when it was trained
    # (not measured in this run; the lines above are)
    out["recorded_at_training"] = {"source": "models/warmstart.npz metadata, not this run",
                                   "evaluation": ws.meta.get("eval")}
    return out


def cpu_scft_reference(ops_s, wA, wB):
    """Q, F and densities of the first iteration on the CPU: the C restatement
    of the Strang step for the sweeps, Simpson densities and F from the SCFT
    oracle (FP64)."""
    from oracle import cstep
    import cfft_hip as cf
    R, t = cf.ops_arrays(ops_s)
    if not os.path.exists(cstep.LIB):
        cstep.build()
    cs = cstep.CStep(R, t, N, 1.0 / NS, CELL * np.eye(3), nthreads=cpu_threads())
    r = cstep.first_iteration(cs, wA, wB, F_A, CHIN, NS)
    return r["Q"], r["F"], r["phiA"], r["phiB"]


def step_roofline(r, K):
    """Whole-step bandwidth of a timed run of K sweeps: the chain kernels'
    algorithmic bytes per contour step (2 fields) over the wall time per
    contour step."""
    chain = sum(v["bytes_per_launch"] for v in r["chain"].values())
    t = r["el"] / (K * NS)
    return {"bytes_per_field_step": chain / 2, "achieved_GBs": chain / t / 1e9,
            "frac": chain / t / 1e9 / HBM_PEAK_GBS,
            "centred_formulation_bytes_per_field_step": r["step_bytes"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-scft", action="store_true")
    ap.add_argument("--scft-iters", type=int, default=600)
    ap.add_argument("--batch", type=int, default=64, help="batched SCFT instances (0: skip)")
    ap.add_argument("--warm-eval", type=int, default=12, help="held-out sets for the warm start")
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--wide", type=int, default=16, help="fields per launch of the batched measurement")
    ap.add_argument("--ensemble-iters", type=int, default=0,
                    help="profiling aid: only build the 64-instance batch as one ensemble and run this many "
                         "iterations in FP64, then in mixed precision, from one host thread")
    ap.add_argument("--strategy", default=None, choices=sorted(STRATEGY_PATH),
                    help="profiling aid: use this launch strategy for the timed region instead of the fastest "
                         "(under a kernel tracer the strategy trials and the tuner are distorted)")
    ap.add_argument("--only-families", action="store_true",
                    help="profiling aid: only time the sparse-stage KRFFT families on Fm-3m 128^3")
    args = ap.parse_args()
    if args.gpus > 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a single field is not sharded; independent instances would be spread one "
              "process per GPU)", file=sys.stderr)
        sys.exit(2)

    import torch
    import cfft_hip as cf

    torch.cuda.set_device(0)
    if args.ensemble_iters > 0:
        print(json.dumps({"ensemble_iterations": ensemble_iterations(cf, args.batch, 4, args.ensemble_iters)}),
              flush=True)
        return
    if args.only_families:
        print(json.dumps({"p3c_families_fm3m_128": families_fm3m(cf, torch)}), flush=True)
        return
    # one plain 64 MiB device copy: the reference point that scales the HBM
    # byte counters of a profiler run of this script (traffic per launch)
    src = torch.empty(8 << 20, dtype=torch.float64, device="cuda")
    torch.empty_like(src).copy_(src)
    torch.cuda.synchronize()
    del src
    ops = cf.get_ops(SG, 3, N)
    shift, ops_s = cf.find_optimal_shift(ops, N)
    Lv = cf.auto_L(ops_s)
    wA, wB = make_fields(ops_s, shift, Lv)
    lat = CELL * np.eye(3)

    res = {}
    variants = {}
    for name, prec, e in (("fp64", cf.FP64, 8), ("mixed", cf.MIXED, 4)):
        # launch strategies: the centred fwd+K+bwd chain (one launch per phase,
        # hipGraph replay) or one cooperative launch per sweep, and the
        # orbit-reduced Q chain; the fastest one is used for the timed region
        trial = {}
        sweeps = {}
        # the plane-pair chain's schedule (tile rows, graph replay, pair
        # workgroup size, lane split) tuned on this workload (cfft_hip.tune)
        from cfft_hip.tune import tune_schedule
        tuned, tuned_times = tune_schedule(N, ops_s, lat, NS, prec, 2, cf.PATH_QS, lanes=True, axis=True)
        for label, path, persist, graph in (("m7_chain", cf.PATH_AUTO, 0, 1), ("m7_persistent", cf.PATH_AUTO, 1, 1),
                                            ("qr_chain", cf.PATH_QR, 0, 1), ("qs_chain", cf.PATH_QS, 0, 1),
                                            ("qs_chain_eager", cf.PATH_QS, 0, 0), ("qs_chain_axis", cf.PATH_QS, 0, 1),
                                            ("qs_chain_tuned", cf.PATH_QS, 0, 1)):
            sw = cf.scft.PropagatorSweep(N, ops_s, lat, F_A, NS, prec, path=path)
            sw.set_fields(wA, wB)
            sw.plan.set_option(cf._lib.OPT_PERSISTENT, persist)
            sw.plan.set_option(cf._lib.OPT_GRAPH, graph)
            if label == "qs_chain_axis":  # the per-axis chain of path QS
                sw.plan.set_option(cf._lib.OPT_AXIS_SPLIT, 1)
            if label == "qs_chain_tuned":
                for k, v in tuned.items():
                    sw.plan.set_option(k, v)
            trial[label] = time_sweeps(sw, 3, 1, torch)
            sweeps[label] = sw
        best = args.strategy if args.strategy else min(trial, key=trial.get)
        sw = sweeps[best]
        for label in list(sweeps):
            if label != best:
                del sweeps[label]
        variants[name] = {k + "_ms_per_step": v / (3 * NS) * 1e3 for k, v in trial.items()}
        variants[name]["used"] = best
        variants[name]["tuned_schedule"] = {"options": {str(k): int(v) for k, v in tuned.items()},
                                            "us_per_step": {k: round(v, 3) for k, v in tuned_times.items()}}
        # full sweep once for Q
        sw.reset()
        sw.run()
        torch.cuda.synchronize()
        Qq, Qd = sw.Q()
        qfields = sw.fields()
        el = time_sweeps(sw, args.steps, args.warmup, torch)
        med = median_sweep_ms(sw, torch)
        stages = profile_stages(sw, NS, torch)
        dom = dominant_stage(stages)
        if dom in TIMED_STAGES:
            scratch = sw.q.clone()
            avg_ms = live_launch_ms(sw.plan, dom, scratch.data_ptr(), sw._E1, sw.stream())
            del scratch
        else:  # persistent sweep: one event pair around each cooperative launch
            avg_ms = stages[dom][0] / stages[dom][1]
        nb = launch_bytes(sw.plan, e, dom, NS)
        achieved = nb / (avg_ms * 1e-3) / 1e9
        # every per-step chain kernel beside the dominant one, timed the same way
        chain = {}
        for st in stages:
            if st not in TIMED_STAGES:
                continue
            scratch = sw.q.clone()
            ms = avg_ms if st == dom else live_launch_ms(sw.plan, st, scratch.data_ptr(), sw._E1, sw.stream())
            del scratch
            b = launch_bytes(sw.plan, e, st, NS)
            chain[st] = {"avg_launch_ms": ms, "bytes_per_launch": b, "achieved": b / (ms * 1e-3) / 1e9,
                         "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        res[name] = dict(el=el, med=med, Qq=Qq, Qd=Qd, q=qfields, stages=stages, dom=dom, avg_ms=avg_ms,
                         achieved=achieved, plan=sw.plan, sbytes=nb, chain=chain,
                         step_bytes=step_bytes(sw.plan, e),
                         axis=bool(sw.plan.options.get(cf._lib.OPT_AXIS_SPLIT)),
                         options={k: v for k, v in sw.plan.options.items()
                                  if k in (cf._lib.OPT_AXIS_SPLIT, cf._lib.OPT_FFT_LANES,
                                           cf._lib.OPT_PAIR_THREADS)})
        del sw

    # ---- many fields per launch (bandwidth regime of the same kernels) ----
    wide = None
    if args.wide > 0:
        wide = {}
        for name, prec, e in (("fp64", cf.FP64, 8), ("mixed", cf.MIXED, 4)):
            wide[name] = batched_fields(cf, torch, ops_s, wA, prec, e, args.wide, 50,
                                        STRATEGY_PATH[variants[name]["used"]], res[name]["options"])

    # ---- the Pmmm 64^3 configuration (one field) ----
    pmmm = config_pmmm_64(cf, torch)
    # ---- the P3c forward families on Fm-3m 128^3 ----
    fams = families_fm3m(cf, torch)

    # ---- SCFT: Im-3m 128^3 plain cold start, Ia-3d 128^3 mixed vs FP64 ----
    scft = None
    if not args.no_scft:
        first, phi0, it, sout, el = scft_im3m(cf, ops_s, wA, wB, 1e-8, args.scft_iters)
        scft = {"im3m_128": {"first_iteration": {"Q": first["Q"], "F": first["F"]},
                             "cold_start": {"iters": it, "res": sout["res"], "F": sout["F"],
                                            "wall_s": el}}}
        if not args.no_cpu:
            Qc, Fc, pA, pB = cpu_scft_reference(ops_s, wA, wB)
            scft["im3m_128"]["vs_cpu_reference"] = {
                "Q_cpu": Qc, "F_cpu": Fc, "rel_dQ": abs(first["Q"] - Qc) / Qc,
                "abs_dF": abs(first["F"] - Fc),
                "max_abs_dphi": float(max(np.max(np.abs(phi0[0] - pA)), np.max(np.abs(phi0[1] - pB))))}
        scft["ia3d_128_mixed"] = scft_ia3d(cf, 1e-8, 1e-6, args.scft_iters)
        # the same melt in a cell near the gyroid's own period, where it stays
        # ordered (at 3.8 it relaxes to the disordered state)
        scft["ia3d_128_ordered_mixed"] = scft_ia3d(cf, 1e-8, 1e-6, args.scft_iters, cell=9.0)
        if args.batch > 0:
            scft["batched"] = batched_instances(cf, args.batch, 4, args.workers)
        scft["learned_warm_start"] = learned_warm_start(args.warm_eval, args.workers)

    r64, r32 = res["fp64"], res["mixed"]
    K = args.steps  # bench steps = full contour sweeps of Ns Strang steps of q and q+
    value = 2 * NS * K / r64["el"]
    mixed_value = 2 * NS * K / r32["el"]
    dq = abs(r32["Qq"] - r64["Qq"])
    qerr = max(float(np.max(np.abs(r32["q"][b] - r64["q"][b]) / np.max(np.abs(r64["q"][b]))))
               for b in range(2))

    cpu = None
    if not args.no_cpu:
        rates = cpu_baseline(ops_s, wA, Lv, args.cpu_seconds)
        nt = max(rates)
        r1, rn = rates[1], rates[nt]
        cpu = {"value": rn[0], "unit": "diffusion steps/s", "cores": nt, "kind": "port",
               "value_1thread": r1[0], "median_step_rate": rn[3], "median_step_rate_1thread": r1[3],
               "sample": f"C restatement of the oracle's Strang step (oracle/cstep.c, FP64, "
                         f"OpenMP): {rn[1]} steps at {nt} threads and {r1[1]} steps at 1 thread "
                         f"of q on the A block, ~{args.cpu_seconds:.0f}s each; table setup "
                         f"{r1[2]:.1f}s excluded"}

    line = {
        "metric": "diffusion steps/s 128^3 Im-3m",
        "value": value,
        "unit": "diffusion steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": r64["el"] / K * 1e3,
        "ms_per_contour_step": r64["el"] / (K * NS) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded BCC first-star w fields + symmetrised noise)",
        "config": {"workload": "Im-3m #229 128^3 cell 4.0 AB diblock f=0.3 chiN=20 Ns=100, "
                               "q and q+ batched; FP64 launch strategy: " + variants["fp64"]["used"],
                   "grid": list(N), "space_group": SG, "batch_fields": 2, "contour_steps": NS,
                   "bench_step": "one full contour sweep: Ns Strang steps of q and of q+ "
                                 "(2 Ns diffusion steps)"},
        "roofline": {"bound": "hbm", "kernel": r64["dom"], "achieved": r64["achieved"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": r64["achieved"] / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(variants["fp64"]["used"], r64["dom"], "fp64", r64["axis"]),
                     "bytes_per_launch": r64["sbytes"],
                     "avg_launch_ms": r64["avg_ms"], "chain_kernels": r64["chain"]},
        "mixed": {"value": mixed_value, "dtype": "f32 data/FFT/propagator, f64 tables+reductions",
                  "speedup_vs_fp64": mixed_value / value, "abs_dQ_vs_fp64": dq,
                  "max_rel_q_err_vs_fp64": qerr,
                  "roofline": {"kernel": r32["dom"], "achieved": r32["achieved"],
                               "frac": r32["achieved"] / HBM_PEAK_GBS,
                               "avg_launch_ms": r32["avg_ms"], "chain_kernels": r32["chain"]}},
        "Q": {"fp64_q": r64["Qq"], "fp64_qdagger": r64["Qd"], "mixed_q": r32["Qq"]},
        # the reference's bench method on full sweeps: median of 30 after 5 warm-ups
        "median_of_30": {"fp64_ms_per_step": r64["med"], "fp64_steps_per_s": 2e3 / r64["med"],
                         "mixed_ms_per_step": r32["med"], "mixed_steps_per_s": 2e3 / r32["med"]},
        # one event pair around every launch of a non-graph sweep (includes
        # the events' own overhead); picks the dominant kernel
        "stage_ms_per_launch_events": {k: v[0] / v[1] for k, v in r64["stages"].items()},
        "launch_strategy": variants,
        "step_alg_bytes_fp64": r64["step_bytes"],
        # the whole Strang step against the same roofline, with the bytes the
        # chain in use moves per field and step (its kernels' algorithmic
        # bytes per launch over the fields per launch); the per-step bytes of
        # the centred fwd+K+bwd formulation beside it for comparison
        "step_roofline": step_roofline(r64, K),
        "step_roofline_mixed": step_roofline(r32, K),
        "cpu_baseline": cpu,
        "many_fields_per_launch": wide,
        "config_pmmm_64": pmmm,
        "p3c_families_fm3m_128": fams,
        "scft": scft,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
