"""GPU parity: HIP pipelines through the C ABI vs the oracle.

FP64 bar: 1e-12 relative (to the field's max magnitude) against the oracle and
the full-grid FFT reference.  Mixed precision (FP32 data / FFT / propagator):
2e-5 relative.
"""
import numpy as np
import pytest

import cfft_hip as cf
import oracle as O
from conftest import group_setup

pytestmark = pytest.mark.gpu

LAT4 = 4.0 * np.eye(3)
DS = 0.01
TOL64 = 1e-12
TOL32 = 2e-5

CENTRED = [229, 230, 70, 72, 225, 227, 63, 139, 74, 69]
PRIMITIVE = [47, 221, 223, 224, 123, 136, 200, 10, 2, 25]


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def field(sg, N, seed=1):
    ops, ops_s, shift, R, t = group_setup(sg, N)
    u = O.make_symmetric(R, t, N, seed=seed)
    Lv = O.auto_L(t)
    return ops_s, R, t, u, Lv


# M7 folds a centred lattice on the L = (2,2,2) subgrid: run it on the centred groups only.
@pytest.mark.parametrize("path,sg", [(cf.PATH_M2, g) for g in PRIMITIVE + CENTRED]
                         + [(cf.PATH_M7, g) for g in CENTRED])
def test_forward_backward_diffuse_fp64(sg, path):
    N = (16, 16, 16)
    ops_s, R, t, u, Lv = field(sg, N)
    cent = O.detect_centering_type(R, t, N)
    if path == cf.PATH_M7:
        assert cent != O.CENT_P and np.all(Lv == 2), sg
    print(f"sg={sg} N={N} path={path}", flush=True)
    plan = cf.DevicePlan(N, ops_s, DS, LAT4, path)
    spec = O.calc_spectral_asu(R, t, N)
    idx, mult, K, rep = plan.tables()
    assert np.array_equal(idx, spec.idx)
    f0 = O.extract_subgrid(u, Lv)
    # forward vs full-grid FFT at the spectral-ASU points
    F = plan.forward(f0)
    ref = np.fft.fftn(u)[tuple(spec.idx[:, d] for d in range(3))]
    assert rel(F, ref) < TOL64
    # backward round trip
    assert rel(plan.backward(F), f0) < TOL64
    # diffusion vs full-grid FFT.K.IFFT and vs the oracle
    d = plan.diffuse(f0)
    assert rel(d, O.fullgrid_reference(u, N, DS, LAT4, Lv)) < TOL64
    assert rel(d, O.m2_scft(f0, spec, R, t, N, DS, LAT4)) < TOL64
    # G0 = FFT_M(f0) (fold + channel FFTs + assemble)
    G0 = plan.forward_g0(f0)
    if path == cf.PATH_M2:
        assert rel(G0, np.fft.fftn(f0)) < TOL64
    else:
        assert rel(G0, O.assemble_G0(O.fft_channels(O.centering_fold(f0, cent)), cent, f0.shape)) < TOL64


@pytest.mark.parametrize("path,sg", [(cf.PATH_M2Q, g) for g in (47, 221, 229, 230, 70, 10)]
                         + [(cf.PATH_M7Q, g) for g in (229, 230, 70)])
def test_q_fused_fp64(sg, path):
    N = (16, 16, 16)
    ops_s, R, t, u, Lv = field(sg, N)
    cent = O.detect_centering_type(R, t, N)
    if path == cf.PATH_M7Q:
        assert cent != O.CENT_P, sg
    plan = cf.DevicePlan(N, ops_s, DS, LAT4, path)
    f0 = O.extract_subgrid(u, Lv)
    qp = O.m2q_plan(R, t, N, DS, LAT4)
    ref = O.m2q_execute(f0, qp) if path == cf.PATH_M2Q else O.m7_scft(f0, qp, cent)
    assert rel(plan.diffuse(f0), ref) < 1e-11
    assert rel(plan.diffuse(f0), O.fullgrid_reference(u, N, DS, LAT4, Lv)) < 1e-11


@pytest.mark.parametrize("sg,N", [(229, (16, 16, 16)), (229, (32, 32, 32)), (230, (24, 24, 24)),
                                  (70, (16, 16, 16)), (47, (16, 16, 16)), (221, (24, 24, 24)),
                                  (225, (32, 32, 32)), (10, (16, 16, 16)), (71, (24, 32, 48))])
def test_orbit_reduced_q_fp64(sg, N):
    """Path QR (Q multiply at the subgrid orbit representatives + orbit
    expansion) = the oracle's restatement = full-grid FFT.K.IFFT; Strang steps
    (fused chain) = the oracle's steps; mixed precision within 2e-5."""
    print(f"QR sg={sg} N={N}", flush=True)
    ops_s, R, t, u, Lv = field(sg, N, seed=4)
    f0 = O.extract_subgrid(u, Lv)
    plan = cf.DevicePlan(N, ops_s, DS, LAT4, cf.PATH_QR)
    qp = O.m2q_plan(R, t, N, DS, LAT4)
    orb = O.subgrid_orbits(R, t, N, Lv)
    d = plan.diffuse(f0)
    assert rel(d, O.qr_scft(f0, qp, orb)) < TOL64
    assert rel(d, O.fullgrid_reference(u, N, DS, LAT4, Lv)) < TOL64
    w = 2.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=8), Lv)
    q = f0 + 1.0
    out = plan.step(q, w, DS, nsteps=4)
    ref = q.copy()
    for _ in range(4):
        ref = O.strang_step(ref, w, DS, lambda x: O.qr_scft(x, qp, orb))
    assert rel(out, ref) < TOL64
    p32 = cf.DevicePlan(N, ops_s, DS, LAT4, cf.PATH_QR, cf.MIXED)
    assert rel(p32.diffuse(f0), d) < TOL32
    spec = O.calc_spectral_asu(R, t, N)
    F = plan.forward(f0)  # the forward transform is shared with the other paths
    assert rel(F, np.fft.fftn(u)[tuple(spec.idx[:, k] for k in range(3))]) < TOL64
    with pytest.raises(cf.CFFTError):
        plan.backward(F)


@pytest.mark.parametrize("sg,N", [(229, (16, 16, 16)), (229, (32, 32, 32)), (230, (24, 24, 24)),
                                  (70, (16, 16, 16)), (47, (16, 16, 16)), (221, (24, 24, 24)),
                                  (225, (32, 32, 32)), (223, (24, 24, 24)), (227, (32, 32, 32)),
                                  (47, (12, 12, 12)), (229, (12, 12, 12)), (230, (48, 48, 48)),
                                  (70, (32, 32, 32)), (227, (16, 16, 16))])
def test_plane_pair_q_fp64(sg, N):
    """Path QS (FFT y, z + separable Q multiply over the sign images + IFFT z, y
    in one kernel per plane pair) = full-grid FFT.K.IFFT and = the oracle's
    Q-fused restatement; Strang steps = the oracle's; mixed precision within
    2e-5; a kernel update equals a fresh plan.  H = 3 exercises the odd-length
    orbit counts."""
    print(f"QS sg={sg} N={N}", flush=True)
    ops_s, R, t, u, Lv = field(sg, N, seed=5)
    f0 = O.extract_subgrid(u, Lv)
    plan = cf.DevicePlan(N, ops_s, DS, LAT4, cf.PATH_QS)
    d = plan.diffuse(f0)
    assert rel(d, O.fullgrid_reference(u, N, DS, LAT4, Lv)) < TOL64
    qp = O.m2q_plan(R, t, N, DS, LAT4)
    orb = O.subgrid_orbits(R, t, N, Lv)
    assert rel(d, O.qr_scft(f0, qp, orb)) < TOL64
    w = 2.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=9), Lv)
    q = f0 + 1.0
    out = plan.step(q, w, DS, nsteps=5)
    ref = q.copy()
    for _ in range(5):
        ref = O.strang_step(ref, w, DS, lambda x: O.qr_scft(x, qp, orb))
    assert rel(out, ref) < TOL64
    p32 = cf.DevicePlan(N, ops_s, DS, LAT4, cf.PATH_QS, cf.MIXED)
    assert rel(p32.diffuse(f0), d) < TOL32
    spec = O.calc_spectral_asu(R, t, N)
    F = plan.forward(f0)
    assert rel(F, np.fft.fftn(u)[tuple(spec.idx[:, k] for k in range(3))]) < TOL64
    with pytest.raises(cf.CFFTError):
        plan.backward(F)
    lat2 = 3.7 * np.eye(3)
    plan.update_kernel(0.02, lat2)
    fresh = cf.DevicePlan(N, ops_s, 0.02, lat2, cf.PATH_QS)
    assert np.array_equal(plan.diffuse(f0), fresh.diffuse(f0))
    assert rel(fresh.diffuse(f0), O.fullgrid_reference(u, N, 0.02, lat2, Lv)) < TOL64


@pytest.mark.parametrize("sg,N,lat", [(10, (16, 16, 16), LAT4), (71, (24, 32, 48), LAT4),
                                      (229, (32, 32, 32), np.array([[4.0, 0.5, 0.0], [0.0, 4.0, 0.0],
                                                                    [0.0, 0.0, 4.0]]))])
def test_plane_pair_q_rejects_unsupported(sg, N, lat):
    """L != (2,2,2), a non-cubic grid and a non-orthogonal cell are refused at
    plan time with an error, not run."""
    ops_s, R, t, u, Lv = field(sg, N, seed=5)
    with pytest.raises(cf.CFFTError):
        cf.DevicePlan(N, ops_s, DS, lat, cf.PATH_QS)


@pytest.mark.parametrize("sg,N", [(229, (32, 32, 32)), (230, (32, 32, 32)), (70, (32, 32, 32)),
                                  (47, (32, 32, 32)), (223, (48, 48, 48)), (229, (48, 48, 48)),
                                  (47, (24, 24, 24)), (230, (96, 96, 96))])
def test_sizes_fp64(sg, N):
    """Mixed radices (H = 6, 8, 12, 16, 24) and both paths."""
    ops_s, R, t, u, Lv = field(sg, N, seed=3)
    f0 = O.extract_subgrid(u, Lv)
    spec = O.calc_spectral_asu(R, t, N)
    ref = O.fullgrid_reference(u, N, DS, LAT4, Lv)
    for path in (cf.PATH_AUTO, cf.PATH_M2):
        print(f"sg={sg} N={N} path={path}", flush=True)
        plan = cf.DevicePlan(N, ops_s, DS, LAT4, path)
        F = plan.forward(f0)
        Fref = np.fft.fftn(u)[tuple(spec.idx[:, d] for d in range(3))]
        assert rel(F, Fref) < TOL64
        assert rel(plan.diffuse(f0), ref) < TOL64


@pytest.mark.parametrize("sg", [229, 47, 230])
def test_step_and_batch(sg):
    """step_diffusion (Strang) vs the oracle, batch of 3 fields = 3 singles."""
    N = (16, 16, 16)
    ops_s, R, t, u, Lv = field(sg, N)
    spec = O.calc_spectral_asu(R, t, N)
    w = O.extract_subgrid(O.make_symmetric(R, t, N, seed=7), Lv) * 3.0
    q = O.extract_subgrid(u, Lv) + 1.0
    plan = cf.DevicePlan(N, ops_s, DS, LAT4, cf.PATH_AUTO)
    out = plan.step(q, w, DS, nsteps=5)
    ref = q.copy()
    for _ in range(5):
        ref = O.strang_step(ref, w, DS, lambda f: O.m2_scft(f, spec, R, t, N, DS, LAT4))
    assert rel(out, ref) < TOL64
    pb = cf.DevicePlan(N, ops_s, DS, LAT4, cf.PATH_AUTO, batch=3)
    qs = np.stack([q, 2 * q, q - 0.5])
    ws = np.stack([w, w, 0.5 * w])
    outb = pb.step(qs, ws, DS, nsteps=2)
    for b in range(3):
        single = plan.step(qs[b], ws[b], DS, nsteps=2)
        assert rel(outb[b], single) < 1e-15


@pytest.mark.parametrize("sg", [229, 230, 47, 70])
def test_mixed_precision(sg):
    N = (32, 32, 32)
    ops_s, R, t, u, Lv = field(sg, N)
    f0 = O.extract_subgrid(u, Lv)
    p64 = cf.DevicePlan(N, ops_s, DS, LAT4, cf.PATH_AUTO, cf.FP64)
    p32 = cf.DevicePlan(N, ops_s, DS, LAT4, cf.PATH_AUTO, cf.MIXED)
    assert rel(p32.forward(f0), p64.forward(f0)) < TOL32
    assert rel(p32.diffuse(f0), p64.diffuse(f0)) < TOL32


def test_update_kernel_equals_replan():
    """test/runtests.jl:344-362: update == replan (< 1e-14), and it changes."""
    N = (16, 16, 16)
    ops_s, R, t, u, Lv = field(221, N)
    f0 = O.extract_subgrid(u, Lv)
    p = cf.plan_m2_scft(N, 221, 3, 0.05, LAT4)
    a = f0.copy()
    cf.execute_m2_scft(p, a)
    cf.update_m2_kernel(p, N, 221, 3, 0.10, LAT4)
    b = f0.copy()
    cf.execute_m2_scft(p, b)
    p2 = cf.plan_m2_scft(N, 221, 3, 0.10, LAT4)
    c = f0.copy()
    cf.execute_m2_scft(p2, c)
    assert np.max(np.abs(b - c)) < 1e-14
    assert np.max(np.abs(a - b)) > 1e-4


def test_ds_zero_identity():
    """test/test_m2_scft.jl:99-116, test/test_centered_scft.jl:126-153."""
    N = (32, 32, 32)
    for sg in (47, 229, 230):
        ops_s, R, t, u, Lv = field(sg, N)
        f0 = O.extract_subgrid(u, Lv)
        p = cf.DevicePlan(N, ops_s, 0.0, LAT4, cf.PATH_AUTO)
        assert np.max(np.abs(p.diffuse(f0) - f0)) < 1e-12


def test_reference_shaped_api():
    """plan_*/execute_* names of the reference (test/runtests.jl:196-341)."""
    N = (16, 16, 16)
    ops_s, R, t, u, Lv = field(229, N)
    spec = cf.calc_spectral_asu(ops_s, 3, N)
    f0 = O.extract_subgrid(u, Lv)
    ref = O.fullgrid_reference(u, N, 0.05, np.eye(3), Lv)
    scft = cf.plan_centered_scft(spec, ops_s, N, 0.05, np.eye(3))
    a = f0.copy()
    cf.execute_centered_scft(scft, a)
    assert np.max(np.abs(a - ref)) < 1e-12
    fwd = cf.plan_krfft_centered(spec, ops_s)
    assert isinstance(fwd, cf.CenteredKRFFTPlan)
    Fs = cf.execute_centered_krfft(fwd, u).copy()
    bwd = cf.plan_centered_ikrfft(spec, ops_s, fwd)
    out = np.zeros_like(f0)
    cf.execute_centered_ikrfft(bwd, Fs, out)
    assert np.max(np.abs(out - f0)) < 1e-12
    pk = cf.plan_krfft(spec, ops_s)
    Fk = cf.execute_krfft(pk, u)
    assert np.max(np.abs(Fk - Fs)) < 1e-11
    bp = cf.plan_m2_backward(spec, ops_s)
    assert np.max(np.abs(cf.execute_m2_backward(bp, Fk).real - f0)) < 1e-12
    q = cf.plan_m2_q(N, 229, 3, 0.05, np.eye(3))
    b = f0.copy()
    cf.execute_m2_q(q, b)
    assert np.max(np.abs(b - ref)) < 1e-10
    m7 = cf.plan_m7_scft(N, 229, 3, 0.05, np.eye(3))
    c = f0.copy()
    cf.execute_m7_scft(m7, c)
    assert np.max(np.abs(c - b)) < 1e-12
    full = np.zeros(N)
    cf.subgrid_to_fullgrid(full, f0, q)
    assert np.max(np.abs(full - u)) < 1e-14
    sub = np.zeros_like(f0)
    cf.fullgrid_to_subgrid(sub, u, q)
    assert np.array_equal(sub, f0)
    solver = cf.plan_diffusion(N, np.eye(3), 229, 3, 0.05, method="qfused")
    d = f0.copy()
    cf.apply_diffusion_operator(solver, d)
    assert np.max(np.abs(d - ref)) < 1e-10
    w = 0.3 * f0
    s = f0.copy()
    cf.step_diffusion(solver, s, w, 0.05)
    e = np.exp(-w * 0.025)
    assert np.max(np.abs(s - e * O.fullgrid_reference(
        O.subgrid_to_fullgrid(e * f0, O.fill_map(R, t, Lv, N), N), N, 0.05, np.eye(3), Lv))) < 1e-10


def test_full_size_im3m_128_properties():
    """Config grid (Im-3m 128^3, cell 4.0): forward vs the oracle forward
    (the spectral-ASU points are checked bit-exact on the CPU), backward round
    trip, ds = 0 identity and linearity."""
    from cfft_hip.scft import symmetrize_subgrid
    N = (128, 128, 128)
    print("Im-3m 128^3 properties", flush=True)
    ops, ops_s, shift, R, t = group_setup(229, N)
    Lv = O.auto_L(t)
    rng = np.random.default_rng(11)
    f0 = symmetrize_subgrid(rng.standard_normal(N), ops_s, Lv)
    g0 = symmetrize_subgrid(rng.standard_normal(N), ops_s, Lv)
    p = cf.DevicePlan(N, ops_s, 0.0, LAT4, cf.PATH_M7)
    idx, mult, _, _ = p.tables()
    spec = O.SpectralASU(idx, mult, N)
    rep = O.select_rep_ops(R, t, Lv, N)
    Fref = O.krfft_forward(np.fft.fftn(f0), spec, R, t, rep, N)
    for prec, tol in ((cf.FP64, TOL64), (cf.MIXED, TOL32)):
        p = cf.DevicePlan(N, ops_s, 0.0, LAT4, cf.PATH_M7, prec)
        F = p.forward(f0)
        assert rel(F, Fref) < tol
        assert rel(p.backward(F), f0) < tol
        assert rel(p.diffuse(f0), f0) < tol
    p = cf.DevicePlan(N, ops_s, DS, LAT4, cf.PATH_M7)
    lhs = p.diffuse(2.0 * f0 - 3.0 * g0)
    rhs = 2.0 * p.diffuse(f0) - 3.0 * p.diffuse(g0)
    assert rel(lhs, rhs) < TOL64
    # the M7 and M2 routes agree on the config grid
    p2 = cf.DevicePlan(N, ops_s, DS, LAT4, cf.PATH_M2)
    assert rel(p.diffuse(f0), p2.diffuse(f0)) < TOL64


@pytest.mark.parametrize("path,axis", [(cf.PATH_M2, 0), (cf.PATH_QS, 0), (cf.PATH_QS, 1)])
def test_config_pmmm_64(path, axis):
    """Pmmm (#47) 64^3, cell 4.0, ds 0.01 (the reference's CPU-runnable
    configuration) at its full size: u ~ U(-1, 1) symmetrised; forward KRFFT
    at the 32768 spectral-ASU points against the full-grid FFT and the backward
    round trip (path M2; the plane-pair chain has no backward); then 100
    Strang steps from q = 1 with w = 5 cos(2 pi x/a) cos(2 pi y/a)
    cos(2 pi z/a) + N(0, 0.1) symmetrised, against the same steps on the full
    grid (NumPy FFT, K, IFFT between the two half-step factors); path QS also
    as the per-axis chain (CFFT_OPT_AXIS_SPLIT)."""
    N = (64, 64, 64)
    print(f"Pmmm 64^3 path={path} axis={axis}", flush=True)
    ops, ops_s, shift, R, t = group_setup(47, N)
    Lv = O.auto_L(t)
    u = O.make_symmetric(R, t, N, seed=47, dist="uniform")
    f0 = O.extract_subgrid(u, Lv)
    plan = cf.DevicePlan(N, ops_s, DS, LAT4, path)
    if axis:
        plan.set_option(cf._lib.OPT_AXIS_SPLIT, 1)
    spec = O.calc_spectral_asu(R, t, N)
    assert spec.idx.shape[0] == 32768
    F = plan.forward(f0)
    assert rel(F, np.fft.fftn(u)[tuple(spec.idx[:, d] for d in range(3))]) < TOL64
    if path == cf.PATH_M2:
        assert rel(plan.backward(F), f0) < TOL64
    a = [2 * np.pi * (np.arange(n) + s * n) / n for n, s in zip(N, shift)]
    X, Y, Z = np.meshgrid(*a, indexing="ij")
    rng = np.random.default_rng(48)
    w = O.symmetrize(5.0 * np.cos(X) * np.cos(Y) * np.cos(Z) + 0.1 * rng.standard_normal(N), R, t)
    K = O.fullgrid_kernel(N, DS, LAT4)
    E = np.exp(-w * DS / 2.0)
    q = np.ones(N)
    for _ in range(100):
        q = E * np.real(np.fft.ifftn(K * np.fft.fftn(E * q)))
    got = plan.step(np.ones(plan.M), O.extract_subgrid(w, Lv), DS, 100)
    assert rel(got, O.extract_subgrid(q, Lv)) < TOL64


@pytest.mark.parametrize("sg,N", [(229, (32, 32, 32)), (230, (32, 32, 32)), (70, (32, 32, 32))])
def test_star_coefficients(sg, N):
    """Star coefficients: c_0 = cell mean; encode/decode of all stars is the
    identity; truncation to the first n stars equals the full-grid FFT with
    every wave vector beyond those stars' |k|^2 shells zeroed."""
    from cfft_hip import stars
    ops_s, R, t, u, Lv = field(sg, N)
    plan = cf.DevicePlan(N, ops_s, 0.0, LAT4, cf.PATH_AUTO)
    f0 = O.extract_subgrid(u, Lv)
    order, k2 = stars.star_order(plan)
    c_all = stars.encode(plan, f0, plan.n_spec, order)
    assert abs(c_all[0] - np.mean(u)) < 1e-12 * np.max(np.abs(u))
    assert rel(stars.decode(plan, c_all, order), f0) < TOL64
    # truncate at a shell boundary: keep all stars with |k|^2 <= k2[n-1]
    n = int(np.searchsorted(k2, k2[min(40, len(k2) - 1)], side="right"))
    approx = stars.decode(plan, stars.encode(plan, f0, n, order), order)
    Fu = np.fft.fftn(u)
    B = 2 * np.pi * np.linalg.inv(LAT4).T
    kk = [np.fft.fftfreq(m, d=1.0 / m) for m in N]
    KX, KY, KZ = np.meshgrid(*kk, indexing="ij")
    kv = np.stack([KX, KY, KZ], axis=-1) @ B.T
    K2 = np.sum(kv * kv, axis=-1)
    ref = np.real(np.fft.ifftn(np.where(K2 <= k2[n - 1] + 1e-9, Fu, 0.0)))
    assert rel(approx, O.extract_subgrid(ref, Lv)) < TOL64


def test_mixed_chain_paired_tiles_im3m_64():
    """Plane-pair chain at Im-3m 64^3 (H = 16: two y-blocks of 8 rows per
    (b, z) and 32 unfold/fold workgroups per field), where the unfold/fold
    kernel runs the two y-blocks sharing the 128-byte lines of FP32 channel
    rows as blocks i and i + 8 of each 16.  Four Strang steps: mixed precision
    within 2e-5 of FP64 and bit-identical across tile heights 8, 4 and 2 (a
    different pairing each); FP64 within 1e-12 of the oracle's steps."""
    N = (64, 64, 64)
    print("Im-3m 64^3 mixed plane-pair chain, tile heights 8/4/2", flush=True)
    ops, ops_s, shift, R, t = group_setup(229, N)
    Lv = O.auto_L(t)
    q = O.extract_subgrid(O.make_symmetric(R, t, N, seed=5), Lv) + 1.0
    w = 2.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=6), Lv)
    p64 = cf.DevicePlan(N, ops_s, DS, LAT4, cf.PATH_QS)
    assert p64.H[1] == 16
    ref = p64.step(q, w, DS, nsteps=4)
    qp = O.m2q_plan(R, t, N, DS, LAT4)
    orb = O.subgrid_orbits(R, t, N, Lv)
    o = q.copy()
    for _ in range(4):
        o = O.strang_step(o, w, DS, lambda x: O.qr_scft(x, qp, orb))
    assert rel(ref, o) < TOL64
    outs = []
    for rows in (8, 4, 2):
        p32 = cf.DevicePlan(N, ops_s, DS, LAT4, cf.PATH_QS, cf.MIXED)
        p32.set_option(cf._lib.OPT_TILE_ROWS, rows)
        outs.append(p32.step(q, w, DS, nsteps=4))
        assert rel(outs[-1], ref) < TOL32, rows
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
