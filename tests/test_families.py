"""The reference's other KRFFT families (recursive P3c, sparse, selective,
G0-ASU forward/backward, general G0-ASU, fractal, staged Pm-3m) on the device
(the four P3c forward families through their own tables on cubic groups,
the rest on the device pipeline): the plan builders' preconditions on the CPU (test/test_g0_asu.jl,
test/test_g0_asu_backward.jl, test/test_staged_pm3m.jl check the same
assertions), and on the GPU the outputs against the full-grid FFT of the
symmetrised field at the spectral-ASU points (the check of
test/test_selective_g0.jl, test/test_general_g0_asu.jl, test/test_fractal_krfft.jl)."""
import numpy as np
import pytest

import cfft_hip as cf
import oracle as O
from conftest import group_setup


def _spec(sg, N):
    ops, ops_s, shift, R, t = group_setup(sg, N)
    return ops_s, R, t, cf.calc_spectral_asu(ops_s, 3, N)


def _spec_half(sg, N):
    """The fractal plans' convention (test/test_fractal_krfft.jl): the
    group's own ops and spectral ASU, a field symmetric under the half-grid
    shifted ops."""
    ops = cf.get_ops(sg, 3, N)
    R, t = cf.ops_arrays(cf.shift_ops_half_grid(ops, N))
    return ops, R, t, cf.calc_spectral_asu(ops, 3, N)


@pytest.mark.parametrize("sg,expect", [(221, True), (229, True), (230, True), (223, True),
                                       (47, False), (70, False), (139, False)])
def test_has_cubic_p3c_symmetry(sg, expect):
    """src/recursive_blocks.jl:974-982: only cubic groups carry the
    body-diagonal 3-fold axis."""
    ops_s, R, t, spec = _spec(sg, (16, 16, 16))
    assert cf.has_cubic_p3c_symmetry(ops_s) is expect


def test_p3c_plans_require_even_subgrid():
    """N = 18: N is even but M = 9 is odd (the P3c assertion M2 * 2 == M)."""
    ops_s, R, t, spec = _spec(229, (18, 18, 18))
    for planner in (cf.plan_krfft_recursive, cf.plan_krfft_selective, cf.plan_krfft_sparse,
                    cf.plan_krfft_g0asu, cf.plan_krfft_g0asu_backward):
        with pytest.raises(cf.CFFTError, match="divisible by 2 for P3c"):
            planner(spec, ops_s)


def test_g0asu_backward_requires_cubic_symmetry():
    ops_s, R, t, spec = _spec(47, (16, 16, 16))
    with pytest.raises(cf.CFFTError, match="requires cubic symmetry"):
        cf.plan_krfft_g0asu_backward(spec, ops_s)


def test_staged_plan_preconditions():
    """src/staged_krfft.jl:116-117."""
    with pytest.raises(cf.CFFTError, match="even"):
        cf.plan_staged_pm3m((18, 18, 17))
    with pytest.raises(cf.CFFTError, match="cubic"):
        cf.plan_staged_pm3m((16, 16, 32))


FORWARD = [(cf.plan_krfft_recursive, cf.execute_recursive_krfft),
           (cf.plan_krfft_sparse, cf.execute_sparse_krfft),
           (cf.plan_krfft_selective, cf.execute_selective_krfft),
           (cf.plan_krfft_g0asu, cf.execute_g0asu_krfft),
           (cf.plan_krfft_g0asu_general, cf.execute_general_g0asu_krfft)]


P3C_FAMILIES = (cf.plan_krfft_recursive, cf.plan_krfft_sparse, cf.plan_krfft_selective, cf.plan_krfft_g0asu)


@pytest.mark.gpu
@pytest.mark.parametrize("sg,N", [(229, (32, 32, 32)), (221, (16, 16, 16)), (47, (16, 16, 16)),
                                  (230, (24, 24, 24)), (70, (16, 16, 16)), (225, (64, 64, 64)),
                                  (227, (32, 32, 32)), (223, (48, 48, 48))])
def test_forward_families_match_full_fft(sg, N):
    """Every forward family against the full FFT at the spectral ASU (1e-12
    relative); on cubic groups the four P3c families run their own tables on
    the device (cfft_p3c_*), on the others the device pipeline."""
    print(f"families sg={sg} N={N}", flush=True)
    ops_s, R, t, spec = _spec(sg, N)
    u = O.make_symmetric(R, t, N, seed=21)
    ref = np.fft.fftn(u)[tuple(spec.idx[:, d] for d in range(3))]
    scale = np.max(np.abs(ref))
    cubic = cf.has_cubic_p3c_symmetry(ops_s)
    for planner, execute in FORWARD:
        plan = planner(spec, ops_s)
        if planner in P3C_FAMILIES:
            assert (getattr(plan, "p3c", None) is not None) == cubic, planner.__name__
        out = execute(plan, spec, u)
        assert out is plan.output_buffer
        assert np.max(np.abs(out - ref)) / scale < 1e-12, planner.__name__
    ops, Rh, th, spec_h = _spec_half(sg, N)
    uh = O.make_symmetric(Rh, th, N, seed=21)
    refh = np.fft.fftn(uh)[tuple(spec_h.idx[:, d] for d in range(3))]
    fp = cf.plan_fractal_krfft(spec_h, ops)
    assert fp.p3c is not None  # the decimation tree on the device (cfft_p3c_create_strided)
    assert np.max(np.abs(cf.execute_fractal_krfft(fp, uh) - refh)) / np.max(np.abs(refh)) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("sg,n,depths", [(229, 32, (1, 2, 3, 4)), (230, 24, (1, 2, 3)), (139, 32, (1, 2, 3, 4)),
                                         (2, 16, (1,)), (229, 128, (None, 1, 2)), (230, 96, (1,))])
def test_fractal_depths_on_device(sg, n, depths):
    """The fractal plan at each depth of its tree (None: the depth it picks,
    6 for Im-3m 128^3 with 5984 leaves of 2^3) against the full FFT at the
    spectral ASU, 1e-12 relative; the leaf lengths 64, 48, 32 and 16 run the
    fused pack + x + y transform, the shorter ones the separate passes."""
    N = (n, n, n)
    ops, R, t, spec = _spec_half(sg, N)
    u = O.make_symmetric(R, t, N, seed=24)
    ref = np.fft.fftn(u)[tuple(spec.idx[:, d] for d in range(3))]
    for depth in depths:
        print(f"fractal sg={sg} N={N} depth={depth}", flush=True)
        fp = cf.plan_fractal_krfft(spec, ops, depth)
        assert fp.p3c is not None
        out = cf.execute_fractal_krfft(fp, u)
        assert np.max(np.abs(out - ref)) / np.max(np.abs(ref)) < 1e-12, fp.fractal


@pytest.mark.gpu
@pytest.mark.parametrize("sg,N", [(229, (32, 32, 32)), (221, (16, 16, 16)), (230, (24, 24, 24))])
def test_g0asu_backward_round_trip(sg, N):
    """F_spec -> u on the full grid reproduces the symmetric field
    (test/test_g0_asu_backward.jl)."""
    ops_s, R, t, spec = _spec(sg, N)
    u = O.make_symmetric(R, t, N, seed=22)
    F = np.fft.fftn(u)[tuple(spec.idx[:, d] for d in range(3))]
    bp = cf.plan_krfft_g0asu_backward(spec, ops_s)
    out = np.zeros(N)
    res = cf.execute_g0asu_ikrfft(bp, spec, F, out)
    assert res is out
    assert np.max(np.abs(out - u)) / np.max(np.abs(u)) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("n", [16, 32])
def test_staged_pm3m(n):
    """F000 = FFT of u(2x), through the P3c engine; the full spectrum at
    random frequencies and at the spectral ASU from F000
    (test/test_staged_pm3m.jl)."""
    N = (n, n, n)
    ops_s, R, t, spec = _spec(221, N)
    u = O.make_symmetric(R, t, N, seed=23)
    plan = cf.plan_staged_pm3m(N)
    assert plan.p3c is not None  # F000 from the P3c engine (cfft_p3c_*)
    F000 = cf.execute_staged(plan, u)
    ref000 = np.fft.fftn(u[::2, ::2, ::2])
    assert np.max(np.abs(F000 - ref000)) / np.max(np.abs(ref000)) < 1e-12
    full = np.fft.fftn(u)
    rng = np.random.default_rng(5)
    h = rng.integers(0, n, size=(50, 3))
    got = cf.lookup_full_spectrum(plan, F000, h[:, 0], h[:, 1], h[:, 2])
    assert np.max(np.abs(got - full[h[:, 0], h[:, 1], h[:, 2]])) / np.max(np.abs(full)) < 1e-12
    res = np.zeros(len(spec), dtype=np.complex128)
    cf.compute_spectral_asu(res, plan, F000, spec)
    ref = full[tuple(spec.idx[:, d] for d in range(3))]
    assert np.max(np.abs(res - ref)) / np.max(np.abs(ref)) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("sg,N", [(47, (16, 16, 32)), (139, (16, 16, 32)), (70, (16, 32, 32))])
def test_fractal_fallback_to_device_pipeline(sg, N):
    """Without a tree (a non-cubic grid; the leaf transforms are cubic) the
    fractal plan runs the device pipeline on the shifted ops' spectral ASU
    and reads the caller's points through the ops: the full FFT at the
    unshifted ops' spectral ASU, 1e-12."""
    print(f"fractal fallback sg={sg} N={N}", flush=True)
    ops, R, t, spec = _spec_half(sg, N)
    u = O.make_symmetric(R, t, N, seed=25)
    ref = np.fft.fftn(u)[tuple(spec.idx[:, d] for d in range(3))]
    fp = cf.plan_fractal_krfft(spec, ops)
    assert fp.p3c is None
    out = cf.execute_fractal_krfft(fp, u)
    assert out.shape == (len(spec.points),)
    assert np.max(np.abs(out - ref)) / np.max(np.abs(ref)) < 1e-12


@pytest.mark.gpu
def test_engine_device_calls():
    """cfft_p3c_forward_dev on device buffers, called repeatedly: new field
    values in the same buffers, a second output buffer and torch work between
    the calls all give the full FFT at the spectral ASU (sparse P3c plan and
    fractal tree, Im-3m 32^3)."""
    import torch
    N = (32, 32, 32)
    s = torch.cuda.current_stream().cuda_stream
    ops_s, R_s, t_s, spec_s = _spec(229, N)
    ops, R_h, t_h, spec_h = _spec_half(229, N)
    for plan, R, t, spec in ((cf.plan_krfft_sparse(spec_s, ops_s), R_s, t_s, spec_s),
                             (cf.plan_fractal_krfft(spec_h, ops), R_h, t_h, spec_h)):
        u = torch.empty(32 ** 3, dtype=torch.float64, device="cuda")
        Fs = [torch.empty(2 * len(spec.points), dtype=torch.float64, device="cuda") for _ in range(2)]
        for rep in range(4):
            uh = O.make_symmetric(R, t, N, seed=40 + rep)
            u.copy_(torch.from_numpy(np.ascontiguousarray(uh.ravel(order="F"))))
            F = Fs[rep % 2]
            F.fill_(np.nan)
            plan.p3c.forward_dev(u.data_ptr(), F.data_ptr(), s)
            torch.cuda.synchronize()
            got = F.cpu().numpy().view(np.complex128)
            ref = np.fft.fftn(uh)[tuple(spec.idx[:, d] for d in range(3))]
            assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12, rep


@pytest.mark.gpu
def test_engine_refuses_bad_tables():
    """cfft_p3c_create_strided refuses a stride that does not divide N, a leaf
    length the device does not transform, offsets outside [0, stride), more
    than eight stages and a column past the leaf outputs; the host forward
    refuses an output buffer of the wrong size."""
    import ctypes as C
    from cfft_hip import p3c as P
    lib = cf._lib.lib()

    def create(N, S, offs, stages):
        rp = (C.c_void_p * len(stages))()
        cs = (C.c_void_p * len(stages))()
        vs = (C.c_void_p * len(stages))()
        keep = []
        for i, (rowptr, cols, vals) in enumerate(stages):
            rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
            cols = np.ascontiguousarray(cols, dtype=np.int32)
            vals = np.ascontiguousarray(np.asarray(vals, dtype=np.complex128).view(np.float64))
            keep += [rowptr, cols, vals]
            rp[i], cs[i], vs[i] = rowptr.ctypes.data, cols.ctypes.data, vals.ctypes.data
        off = np.ascontiguousarray(np.asarray(offs, dtype=np.int32).ravel())
        nr = np.array([len(s[0]) - 1 for s in stages], dtype=np.int64)
        h = C.c_void_p()
        rc = lib.cfft_p3c_create_strided(N, S, len(off) // 3, cf._lib.i32p(off), len(stages), cf._lib.i64p(nr),
                                         rp, cs, vs, C.byref(h))
        if rc == 0:
            lib.cfft_p3c_destroy(h)
        return rc

    one = (np.array([0, 1]), np.array([0]), np.array([1.0 + 0j]))
    assert create(16, 2, [[0, 0, 0]], [one]) == 0
    assert create(16, 3, [[0, 0, 0]], [one]) == cf._lib.CFFT_ERR_NOT_DIVISIBLE
    assert create(20, 2, [[0, 0, 0]], [one]) == cf._lib.CFFT_ERR_UNSUPPORTED  # leaf length 10
    assert create(16, 2, [[0, 2, 0]], [one]) == cf._lib.CFFT_ERR_INVALID
    assert create(16, 2, [[0, 0, 0]], [one] * 9) == cf._lib.CFFT_ERR_INVALID
    bad = (np.array([0, 1]), np.array([8 ** 3]), np.array([1.0 + 0j]))  # one leaf of 8^3 outputs
    assert create(16, 2, [[0, 0, 0]], [bad]) == cf._lib.CFFT_ERR_BOUNDS
    N = (16, 16, 16)
    ops_s, R, t, spec = _spec(229, N)
    plan = cf.plan_krfft_sparse(spec, ops_s)
    u = O.make_symmetric(R, t, N, seed=26)
    with pytest.raises(cf.CFFTError) as e:
        plan.p3c.forward(u, np.zeros(plan.p3c.n_out + 1, dtype=np.complex128))
    assert e.value.code == cf._lib.CFFT_ERR_SIZE
    assert P.P3cEngine is type(plan.p3c)
