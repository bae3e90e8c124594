"""Pin the NumPy oracle (oracle/krfft.py) before trusting it.

Checks (each chosen so a dropped term, wrong sign/index or transposed operand
fails): full-grid FFT of symmetrised fields at every spectral-ASU point
(test/runtests.jl:130-155), round trips (test/test_m2_backward.jl:98-107),
diffusion vs the full-grid FFT.K.IFFT reference (test/test_helpers.jl:56-73),
the closed-form / structural counts of the reference's tests, and brute-force
orbit enumeration at tiny N.
"""
import itertools

import numpy as np
import pytest

import oracle as O
from conftest import group_setup

N16 = (16, 16, 16)
LAT = np.eye(3)
DS = 0.05
GROUPS = [221, 225, 229, 63, 47, 227, 230, 70, 72, 10, 2, 123, 136, 139, 223, 224, 74, 69, 200]


@pytest.fixture(scope="module")
def prep():
    out = {}
    for sg in GROUPS:
        ops, ops_s, shift, R, t = group_setup(sg, N16)
        u = O.make_symmetric(R, t, N16, seed=sg)
        spec = O.calc_spectral_asu(R, t, N16)
        out[sg] = dict(R=R, t=t, u=u, spec=spec, L=O.auto_L(t))
    return out


@pytest.mark.parametrize("sg", GROUPS)
def test_forward_vs_fullgrid_fft(prep, sg):
    p = prep[sg]
    F = np.fft.fftn(p["u"])
    f0 = O.extract_subgrid(p["u"], p["L"])
    rep = O.select_rep_ops(p["R"], p["t"], p["L"], N16)
    Fs = O.krfft_forward(np.fft.fftn(f0), p["spec"], p["R"], p["t"], rep, N16)
    ref = F[tuple(p["spec"].idx[:, d] for d in range(3))]
    assert np.max(np.abs(Fs - ref)) / np.max(np.abs(ref)) < 1e-12


@pytest.mark.parametrize("sg", GROUPS)
def test_m2_backward_roundtrip(prep, sg):
    p = prep[sg]
    f0 = O.extract_subgrid(p["u"], p["L"])
    rep = O.select_rep_ops(p["R"], p["t"], p["L"], N16)
    Fs = O.krfft_forward(np.fft.fftn(f0), p["spec"], p["R"], p["t"], rep, N16)
    tab = O.m2_backward_table(p["spec"], p["R"], p["t"], N16)
    assert np.max(np.abs(np.real(O.m2_backward(Fs, tab)) - f0)) < 1e-12


@pytest.mark.parametrize("sg", GROUPS)
def test_m2_scft_vs_fullgrid(prep, sg):
    p = prep[sg]
    f0 = O.extract_subgrid(p["u"], p["L"])
    ref = O.fullgrid_reference(p["u"], N16, 0.002, LAT * 0.5, p["L"])
    out = O.m2_scft(f0, p["spec"], p["R"], p["t"], N16, 0.002, LAT * 0.5)
    assert np.max(np.abs(out - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(out - f0)) > 1e-3  # the operator did something


@pytest.mark.parametrize("sg", [229, 230, 70, 72, 225, 227, 63, 139, 74, 69])
def test_centered_vs_fullgrid(prep, sg):
    p = prep[sg]
    cent = O.detect_centering_type(p["R"], p["t"], N16)
    f0 = O.extract_subgrid(p["u"], p["L"])
    ref = O.fullgrid_reference(p["u"], N16, 0.002, LAT * 0.5, p["L"])
    out = O.centered_scft(f0, p["spec"], p["R"], p["t"], N16, 0.002, LAT * 0.5, cent)
    assert np.max(np.abs(out - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    # fold + FFT + assemble = fft(f0)  (test/runtests.jl:177-192)
    G0 = O.assemble_G0(O.fft_channels(O.centering_fold(f0, cent)), cent, f0.shape)
    assert np.max(np.abs(G0 - np.fft.fftn(f0))) < 1e-10
    # fold o unfold = id (test/runtests.jl:258-269)
    back = O.centering_unfold(O.ifft_channels(O.fft_channels(O.centering_fold(f0, cent))), cent, f0.shape)
    assert np.max(np.abs(back - f0)) < 1e-12


@pytest.mark.parametrize("sg", [47, 221, 225, 229, 230, 10])
def test_q_fused_vs_fullgrid(prep, sg):
    p = prep[sg]
    f0 = O.extract_subgrid(p["u"], p["L"])
    ref = O.fullgrid_reference(p["u"], N16, 0.002, LAT * 0.5, p["L"])
    qp = O.m2q_plan(p["R"], p["t"], N16, 0.002, LAT * 0.5)
    assert np.max(np.abs(O.m2q_execute(f0, qp) - ref)) < 1e-10
    cent = O.detect_centering_type(p["R"], p["t"], N16)
    if cent != O.CENT_P:
        # M7 vs M2+Q (test/runtests.jl:271-284)
        assert np.max(np.abs(O.m7_scft(f0, qp, cent) - O.m2q_execute(f0, qp))) < 1e-12


def test_ds_zero_identity(prep):
    """test/test_m2_scft.jl:99-116: K = 1 reproduces f0."""
    for sg in (47, 229, 230):
        p = prep[sg]
        f0 = O.extract_subgrid(p["u"], p["L"])
        out = O.m2_scft(f0, p["spec"], p["R"], p["t"], N16, 0.0, LAT)
        assert np.max(np.abs(out - f0)) < 1e-12


def test_fill_map_roundtrip(prep):
    """subgrid_to_fullgrid(fullgrid_to_subgrid(u)) = u (test/test_q_fused.jl:175-198)."""
    for sg in (47, 229, 70):
        p = prep[sg]
        fm = O.fill_map(p["R"], p["t"], p["L"], N16)
        f0 = O.fullgrid_to_subgrid(p["u"], p["L"])
        assert np.max(np.abs(O.subgrid_to_fullgrid(f0, fm, N16) - p["u"])) < 1e-14


def test_prefold_vs_fft(prep):
    """Centring pre-fold on the full grid (src/centering_prefold.jl) reproduces
    the full FFT at the spectral-ASU points (test/test_centering_prefold.jl)."""
    for sg in (229, 225, 63):
        p = prep[sg]
        cent = O.detect_centering_type(p["R"], p["t"], N16)
        chans, pars, Nsub = O.centering_prefold(p["u"], cent)
        hat = [np.fft.fftn(c) for c in chans]
        out = O.prefold_merge(p["spec"], hat, pars, Nsub, N16)
        F = np.fft.fftn(p["u"])
        ref = F[tuple(p["spec"].idx[:, d] for d in range(3))]
        assert np.max(np.abs(out - ref)) < 1e-12 * np.max(np.abs(ref))


# ---------------------------------------------------------------------------
# known answers and brute force
# ---------------------------------------------------------------------------

def test_shift_p2mm():
    """test/runtests.jl:42-52: p2mm 8x8 shift = (0.5/8, 0.5/8)."""
    import cfft_hip as cf
    R, t = cf.ops_arrays(cf.get_ops(6, 2, (8, 8)))
    s, ts = O.find_optimal_shift(R, t, (8, 8))
    assert np.allclose(s, [0.5 / 8, 0.5 / 8])
    ok, _ = O.check_shift_invariance(R, t, s, (8, 8))
    assert ok


def test_spectral_counts_known_answers():
    """test/test_spectral_indexing.jl:43-71 and the Pmmm precondition of
    src/krfft.jl:465-466 (n_spec = M^3 at 64^3)."""
    import cfft_hip as cf
    R, t = cf.ops_arrays(cf.get_ops(1, 2, (4, 4)))
    s = O.calc_spectral_asu(R, t, (4, 4))
    assert s.idx.shape[0] == 16 and s.mult.sum() == 16
    R, t = cf.ops_arrays(cf.get_ops(6, 2, (16, 16)))
    s = O.calc_spectral_asu(R, t, (16, 16))
    assert s.mult.sum() == 256
    g = np.nonzero(np.all(O.get_k_vector(s.idx, (16, 16)) == 0, axis=1))[0]
    assert len(g) == 1 and s.mult[g[0]] == 1
    ops, ops_s, shift, R, t = group_setup(47, (64, 64, 64))
    s = O.calc_spectral_asu(R, t, (64, 64, 64))
    assert s.idx.shape[0] == 32 ** 3


def test_get_k_vector_known_answers():
    """test/test_spectral_indexing.jl:73-99."""
    assert O.get_k_vector(np.arange(4)[:, None], (4,)).ravel().tolist() == [0, 1, -2, -1]
    assert O.get_k_vector(np.arange(5)[:, None], (5,)).ravel().tolist() == [0, 1, 2, -2, -1]


def test_kernel_value_known_answer():
    """test/test_matrix_q.jl:24-37: K(1,0,0) = exp(-(2 pi/10)^2 0.01), L = 10."""
    spec = O.SpectralASU(np.array([[1, 0, 0]]), np.array([1]), (4, 4, 4))
    K = O.kernel_spec(spec, 0.01, 10.0 * np.eye(3))
    assert K[0] == pytest.approx(np.exp(-(2 * np.pi / 10) ** 2 * 0.01), rel=1e-15)


def test_auto_L_known_answers():
    """src/krfft.jl:219-222: Pmmm [2,2,2], Pmm2 [2,2,1]; P2/m != [2,2,2]
    (test/runtests.jl:394)."""
    for sg, L in [(47, [2, 2, 2]), (25, [2, 2, 1])]:
        _, _, _, R, t = group_setup(sg, (16, 16, 16))
        assert O.auto_L(t).tolist() == L
    _, _, _, R, t = group_setup(10, (16, 16, 16))
    assert O.auto_L(t).tolist() != [2, 2, 2]


def _brute_orbits(R, t, N):
    """Orbit of every point by repeated application (pure Python)."""
    pts = list(itertools.product(*[range(n) for n in N]))
    seen = {}
    orbits = []
    for p in pts:
        if p in seen:
            continue
        orb = {p}
        stack = [p]
        while stack:
            c = stack.pop()
            for g in range(R.shape[0]):
                y = tuple(int(v) for v in np.mod(R[g] @ np.array(c) + t[g], N))
                if y not in orb:
                    orb.add(y)
                    stack.append(y)
        for q in orb:
            seen[q] = len(orbits)
        orbits.append(orb)
    return orbits


@pytest.mark.parametrize("sg", [47, 229, 230, 70, 10])
def test_spectral_asu_brute_force(sg):
    """Brute-force orbits + extinction closed form at N = 4^3 / 8^3."""
    N = (8, 8, 8)
    _, _, _, R, t = group_setup(sg, N)
    s = O.calc_spectral_asu(R, t, N)
    Rs = O.dual_R(R)
    orbits = _brute_orbits(Rs, np.zeros_like(t), N)
    exp = []
    for orb in orbits:
        rep = min(orb, key=lambda k: k[0] + N[0] * (k[1] + N[1] * k[2]))
        k = np.array(rep)
        acc = 0j
        for g in range(R.shape[0]):
            if np.all(np.mod(Rs[g] @ k - k, N) == 0):
                acc += np.exp(-2j * np.pi * sum(k[d] * t[g, d] / N[d] for d in range(3)))
        if abs(acc) > 1e-5:
            exp.append((rep, len(orb)))
    exp.sort()
    assert [tuple(int(v) for v in r) for r in s.idx] == [e[0] for e in exp]
    assert s.mult.tolist() == [e[1] for e in exp]
    # orbit-size sum over ALL orbits = N^3
    assert sum(len(o) for o in orbits) == 512


def test_special_positions_brute_force():
    N = (12, 12, 12)
    for sg in (47, 229):
        _, ops_s, _, R, t = group_setup(sg, (12, 12, 12))
        orbits = _brute_orbits(R, t, N)
        exp = sum(1 for o in orbits if len(o) < R.shape[0])
        assert O.count_special_positions_fast(N, R, t) == exp


def test_symmetrize_gather_equals_scatter():
    import cfft_hip as cf
    for sg in (229, 70, 10):
        _, _, _, R, t = group_setup(sg, (8, 8, 8))
        u = np.random.default_rng(sg).standard_normal((8, 8, 8))
        a = O.symmetrize(u, R, t)
        b = O.symmetrize_scatter(u, R, t)
        assert np.max(np.abs(a - b)) < 1e-14
        # invariance under every op
        for g in range(R.shape[0]):
            X = np.stack(np.meshgrid(*[np.arange(8)] * 3, indexing="ij"), -1).reshape(-1, 3)
            y = np.mod(X @ R[g].T + t[g], 8)
            assert np.allclose(a[y[:, 0], y[:, 1], y[:, 2]], a[X[:, 0], X[:, 1], X[:, 2]], atol=1e-14)
