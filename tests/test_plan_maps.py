"""Plan-time index maps of libcfft_hip.so (C++) vs the oracle, bit-exact.

No GPU needed: these entry points are host-only.  Also checks that the library
loads and exports every symbol declared in include/cfft_hip.h.
"""
import os
import re

import numpy as np
import pytest

import cfft_hip as cf
import oracle as O
from cfft_hip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GROUPS = [2, 10, 25, 47, 63, 69, 70, 72, 74, 123, 136, 139, 200, 221, 223, 224, 225, 227, 229, 230]


def test_library_exports_header_symbols():
    hdr = open(os.path.join(ROOT, "include", "cfft_hip.h")).read()
    names = sorted(set(re.findall(r"^int (cfft_\w+)\(", hdr, re.M)))
    assert len(names) >= 20
    lib = _lib.lib()
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTED) == names


@pytest.mark.parametrize("sg", GROUPS)
@pytest.mark.parametrize("N", [(16, 16, 16), (24, 24, 24)])
def test_shift_and_spectral_asu_bit_exact(sg, N):
    try:
        ops = cf.get_ops(sg, 3, N)
    except ValueError:
        pytest.skip("grid not commensurate with the group")
    R0, t0 = cf.ops_arrays(ops)
    shift, ops_s = cf.find_optimal_shift(ops, N)
    s_ref, t_ref = O.find_optimal_shift(R0, t0, N)
    assert np.array_equal(shift, s_ref)
    R, t = cf.ops_arrays(ops_s)
    assert np.array_equal(R, R0) and np.array_equal(t, t_ref)
    spec = cf.calc_spectral_asu(ops_s, 3, N)
    ref = O.calc_spectral_asu(R, t, N)
    assert np.array_equal(spec.idx, ref.idx)
    assert np.array_equal(spec.mult, ref.mult)
    Lv = cf.auto_L(ops_s)
    assert np.array_equal(Lv, O.auto_L(t))
    assert np.array_equal(cf.select_rep_ops(ops_s, Lv, N), O.select_rep_ops(R, t, Lv, N))
    assert int(cf.detect_centering(ops_s, N)) == O.detect_centering_type(R, t, N)
    if all(N[d] % Lv[d] == 0 for d in range(3)):
        assert np.array_equal(cf.build_fill_map(ops_s, Lv, N), O.fill_map(R, t, Lv, N))


@pytest.mark.parametrize("sg", [47, 229, 230, 70])
def test_special_positions_bit_exact(sg):
    N = (12, 12, 12)
    ops = cf.get_ops(sg, 3, N)
    _, ops_s = cf.find_optimal_shift(ops, N)
    R, t = cf.ops_arrays(ops_s)
    assert cf.count_special_positions(N, ops_s) == O.count_special_positions_fast(N, R, t)


def test_2d_known_answers():
    """p2mm shift (test/runtests.jl:42-52) and spectral counts
    (test/test_spectral_indexing.jl:43-71) through the C ABI."""
    ops = cf.get_ops(6, 2, (8, 8))
    shift, _ = cf.find_optimal_shift(ops, (8, 8))
    assert np.allclose(shift, [0.5 / 8, 0.5 / 8])
    s = cf.calc_spectral_asu(cf.get_ops(1, 2, (4, 4)), 2, (4, 4))
    assert len(s) == 16 and s.mult.sum() == 16
    s = cf.calc_spectral_asu(cf.get_ops(6, 2, (16, 16)), 2, (16, 16))
    assert s.mult.sum() == 256
    gam = [i for i in range(len(s)) if np.all(cf.get_k_vector(s, i) == 0)]
    assert len(gam) == 1 and s.mult[gam[0]] == 1
    with pytest.raises(IndexError, match="out of bounds"):
        cf.get_k_vector(s, len(s))


@pytest.mark.parametrize("sg,N,n_exp", [(47, (64, 64, 64), 32768)])
def test_pmmm_64_nspec(sg, N, n_exp):
    """Precondition of the Pmmm fast path, src/krfft.jl:465-466."""
    _, ops_s = cf.find_optimal_shift(cf.get_ops(sg, 3, N), N)
    assert len(cf.calc_spectral_asu(ops_s, 3, N)) == n_exp


def test_im3m_128_spectral_asu_matches_oracle():
    """Full-size config grid: Im-3m 128^3 spectral ASU, bit-exact."""
    N = (128, 128, 128)
    _, ops_s = cf.find_optimal_shift(cf.get_ops(229, 3, N), N)
    R, t = cf.ops_arrays(ops_s)
    spec = cf.calc_spectral_asu(ops_s, 3, N)
    ref = O.calc_spectral_asu(R, t, N)
    assert np.array_equal(spec.idx, ref.idx) and np.array_equal(spec.mult, ref.mult)
    # orbit sizes of the kept points + extinct orbits cover the grid once
    assert spec.mult.sum() <= 128 ** 3


def test_plan_errors_before_device_use():
    """Host-side validation maps to the reference's error sites."""
    N = (16, 16, 16)
    _, ops_s = cf.find_optimal_shift(cf.get_ops(47, 3, N), N)
    # P-centring passed to the centred path (src/q_fused_krfft.jl:833-836)
    with pytest.raises(cf.CFFTError) as e:
        cf.DevicePlan(N, ops_s, 0.01, None, cf.PATH_M7)
    assert e.value.code == _lib.CFFT_ERR_CENTERING
    # odd M (src/centering_fold.jl:61): N = 18 -> M = 9
    _, ops18 = cf.find_optimal_shift(cf.get_ops(47, 3, (18, 18, 18)), (18, 18, 18))
    with pytest.raises(cf.CFFTError) as e:
        cf.DevicePlan((18, 18, 18), ops18, 0.01, None, cf.PATH_M2)
    assert e.value.code == _lib.CFFT_ERR_ODD_M
    # unknown method (src/diffusion_solver.jl:250)
    with pytest.raises(cf.CFFTError) as e:
        cf.DevicePlan(N, ops_s, 0.01, None, 17)
    assert e.value.code == _lib.CFFT_ERR_UNKNOWN_METHOD
    with pytest.raises(cf.CFFTError) as e:
        cf.plan_diffusion(N, np.eye(3), 47, 3, 0.01, method="krfft")
    assert e.value.code == _lib.CFFT_ERR_UNKNOWN_METHOD
    assert "Unknown method: krfft" in str(e.value)
    # N not divisible by L (src/krfft.jl:309-311)
    with pytest.raises(cf.CFFTError) as e:
        cf.build_fill_map(ops_s, [2, 2, 2], (15, 16, 16))
    assert e.value.code == _lib.CFFT_ERR_NOT_DIVISIBLE


def test_plan_m7_rejects_primitive():
    with pytest.raises(cf.CFFTError) as e:
        cf.plan_m7_scft((16, 16, 16), 221, 3, 0.01, np.eye(3))
    assert e.value.code == _lib.CFFT_ERR_CENTERING
