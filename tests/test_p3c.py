"""Tables of the P3c KRFFT families (cfft_hip.p3c), on the CPU: applied to
NumPy FFTs of the stride-4 sub-sub-grids of a symmetrised field, each
family's sparse stages give the full-grid FFT at every spectral-ASU point
(the check the reference's family tests make, test/test_selective_g0.jl,
test/test_g0_asu.jl), for cubic groups whose 3-fold carries a translation
after the magic shift as well as for those where it does not.  The class
structure of the 8 sub-sub-grid offsets under a translation-free 3-fold is
the reference's four buffers (000, 001-class, 110-class, 111)."""
import numpy as np
import pytest

import cfft_hip as cf
import oracle as O
from cfft_hip import p3c
from conftest import group_setup


def _spmv(stage, x):
    rowptr, cols, vals = stage
    y = np.zeros(len(rowptr) - 1, dtype=np.complex128)
    rows = np.repeat(np.arange(len(rowptr) - 1), np.diff(rowptr))
    np.add.at(y, rows, vals * x[cols])
    return y


def _rep_ffts(u, cls):
    eps, orbit_rep, chain, reps = cls
    return np.concatenate([np.fft.fftn(u[2 * eps[r][0]::4, 2 * eps[r][1]::4, 2 * eps[r][2]::4]).ravel(order="F")
                           for r in reps])


@pytest.mark.parametrize("sg", [200, 221, 223, 225, 227, 229, 230])
@pytest.mark.parametrize("n", [16, 24, 32])
def test_family_tables_give_full_fft(sg, n):
    N = (n, n, n)
    ops, ops_s, shift, R, t = group_setup(sg, N)
    u = O.make_symmetric(R, t, N, seed=sg + n)
    spec = cf.calc_spectral_asu(ops_s, 3, N)
    ref = np.fft.fftn(u)[tuple(spec.idx[:, d] for d in range(3))]
    scale = np.max(np.abs(ref))
    sizes = {}
    for kind in ("sparse", "recursive", "selective", "g0asu"):
        cls, stages, info = p3c.family_tables(kind, spec, ops_s)
        x = _rep_ffts(u, cls)
        for st in stages:
            assert st[1].min() >= 0 and st[1].max() < len(x)
            x = _spmv(st, x)
        assert np.max(np.abs(x - ref)) / scale < 1e-12, kind
        sizes[kind] = info
    # the selective positions are a subset of the M-grid, the G0-ASU
    # representatives a subset of those
    M3 = (n // 2) ** 3
    assert sizes["recursive"]["g0_positions"] == M3
    assert sizes["g0asu"]["g0_reps"] < sizes["selective"]["g0_positions"] <= M3


def test_sparse_rows_and_classes():
    """Sparse rows merge to at most 64 entries (8 ops x 8 offsets); a
    translation-free 3-fold splits the offsets into the reference's four
    classes {000}, {100, 010, 001}, {110, 011, 101}, {111}."""
    N = (32, 32, 32)
    ops = cf.get_ops(225, 3, N)  # Fm-3m: the 3-fold through the origin, no shift
    spec = cf.calc_spectral_asu(ops, 3, N)
    eps, orbit_rep, chain, reps = p3c.p3c_classes(ops, 32)
    assert reps == [0, 1, 3, 7]
    classes = {}
    for e in range(8):
        classes.setdefault(orbit_rep[e], set()).add(tuple(int(v) for v in eps[e]))
    assert classes[1] == {(1, 0, 0), (0, 1, 0), (0, 0, 1)}
    assert classes[3] == {(1, 1, 0), (0, 1, 1), (1, 0, 1)}
    _, (stage,), _ = p3c.family_tables("sparse", spec, ops)
    assert np.diff(stage[0]).max() <= 64


def test_non_cubic_groups_refused():
    """Without the body-diagonal 3-fold (Pmmm, Fddd, I4/mmm) the P3c tables
    are refused; the families then run the device pipeline instead."""
    for sg in (47, 70, 139):
        N = (16, 16, 16)
        _, ops_s = cf.find_optimal_shift(cf.get_ops(sg, 3, N), N)
        with pytest.raises(cf.CFFTError) as e:
            p3c.p3c_classes(ops_s, 16)
        assert e.value.code == cf._lib.CFFT_ERR_UNSUPPORTED
