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


@pytest.mark.parametrize("n", [16, 24, 32])
def test_staged_g0_table(n):
    """The staged Pm-3m table (G0 at every M-grid point) gives the FFT of the
    stride-2 subgrid, fftn(u[::2, ::2, ::2]), for a Pm-3m-symmetric field."""
    N = (n, n, n)
    ops, ops_s, shift, R, t = group_setup(221, N)
    u = O.make_symmetric(R, t, N, seed=n)
    cls, (stage,) = p3c.g0_table(ops_s, n)
    got = _spmv(stage, _rep_ffts(u, cls))
    ref = np.fft.fftn(u[::2, ::2, ::2]).ravel(order="F")
    assert np.diff(stage[0]).max() <= 8
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12


@pytest.mark.parametrize("sg,n", [(229, 32), (221, 16), (225, 32), (230, 24), (47, 16), (70, 32),
                                  (139, 32), (2, 16), (227, 32), (200, 16), (123, 16)])
def test_fractal_tables_give_full_fft(sg, n):
    """The fractal decimation tree (cfft_hip.p3c.fractal_tables), at every
    depth up to the one the plan picks: NumPy FFTs of the leaves, run through
    the stages, give the full-grid FFT at every spectral-ASU point (the check
    of test/test_fractal_krfft.jl), for cubic and non-cubic groups."""
    N = (n, n, n)
    ops, ops_s, shift, R, t = group_setup(sg, N)
    u = O.make_symmetric(R, t, N, seed=sg + 2 * n)
    spec = cf.calc_spectral_asu(ops_s, 3, N)
    ref = np.fft.fftn(u)[tuple(spec.idx[:, d] for d in range(3))]
    D = p3c.fractal_depth(ops_s, n)
    assert D >= 1
    leaves = []
    for depth in range(1, D + 1):
        S, offs, stages, info = p3c.fractal_tables(spec, ops_s, depth)
        assert S == 2 ** depth and len(stages) == depth and info["leaves"] == len(offs)
        x = np.concatenate([np.fft.fftn(u[o[0]::S, o[1]::S, o[2]::S]).ravel(order="F") for o in offs])
        for st in stages:
            assert st[1].min() >= 0 and st[1].max() < len(x)
            assert np.diff(st[0]).max() <= 8
            x = _spmv(st, x)
        assert np.max(np.abs(x - ref)) / np.max(np.abs(ref)) < 1e-12, depth
        leaves.append(len(offs) * (n // S) ** 3)
    # every level the plan takes shrinks the leaves
    assert all(b < a for a, b in zip([n ** 3] + leaves, leaves))


def test_fractal_depth_and_orbits():
    """Split depth: the magic-shifted groups tested here relate all 8 parity
    classes at the first split; Pmmm and P-1 stop there (the next split
    gives 8x the leaves), Im-3m goes on while the leaves shrink; every
    offset's (R, d) maps it onto its representative."""
    for sg, n, want in ((47, 16, 1), (2, 16, 1), (229, 32, 4), (230, 24, 3)):
        N = (n, n, n)
        _, ops_s = cf.find_optimal_shift(cf.get_ops(sg, 3, N), N)
        assert p3c.fractal_depth(ops_s, n) == want, sg
        S = 4
        E, rep, Rs, ds = p3c._offset_orbits(ops_s, n, S)
        key = E[:, 0] + S * (E[:, 1] + S * E[:, 2])
        img = np.einsum("eij,ej->ei", Rs, E)
        # R e + t = rep + S d for the op chosen per offset (t from that op)
        for i in range(len(E)):
            t = [np.mod(np.rint(op.tv).astype(np.int64), n) for op in ops_s
                 if np.array_equal(np.rint(op.Rm).astype(np.int64), Rs[i])]
            assert any(np.array_equal(img[i] + tt, E[key == rep[i]][0] + S * ds[i]) for tt in t)
        assert np.all(rep <= key)
    # P1: nothing to split
    N = (16, 16, 16)
    assert p3c.fractal_depth(cf.get_ops(1, 3, N), 16) == 0


def test_half_grid_shift():
    """shift_ops_half_grid (src/fractal_krfft.jl:265-287): t' = t + (I - R)/2
    summed over each row; a mirror x -> -x gains t = 1, a 3-fold permutation
    nothing, the identity nothing; the shifted ops of Im-3m form a group (the
    original conjugated by the half-cell translation); the fractal tables over
    them give the full FFT of a field symmetric under them with the spectral
    ASU of the unshifted ops (the reference's fractal test convention)."""
    N = (16, 16, 16)
    ops = cf.get_ops(221, 3, N)
    sh = cf.shift_ops_half_grid(ops, N)
    for op, o2 in zip(ops, sh):
        R = op.Rm
        assert np.array_equal(o2.Rm, R)
        want = np.mod(op.tv + ((np.eye(3, dtype=np.int64) - R).sum(axis=1)) // 2, 16)
        assert np.array_equal(o2.tv, want)
    mirror = [o2 for op, o2 in zip(ops, sh) if np.array_equal(op.Rm, np.diag([-1, 1, 1])) and not op.tv.any()]
    assert mirror and np.array_equal(mirror[0].tv, [1, 0, 0])
    for sg, n in ((229, 16), (47, 16), (230, 24)):
        N = (n, n, n)
        ops = cf.get_ops(sg, 3, N)
        sh = cf.shift_ops_half_grid(ops, N)
        keys = {(tuple(o.Rm.ravel()), tuple(np.mod(o.tv, n))) for o in sh}
        for a in sh[:8]:
            for b in sh[:8]:
                R = a.Rm @ b.Rm
                t = np.mod(a.Rm @ b.tv + a.tv, n)
                assert (tuple(R.ravel()), tuple(t)) in keys
        R, t = cf.ops_arrays(sh)
        u = O.make_symmetric(R, t, N, seed=n)
        spec = cf.calc_spectral_asu(ops, 3, N)
        ref = np.fft.fftn(u)[tuple(spec.idx[:, d] for d in range(3))]
        D = p3c.fractal_depth(sh, n)
        S, offs, stages, info = p3c.fractal_tables(spec, sh, D)
        x = np.concatenate([np.fft.fftn(u[o[0]::S, o[1]::S, o[2]::S]).ravel(order="F") for o in offs])
        for st in stages:
            x = _spmv(st, x)
        assert np.max(np.abs(x - ref)) / np.max(np.abs(ref)) < 1e-12, sg
