"""Real-space ASU, block packing, block CFFT and the full-grid centring pre-fold.

CPU part: the oracle against the reference's known answers
(test/test_asu.jl, test/test_packing.jl) and brute-force orbit coverage, and
the library's host-side calc_asu / pack_asu bit-exact against the oracle.
GPU part (marked ``gpu``): plan_cfft mul/ldiv against NumPy's FFT of every
block, and the pre-fold KRFFT against the full-grid FFT
(test/test_cfft.jl, test/test_centering_prefold.jl).  FP64 bar 1e-12
relative.
"""
import numpy as np
import pytest

import cfft_hip as cf
import oracle as O
from conftest import group_setup

CASES = [(6, (8, 8)), (7, (8, 8)), (47, (8, 8, 8)), (6, (16, 16)), (229, (16, 16, 16)),
         (230, (16, 16, 16)), (225, (16, 16, 16)), (70, (16, 16, 16)), (221, (12, 12, 12)),
         (1, (6, 6, 6)), (2, (8, 8, 8)), (139, (16, 16, 16)), (227, (16, 16, 16))]


def _orbit(p, R, t, N):
    orb = {tuple(p)}
    stack = [tuple(p)]
    Nv = np.asarray(N)
    while stack:
        c = np.asarray(stack.pop())
        for g in range(R.shape[0]):
            y = tuple(int(v) for v in np.mod(R[g] @ c + t[g], Nv))
            if y not in orb:
                orb.add(y)
                stack.append(y)
    return orb


@pytest.mark.parametrize("sg,N,expected", [(6, (8, 8), 16), (7, (8, 8), 16), (47, (8, 8, 8), 64)])
def test_oracle_asu_known_sizes(sg, N, expected):
    """test/test_asu.jl:42-72: sizes, coverage and disjointness."""
    _, _, _, R, t = group_setup(sg, N)
    pts = O.calc_asu(N, R, t)
    assert len(pts) == expected
    seen = set()
    for idx, _, mult in pts:
        orb = _orbit(idx, R, t, N)
        assert not (seen & orb)
        assert len(orb) == mult
        seen |= orb
    assert len(seen) == int(np.prod(N))


@pytest.mark.parametrize("sg,N", [(229, (16, 16, 16)), (230, (16, 16, 16)), (70, (8, 8, 8)),
                                  (221, (12, 12, 12)), (2, (8, 8, 8))])
def test_oracle_asu_partitions_grid(sg, N):
    """Orbits of the ASU points partition the grid; multiplicities sum to N^3."""
    _, _, _, R, t = group_setup(sg, N)
    pts = O.calc_asu(N, R, t)
    seen = set()
    for idx, _, mult in pts:
        orb = _orbit(idx, R, t, N)
        assert not (seen & orb)
        assert len(orb) == mult
        seen |= orb
    assert len(seen) == int(np.prod(N))
    assert sum(m for _, _, m in pts) == int(np.prod(N))


@pytest.mark.parametrize("sg,N", [(6, (8, 8)), (7, (8, 8)), (47, (8, 8, 8)), (229, (16, 16, 16))])
def test_oracle_pack_covers_points(sg, N):
    """test/test_packing.jl:7-24: packed sizes sum to the point count; blocks
    are disjoint boxes of points of one depth."""
    _, _, _, R, t = group_setup(sg, N)
    pts = O.calc_asu(N, R, t)
    blocks = O.pack_asu(pts, N)
    depth_of = {idx: dep for idx, dep, _ in pts}
    covered = set()
    for start, shape, depth in blocks:
        import itertools
        box = set(itertools.product(*[range(s, s + n) for s, n in zip(start, shape)]))
        assert not (covered & box)
        assert all(depth_of.get(p) == depth for p in box)
        covered |= box
    assert covered == set(depth_of)
    assert sum(int(np.prod(s)) for _, s, _ in blocks) == len(pts)


def test_oracle_pack_p2mm_one_block():
    """p2mm 8x8 (half-pixel shift): the 16 points form one 4x4 block."""
    _, _, _, R, t = group_setup(6, (8, 8))
    blocks = O.pack_asu(O.calc_asu((8, 8), R, t), (8, 8))
    assert len(blocks) == 1 and blocks[0][1] == (4, 4)


@pytest.mark.parametrize("sg,N", CASES)
def test_calc_asu_bit_exact(sg, N):
    _, ops_s, _, R, t = group_setup(sg, N)
    pts = cf.calc_asu(N, ops_s)
    ref = O.calc_asu(N, R, t)
    assert len(pts) == len(ref)
    for p, (idx, dep, mult) in zip(pts, ref):
        assert tuple(int(v) for v in p.idx) == idx
        assert tuple(int(v) for v in p.depth) == dep
        assert p.multiplicity == mult
    pts2, shift = cf.calc_asu(sg, len(N), N)
    assert len(pts2) == len(pts)
    assert np.allclose(shift, group_setup(sg, N)[2])


@pytest.mark.parametrize("sg,N", CASES)
def test_pack_asu_bit_exact(sg, N):
    _, ops_s, _, R, t = group_setup(sg, N)
    pts = cf.calc_asu(N, ops_s)
    asu = cf.pack_asu(pts, N)
    ref = O.pack_asu(O.calc_asu(N, R, t), N)
    blocks = asu.blocks()
    assert len(blocks) == len(ref)
    for b, (start, shape, depth) in zip(blocks, ref):
        assert tuple(r.start for r in b.range) == start
        assert b.data.shape == shape
        assert tuple(int(v) for v in b.depth) == depth
    # grouping by effective dimension (src/pack_asu.jl:92-98)
    for eff, bl in asu.dim_blocks.items():
        assert all(sum(1 for r in b.range if len(r) > 1) == eff for b in bl)
    assert sum(b.data.size for b in blocks) == len(pts)


def test_pack_asu_empty():
    asu = cf.pack_asu([], (8, 8, 8))
    assert asu.dim_blocks == {} and asu.blocks() == []


# ---------------------------------------------------------------------------
# GPU: block CFFT and pre-fold KRFFT
# ---------------------------------------------------------------------------

def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.gpu
@pytest.mark.parametrize("sg,N", [(6, (16, 16)), (7, (8, 8)), (47, (8, 8, 8)), (229, (16, 16, 16)),
                                  (230, (16, 16, 16)), (1, (6, 6, 6)), (221, (12, 12, 12)),
                                  (225, (32, 32, 32))])
def test_block_cfft_vs_numpy(sg, N):
    """test/test_cfft.jl:70-110: mul!/ldiv! per block vs FFT, and round trip."""
    print(f"block cfft sg={sg} N={N}", flush=True)
    plan = cf.plan_cfft(N, sg)
    rng = np.random.default_rng(sg)
    x = plan.asu.similar(np.complex128)
    for b in x.blocks():
        b.data[...] = rng.standard_normal(b.data.shape) + 1j * rng.standard_normal(b.data.shape)
    y = plan.asu.similar(np.complex128)
    cf.mul(y, plan, x)
    ref = O.block_fft([b.data for b in x.blocks()])
    for b, r in zip(y.blocks(), ref):
        assert _rel(b.data, r) < 1e-12
    z = plan.asu.similar(np.complex128)
    cf.ldiv(z, plan, y)
    for b, r in zip(z.blocks(), x.blocks()):
        assert _rel(b.data, r.data) < 1e-12
    ref_i = O.block_ifft([b.data for b in y.blocks()])
    for b, r in zip(z.blocks(), ref_i):
        assert _rel(b.data, r) < 1e-12


@pytest.mark.gpu
def test_block_cfft_parseval_and_linearity():
    """Size-independent properties at 64^3 Im-3m: Parseval per block
    (sum |X|^2 = n sum |x|^2) and linearity of mul."""
    N = (64, 64, 64)
    plan = cf.plan_cfft(N, 229)
    rng = np.random.default_rng(5)
    x1, x2 = plan.asu.similar(np.complex128), plan.asu.similar(np.complex128)
    for b1, b2 in zip(x1.blocks(), x2.blocks()):
        b1.data[...] = rng.standard_normal(b1.data.shape)
        b2.data[...] = rng.standard_normal(b2.data.shape) * 1j
    xs = plan.asu.similar(np.complex128)
    for bs, b1, b2 in zip(xs.blocks(), x1.blocks(), x2.blocks()):
        bs.data[...] = 2.0 * b1.data - 3.0 * b2.data
    y1, y2, ys = (plan.asu.similar(np.complex128) for _ in range(3))
    cf.mul(y1, plan, x1)
    cf.mul(y2, plan, x2)
    cf.mul(ys, plan, xs)
    for a, b, c, x in zip(y1.blocks(), y2.blocks(), ys.blocks(), x1.blocks()):
        assert _rel(c.data, 2.0 * a.data - 3.0 * b.data) < 1e-12
        n = x.data.size
        assert abs(np.sum(np.abs(a.data) ** 2) - n * np.sum(np.abs(x.data) ** 2)) <= 1e-10 * n * np.sum(np.abs(x.data) ** 2)


@pytest.mark.gpu
def test_block_cfft_errors():
    plan = cf.plan_cfft((8, 8, 8), 47)
    other = cf.pack_asu(cf.calc_asu((8, 8, 8), group_setup(229, (8, 8, 8))[1]), (8, 8, 8))
    y = plan.asu.similar(np.complex128)
    with pytest.raises(cf.CFFTError):
        cf.mul(y, plan, other)


PREFOLD = [(229, (16, 16, 16), cf.CentI, 4, (8, 8, 8)), (139, (16, 16, 16), cf.CentI, 4, (8, 8, 8)),
           (225, (16, 16, 16), cf.CentF, 2, (8, 8, 8)), (70, (16, 16, 16), cf.CentF, 2, (8, 8, 8)),
           (69, (16, 16, 16), cf.CentF, 2, (8, 8, 8)), (35, (16, 16, 16), cf.CentC, 2, (8, 8, 16)),
           (63, (16, 16, 16), cf.CentC, 2, (8, 8, 16)), (229, (32, 32, 32), cf.CentI, 4, (16, 16, 16)),
           (225, (32, 32, 32), cf.CentF, 2, (16, 16, 16)), (230, (24, 24, 24), cf.CentI, 4, (12, 12, 12))]


@pytest.mark.gpu
@pytest.mark.parametrize("sg,N,cent,nch,Nsub", PREFOLD)
def test_prefold_krfft(sg, N, cent, nch, Nsub):
    """test/test_centering_prefold.jl:104-160: pre-fold KRFFT = full FFT at the
    spectral-ASU points (< 1e-10 there; 1e-12 relative here), and equal to the
    oracle's pre-fold + merge."""
    print(f"prefold sg={sg} N={N}", flush=True)
    ops = cf.get_ops(sg, 3, N)
    R, t = cf.ops_arrays(ops)
    plan = cf.plan_krfft_centered(sg, N)
    assert isinstance(plan, cf.CenteringPreFoldKRFFTPlan)
    assert plan.centering == cent and plan.n_channels == nch and plan.N_sub == Nsub
    u = O.make_symmetric(R, t, N, seed=sg)
    out = cf.execute_centered_krfft(plan, u).copy()
    spec = plan.full_spec_asu
    ref = np.fft.fftn(u)[tuple(spec.idx[:, d] for d in range(3))]
    assert _rel(out, ref) < 1e-12
    chans, pars, Ns = O.centering_prefold(u, int(cent))
    hat = [np.fft.fftn(c) for c in chans]
    assert _rel(out, O.prefold_merge(spec, hat, pars, Ns, N)) < 1e-12


@pytest.mark.gpu
def test_prefold_from_ops_and_centp_error():
    N = (16, 16, 16)
    plan = cf.plan_krfft_centered(cf.get_ops(229, 3, N), N)
    assert plan.n_channels == 4
    with pytest.raises(cf.CFFTError) as e:
        cf.plan_krfft_centered(221, N)
    assert e.value.code == 6
