"""Interleaved packing (Mode B): analyze_interleaved_orbits and
pack_asu_interleaved against the reference's own known answers
(test/test_interleaved_orbits.jl, test/test_packing_mode_b.jl) and against
what the group fixes: for a closed group the orbit of a coset shift p is
{R p + t mod L : (R, t) in G}, the orbits partition the L-cosets, and the
ops listed for a member are exactly those sending the representative
there.  The plan built from a Mode B real ASU (src/krfft.jl:119) is checked
against the full FFT on the GPU."""
import itertools

import numpy as np
import pytest

import cfft_hip as cf


def _coset_orbit(p, ops, Lv):
    return {tuple(int(v) for v in np.mod(op.Rm @ np.asarray(p) + op.tv, Lv)) for op in ops}


def test_reference_known_answers():
    """test/test_interleaved_orbits.jl: P1 8 orbits of 1; P4 6 orbits, two of
    size 2, multiplicities summing to 8; Pmmm 8 orbits of 1; P4mm at most 6."""
    N = (16, 16, 16)
    o = cf.analyze_interleaved_orbits(N, cf.get_ops(1, 3, N), L=(2, 2, 2))
    assert len(o) == 8 and all(x.multiplicity == 1 for x in o)
    o = cf.analyze_interleaved_orbits(N, cf.get_ops(75, 3, N), L=(2, 2, 2))
    assert len(o) == 6 and sum(x.multiplicity for x in o) == 8
    assert sum(x.multiplicity == 2 for x in o) == 2
    o = cf.analyze_interleaved_orbits(N, cf.get_ops(47, 3, N), L=(2, 2, 2))
    assert len(o) == 8 and all(x.multiplicity == 1 for x in o)
    o = cf.analyze_interleaved_orbits(N, cf.get_ops(99, 3, N), L=(2, 2, 2))
    assert len(o) <= 6
    # the default stride is 2 on every axis
    assert len(cf.analyze_interleaved_orbits(N, cf.get_ops(75, 3, N))) == 6


@pytest.mark.parametrize("sg", [1, 2, 47, 70, 75, 99, 123, 139, 221, 223, 225, 227, 229, 230])
@pytest.mark.parametrize("shifted", [False, True])
def test_orbits_are_group_orbits(sg, shifted):
    """Members = the group orbit of the representative mod L; orbits
    partition the 8 cosets; the representative is the first coset of its
    orbit in product order (first axis fastest); members sorted; each
    member's op list is every op that sends the representative there."""
    N = (24, 24, 24)
    ops = cf.get_ops(sg, 3, N)
    if shifted:
        _, ops = cf.find_optimal_shift(ops, N)
    Lv = np.array([2, 2, 2])
    orbits = cf.analyze_interleaved_orbits(N, ops, L=(2, 2, 2))
    order = [tuple(p[::-1]) for p in itertools.product(range(2), repeat=3)]  # first axis fastest
    seen = set()
    for o in orbits:
        rep = tuple(int(v) for v in o.representative)
        mem = [tuple(int(v) for v in m) for m in o.members]
        assert set(mem) == _coset_orbit(rep, ops, Lv)
        assert mem == sorted(mem) and o.multiplicity == len(mem)
        assert min(order.index(m) for m in mem) == order.index(rep)
        assert not (set(mem) & seen)
        seen |= set(mem)
        for m, mops in zip(mem, o.ops):
            want = [op for op in ops if tuple(int(v) for v in np.mod(op.Rm @ np.asarray(rep) + op.tv, Lv)) == m]
            assert mops == want
    assert seen == set(order)


def test_pack_interleaved_blocks():
    """asu_only (default): one block, the zero-shift subgrid u[0::2, ...],
    no orbit; asu_only=False: one block per orbit representative with its
    orbit, data u[r::2, ...], 0-based strided ranges, zero depths, all under
    dim_blocks[3]; test/test_packing_mode_b.jl's P4 and Pmmm counts."""
    N = (16, 16, 16)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    for sg, n_orb in ((75, 6), (47, 8), (229, None)):
        ops = cf.get_ops(sg, 3, N)
        a = cf.pack_asu_interleaved(u, N, ops, L=(2, 2, 2))
        assert list(a.dim_blocks) == [3] and len(a.dim_blocks[3]) == 1
        b = a.dim_blocks[3][0]
        assert b.orbit is None and np.array_equal(b.data, u[0::2, 0::2, 0::2])
        assert [(r.start, r.step, len(r)) for r in b.range] == [(0, 2, 8)] * 3
        assert np.array_equal(b.depth, np.zeros(3)) and a.shift == (0.0, 0.0, 0.0)
        full = cf.pack_asu_interleaved(u, N, ops, L=(2, 2, 2), asu_only=False)
        blocks = full.dim_blocks[3]
        orbits = cf.analyze_interleaved_orbits(N, ops, L=(2, 2, 2))
        assert len(blocks) == len(orbits)
        if n_orb is not None:
            assert len(blocks) == n_orb
        for blk, o in zip(blocks, orbits):
            r = [int(v) for v in o.representative]
            assert blk.orbit is not None and np.array_equal(blk.orbit.representative, o.representative)
            assert np.array_equal(blk.data, u[r[0]::2, r[1]::2, r[2]::2])
            assert [rg.start for rg in blk.range] == r and all(rg.step == 2 for rg in blk.range)
        assert sum(blk.data.size for blk in blocks) == len(orbits) * (16 ** 3) // 8
        if sg == 47:  # Pmmm: no reduction, the blocks cover the grid
            assert sum(blk.data.size for blk in blocks) == 16 ** 3
    with pytest.raises(cf.CFFTError):
        cf.pack_asu_interleaved(u[:8], N, cf.get_ops(47, 3, N))


@pytest.mark.gpu
@pytest.mark.parametrize("sg", [47, 229])
def test_plan_from_interleaved_asu(sg):
    """plan_krfft(real_asu, spec, ops) with the Mode B ASU: the forward
    reconstruction from input_buffer = the stride-2 subgrid matches the full
    FFT of the symmetric field at every spectral-ASU point (1e-12 relative),
    as the reference's Mode B benchmark checks; a stride that is not auto_L's
    is refused."""
    from conftest import group_setup
    import oracle as O
    N = (16, 16, 16)
    ops, ops_s, shift, R, t = group_setup(sg, N)
    u = O.make_symmetric(R, t, N, seed=sg)
    asu = cf.pack_asu_interleaved(u, N, ops_s, L=(2, 2, 2), asu_only=True)
    spec = cf.calc_spectral_asu(ops_s, 3, N)
    plan = cf.plan_krfft(asu, spec, ops_s)
    plan.input_buffer[...] = u[0::2, 0::2, 0::2]
    F = cf.fft_reconstruct(plan)
    ref = np.fft.fftn(u)
    for i in range(len(spec.points)):
        h = np.mod(cf.get_k_vector(spec, i), N)
        assert abs(F[i] - ref[tuple(h)]) <= 1e-12 * np.max(np.abs(ref))
    bad = cf.pack_asu_interleaved(u, N, ops_s, L=(4, 4, 4))
    with pytest.raises(cf.CFFTError):
        cf.plan_krfft(bad, spec, ops_s)
