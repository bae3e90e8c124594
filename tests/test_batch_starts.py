"""Start fields of the batched-instance workload: the lowest star that the
group does not extinguish (host code; the symmetrisation is the gather form of
test/test_helpers.jl:11-29)."""
import numpy as np
import pytest

from cfft_hip import batch
from cfft_hip.plans import auto_L, find_optimal_shift
from cfft_hip.scft import symmetrize_subgrid
from cfft_hip.symmetry_ops import get_ops

# {100} survives only in Pmmm; {110} is allowed in Pm-3n (hhl: l even) and
# Im-3m (h+k+l even); {111} is the first Fddd star (all odd); Ia-3d extinguishes
# {110} (hhl: 2h+l = 4n) and {200} (00l: l = 4n), leaving {211}
EXPECTED = {47: (1, 0, 0), 223: (1, 1, 0), 229: (1, 1, 0), 70: (1, 1, 1), 230: (2, 1, 1)}


@pytest.mark.parametrize("sg", sorted(EXPECTED))
def test_lowest_star_is_first_allowed_reflection(sg):
    N = (16, 16, 16)
    ops = get_ops(sg, 3, N)
    shift, ops_s = find_optimal_shift(ops, N)
    Lv = auto_L(ops_s)
    u, h = batch.lowest_star(N, ops_s, shift, Lv)
    assert tuple(h) == EXPECTED[sg]
    assert abs(np.sqrt(np.mean(u * u)) - 1.0) < 1e-12
    # every smaller candidate star symmetrises to zero (extinct), cos and sin
    x = [(np.arange(n) + s * n) / n for n, s in zip(N, shift)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    for k in batch._STARS[:batch._STARS.index(tuple(h))]:
        arg = 2 * np.pi * (k[0] * X + k[1] * Y + k[2] * Z)
        for fn in (np.cos, np.sin):
            assert np.max(np.abs(symmetrize_subgrid(fn(arg), ops_s, Lv))) < 1e-12


def test_start_fields_seeded_and_antisymmetric_in_the_harmonic():
    inst = batch.Instance((16, 16, 16), 229, 4.0, 0.3, 15.0, 20, 7)
    ops = get_ops(inst.sg, 3, inst.N)
    shift, ops_s = find_optimal_shift(ops, inst.N)
    a1, b1 = batch.start_fields(inst, ops_s, shift, amp=2.0, noise=0.0)
    assert np.allclose(a1, -b1, atol=1e-14)
    a2, b2 = batch.start_fields(inst, ops_s, shift)
    a3, b3 = batch.start_fields(inst, ops_s, shift)
    assert np.array_equal(a2, a3) and np.array_equal(b2, b3)
    assert 0.0 < np.max(np.abs(a2 - a1)) < 1.0


@pytest.mark.parametrize("sg,N,lat,expect", [
    (229, (32, 32, 32), 4.0 * np.eye(3), True), (230, (32, 32, 32), 3.8 * np.eye(3), True),
    (70, (32, 32, 32), np.diag([4.0, 5.0, 6.0]), True), (47, (24, 24, 24), 4.0 * np.eye(3), True),
    (223, (48, 48, 48), 4.0 * np.eye(3), True),
    (10, (16, 16, 16), 4.0 * np.eye(3), False),            # L = (2,2,1)
    (71, (24, 32, 48), 4.0 * np.eye(3), False),            # not cubic
    (229, (32, 32, 32), np.array([[4.0, 0.5, 0], [0, 4.0, 0], [0, 0, 4.0]]), False)])  # oblique
def test_plane_pair_path_eligibility(sg, N, lat, expect):
    """Host check used to pick path QS: cubic grid, L = (2,2,2), orthogonal
    cell, diagonal representative ops with 8 distinct sign patterns."""
    from cfft_hip.plans import qs_eligible
    ops = get_ops(sg, 3, N)
    shift, ops_s = find_optimal_shift(ops, N)
    assert qs_eligible(N, ops_s, lat) is expect


def test_groups_of_partitions_by_grid_group_cell_ns():
    """Grouping for shared contexts: every instance once, each group one
    (grid, group, cell, Ns), at most max_group long, largest work first."""
    from cfft_hip import batch
    insts = batch.draw_instances(64, 11)
    gs = batch.groups_of(insts)
    assert sorted(k for g in gs for k in g) == list(range(64))
    for g in gs:
        keys = {(tuple(insts[k].N), insts[k].sg, insts[k].cell, insts[k].ns) for k in g}
        assert len(keys) == 1
    work = [len(g) * int(np.prod(insts[g[0]].N)) for g in gs]
    assert work == sorted(work, reverse=True)
    keys = {(tuple(i.N), i.sg) for i in insts}
    assert len(gs) == len(keys)
    small = batch.groups_of(insts, max_group=2)
    assert max(len(g) for g in small) == 2
    assert sorted(k for g in small for k in g) == list(range(64))


def test_default_path_by_precision_and_tolerance():
    """FP64 takes the plane-pair chain where it applies; mixed precision takes
    it only for a stage that stops at residual >= QS_MIXED_TOL (above that
    chain's mixed-precision floor), otherwise the default path."""
    import cfft_hip as cf
    from cfft_hip.scft import QS_MIXED_TOL, default_path
    N = (16, 16, 16)
    _, ops_s = find_optimal_shift(get_ops(229, 3, N), N)
    lat = 4.0 * np.eye(3)
    assert default_path(N, ops_s, lat, cf.FP64) == cf.PATH_QS
    assert default_path(N, ops_s, lat, cf.MIXED) == cf.PATH_AUTO
    assert default_path(N, ops_s, lat, cf.MIXED, QS_MIXED_TOL) == cf.PATH_QS
    assert default_path(N, ops_s, lat, cf.MIXED, QS_MIXED_TOL / 10) == cf.PATH_AUTO
    # a non-orthogonal cell is not eligible in either precision
    skew = lat + 0.3 * np.eye(3, k=1)
    assert default_path(N, ops_s, skew, cf.FP64) == cf.PATH_AUTO
    assert default_path(N, ops_s, skew, cf.MIXED, 1e-3) == cf.PATH_AUTO


def test_schedule_candidates():
    """Tuner candidates: tile heights 8, 4, 2 dividing H1 (1 when none does),
    each with and without graph replay."""
    import cfft_hip._lib as L
    from cfft_hip.tune import schedule_candidates
    rows = lambda H1: sorted({c[L.OPT_TILE_ROWS] for c in schedule_candidates(H1)})
    assert rows(32) == [2, 4, 8] and rows(12) == [2, 4] and rows(6) == [2] and rows(3) == [1]
    assert all({c[L.OPT_GRAPH] for c in schedule_candidates(32) if c[L.OPT_TILE_ROWS] == r} == {0, 1}
               for r in (2, 4, 8))


@pytest.mark.parametrize("sg", [47, 70, 223, 229, 230, 225])
def test_star_wave_equals_full_grid_group_average(sg):
    """The star-sum form of the symmetrised plane wave against the brute-force
    group average of the full-grid wave (gather form of
    test/test_helpers.jl:11-29), every candidate star, cos and sin."""
    N = (16, 16, 16)
    ops = get_ops(sg, 3, N)
    shift, ops_s = find_optimal_shift(ops, N)
    Lv = auto_L(ops_s)
    x = [(np.arange(n) + s * n) / n for n, s in zip(N, shift)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    for h in batch._STARS:
        for fn in (np.cos, np.sin):
            ref = symmetrize_subgrid(fn(2 * np.pi * (h[0] * X + h[1] * Y + h[2] * Z)), ops_s, Lv)
            assert np.max(np.abs(batch.star_wave(h, fn, N, ops_s, shift, Lv) - ref)) < 1e-13


@pytest.mark.parametrize("sg,n", [(47, 16), (70, 16), (223, 24), (229, 16), (230, 24)])
def test_subgrid_projection(sg, n):
    """Projection onto the symmetric subgrid fields: the subgrid of a
    symmetric full-grid field is a fixed point, a projected random field is
    invariant under every op that keeps the subgrid (brute force over the
    full-grid ops) and the projection is idempotent and keeps the sum."""
    N = (n, n, n)
    ops = get_ops(sg, 3, N)
    shift, ops_s = find_optimal_shift(ops, N)
    Lv = auto_L(ops_s)
    maps = batch.subgrid_maps(N, ops_s, Lv)
    rng = np.random.default_rng(sg)
    sym = symmetrize_subgrid(rng.standard_normal(N), ops_s, Lv)
    assert np.max(np.abs(batch.project_subgrid(sym, maps) - sym)) < 1e-14
    M = sym.shape
    f = rng.standard_normal(M)
    p = batch.project_subgrid(f, maps)
    assert np.max(np.abs(batch.project_subgrid(p, maps) - p)) < 1e-14
    assert abs(p.sum() - f.sum()) < 1e-10
    # invariance, from the full-grid ops directly
    g = np.meshgrid(*[np.arange(m) * l for m, l in zip(M, Lv)], indexing="ij")
    Xs = np.stack([a.ravel(order="F") for a in g], axis=1)
    pv = p.ravel(order="F")
    n_keep = 0
    for op in ops_s:
        Yf = np.mod(Xs @ op.Rm.T + op.tv, np.asarray(N))
        if np.all(Yf % np.asarray(Lv) == 0):
            n_keep += 1
            y = Yf // np.asarray(Lv)
            img = y[:, 0] + M[0] * (y[:, 1] + M[1] * y[:, 2])
            assert np.max(np.abs(pv[img] - pv)) < 1e-14
    assert n_keep >= len(maps) >= 1
