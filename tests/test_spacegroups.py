"""Operation tables (the role of Crystalline.jl) pinned against the reference's
own fixture rows (src/mirror_table.jl) and ITA known answers."""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import cfft_hip as cf
from cfft_hip.spacegroups import GROUP_ORDER_3D, HALL_3D, space_group_ops12

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "mirror_table_rows.json")))

_NORMAL_TO_R = {(0, 0, 1): np.diag([1, 1, -1]), (0, -1, 0): np.diag([1, -1, 1]),
                (1, 0, 0): np.diag([-1, 1, 1])}


@pytest.mark.parametrize("sg", sorted(HALL_3D))
def test_group_order_and_closure(sg):
    ops = space_group_ops12(sg)
    assert len(ops) == GROUP_ORDER_3D[sg]
    keys = {(tuple(R.ravel()), tuple(t % 12)) for R, t in ops}
    assert len(keys) == len(ops)
    # closed under composition modulo integer translations
    for R1, t1 in ops[:12]:
        for R2, t2 in ops:
            k = (tuple((R1 @ R2).ravel()), tuple((R1 @ t2 + t1) % 12))
            assert k in keys


@pytest.mark.parametrize("sg", sorted(int(k) for k in FIX["rows"]))
def test_mirror_table_rows(sg):
    """Centring vectors and the first mirror/glide translation of each
    axis-normal mirror listed by src/mirror_table.jl are in our op set."""
    row = FIX["rows"][str(sg)]
    ops = space_group_ops12(sg)
    cents = [np.zeros(3, dtype=int)] + [np.array([int(Fraction(x) * 12) for x in v]) for v in
                                        row["centering_vecs"]]
    # centring vectors: exactly the pure translations
    pure = sorted(tuple(t) for R, t in ops if np.array_equal(R, np.eye(3)) and np.any(t))
    assert pure == sorted(tuple(c % 12) for c in cents[1:])
    for normal, tr in zip(row["normals"], row["translations"]):
        Rm = _NORMAL_TO_R[tuple(normal)]
        t12 = np.array([int(Fraction(x) * 12) for x in tr])
        cand = [t for R, t in ops if np.array_equal(R, Rm)]
        assert any(np.array_equal((t - t12) % 12, np.zeros(3)) for t in cand), (sg, normal, tr)
    # number of axis-normal mirror rotations present
    n_m = sum(any(np.array_equal(R, Rm) for R, _ in ops) for Rm in _NORMAL_TO_R.values())
    assert n_m == row["n_mirrors"]


def test_known_ita_positions():
    """A few general positions written out in the International Tables."""
    def has(sg, R, t):
        t12 = (np.array([int(Fraction(x) * 12) for x in t])) % 12
        return any(np.array_equal(Ro, np.array(R)) and np.array_equal(to, t12)
                   for Ro, to in space_group_ops12(sg))
    # Ia-3d (230): -x+1/2, -y, z+1/2 and -y+1/4, x+3/4, z+1/4
    assert has(230, [[-1, 0, 0], [0, -1, 0], [0, 0, 1]], ["1/2", "0", "1/2"])
    assert has(230, [[0, -1, 0], [1, 0, 0], [0, 0, 1]], ["1/4", "3/4", "1/4"])
    # Fddd origin 2 (70): -x+3/4, -y+3/4, z
    assert has(70, [[-1, 0, 0], [0, -1, 0], [0, 0, 1]], ["3/4", "3/4", "0"])
    # Pm-3n (223): -y+1/2, x+1/2, z+1/2
    assert has(223, [[0, -1, 0], [1, 0, 0], [0, 0, 1]], ["1/2", "1/2", "1/2"])
    # Im-3m (229): z, x, y
    assert has(229, [[0, 0, 1], [1, 0, 0], [0, 1, 0]], ["0", "0", "0"])


def test_detect_centering_known_answers():
    """test/test_centering_prefold.jl:45-55."""
    N = (16, 16, 16)
    exp = {1: cf.CentP, 221: cf.CentP, 229: cf.CentI, 225: cf.CentF, 70: cf.CentF,
           35: cf.CentC, 139: cf.CentI, 69: cf.CentF}
    for sg, c in exp.items():
        assert cf.detect_centering_type(cf.get_ops(sg, 3, N), N) == c
        assert cf.detect_centering(cf.get_ops(sg, 3, N), N) == c


def test_strip_centering_counts():
    """test/test_centering_prefold.jl:57-77."""
    N = (16, 16, 16)
    sub, Ns = cf.strip_centering(cf.get_ops(229, 3, N), cf.CentI, N)
    assert Ns == (8, 8, 8) and len(sub) == 48
    sub, Ns = cf.strip_centering(cf.get_ops(225, 3, N), cf.CentF, N)
    assert Ns == (8, 8, 8) and len(sub) == 48
    sub, Ns = cf.strip_centering(cf.get_ops(35, 3, N), cf.CentC, N)
    assert Ns == (8, 8, 16)


def test_not_commensurate():
    """src/symmetry_ops.jl:41-43: Fd-3m quarter translations need N % 4 == 0."""
    with pytest.raises(ValueError, match="not commensurate"):
        cf.get_ops(227, 3, (6, 6, 6))


def test_dual_ops():
    """test/test_spectral_indexing.jl:11-40."""
    ops = [cf.SymOp.make([[0, -1], [1, 0]], [0, 0])]
    assert np.array_equal(cf.dual_ops(ops)[0].Rm, np.array([[0, -1], [1, 0]]))
    R6 = np.array([[1, -1], [1, 0]])
    d = cf.dual_ops([cf.SymOp.make(R6, [0, 0])])[0]
    assert np.array_equal(d.Rm, np.array([[0, 1], [-1, 1]]).T)
