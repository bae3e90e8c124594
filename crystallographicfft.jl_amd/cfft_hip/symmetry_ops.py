"""Symmetry operations on integer grids (src/symmetry_ops.jl).

Host-side mirror of the reference's ``SymmetryOps`` module: ``SymOp`` with an
integer rotation R and an integer translation t in grid units, ``get_ops``
(src/symmetry_ops.jl:51-53), ``convert_op`` (:38-46), ``apply_op`` (:61-79),
``check_shift_invariance`` (:81-94), ``dual_ops`` (:103-112),
``detect_centering_type`` (:133-179) and ``strip_centering`` (:192-222).

Ops travel to the C ABI as a flat int32 array of ``nops x 12`` entries:
R row-major (R[0,0], R[0,1], ..., R[2,2]) followed by t (3 entries).
"""
from __future__ import annotations

from dataclasses import dataclass
from enum import IntEnum

import numpy as np

from .spacegroups import space_group_ops12


class CenteringType(IntEnum):
    """src/symmetry_ops.jl:123 (@enum CenteringType)."""
    CentP = 0
    CentC = 1
    CentA = 2
    CentI = 3
    CentF = 4


CentP, CentC, CentA, CentI, CentF = (CenteringType.CentP, CenteringType.CentC,
                                     CenteringType.CentA, CenteringType.CentI,
                                     CenteringType.CentF)


@dataclass(frozen=True)
class SymOp:
    """src/symmetry_ops.jl:17-20: x -> R x + t (mod N), grid units."""
    R: tuple
    t: tuple

    @property
    def Rm(self) -> np.ndarray:
        D = len(self.t)
        return np.array(self.R, dtype=np.int64).reshape(D, D)

    @property
    def tv(self) -> np.ndarray:
        return np.array(self.t, dtype=np.int64)

    @staticmethod
    def make(R, t) -> "SymOp":
        R = np.asarray(np.rint(R), dtype=np.int64)
        t = np.asarray(np.rint(t), dtype=np.int64)
        return SymOp(tuple(int(v) for v in R.ravel()), tuple(int(v) for v in t))


def convert_op(R, t12, N) -> SymOp:
    """src/symmetry_ops.jl:38-46: translation x N must be integral (atol 1e-5)."""
    t_grid = np.asarray(t12, dtype=np.float64) / 12.0 * np.asarray(N, dtype=np.float64)
    t_int = np.rint(t_grid)
    if not np.all(np.abs(t_grid - t_int) <= 1e-5):
        raise ValueError(f"Symmetry operation not commensurate with grid {tuple(N)}")
    return SymOp.make(R, t_int)


def get_ops(sg_num: int, dim: int, N) -> list[SymOp]:
    """src/symmetry_ops.jl:51-53."""
    N = tuple(int(n) for n in N)
    if len(N) != dim:
        raise ValueError("len(N) must equal dim")
    return [convert_op(R, t, N) for R, t in space_group_ops12(sg_num, dim)]


def ops_arrays(ops):
    """(R (n,D,D) int64, t (n,D) int64) views of a list of SymOp."""
    R = np.stack([op.Rm for op in ops])
    t = np.stack([op.tv for op in ops])
    return R, t


def ops_from_arrays(R, t) -> list[SymOp]:
    return [SymOp.make(R[i], t[i]) for i in range(R.shape[0])]


def pack_ops(ops) -> np.ndarray:
    """Flat int32 array (nops x 12): R row-major, then t (3-D only)."""
    R, t = ops_arrays(ops)
    if R.shape[1] != 3:
        raise ValueError("the C ABI takes 3-D operations")
    return np.ascontiguousarray(np.concatenate([R.reshape(-1, 9), t], axis=1).astype(np.int32))


def apply_op(op: SymOp, x, N):
    """src/symmetry_ops.jl:61-79."""
    return np.mod(op.Rm @ np.asarray(x, dtype=np.int64) + op.tv, np.asarray(N))


def check_shift_invariance(ops, shift, N):
    """src/symmetry_ops.jl:81-94."""
    shift = np.asarray(shift, dtype=np.float64)
    deltas = []
    for op in ops:
        df = (op.Rm.astype(np.float64) @ shift - shift) * np.asarray(N, dtype=np.float64)
        di = np.rint(df)
        if not np.all(np.abs(df - di) <= 1e-5):
            return False, []
        deltas.append(di.astype(np.int64))
    return True, deltas


def dual_ops(ops):
    """src/symmetry_ops.jl:103-112: R* = (R^-1)^T, t = 0."""
    out = []
    for op in ops:
        Rinv = np.rint(np.linalg.inv(op.Rm.astype(np.float64))).astype(np.int64)
        out.append(SymOp.make(Rinv.T, np.zeros(len(op.t))))
    return out


def detect_centering_type(ops, N) -> CenteringType:
    """src/symmetry_ops.jl:133-179."""
    D = len(N)
    pure = []
    for op in ops:
        if not np.array_equal(op.Rm, np.eye(D, dtype=np.int64)):
            continue
        if all(abs(v) < 0.1 for v in op.t):
            continue
        pure.append(tuple(abs(abs(op.t[d]) / N[d] - 0.5) < 0.01 for d in range(D)))
    if not pure or D < 3:
        return CentP
    has_xyz = any(all(p) for p in pure)
    has_0yz = any((not p[0]) and p[1] and p[2] for p in pure)
    has_x0z = any(p[0] and (not p[1]) and p[2] for p in pure)
    has_xy0 = any(p[0] and p[1] and (not p[2]) for p in pure)
    if has_0yz + has_x0z + has_xy0 >= 2:
        return CentF
    if has_xyz:
        return CentI
    if has_xy0:
        return CentC
    if has_0yz or has_x0z:
        return CentA
    return CentP


def strip_centering(ops, centering: CenteringType, N):
    """src/symmetry_ops.jl:192-222."""
    D = len(N)
    if centering in (CentI, CentF):
        halve = [True] * D
    elif centering == CentC:
        halve = [d < 2 for d in range(D)]
    elif centering == CentA:
        halve = [d >= 1 for d in range(D)]
    else:
        halve = [False] * D
    N_sub = tuple(N[d] // 2 if halve[d] else N[d] for d in range(D))
    seen = set()
    out = []
    for op in ops:
        t_sub = tuple(int(op.t[d]) % N_sub[d] for d in range(D))
        key = (op.R, t_sub)
        if key not in seen:
            seen.add(key)
            out.append(SymOp(op.R, t_sub))
    return out, N_sub
