"""Real-space ASU, block packing and the block CFFT over the C ABI.

Reference counterparts:
  calc_asu(N, ops), calc_asu(sg_num, dim, N)   src/asu.jl:51-55, 216-325
  ASUBlock / CrystallographicASU                src/asu.jl:29-44
  pack_asu                                      src/pack_asu.jl:6-102
  InterleavedOrbit / analyze_interleaved_orbits src/asu.jl:16-21, 331-431
  pack_asu_interleaved (Mode B)                 src/pack_asu.jl:113-202
  CFFTPlan / plan_cfft                          src/cfft_plan.jl:4-40
  mul! / ldiv!  (here mul / ldiv)               src/cfft_plan.jl:43-66,
                                                src/execution.jl:10-64

Block data is moved to the GPU as one packed buffer (all blocks one after
another, column-major inside a block) and transformed there by one batched DFT
launch per axis.  Depth groups are taken in sorted order (the reference walks a
Julia Dict, whose order is not specified); within a group, blocks come in the
order the greedy packer creates them, as in the reference.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from . import _lib as L_
from .plans import ASUPoint, _N3, _pack_any, find_optimal_shift
from .symmetry_ops import get_ops

__all__ = ["calc_asu", "ASUBlock", "CrystallographicASU", "pack_asu", "CFFTPlan", "plan_cfft",
           "mul", "ldiv", "InterleavedOrbit", "analyze_interleaved_orbits", "pack_asu_interleaved"]


def _calc_asu_ops(N, ops):
    N3, dim = _N3(N)
    packed = _pack_any(ops, dim)
    lib = L.lib()
    n = C.c_int64()
    Narr = np.asarray(N3, dtype=np.int32)
    L.check(lib.cfft_calc_asu(dim, L.i32p(Narr), L.i32p(packed), len(ops), None, None, None, 0,
                              C.byref(n)))
    cnt = n.value
    idx = np.zeros((max(cnt, 1), 3), dtype=np.int32)
    dep = np.zeros((max(cnt, 1), 3), dtype=np.int32)
    mult = np.zeros(max(cnt, 1), dtype=np.int32)
    L.check(lib.cfft_calc_asu(dim, L.i32p(Narr), L.i32p(packed), len(ops), L.i32p(idx), L.i32p(dep),
                              L.i32p(mult), cnt, C.byref(n)))
    return [ASUPoint(idx[i, :dim].astype(np.int64), dep[i, :dim].astype(np.int64), int(mult[i]))
            for i in range(cnt)]


def calc_asu(*args):
    """``calc_asu(N, ops)`` -> list of ASUPoint (ops already shifted), or
    ``calc_asu(sg_num, dim, N)`` -> (points, shift) with the magic shift,
    src/asu.jl:51-55 and 216-325."""
    if len(args) == 2:
        return _calc_asu_ops(args[0], args[1])
    if len(args) == 3:
        sg, dim, N = args
        ops = get_ops(int(sg), int(dim), tuple(N))
        shift, shifted = find_optimal_shift(ops, tuple(N))
        return _calc_asu_ops(N, shifted), tuple(float(s) for s in shift)
    raise TypeError("calc_asu(N, ops) or calc_asu(sg_num, dim, N)")


@dataclass
class ASUBlock:
    """src/asu.jl:29-38: data array, per-dimension index ranges, depth."""
    data: np.ndarray
    range: list
    depth: np.ndarray
    orbit: object = None


@dataclass
class CrystallographicASU:
    """src/asu.jl:40-44: blocks grouped by effective dimension (number of
    block axes longer than 1) and the magic shift."""
    dim_blocks: dict
    shift: tuple
    _order: list = field(default_factory=list, repr=False)  # (eff_dim, i) in packed order

    def blocks(self):
        """Blocks in packed (device) order."""
        return [self.dim_blocks[d][i] for d, i in self._order]

    def similar(self, dtype=None):
        """A CrystallographicASU of the same layout with zeroed data."""
        out = {d: [ASUBlock(np.zeros_like(b.data, dtype=dtype or b.data.dtype), list(b.range),
                            b.depth.copy(), b.orbit) for b in bl]
               for d, bl in self.dim_blocks.items()}
        return CrystallographicASU(out, self.shift, list(self._order))


def _pack_blocks(points, dim):
    lib = L.lib()
    n = len(points)
    idx = np.zeros((max(n, 1), 3), dtype=np.int32)
    dep = np.zeros((max(n, 1), 3), dtype=np.int32)
    for i, p in enumerate(points):
        idx[i, :dim] = np.asarray(p.idx)[:dim]
        dep[i, :dim] = np.asarray(p.depth)[:dim]
    cnt = C.c_int64()
    L.check(lib.cfft_pack_asu(dim, L.i32p(idx), L.i32p(dep), n, None, 0, C.byref(cnt)))
    out = np.zeros((max(cnt.value, 1), 9), dtype=np.int32)
    L.check(lib.cfft_pack_asu(dim, L.i32p(idx), L.i32p(dep), n, L.i32p(out), cnt.value,
                              C.byref(cnt)))
    return out[:cnt.value]


def pack_asu(points, N, dtype=np.float64, shift=None):
    """src/pack_asu.jl:6-102: greedy unit-stride boxes per depth group."""
    dim = len(tuple(N))
    tab = _pack_blocks(points, dim)
    dim_blocks: dict = {}
    order = []
    for row in tab:
        start, shape, depth = row[0:dim], row[3:3 + dim], row[6:6 + dim]
        rng = [range(int(start[d]), int(start[d]) + int(shape[d])) for d in range(dim)]
        eff = int(np.count_nonzero(shape > 1))
        blk = ASUBlock(np.zeros(tuple(int(s) for s in shape), dtype=dtype), rng,
                       depth.astype(np.int64))
        dim_blocks.setdefault(eff, []).append(blk)
        order.append((eff, len(dim_blocks[eff]) - 1))
    sh = tuple(shift) if shift is not None else (0.0,) * dim
    return CrystallographicASU(dim_blocks, sh, order)


@dataclass
class InterleavedOrbit:
    """src/asu.jl:16-21: the coset shifts x_s in [0, L) of one orbit under the
    ops taken mod L; ops[i] lists the ops g with g(representative) = members[i]."""
    representative: np.ndarray
    members: list
    ops: list
    multiplicity: int


def analyze_interleaved_orbits(N, ops, L=None):
    """src/asu.jl:331-431: orbits of the stride-L subgrids (coset shifts
    p in [0, L)^D, first axis fastest) under x -> (R x + t) mod L, t in grid
    units; each orbit is found by closing the representative under the ops,
    then every member lists the ops that send the representative there.
    Members are sorted lexicographically (first component most significant)."""
    D = len(tuple(N))
    Lv = np.array([2] * D if L is None else [int(v) for v in L], dtype=np.int64)
    Rs = [op.Rm for op in ops]
    ts = [op.tv for op in ops]
    # coset shifts in the reference's product order (first axis fastest)
    coarse = [np.array(p[::-1], dtype=np.int64) for p in np.ndindex(*Lv[::-1])]
    visited = set()
    orbits = []
    for p in coarse:
        if tuple(p) in visited:
            continue
        members = {tuple(p)}
        queue = [p]
        while queue:
            cur = queue.pop(0)
            for R, t in zip(Rs, ts):
                x = tuple(int(v) for v in np.mod(R @ cur + t, Lv))
                if x not in members:
                    members.add(x)
                    queue.append(np.array(x, dtype=np.int64))
        final = sorted(members)
        mops = [[op for op, R, t in zip(ops, Rs, ts) if tuple(int(v) for v in np.mod(R @ p + t, Lv)) == m]
                for m in final]
        orbits.append(InterleavedOrbit(p.copy(), [np.array(m, dtype=np.int64) for m in final], mops,
                                       len(final)))
        visited |= members
    return orbits


def pack_asu_interleaved(u, N, ops, L=None, asu_only=True):
    """src/pack_asu.jl:113-202 (Mode B): the stride-L subgrids of a full grid
    as blocks.  asu_only (default): the one subgrid of the orbit whose
    representative is the zero shift (Gamma_0; the first orbit if none is),
    without orbit information; otherwise one block per orbit representative,
    each carrying its orbit.  Block data are copies of u[r_0::L_0, ...]
    (host arrays), ranges the 0-based strided index ranges, depths zero;
    all blocks sit under dim_blocks[D] and the shift is zero."""
    u = np.asarray(u)
    N = tuple(int(n) for n in N)
    D = len(N)
    if u.shape != N:
        raise L_.CFFTError(L_.CFFT_ERR_INVALID, "u must have the grid's shape")
    orbits = analyze_interleaved_orbits(N, ops, L)
    Lv = [2] * D if L is None else [int(v) for v in L]
    if asu_only:
        zero = [o for o in orbits if not np.any(o.representative)]
        orbits = zero[:1] if zero else orbits[:1]
    blocks = []
    for o in orbits:
        rep = [int(v) for v in o.representative]
        sl = tuple(slice(rep[d], None, Lv[d]) for d in range(D))
        rng = [range(rep[d], N[d], Lv[d]) for d in range(D)]
        blocks.append(ASUBlock(np.array(u[sl], copy=True), rng, np.zeros(D, dtype=np.int64),
                               None if asu_only else o))
    return CrystallographicASU({D: blocks}, (0.0,) * D, [(D, i) for i in range(len(blocks))])


class CFFTPlan:
    """src/cfft_plan.jl:4-12: the packed ASU and one device plan running the
    DFT of every block."""

    def __init__(self, asu: CrystallographicASU, dim: int):
        self.asu = asu
        self.dim = dim
        blocks = asu.blocks()
        tab = np.zeros((max(len(blocks), 1), 9), dtype=np.int32)
        for i, b in enumerate(blocks):
            for d in range(dim):
                tab[i, d] = b.range[d].start
                tab[i, 3 + d] = len(b.range[d])
                tab[i, 6 + d] = b.depth[d]
            for d in range(dim, 3):
                tab[i, 3 + d] = 1
        self._sizes = [int(np.prod(b.data.shape)) for b in blocks]
        self.total = int(sum(self._sizes))
        lib = L.lib()
        h = C.c_void_p()
        L.check(lib.cfft_block_plan_create(dim, L.i32p(tab), len(blocks), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            L.lib().cfft_block_plan_destroy(h)
            self._h = None

    def _gather(self, x: CrystallographicASU):
        buf = np.empty(self.total, dtype=np.complex128)
        off = 0
        for b, n in zip(x.blocks(), self._sizes):
            if b.data.size != n:
                raise L.CFFTError(L.CFFT_ERR_INVALID, "block shape differs from the plan")
            buf[off:off + n] = np.ravel(b.data, order="F")
            off += n
        return buf

    def _scatter(self, buf, y: CrystallographicASU):
        off = 0
        for b, n in zip(y.blocks(), self._sizes):
            if b.data.size != n:
                raise L.CFFTError(L.CFFT_ERR_INVALID, "block shape differs from the plan")
            if not np.iscomplexobj(b.data):
                raise L.CFFTError(L.CFFT_ERR_INVALID, "output blocks must be complex")
            b.data[...] = buf[off:off + n].reshape(b.data.shape, order="F")
            off += n

    def _run(self, y, x, fwd: bool):
        if len(x.blocks()) != len(self._sizes) or len(y.blocks()) != len(self._sizes):
            raise L.CFFTError(L.CFFT_ERR_INVALID, "mismatched block counts")
        src = self._gather(x)
        dst = np.empty_like(src)
        fn = L.lib().cfft_block_forward if fwd else L.lib().cfft_block_inverse
        L.check(fn(self._h, L.f64p(src.view(np.float64)), L.f64p(dst.view(np.float64))))
        self._scatter(dst, y)
        return y


def plan_cfft(N, sg_num: int, dtype=np.complex128, dim=None) -> CFFTPlan:
    """src/cfft_plan.jl:15-40: calc_asu (magic shift) -> pack_asu -> plan."""
    N = tuple(int(n) for n in N)
    dim = len(N) if dim is None else int(dim)
    points, shift = calc_asu(sg_num, dim, N)
    return CFFTPlan(pack_asu(points, N, dtype, shift), dim)


def mul(y: CrystallographicASU, p: CFFTPlan, x: CrystallographicASU) -> CrystallographicASU:
    """mul!(y, p, x): forward DFT of every block (src/execution.jl:10-31)."""
    return p._run(y, x, True)


def ldiv(y: CrystallographicASU, p: CFFTPlan, x: CrystallographicASU) -> CrystallographicASU:
    """ldiv!(y, p, x): normalised inverse DFT of every block
    (src/execution.jl:37-64)."""
    return p._run(y, x, False)
