"""Tables of the sparse-stage KRFFT families (P3c and fractal) and their
device engine (cfft_p3c_*).

Reference: src/recursive_blocks.jl — plan_krfft_recursive (:64-144),
plan_krfft_sparse (:514-629), plan_krfft_selective (:756-888),
plan_krfft_g0asu (:1034-1244) and their execute functions;
src/fractal_krfft.jl — build_recursive_tree (:300-600), plan_fractal_krfft
(:678-722), execute_fractal_krfft! (:869-901).

For a cubic grid N (M = N/2, M2 = N/4) and shifted ops holding the
body-diagonal 3-fold C: (x, y, z) -> (z, x, y) with an even translation
t = 2 tau, write the stride-2 subgrid spectrum as
    G0(q) = sum_e exp(-2 pi i q.e / M) G_e(q mod M2),
with G_e the M2^3 FFT of the sub-sub-grid g_e(k) = u(4 k + 2 e), e in {0,1}^3.
The 3-fold maps g_e(k) = g_{e'}(C k + d) with e' = (C e + tau) mod 2 and
d = (C e + tau - e') / 2, so G_e(q) = exp(2 pi i (C q).d / M2) G_{e'}(C q):
only one offset per orbit of e -> e' needs an FFT (4 when tau = 0: 000,
100, 110, 111), the others follow through the chain e -> e' -> ... to the
orbit's representative.  The spectral-ASU value is
    F(h) = sum_a exp(-2 pi i h.t_a / N) G0(R_a^T h mod M)
over the 8 representative ops (the A8 sum).  The reference's P3c formula
hard-codes a translation-free 3-fold; the chain above also covers the
3-folds the magic shift gives an even translation.

The families differ in what they tabulate: sparse, one CSR matrix from the
FFT outputs straight to the spectral ASU (about 40 entries per row after
merging); recursive, G0 at every M-grid point (8 entries per point) then
the A8 sum; selective, G0 only at the positions the A8 sum reads; G0-ASU,
G0 only at orbit representatives of those positions under the ops with even
translations, G0(p) = exp(-2 pi i p.tau/M) G0(R^T p) with tau = t/2, the
phases merged into the A8 weights.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .plans import SpectralIndexing, auto_L, get_k_vector, select_rep_ops

C3 = np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]], dtype=np.int64)


def p3c_classes(ops_shifted, n):
    """(offsets e (8 x 3), orbit representative of each e, chain steps d of
    each e to its representative, sorted representatives) for the first op
    with R = C and an even translation; CFFTError(UNSUPPORTED) if none."""
    g3 = [op for op in ops_shifted if np.array_equal(op.Rm, C3) and np.all(np.mod(op.tv, n) % 2 == 0)]
    if not g3:
        raise L.CFFTError(L.CFFT_ERR_UNSUPPORTED,
                          "P3c needs the body-diagonal 3-fold rotation with an even translation")
    tau = np.mod(g3[0].tv, n) // 2
    eps = np.array([[(e >> 0) & 1, (e >> 1) & 1, (e >> 2) & 1] for e in range(8)], dtype=np.int64)
    v = eps @ C3.T + tau
    nxt = v % 2
    phi = nxt[:, 0] + 2 * nxt[:, 1] + 4 * nxt[:, 2]
    d = (v - nxt) // 2
    orbit_rep, chain = {}, {}
    for e in range(8):
        seen, x = [e], e
        while phi[x] not in seen:
            x = int(phi[x])
            seen.append(x)
        orbit_rep[e] = min(seen)
        steps, x = [], e
        while x != orbit_rep[e]:
            steps.append(d[x])
            x = int(phi[x])
        chain[e] = steps
    reps = sorted(set(orbit_rep.values()))
    return eps, orbit_rep, chain, reps


def g0_rows(Q, n, cls):
    """COO rows of G0 at the M-grid positions Q (k x 3) from the
    representatives' FFT outputs: 8 (row, column, value) per position."""
    eps, orbit_rep, chain, reps = cls
    M, M2 = n // 2, n // 4
    vol = M2 ** 3
    slot = {r: i for i, r in enumerate(reps)}
    Q = np.asarray(Q, dtype=np.int64)
    rows, cols, vals = [], [], []
    for e in range(8):
        qq = Q % M2
        ph = np.zeros(len(Q))
        for dv in chain[e]:
            qq = (qq @ C3.T) % M2
            ph += (qq @ dv) / M2
        cols.append(slot[orbit_rep[e]] * vol + qq[:, 0] + M2 * (qq[:, 1] + M2 * qq[:, 2]))
        vals.append(np.exp(-2j * np.pi * (Q @ eps[e]) / M) * np.exp(2j * np.pi * ph))
        rows.append(np.arange(len(Q)))
    return np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)


def a8_terms(spec: SpectralIndexing, ops_shifted, n):
    """Per spectral point h and representative op a: the M-grid position
    R_a^T h mod M (n_spec x 8 x 3) and the weight exp(-2 pi i h.t_a / N)."""
    rep = select_rep_ops(ops_shifted, auto_L(ops_shifted), tuple(spec.N))
    H = np.array([get_k_vector(spec, i) for i in range(len(spec.points))], dtype=np.int64).reshape(-1, 3)
    M = n // 2
    Q = np.stack([(H @ ops_shifted[a].Rm) % M for a in rep], 1)
    W = np.stack([np.exp(-2j * np.pi * (H @ ops_shifted[a].tv) / n) for a in rep], 1)
    return Q, W


def csr(rows, cols, vals, n_rows):
    """(rowptr int64, cols int32, vals complex128), duplicate (row, column)
    entries summed, columns sorted within a row."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    key = rows * (int(cols.max()) + 1 if cols.size else 1) + cols
    uk, inv = np.unique(key, return_inverse=True)
    v = np.zeros(len(uk), dtype=np.complex128)
    np.add.at(v, inv, vals)
    ncol = int(cols.max()) + 1 if cols.size else 1
    r, c = uk // ncol, uk % ncol
    rowptr = np.zeros(n_rows + 1, dtype=np.int64)
    np.add.at(rowptr, r + 1, 1)
    return np.cumsum(rowptr), c.astype(np.int32), v


def _lin(q, M):
    return q[..., 0] + M * (q[..., 1] + M * q[..., 2])


def g0_table(ops_shifted, n):
    """(classes, [stage]): G0 at every M-grid point (x fastest), the F000 =
    FFT of u(2 x) of the staged Pm-3m plan (src/staged_krfft.jl:562-567),
    from the representatives' sub-sub-grid FFTs."""
    cls = p3c_classes(ops_shifted, n)
    M = n // 2
    g = np.arange(M ** 3)
    allq = np.stack([g % M, (g // M) % M, g // (M * M)], 1)
    return cls, [csr(*g0_rows(allq, n, cls), M ** 3)]


def family_tables(kind, spec: SpectralIndexing, ops_shifted):
    """(classes, stages): the CSR stages of one family ("sparse",
    "recursive", "selective", "g0asu") and a summary of its table sizes."""
    n = int(spec.N[0])
    cls = p3c_classes(ops_shifted, n)
    M = n // 2
    Q, W = a8_terms(spec, ops_shifted, n)
    ns = Q.shape[0]
    if kind == "sparse":
        r, c, v = g0_rows(Q.reshape(-1, 3), n, cls)
        v = v * W.reshape(-1)[r]
        return cls, [csr(r // 8, c, v, ns)], {}
    a_rows = np.repeat(np.arange(ns), 8)
    if kind == "recursive":
        g = np.arange(M ** 3)
        allq = np.stack([g % M, (g // M) % M, g // (M * M)], 1)
        B = csr(*g0_rows(allq, n, cls), M ** 3)
        A = csr(a_rows, _lin(Q, M).reshape(-1), W.reshape(-1), ns)
        return cls, [B, A], {"g0_positions": M ** 3}
    S, inv = np.unique(_lin(Q, M).reshape(-1), return_inverse=True)
    Sq = np.stack([S % M, (S // M) % M, S // (M * M)], 1)
    if kind == "selective":
        B = csr(*g0_rows(Sq, n, cls), len(S))
        A = csr(a_rows, inv, W.reshape(-1), ns)
        return cls, [B, A], {"g0_positions": len(S)}
    if kind != "g0asu":
        raise ValueError(kind)
    # orbits of S under the ops with even translation: G0(p) = e^{-2 pi i p.tau/M} G0(R^T p)
    even = [op for op in ops_shifted if np.all(np.mod(op.tv, n) % 2 == 0)]
    maps = [(np.rint(np.linalg.inv(op.Rm).T).astype(np.int64), np.mod(op.tv, n) // 2) for op in even]
    pos = {int(s): i for i, s in enumerate(S)}
    rep_of = np.full(len(S), -1, dtype=np.int64)
    phase = np.zeros(len(S), dtype=np.complex128)
    reps = []
    for i in range(len(S)):
        if rep_of[i] >= 0:
            continue
        p0 = Sq[i]
        k = len(reps)
        reps.append(i)
        for Rit, tau in maps:
            p = (Rit @ p0) % M
            j = pos.get(int(_lin(p, M)))
            if j is not None and rep_of[j] < 0:
                rep_of[j] = k
                phase[j] = np.exp(-2j * np.pi * (p @ tau) / M)
    B = csr(*g0_rows(Sq[reps], n, cls), len(reps))
    A = csr(a_rows, rep_of[inv], W.reshape(-1) * phase[inv], ns)
    return cls, [B, A], {"g0_positions": len(S), "g0_reps": len(reps)}


def _offset_orbits(ops_shifted, n, S):
    """Orbits of the stride-S offsets e in [0, S)^3 (x fastest) under the
    group's action e -> (R e + t) mod S: per offset its representative (the
    smallest key in the orbit, which is the orbit {g e}) and the (R, d) of an
    op that maps it there, R e + t = rep + S d, so g_e(k) = g_rep(R k + d)
    for g_e(k) = u(S k + e)."""
    E = np.array([[a, b, c] for c in range(S) for b in range(S) for a in range(S)], dtype=np.int32)
    Rm = np.stack([np.rint(np.asarray(op.Rm)).astype(np.int32) for op in ops_shifted])
    tv = np.stack([np.mod(np.rint(np.asarray(op.tv)).astype(np.int64), n) % S for op in ops_shifted]).astype(np.int32)
    rep = np.full(len(E), np.iinfo(np.int32).max, dtype=np.int32)
    best = np.zeros(len(E), dtype=np.int32)
    for g in range(len(Rm)):  # one op at a time: memory stays O(S^3)
        W = (E @ Rm[g].T + tv[g]) % S
        k = W[:, 0] + S * (W[:, 1] + S * W[:, 2])
        m = k < rep
        rep[m], best[m] = k[m], g
    tfull = np.stack([np.mod(np.rint(np.asarray(op.tv)).astype(np.int64), n) for op in ops_shifted])
    R = Rm[best].astype(np.int64)
    V = np.einsum("eij,ej->ei", R, E.astype(np.int64)) + tfull[best]
    Erep = np.stack([rep % S, (rep // S) % S, rep // (S * S)], 1).astype(np.int64)
    return E.astype(np.int64), rep.astype(np.int64), R, (V - Erep) // S


def fractal_depth(ops_shifted, n, lens=None, max_depth=6):
    """Depth of the decimation tree: split by parity while it still reduces
    the leaves (fewer than 8x the orbits one level down, the point where the
    reference's recursion finds no symmetry left to use), with leaf length
    n / 2^depth >= 2 and, if given, in `lens`."""
    d, orbits = 0, 1
    while d < max_depth and n % (2 << d) == 0 and n // (2 << d) >= 2:
        S = 2 << d
        if lens is not None and n // S not in lens:
            break
        _, rep, _, _ = _offset_orbits(ops_shifted, n, S)
        k = int(np.sum(rep == np.arange(len(rep))))
        if k >= 8 * orbits:
            break
        d, orbits = d + 1, k
    return d


def fractal_tables(spec: SpectralIndexing, ops_shifted, depth):
    """(stride, leaf offsets, stages, summary) of the fractal family for a
    cubic grid: leaves are the stride-2^depth sub-grids g_e(k) = u(2^depth k
    + e) of one offset per orbit; level l's values G^l_e(q) (stride 2^l,
    q mod N/2^l) come from level l+1 by the parity split
        G^l_e(q) = sum_b exp(-2 pi i q.b / M_l) G^{l+1}_{e + 2^l b}(q mod M_{l+1}),
    each child through its orbit representative,
        G_c(p) = exp(2 pi i (R^-T p).d / M) G_rep(R^-T p mod M).
    Only the values the spectral ASU needs are formed at each level, top
    down; the stages run bottom up (leaf FFTs -> level depth-1 -> ... ->
    spectral ASU), 8 entries per row before merging."""
    n = int(spec.N[0])
    if depth < 1:
        raise ValueError("depth >= 1")
    H = np.array([get_k_vector(spec, i) for i in range(len(spec.points))], dtype=np.int64).reshape(-1, 3)
    B = np.array([[(b >> 0) & 1, (b >> 1) & 1, (b >> 2) & 1] for b in range(8)], dtype=np.int64)
    # rows of level 0: (offset 0, h)
    rows_e = np.zeros((len(H), 3), dtype=np.int64)
    rows_q = H % n
    stages, sizes = [], [len(H)]
    for lev in range(depth):
        S1, M0, M1 = 2 << lev, n >> lev, n >> (lev + 1)
        E, rep, Rs, ds = _offset_orbits(ops_shifted, n, S1)
        reps = np.nonzero(rep == np.arange(len(rep)))[0]
        slot = np.full(len(rep), -1, dtype=np.int64)
        slot[reps] = np.arange(len(reps))
        r_idx, c_key, vals = [], [], []
        for b in B:
            c = rows_e + (1 << lev) * b
            ci = c[:, 0] + S1 * (c[:, 1] + S1 * c[:, 2])
            p = rows_q % M1
            Rit = np.rint(np.linalg.inv(Rs[ci]).transpose(0, 2, 1)).astype(np.int64)
            pp = np.einsum("rij,rj->ri", Rit, p) % M1
            v = np.exp(-2j * np.pi * (rows_q @ b) / M0) * np.exp(2j * np.pi * np.sum(pp * ds[ci], 1) / M1)
            r_idx.append(np.arange(len(rows_q)))
            c_key.append(slot[rep[ci]] * M1 ** 3 + pp[:, 0] + M1 * (pp[:, 1] + M1 * pp[:, 2]))
            vals.append(v)
        r_idx, c_key, vals = np.concatenate(r_idx), np.concatenate(c_key), np.concatenate(vals)
        if lev + 1 < depth:
            uk, cols = np.unique(c_key, return_inverse=True)
            sl, q = uk // M1 ** 3, uk % M1 ** 3
            rows_e = E[reps[sl]]
            rows_q = np.stack([q % M1, (q // M1) % M1, q // (M1 * M1)], 1)
            sizes.append(len(uk))
        else:
            cols, leaf_offsets = c_key, E[reps]
        stages.append(csr(r_idx, cols, vals, len(r_idx) // 8))
    return 1 << depth, leaf_offsets, stages[::-1], {"depth": depth, "leaves": len(leaf_offsets),
                                                     "level_rows": sizes}


class P3cEngine:
    """Device side of a sparse-stage family plan (cfft_p3c_*): the leaf pack,
    the leaf FFTs and the family's sparse stages.  P3c plans pass the class
    structure `cls` (stride-4 sub-sub-grids at 2 e); fractal plans pass
    `stride` and the raw leaf offsets in [0, stride)."""

    def __init__(self, n, cls, stages, stride=None, offsets=None):
        self.n = int(n)
        self.n_out = int(len(stages[-1][0]) - 1)
        if stride is None:
            eps, orbit_rep, chain, reps = cls
            off = np.ascontiguousarray(eps[reps].astype(np.int32).ravel())
        else:
            off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int32).ravel())
        self._keep = []
        rp = (C.c_void_p * len(stages))()
        cs = (C.c_void_p * len(stages))()
        vs = (C.c_void_p * len(stages))()
        nr = np.array([len(s[0]) - 1 for s in stages], dtype=np.int64)
        for i, (rowptr, cols, vals) in enumerate(stages):
            rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
            cols = np.ascontiguousarray(cols, dtype=np.int32)
            vals = np.ascontiguousarray(np.asarray(vals, dtype=np.complex128).view(np.float64))
            self._keep += [rowptr, cols, vals]
            rp[i], cs[i], vs[i] = rowptr.ctypes.data, cols.ctypes.data, vals.ctypes.data
        h = C.c_void_p()
        if stride is None:
            L.check(L.lib().cfft_p3c_create(self.n, len(off) // 3, L.i32p(off), len(stages), L.i64p(nr),
                                            rp, cs, vs, C.byref(h)))
        else:
            L.check(L.lib().cfft_p3c_create_strided(self.n, int(stride), len(off) // 3, L.i32p(off), len(stages),
                                                    L.i64p(nr), rp, cs, vs, C.byref(h)))
        self._keep = []
        self._h = h
        self.nnz = [int(s[0][-1]) for s in stages]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.lib().cfft_p3c_destroy(h)
            except Exception:
                pass
            self._h = None

    def forward_dev(self, u_ptr, F_ptr, stream=None):
        """Device-resident forward: u (N^3 FP64, x fastest) -> F (n_spec
        complex FP64), on `stream`, without synchronising."""
        L.check(L.lib().cfft_p3c_forward_dev(self._h, C.c_void_p(u_ptr), C.c_void_p(F_ptr), C.c_void_p(stream)))

    def forward(self, u, out=None):
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).ravel(order="F"))
        if u.size != self.n ** 3:
            raise L.CFFTError(L.CFFT_ERR_SIZE, "field must have N^3 points")
        if out is None:
            F = np.empty(self.n_out, dtype=np.complex128)
        else:
            F = out
            if not (isinstance(F, np.ndarray) and F.dtype == np.complex128 and F.shape == (self.n_out,)
                    and F.flags.c_contiguous):
                raise L.CFFTError(L.CFFT_ERR_SIZE, f"output must be a contiguous complex128 array of "
                                                   f"{self.n_out} values")
        L.check(L.lib().cfft_p3c_forward(self._h, L.f64p(u), L.f64p(F.view(np.float64))))
        return F
