"""NumPy FP64 restatement of the reference KRFFT algorithms (test oracle).

Conventions (identical to the reference):
  * grid coordinates are 0-based; arrays are indexed a[x, y, z] with shape
    (N1, N2, N3); linear indices are column-major (x fastest), as in Julia;
  * an operation is (R, t): x -> R x + t (mod N), R an integer matrix, t an
    integer vector in grid units;
  * the FFT is unnormalised forward, 1/n inverse (FFTW / numpy convention).

Ops are passed as two arrays: R with shape (n_ops, D, D) and t with shape
(n_ops, D).

Pinning (tests/test_oracle_*.py): the forward transform is checked against the
full-grid FFT of a symmetrised field at every spectral-ASU point (the
reference's own oracle, test/runtests.jl:130-155); the backward transform by
the round trip (test/test_m2_backward.jl:98-107); the diffusion pipelines by the
full-grid FFT·K·IFFT reference (test/test_helpers.jl:56-73); the index maps by
brute-force orbit enumeration, the closed-form spectral-ASU counts of the
reference tests (test/test_spectral_indexing.jl) and the structural counts of
test/debug_tree_output.log.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "CENT_P", "CENT_C", "CENT_A", "CENT_I", "CENT_F", "symmetrize", "symmetrize_scatter",
    "trunc_div", "apply_ops", "check_shift_invariance", "rescale_ops",
    "count_special_positions_fast", "find_optimal_shift", "dual_R",
    "SpectralASU", "calc_spectral_asu", "get_k_vector", "auto_L",
    "select_rep_ops", "detect_centering_type", "strip_centering",
    "alive_offsets", "lin_index", "make_symmetric", "fullgrid_reference", "fullgrid_kernel",
    "extract_subgrid", "kernel_spec", "krfft_forward", "reverse_lookup",
    "m2_backward_table", "m2_backward", "m2_scft", "centering_fold",
    "centering_unfold", "fft_channels", "ifft_channels", "assemble_G0",
    "disassemble_G0", "centered_krfft_forward", "CenteredBackward",
    "centered_backward_plan", "centered_backward", "centered_scft",
    "m2q_plan", "m2q_execute", "m7_scft", "fill_map", "subgrid_to_fullgrid",
    "fullgrid_to_subgrid", "centering_prefold", "prefold_merge", "subgrid_orbits", "qr_scft",
    "strang_step",
]

CENT_P, CENT_C, CENT_A, CENT_I, CENT_F = 0, 1, 2, 3, 4


def trunc_div(a, b):
    """Integer division truncating toward zero (Julia's ``÷``)."""
    a = np.asarray(a)
    q = np.abs(a) // abs(b)
    return np.where(a < 0, -q, q) * (1 if b > 0 else -1)


def lin_index(coords, dims):
    """Column-major linear index (0-based) of integer coordinates (..., D)."""
    coords = np.asarray(coords)
    li = np.zeros(coords.shape[:-1], dtype=np.int64)
    stride = 1
    for d, n in enumerate(dims):
        li = li + coords[..., d] * stride
        stride *= n
    return li


def _coords(n_total, dims):
    lin = np.arange(n_total, dtype=np.int64)
    out = np.empty((n_total, len(dims)), dtype=np.int64)
    rem = lin
    for d, n in enumerate(dims):
        out[:, d] = rem % n
        rem = rem // n
    return out


def apply_ops(R, t, x, N):
    """Images g(x) = R x + t mod N for every op: shape (n_ops, n_pts, D).

    src/symmetry_ops.jl:61-73 (apply_op!).
    """
    x = np.asarray(x, dtype=np.int64)
    img = np.einsum("gij,pj->gpi", R, x) + t[:, None, :]
    return np.mod(img, np.asarray(N, dtype=np.int64))


def check_shift_invariance(R, t, shift, N):
    """src/symmetry_ops.jl:81-94: delta = (R s - s) N must be integral (atol 1e-5)."""
    shift = np.asarray(shift, dtype=np.float64)
    Nf = np.asarray(N, dtype=np.float64)
    delta_f = (np.einsum("gij,j->gi", R.astype(np.float64), shift) - shift) * Nf
    delta_i = np.rint(delta_f).astype(np.int64)
    if not np.all(np.abs(delta_f - delta_i) <= 1e-5):
        return False, None
    return True, delta_i


def rescale_ops(R, t, N_from, N_to):
    """src/asu.jl:130-138: t_new = (t * N_to) ÷ N_from (truncating)."""
    t_new = np.stack([trunc_div(t[:, d] * N_to[d], N_from[d]) for d in range(len(N_from))], axis=1)
    return R.copy(), t_new.astype(np.int64)


def _orbit_components(R, t, N):
    """Orbit labels on the grid under x -> R x + t (closure = connected
    components, the fixed point of the reference's BFS worklists)."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    n_total = int(np.prod(N))
    X = _coords(n_total, N)
    src = np.arange(n_total, dtype=np.int64)
    rows, cols = [], []
    for g in range(R.shape[0]):
        img = np.mod(X @ R[g].T + t[g], np.asarray(N))
        rows.append(src)
        cols.append(lin_index(img, N))
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    A = sp.coo_matrix((np.ones(rows.size, dtype=np.int8), (rows, cols)), shape=(n_total, n_total))
    _, labels = connected_components(A, directed=True, connection="weak")
    return labels


def count_special_positions_fast(N, R, t):
    """src/asu.jl:148-214: number of orbits with multiplicity < |G|."""
    labels = _orbit_components(R, t, N)
    sizes = np.bincount(labels)
    return int(np.count_nonzero(sizes < R.shape[0]))


def find_optimal_shift(R, t, N):
    """src/asu.jl:57-123 (magic shift search).

    Returns (shift, t_shifted): shift is the fractional shift b/N, and the
    shifted ops are t' = t + (R b - b) N.
    """
    D = len(N)
    Np = (12,) * D
    Rp, tp = rescale_ops(R, t, N, Np)
    cands = [np.zeros(D), np.full(D, 0.5), np.full(D, 1.0 / 3.0), np.full(D, 0.25)]
    best = cands[0]
    min_sp = None
    for frac in cands:
        shift = frac / np.asarray(Np, dtype=np.float64)
        ok, deltas = check_shift_invariance(Rp, tp, shift, Np)
        if not ok:
            continue
        sp_count = count_special_positions_fast(Np, Rp, tp + deltas)
        if min_sp is None or sp_count < min_sp:
            min_sp = sp_count
            best = frac
        if min_sp == 0:
            break
    shift = best / np.asarray(N, dtype=np.float64)
    ok, deltas = check_shift_invariance(R, t, shift, N)
    if ok:
        best_t = t + deltas
    else:
        shift = np.zeros(D)
        best_t = t.copy()
    shift_half = 0.5 / np.asarray(N, dtype=np.float64)
    ok_h, d_h = check_shift_invariance(R, t, shift_half, N)
    if ok_h:
        t_half = t + d_h
        lp_half = int(np.prod([2 if np.any(t_half[:, d] % 2 == 1) else 1 for d in range(D)]))
        lp_best = int(np.prod([2 if np.any(best_t[:, d] % 2 == 1) else 1 for d in range(D)]))
        if lp_half > lp_best:
            shift = shift_half
            best_t = t_half
    return shift, best_t.astype(np.int64)


def dual_R(R):
    """src/symmetry_ops.jl:103-112: R* = (R^-1)^T (integer)."""
    return np.rint(np.transpose(np.linalg.inv(R.astype(np.float64)), (0, 2, 1))).astype(np.int64)


@dataclass
class SpectralASU:
    """src/spectral_indexing.jl:9-13 (points + multiplicities + grid)."""
    idx: np.ndarray      # (n_spec, D) 0-based frequencies in [0, N)
    mult: np.ndarray     # (n_spec,) orbit sizes
    N: tuple


def _rot_closed(Rs):
    """True when the rotation parts form a group (closed under products)."""
    keys = {tuple(r.ravel()) for r in Rs}
    return all(tuple((a @ b).ravel()) in keys for a in Rs for b in Rs)


def _orbit_min_closed(Rs, N):
    """For a closed rotation group: orbit minimum (column-major linear index)
    and orbit size of every grid point, orbit = {R h mod N : R in Rs}."""
    n_total = int(np.prod(N))
    X = _coords(n_total, N)
    Nv = np.asarray(N, dtype=np.int64)
    lin = np.arange(n_total, dtype=np.int64)
    mn = lin.copy()
    stab = np.zeros(n_total, dtype=np.int64)
    for Rg in Rs:
        img = lin_index(np.mod(X @ Rg.T, Nv), N)
        np.minimum(mn, img, out=mn)
        stab += img == lin
    # orbit size = |G| / |stabiliser| (each op counted once per distinct matrix)
    return mn, len(Rs) // stab


def calc_spectral_asu(R, t, N):
    """src/spectral_indexing.jl:22-176 (direct orbit enumeration).

    Orbits of [0,N)^D under R* = (R^-1)^T; representative = smallest
    column-major linear index; kept when |sum_{g in Stab} exp(-2 pi i k.t_g/N)|
    > 1e-5; sorted lexicographically by (k1, k2, ..., kD).  A closed rotation
    set takes the orbit as {R* h} directly; otherwise the closure is the
    connected components of the op graph (as the reference's worklist).
    """
    N = tuple(int(n) for n in N)
    D = len(N)
    Rs = dual_R(R)
    n_total = int(np.prod(N))
    Nv = np.asarray(N, dtype=np.int64)
    X = _coords(n_total, N)
    if _rot_closed(Rs):
        mn, size = _orbit_min_closed(np.unique(Rs, axis=0), N)
        reps = np.flatnonzero(mn == np.arange(n_total))
        mult = size[reps]
    else:
        labels = _orbit_components(Rs, np.zeros_like(t), N)
        sizes = np.bincount(labels)
        minlin = np.full(sizes.size, n_total, dtype=np.int64)
        np.minimum.at(minlin, labels, np.arange(n_total, dtype=np.int64))
        reps = minlin
        mult = sizes
    K = X[reps]
    # extinction filter
    keep = np.zeros(reps.size, dtype=bool)
    for i0 in range(0, reps.size, 4096):
        k = K[i0:i0 + 4096]
        img = np.einsum("gij,pj->gpi", Rs, k)
        stab = np.all(np.mod(img - k[None], Nv) == 0, axis=2)          # (g, p)
        phase = np.zeros((R.shape[0], k.shape[0]))
        for d in range(D):
            phase = phase + k[None, :, d] * t[:, None, d] / N[d]
        s = np.sum(np.where(stab, np.exp(-2j * np.pi * phase), 0.0), axis=0)
        keep[i0:i0 + 4096] = np.abs(s) > 1e-5
    K = K[keep]
    mult = mult[keep]
    order = np.lexsort(tuple(K[:, d] for d in reversed(range(D))))
    return SpectralASU(K[order].copy(), mult[order].copy(), N)


def get_k_vector(idx, N):
    """src/spectral_indexing.jl:195-238: wrap 0..N-1 to (-N/2, N/2]-style
    signed frequencies (val >= N - N÷2 -> val - N)."""
    idx = np.asarray(idx, dtype=np.int64)
    out = idx.copy()
    for d, n in enumerate(N):
        out[..., d] = np.where(idx[..., d] >= n - n // 2, idx[..., d] - n, idx[..., d])
    return out


def _centered(h, N):
    """Centred frequency used by the diffusion kernels (h >= N÷2 -> h - N),
    src/q_fused_krfft.jl:966-973, src/centering_fold.jl:884-891."""
    h = np.asarray(h, dtype=np.int64)
    out = h.copy()
    for d, n in enumerate(N):
        out[..., d] = np.where(h[..., d] >= n // 2, h[..., d] - n, h[..., d])
    return out


def auto_L(t):
    """src/krfft.jl:225-279."""
    n_ops, dim = t.shape
    L_max = np.array([2 if np.any(t[:, d] % 2 == 1) else 1 for d in range(dim)])
    n_sub = int(np.prod(L_max))
    if n_sub <= 1:
        return L_max
    reach = {tuple(int(v) for v in np.mod(t[i], L_max)) for i in range(n_ops)}
    if len(reach) == n_sub:
        return L_max
    best = np.ones(dim, dtype=np.int64)
    for mask in range(1, 2 ** dim):
        L_try = np.ones(dim, dtype=np.int64)
        for d in range(dim):
            if (mask >> d) & 1 and L_max[d] == 2:
                L_try[d] = 2
        n = int(np.prod(L_try))
        r = {tuple(int(v) for v in np.mod(t[i], L_try)) for i in range(n_ops)}
        if len(r) == n and n > int(np.prod(best)):
            best = L_try
    return best


def _cosets(L):
    """Iterators.product order (first dimension fastest)."""
    L = [int(v) for v in L]
    out = []
    for flat in range(int(np.prod(L))):
        rem = flat
        a = []
        for l in L:
            a.append(rem % l)
            rem //= l
        out.append(tuple(a))
    return out


def select_rep_ops(R, t, L, N):
    """src/krfft.jl:958-987: one op per coset x0 = t mod L, preferring diagonal
    R (quality 1) with t_d mod N_d in {0, N_d-1} (quality 2); first wins ties."""
    dim = len(N)
    best = {}
    qual = {}
    for i in range(R.shape[0]):
        x0 = tuple(int(v) for v in np.mod(t[i], L))
        off = R[i] - np.diag(np.diag(R[i]))
        is_diag = not np.any(off)
        simple = all(int(t[i, d]) % N[d] in (0, N[d] - 1) for d in range(dim))
        q = (2 if simple else 1) if is_diag else 0
        if x0 not in best or q > qual[x0]:
            best[x0] = i
            qual[x0] = q
    reps = []
    for x0 in _cosets(L):
        if x0 not in best:
            raise ValueError(f"Subgrid x0={x0} not reachable")
        reps.append(best[x0])
    return np.array(reps, dtype=np.int64)


def detect_centering_type(R, t, N):
    """src/symmetry_ops.jl:133-179."""
    D = R.shape[1]
    pure = []
    I = np.eye(D, dtype=np.int64)
    for i in range(R.shape[0]):
        if not np.array_equal(R[i], I):
            continue
        if np.all(np.abs(t[i]) < 0.1):
            continue
        pure.append(tuple(abs(abs(int(t[i, d])) / N[d] - 0.5) < 0.01 for d in range(D)))
    if not pure or D < 3:
        return CENT_P
    has_xyz = any(all(p) for p in pure)
    has_0yz = any((not p[0]) and p[1] and p[2] for p in pure)
    has_x0z = any(p[0] and (not p[1]) and p[2] for p in pure)
    has_xy0 = any(p[0] and p[1] and (not p[2]) for p in pure)
    if has_0yz + has_x0z + has_xy0 >= 2:
        return CENT_F
    if has_xyz:
        return CENT_I
    if has_xy0:
        return CENT_C
    if has_0yz:
        return CENT_A
    if has_x0z:
        return CENT_A
    return CENT_P


def strip_centering(R, t, cent, N):
    """src/symmetry_ops.jl:192-222."""
    D = len(N)
    if cent in (CENT_I, CENT_F):
        halve = [True] * D
    elif cent == CENT_C:
        halve = [d < 2 for d in range(D)]
    elif cent == CENT_A:
        halve = [d >= 1 for d in range(D)]
    else:
        halve = [False] * D
    Ns = tuple(N[d] // 2 if halve[d] else N[d] for d in range(D))
    seen = set()
    Ro, to = [], []
    for i in range(R.shape[0]):
        ts = tuple(int(t[i, d]) % Ns[d] for d in range(D))
        key = (tuple(R[i].ravel()), ts)
        if key not in seen:
            seen.add(key)
            Ro.append(R[i])
            to.append(ts)
    return np.array(Ro), np.array(to, dtype=np.int64), Ns


def alive_offsets(cent):
    """src/centering_fold.jl:111-127."""
    return {
        CENT_I: [(0, 0, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
        CENT_F: [(0, 0, 0), (1, 1, 1)],
        CENT_C: [(0, 0, 0), (1, 1, 0), (0, 0, 1), (1, 1, 1)],
        CENT_A: [(0, 0, 0), (0, 1, 1), (1, 0, 0), (1, 1, 1)],
    }[cent]


# ---------------------------------------------------------------------------
# test fields (test/test_helpers.jl)
# ---------------------------------------------------------------------------

def make_symmetric(R, t, N, seed=42, dist="normal"):
    """test/test_helpers.jl:11-29 (scatter-average over the ops).

    The random stream is NumPy's PCG64 (Julia's Xoshiro stream cannot be
    reproduced), so fields match the reference's in construction, not bits.
    """
    rng = np.random.default_rng(seed)
    u = rng.standard_normal(N) if dist == "normal" else rng.uniform(-1.0, 1.0, N)
    return symmetrize(u, R, t)


def symmetrize(u, R, t):
    """Group average of u.  test/test_helpers.jl:11-29 scatters u(x) to g(x)
    and divides by |G|; for a group this equals the gather form
    mean_g u(g(x)) computed here (g -> g^-1 permutes the group)."""
    N = u.shape
    n_total = u.size
    X = _coords(n_total, N)
    uf = u.reshape(-1, order="F")
    out = np.zeros(n_total)
    for g in range(R.shape[0]):
        y = np.mod(X @ R[g].T + t[g], np.asarray(N))
        out += uf[lin_index(y, N)]
    out /= R.shape[0]
    return out.reshape(N, order="F")


def symmetrize_scatter(u, R, t):
    """The scatter form of test/test_helpers.jl:11-29 (kept to pin the gather
    form: both agree for a group)."""
    N = u.shape
    n_total = u.size
    X = _coords(n_total, N)
    uf = u.reshape(-1, order="F")
    out = np.zeros(n_total)
    for g in range(R.shape[0]):
        y = np.mod(X @ R[g].T + t[g], np.asarray(N))
        np.add.at(out, lin_index(y, N), uf)
    out /= R.shape[0]
    return out.reshape(N, order="F")


def _recip_B(lattice):
    return 2 * np.pi * np.linalg.inv(np.asarray(lattice, dtype=np.float64)).T


def fullgrid_kernel(N, ds, lattice):
    """K(h) = exp(-|2 pi A^-T h_c|^2 ds) on the full grid, h_c the centred
    frequency (test/test_helpers.jl:56-73)."""
    B = _recip_B(lattice)
    grids = np.meshgrid(*[np.arange(n) for n in N], indexing="ij")
    hc = np.stack([np.where(g >= n // 2, g - n, g) for g, n in zip(grids, N)], axis=-1)
    kv = hc @ B.T
    return np.exp(-np.sum(kv * kv, axis=-1) * ds)


def fullgrid_reference(u, N, ds, lattice, L):
    """test/test_helpers.jl:56-73: stride-L subgrid of IFFT(K . FFT(u))."""
    f = np.real(np.fft.ifftn(fullgrid_kernel(N, ds, lattice) * np.fft.fftn(u)))
    return extract_subgrid(f, L)


def extract_subgrid(u, L):
    """test/test_helpers.jl:76-80."""
    sl = tuple(slice(0, None, int(l)) for l in L)
    return np.ascontiguousarray(u[sl])


def kernel_spec(spec: SpectralASU, ds, lattice):
    """K_spec[i] = exp(-|2 pi A^-T h_c|^2 ds), src/q_fused_krfft.jl:966-973."""
    B = _recip_B(lattice)
    hc = _centered(spec.idx, spec.N).astype(np.float64)
    kv = hc @ B.T
    return np.exp(-np.sum(kv * kv, axis=1) * ds)


# ---------------------------------------------------------------------------
# M2 forward / backward (src/krfft.jl)
# ---------------------------------------------------------------------------

def krfft_forward(Y0, spec: SpectralASU, R, t, rep, N):
    """src/krfft.jl:301-389 (recon table) + :529-545 (general reconstruction).

    F(h) = sum_a exp(-2 pi i h_c.t_a/N) Y0[(R_a^T h_c) mod M], h_c the signed
    frequency of the spectral-ASU point.  (The Pmmm fast path :475-526 is the
    same sum with diagonal R and t in {0,-1}.)
    """
    M = Y0.shape
    hc = get_k_vector(spec.idx, N)
    F = np.zeros(hc.shape[0], dtype=np.complex128)
    for a in rep:
        phase = np.zeros(hc.shape[0])
        for d in range(len(N)):
            phase = phase + hc[:, d] * t[a, d] / N[d]
        w = np.exp(-1j * 2 * np.pi * phase)
        rot = np.mod(hc @ R[a], np.asarray(M))
        F += w * Y0[tuple(rot[:, d] for d in range(len(M)))]
    return F


def reverse_lookup(spec: SpectralASU, R, t, N):
    """src/krfft.jl:1183-1255: every h in [0,N)^D -> (spec_idx, phase, conj),
    first-come in the reference's insertion order.  Missing (extinct) -> -1."""
    N = tuple(N)
    D = len(N)
    Nv = np.asarray(N, dtype=np.int64)
    n_spec = spec.idx.shape[0]
    n_ops = R.shape[0]
    h = spec.idx
    # insertion order per spectral point: identity, then per op (rot, neg), then -h
    keys, sidx, phases, conj = [], [], [], []
    phase_val = np.zeros((n_ops, n_spec))
    for d in range(D):
        phase_val = phase_val + h[None, :, d] * t[:, None, d] / N[d]
    rot = np.mod(np.einsum("gji,pj->gpi", R, h), Nv)          # R^T h
    neg = np.mod(-rot, Nv)
    ar = np.arange(n_spec)
    blocks_k = [lin_index(h, N)[:, None]]
    blocks_p = [np.ones((n_spec, 1), dtype=np.complex128)]
    blocks_c = [np.zeros((n_spec, 1), dtype=bool)]
    for g in range(n_ops):
        blocks_k.append(lin_index(rot[g], N)[:, None])
        blocks_p.append(np.exp(1j * 2 * np.pi * phase_val[g])[:, None])
        blocks_c.append(np.zeros((n_spec, 1), dtype=bool))
        blocks_k.append(lin_index(neg[g], N)[:, None])
        blocks_p.append(np.exp(-1j * 2 * np.pi * phase_val[g])[:, None])
        blocks_c.append(np.ones((n_spec, 1), dtype=bool))
    blocks_k.append(lin_index(np.mod(-h, Nv), N)[:, None])
    blocks_p.append(np.ones((n_spec, 1), dtype=np.complex128))
    blocks_c.append(np.ones((n_spec, 1), dtype=bool))
    # order: spectral point outer, entries inner
    K = np.concatenate(blocks_k, axis=1).ravel()
    P = np.concatenate(blocks_p, axis=1).ravel()
    C = np.concatenate(blocks_c, axis=1).ravel()
    S = np.repeat(ar, len(blocks_k))
    uniq, first = np.unique(K, return_index=True)
    n_total = int(np.prod(N))
    idx_arr = np.full(n_total, -1, dtype=np.int64)
    ph_arr = np.zeros(n_total, dtype=np.complex128)
    cj_arr = np.zeros(n_total, dtype=bool)
    idx_arr[uniq] = S[first]
    ph_arr[uniq] = P[first]
    cj_arr[uniq] = C[first]
    return idx_arr, ph_arr, cj_arr


def m2_backward_table(spec: SpectralASU, R, t, N):
    """src/krfft.jl:997-1068 (+ plan_m2_backward :1101-1170): per fibre q the
    d x d butterfly B(q)[a,b] = exp(-2 pi i h_a.t_b/N), h_a = q + M alpha_a; the
    first row x of B^-1 (solve B^T x = e1) times the reverse-lookup phase.

    Returns (spec_idx (d, Mvol), weight (d, Mvol), conj (d, Mvol), L, M, rep).
    Missing entries have weight 0 and index 0.
    """
    N = tuple(N)
    D = len(N)
    L = auto_L(t)
    M = tuple(int(N[d] // L[d]) for d in range(D))
    if any(L[d] * M[d] != N[d] for d in range(D)):
        raise ValueError(f"Grid size N={N} not divisible by auto L={tuple(L)}.")
    rep = select_rep_ops(R, t, L, N)
    d = int(np.prod(L))
    idx_arr, ph_arr, cj_arr = reverse_lookup(spec, R, t, N)
    Mvol = int(np.prod(M))
    Q = _coords(Mvol, M)
    alphas = np.array(_cosets(L), dtype=np.int64)                    # (d, D)
    H = Q[None, :, :] + alphas[:, None, :] * np.asarray(M)[None, None, :]   # (d, Mvol, D)
    tb = t[rep]                                                      # (d, D)
    phase = np.zeros((d, d, Mvol))                                   # (a, b, q)
    for dd in range(D):
        phase = phase + H[:, None, :, dd] * tb[None, :, None, dd] / N[dd]
    B = np.exp(-2j * np.pi * phase)                                  # B[a,b,q]
    Bq = np.transpose(B, (2, 0, 1))                                  # (q, a, b)
    e1 = np.zeros((Mvol, d, 1), dtype=np.complex128)
    e1[:, 0, 0] = 1.0
    x = np.linalg.solve(np.transpose(Bq, (0, 2, 1)), e1)[:, :, 0]   # (q, a)
    hl = lin_index(np.mod(H, np.asarray(N)), N)                      # (d, Mvol)
    sidx = idx_arr[hl]
    w = x.T * ph_arr[hl]
    w = np.where(sidx >= 0, w, 0.0)
    cj = np.where(sidx >= 0, cj_arr[hl], False)
    sidx = np.where(sidx >= 0, sidx, 0)
    return sidx, w, cj, L, M, rep


def m2_backward(F_spec, table):
    """src/krfft.jl:1269-1311: Y0(q) = sum_a w(q,a) [F; conj F][idx(q,a)],
    then the M-grid IFFT (complex result, as bplan.f0_buf)."""
    sidx, w, cj, L, M, _ = table
    Fv = np.where(cj, np.conj(F_spec[sidx]), F_spec[sidx])
    Y = np.sum(w * Fv, axis=0).reshape(M, order="F")
    return np.fft.ifftn(Y)


def m2_scft(f0, spec, R, t, N, ds, lattice, table=None):
    """src/q_fused_krfft.jl:948-1026 (plan_m2_scft / execute_m2_scft!)."""
    if table is None:
        table = m2_backward_table(spec, R, t, N)
    rep = table[5]
    Y0 = np.fft.fftn(f0.astype(np.complex128))
    F = krfft_forward(Y0, spec, R, t, rep, N)
    F = F * kernel_spec(spec, ds, lattice)
    return np.real(m2_backward(F, table))


# ---------------------------------------------------------------------------
# centering fold on the stride-2 subgrid (src/centering_fold.jl)
# ---------------------------------------------------------------------------

def _fold_twiddle(off, M, H):
    """src/centering_fold.jl:76-88: tw_d[n] = exp(-2 pi i off_d n / M_d)."""
    tws = []
    for d in range(3):
        n = np.arange(H[d])
        tws.append(np.exp(-2j * np.pi * off[d] * n / M[d]) if off[d] else np.ones(H[d], complex))
    return tws[0][:, None, None] * tws[1][None, :, None] * tws[2][None, None, :]


def centering_fold(f0, cent):
    """src/centering_fold.jl:140-257: g_off(n0) = sum_eps (-1)^(off.eps)
    f0(n0 + H eps) . exp(-2 pi i off.n0/M), for the alive offsets."""
    M = f0.shape
    if any(m % 2 for m in M):
        raise ValueError("Subgrid dimensions must be even for centering fold")
    H = tuple(m // 2 for m in M)
    chans = []
    for off in alive_offsets(cent):
        g = np.zeros(H)
        for e in range(8):
            eps = (e & 1, (e >> 1) & 1, (e >> 2) & 1)
            s = -1.0 if (off[0] * eps[0] + off[1] * eps[1] + off[2] * eps[2]) & 1 else 1.0
            g = g + s * f0[eps[0] * H[0]:eps[0] * H[0] + H[0],
                           eps[1] * H[1]:eps[1] * H[1] + H[1],
                           eps[2] * H[2]:eps[2] * H[2] + H[2]]
        chans.append(g * _fold_twiddle(off, M, H))
    return chans


def fft_channels(chans):
    """src/centering_fold.jl:264-268."""
    return [np.fft.fftn(c) for c in chans]


def ifft_channels(chans):
    """src/centering_fold.jl:442-446."""
    return [np.fft.ifftn(c) for c in chans]


def assemble_G0(chan_hat, cent, M):
    """src/centering_fold.jl:277-296."""
    G0 = np.zeros(M, dtype=np.complex128)
    for c, off in enumerate(alive_offsets(cent)):
        G0[off[0]::2, off[1]::2, off[2]::2] = chan_hat[c]
    return G0


def disassemble_G0(G0, cent):
    """src/centering_fold.jl:454-469."""
    return [G0[off[0]::2, off[1]::2, off[2]::2].copy() for off in alive_offsets(cent)]


def centering_unfold(chans, cent, M):
    """src/centering_fold.jl:481-580: undo the twiddle, real part, inverse WHT
    / 8 (extinct channels zero)."""
    H = tuple(m // 2 for m in M)
    offs = alive_offsets(cent)
    c_raw = {}
    for c, off in enumerate(offs):
        c_raw[off] = np.real(chans[c] * np.conj(_fold_twiddle(off, M, H)))
    f0 = np.zeros(M)
    for e in range(8):
        eps = (e & 1, (e >> 1) & 1, (e >> 2) & 1)
        acc = np.zeros(H)
        for off, g in c_raw.items():
            s = -1.0 if (off[0] * eps[0] + off[1] * eps[1] + off[2] * eps[2]) & 1 else 1.0
            acc = acc + s * g
        f0[eps[0] * H[0]:eps[0] * H[0] + H[0],
           eps[1] * H[1]:eps[1] * H[1] + H[1],
           eps[2] * H[2]:eps[2] * H[2] + H[2]] = acc / 8.0
    return f0


def centered_krfft_forward(f0, spec, R, t, N, cent):
    """src/centering_fold.jl:392-431: fold -> FFT channels -> assemble G0 ->
    reconstruct (recon table of plan_krfft)."""
    L = auto_L(t)
    rep = select_rep_ops(R, t, L, N)
    G0 = assemble_G0(fft_channels(centering_fold(f0, cent)), cent, f0.shape)
    return krfft_forward(G0, spec, R, t, rep, N)


@dataclass
class CenteredBackward:
    """src/centering_fold.jl:599-620 (CSR at orbit reps + orbit expand)."""
    offsets: np.ndarray
    spec_idx: np.ndarray
    weight: np.ndarray
    orbit_rep: np.ndarray
    orbit_id: np.ndarray
    orbit_phase: np.ndarray
    M: tuple
    cent: int


def centered_backward_plan(spec, R, t, N, cent):
    """src/centering_fold.jl:630-731."""
    sidx, w, cj, L, M, rep = m2_backward_table(spec, R, t, N)
    M = tuple(M)
    Mvol = int(np.prod(M))
    rem = [(R[i], trunc_div(t[i], 2)) for i in range(R.shape[0]) if np.all(t[i] % 2 == 0)]
    orbit_id = np.full(Mvol, -1, dtype=np.int64)
    orbit_phase = np.zeros(Mvol, dtype=np.complex128)
    reps = []
    Mv = np.asarray(M, dtype=np.int64)
    for m in range(Mvol):
        if orbit_id[m] >= 0:
            continue
        reps.append(m)
        oid = len(reps) - 1
        orbit_id[m] = oid
        orbit_phase[m] = 1.0
        queue = [m]
        qi = 0
        while qi < len(queue):
            cur = queue[qi]
            qi += 1
            qv = np.array([cur % M[0], (cur // M[0]) % M[1], cur // (M[0] * M[1])], dtype=np.int64)
            cph = orbit_phase[cur]
            for Rg, s in rem:
                rq = np.mod(Rg.T @ qv, Mv)
                rl = int(rq[0] + M[0] * rq[1] + M[0] * M[1] * rq[2])
                if orbit_id[rl] < 0:
                    ph = np.exp(1j * np.pi * 2 * float(np.sum(qv * s / Mv)))
                    orbit_id[rl] = oid
                    orbit_phase[rl] = ph * cph
                    queue.append(rl)
    reps = np.array(reps, dtype=np.int64)
    offsets = [0]
    s_out, w_out = [], []
    for q in reps:
        for a in range(sidx.shape[0]):
            if abs(w[a, q]) > 1e-15:
                w_out.append(w[a, q])
                s_out.append(-(sidx[a, q] + 1) if cj[a, q] else sidx[a, q] + 1)
        offsets.append(len(s_out))
    return CenteredBackward(np.array(offsets), np.array(s_out, dtype=np.int64),
                            np.array(w_out, dtype=np.complex128), reps, orbit_id,
                            orbit_phase, M, cent)


def centered_backward(F_spec, bp: CenteredBackward):
    """src/centering_fold.jl:745-813."""
    n_orb = bp.orbit_rep.size
    G_reps = np.zeros(n_orb, dtype=np.complex128)
    for i in range(n_orb):
        lo, hi = bp.offsets[i], bp.offsets[i + 1]
        s = bp.spec_idx[lo:hi]
        vals = np.where(s > 0, F_spec[np.abs(s) - 1], np.conj(F_spec[np.abs(s) - 1]))
        G_reps[i] = np.sum(bp.weight[lo:hi] * vals)
    G0 = (bp.orbit_phase * G_reps[bp.orbit_id]).reshape(bp.M, order="F")
    chans = ifft_channels(disassemble_G0(G0, bp.cent))
    return centering_unfold(chans, bp.cent, bp.M)


def centered_scft(f0, spec, R, t, N, ds, lattice, cent, bp=None):
    """src/centering_fold.jl:868-931 (plan/execute_centered_scft!)."""
    if bp is None:
        bp = centered_backward_plan(spec, R, t, N, cent)
    F = centered_krfft_forward(f0, spec, R, t, N, cent)
    F = F * kernel_spec(spec, ds, lattice)
    return centered_backward(F, bp)


# ---------------------------------------------------------------------------
# Q-fused (src/q_fused_krfft.jl)
# ---------------------------------------------------------------------------

def m2q_plan(R, t, N, ds, lattice):
    """src/q_fused_krfft.jl:83-169, :182-413: per fibre q,
    Q = B^-1 diag(K) B (first row), gather idx(b,q) = lin(R_b^T q mod M)."""
    N = tuple(N)
    D = len(N)
    L = auto_L(t)
    M = tuple(int(N[d] // L[d]) for d in range(D))
    rep = select_rep_ops(R, t, L, N)
    d = len(rep)
    Mvol = int(np.prod(M))
    Q = _coords(Mvol, M)
    alphas = np.array(_cosets(L), dtype=np.int64)
    H = Q[None, :, :] + alphas[:, None, :] * np.asarray(M)[None, None, :]   # (a, q, D)
    Bm = _recip_B(lattice)
    hc = _centered(H, N).astype(np.float64)
    kv = hc @ Bm.T
    Kv = np.exp(-np.sum(kv * kv, axis=-1) * ds)                             # (a, q)
    tb = t[rep]
    phase = np.zeros((d, d, Mvol))
    for dd in range(D):
        phase = phase + H[:, None, :, dd] * tb[None, :, None, dd] / N[dd]
    B = np.transpose(np.exp(-2j * np.pi * phase), (2, 0, 1))                # (q, a, b)
    KB = Kv.T[:, :, None] * B
    Qfull = np.linalg.solve(B, KB)
    Qrow = Qfull[:, 0, :].T                                                  # (b, q)
    gather = np.stack([lin_index(np.mod(Q @ R[b], np.asarray(M)), M) for b in rep])  # (b, q)
    return dict(Q=Qrow, gather=gather, L=L, M=M, rep=rep)


def m2q_execute(f0, plan):
    """src/q_fused_krfft.jl:553-592 (generic Q multiply; the separable Pmmm
    path :606-667 computes the same product)."""
    M = plan["M"]
    Y = np.fft.fftn(f0.astype(np.complex128)).reshape(-1, order="F")
    Yn = np.sum(plan["Q"] * Y[plan["gather"]], axis=0).reshape(M, order="F")
    return np.real(np.fft.ifftn(Yn))


def m7_scft(f0, plan, cent):
    """src/q_fused_krfft.jl:870-895: fold/FFT/assemble -> Q multiply ->
    disassemble/IFFT/unfold."""
    M = f0.shape
    G0 = assemble_G0(fft_channels(centering_fold(f0, cent)), cent, M).reshape(-1, order="F")
    Gn = np.sum(plan["Q"] * G0[plan["gather"]], axis=0).reshape(M, order="F")
    return centering_unfold(ifft_channels(disassemble_G0(Gn, cent)), cent, M)


def subgrid_orbits(R, t, N, L):
    """Orbits of the M-grid under the ops that keep the stride-L subgrid
    (t = 0 mod L, s = t/L): G0(R^T q) = exp(2 pi i q.s/M) G0(q) for a
    symmetric subgrid field (the orbit expansion of
    src/centering_fold.jl:630-731).  Returns (rep (Mvol,), phase (Mvol,),
    reps): every q's representative (linear index) and the phase with
    G0(q) = phase * G0(rep)."""
    N = tuple(N)
    M = tuple(int(N[d] // L[d]) for d in range(3))
    Mv = np.asarray(M, dtype=np.int64)
    Mvol = int(np.prod(M))
    rem = [(R[i], t[i] // np.asarray(L)) for i in range(R.shape[0]) if np.all(t[i] % np.asarray(L) == 0)]
    rep = np.full(Mvol, -1, dtype=np.int64)
    phase = np.zeros(Mvol, dtype=np.complex128)
    reps = []
    for m in range(Mvol):
        if rep[m] >= 0:
            continue
        reps.append(m)
        rep[m] = m
        phase[m] = 1.0
        queue = [m]
        qi = 0
        while qi < len(queue):
            cur = queue[qi]
            qi += 1
            qv = np.array([cur % M[0], (cur // M[0]) % M[1], cur // (M[0] * M[1])], dtype=np.int64)
            for Rg, s_ in rem:
                rq = np.mod(Rg.T @ qv, Mv)
                rl = int(rq[0] + M[0] * (rq[1] + M[1] * rq[2]))
                if rep[rl] < 0:
                    rep[rl] = m
                    phase[rl] = phase[cur] * np.exp(2j * np.pi * float(np.sum(qv * s_ / Mv)))
                    queue.append(rl)
    return rep, phase, np.array(reps, dtype=np.int64)


def qr_scft(f0, qplan, orbits):
    """Q multiply evaluated at the orbit representatives only, then expanded
    with the orbit phases (the device path QR); equals m2q_execute for a
    symmetric f0."""
    M = f0.shape
    rep, phase, reps = orbits
    Y = np.fft.fftn(f0.astype(np.complex128)).reshape(-1, order="F")
    Yr = np.sum(qplan["Q"][:, reps] * Y[qplan["gather"][:, reps]], axis=0)
    full = np.zeros(Y.size, dtype=np.complex128)
    full[reps] = Yr
    Yn = (phase * full[rep]).reshape(M, order="F")
    return np.real(np.fft.ifftn(Yn))


def fill_map(R, t, L, N):
    """src/q_fused_krfft.jl:733-783: for each x the first op (list order)
    landing on the stride-L subgrid; value = linear M-grid index (0-based)."""
    N = tuple(N)
    D = len(N)
    M = tuple(int(N[d] // L[d]) for d in range(D))
    n_total = int(np.prod(N))
    X = _coords(n_total, N)
    out = np.full(n_total, -1, dtype=np.int64)
    for g in range(R.shape[0]):
        todo = out < 0
        if not np.any(todo):
            break
        y = np.mod(X[todo] @ R[g].T + t[g], np.asarray(N))
        on = np.all(y % np.asarray(L) == 0, axis=1)
        sub = lin_index(y[on] // np.asarray(L), M)
        idx = np.nonzero(todo)[0][on]
        out[idx] = sub
    if np.any(out < 0):
        raise ValueError("No symmetry operation maps grid point to the subgrid.")
    return out


def subgrid_to_fullgrid(f0, fmap, N):
    """src/q_fused_krfft.jl:681-688."""
    return f0.reshape(-1, order="F")[fmap].reshape(N, order="F")


def fullgrid_to_subgrid(u, L):
    """src/q_fused_krfft.jl:698-719."""
    return extract_subgrid(u, L)


# ---------------------------------------------------------------------------
# centering pre-fold on the full grid (src/centering_prefold.jl)
# ---------------------------------------------------------------------------

def centering_prefold(u, cent):
    """src/centering_prefold.jl:27-131."""
    Nf = u.shape
    Nh = tuple(n // 2 for n in Nf)
    ix, iy, iz = np.meshgrid(*[np.arange(n) for n in Nh], indexing="ij")
    if cent == CENT_I:
        v0 = u[:Nh[0], :Nh[1], :Nh[2]]
        vx = u[Nh[0]:, :Nh[1], :Nh[2]]
        vy = u[:Nh[0], Nh[1]:, :Nh[2]]
        vz = u[:Nh[0], :Nh[1], Nh[2]:]
        terms = (v0, vx, vy, vz)
        signs = ((1, 1, 1, 1), (1, -1, -1, 1), (1, -1, 1, -1), (1, 1, -1, -1))
        pars = [(0, 0, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1)]
        chans = []
        for s, p in zip(signs, pars):
            tw = np.exp(-2j * np.pi * (p[0] * ix / Nf[0] + p[1] * iy / Nf[1] + p[2] * iz / Nf[2]))
            chans.append(2.0 * tw * sum(si * ti for si, ti in zip(s, terms)))
        return chans, pars, Nh
    if cent == CENT_F:
        v0 = u[:Nh[0], :Nh[1], :Nh[2]]
        vs = u[Nh[0]:, :Nh[1], :Nh[2]]
        tw = np.exp(-2j * np.pi * (ix / Nf[0] + iy / Nf[1] + iz / Nf[2]))
        return [4.0 * (v0 + vs) + 0j, 4.0 * (v0 - vs) * tw], [(0, 0, 0), (1, 1, 1)], Nh
    if cent == CENT_C:
        Ns = (Nh[0], Nh[1], Nf[2])
        ix, iy, iz = np.meshgrid(*[np.arange(n) for n in Ns], indexing="ij")
        v0 = u[:Nh[0], :Nh[1], :]
        vs = u[Nh[0]:, :Nh[1], :]
        tw = np.exp(-2j * np.pi * (ix / Nf[0] + iy / Nf[1]))
        return [2.0 * (v0 + vs) + 0j, 2.0 * (v0 - vs) * tw], [(0, 0, 0), (1, 1, 0)], Ns
    if cent == CENT_A:
        Ns = (Nf[0], Nh[1], Nh[2])
        ix, iy, iz = np.meshgrid(*[np.arange(n) for n in Ns], indexing="ij")
        v0 = u[:, :Nh[1], :Nh[2]]
        vs = u[:, Nh[1]:, :Nh[2]]
        tw = np.exp(-2j * np.pi * (iy / Nf[1] + iz / Nf[2]))
        return [2.0 * (v0 + vs) + 0j, 2.0 * (v0 - vs) * tw], [(0, 0, 0), (0, 1, 1)], Ns
    raise ValueError("CentP has no pre-fold")


def prefold_merge(spec, chans_hat, pars, Nsub, N):
    """src/centering_prefold.jl:207-257 (merge table + merge)."""
    out = np.zeros(spec.idx.shape[0], dtype=np.complex128)
    for i, h in enumerate(spec.idx):
        par = tuple(int(h[d]) % 2 if Nsub[d] < N[d] else 0 for d in range(3))
        if par not in pars:
            continue
        c = pars.index(par)
        k = tuple((int(h[d]) - par[d]) // 2 % Nsub[d] if Nsub[d] < N[d] else int(h[d]) % Nsub[d]
                  for d in range(3))
        out[i] = chans_hat[c][k]
    return out


# ---------------------------------------------------------------------------
# diffusion step (src/diffusion_solver.jl:254-282)
# ---------------------------------------------------------------------------

def strang_step(q, w, dt, diffuse):
    """q <- e^{-w dt/2} D e^{-w dt/2} q (step_diffusion!)."""
    e = np.exp(-w * dt / 2.0)
    return e * diffuse(e * q)
