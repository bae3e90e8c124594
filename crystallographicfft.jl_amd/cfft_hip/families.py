"""Entry points of the reference's other forward/backward KRFFT families.

The reference carries several CPU factorizations of the same transform beside
its main pipeline: the recursive P3c + A8 butterfly (src/recursive_blocks.jl:
64-366), the sparse and selective-G0 variants (:514-697, :756-972), the G0-ASU
forward and backward (:1034-1323, src/recursive_blocks_backward.jl:135-548),
the general stride-L G0-ASU (src/recursive_blocks.jl:1378-1588), the fractal
trees (src/fractal_krfft.jl:678-1279, src/fractal_krfft_v3.jl:205-452) and the
staged Pm-3m decomposition (src/staged_krfft.jl:115-900).  Every one of them
computes the spectral-ASU values F(h) of a symmetric real field u on the full
N^3 grid (or, backward, u from them); they differ in how the work is ordered.
The four P3c forward families (recursive, sparse, selective, G0-ASU) run
their own algorithm on the device when the shifted ops hold the
body-diagonal 3-fold with an even translation (`cfft_hip.p3c`: FFTs of the
stride-4 sub-sub-grids of one offset per 3-fold orbit, then the family's
own tables as sparse stages, `cfft_p3c_*`); other groups, where the
reference's P3c formula does not hold (src/recursive_blocks.jl:958-972),
take the device pipeline of `DevicePlan` (stride-L subgrid, centring fold,
batched LDS FFTs, reconstruction), which is exact for every group.  The
remaining families (general G0-ASU, G0-ASU backward, fractal, staged Pm-3m)
run on the device pipeline.  The plan builders keep the reference's
preconditions and error behaviour, the executes return the same arrays.  The
staged Pm-3m plan returns F000 = FFT of u(2x) on the (N/2)^3 grid, the
quantity its reference counterpart returns.
"""
from __future__ import annotations

import warnings

import numpy as np

from . import _lib as L
from .plans import DevicePlan, SpectralIndexing, auto_L


def has_cubic_p3c_symmetry(ops) -> bool:
    """src/recursive_blocks.jl:974-982: an op whose R is the body-diagonal
    3-fold rotation [0 0 1; 1 0 0; 0 1 0]."""
    target = np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]])
    return any(np.array_equal(np.rint(np.asarray(op.Rm)).astype(int), target) for op in ops)


# sub-sub-grid FFT lengths of the P3c engine (csrc/p3c.hip)
_P3C_LENS = (2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64)


def p3c_applies(ops_shifted, n) -> bool:
    """The 3-fold rotation [0 0 1; 1 0 0; 0 1 0] with an even translation
    (mod n): the condition under which the P3c sub-sub-grid relation holds."""
    target = np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]])
    return any(np.array_equal(np.rint(np.asarray(op.Rm)).astype(int), target)
               and np.all(np.mod(np.asarray(op.tv), n) % 2 == 0) for op in ops_shifted)


def _check_p3c(spec: SpectralIndexing, what: str):
    """The assertions every P3c family opens with (3-D, N and N/2 even)."""
    N = tuple(int(n) for n in spec.N)
    if len(N) != 3:
        raise L.CFFTError(L.CFFT_ERR_INVALID, f"{what} currently supports 3D only")
    if any(n % 2 for n in N):
        raise L.CFFTError(L.CFFT_ERR_NOT_DIVISIBLE, "Grid size must be divisible by 2")
    if any((n // 2) % 2 for n in N):
        raise L.CFFTError(L.CFFT_ERR_NOT_DIVISIBLE, "M must be divisible by 2 for P3c")
    return N


class _ForwardFamilyPlan:
    """Shared state of the forward families: the device plan, the grid and
    the output buffer that the execute functions fill and return (an alias the
    next call overwrites, as in the reference)."""

    def __init__(self, spec: SpectralIndexing, ops_shifted):
        self.spec = spec
        self.grid_N = tuple(int(n) for n in spec.N)
        self.dev = DevicePlan(self.grid_N, ops_shifted, 0.0, None, L.PATH_AUTO)
        self.L = tuple(self.dev.L)
        self.subgrid_dims = tuple(self.dev.M)
        self.sub_sub_dims = tuple(m // 2 for m in self.dev.M)
        self.n_spec = self.dev.n_spec
        self.output_buffer = np.zeros(self.n_spec, dtype=np.complex128)

    def _forward(self, u):
        u = np.asarray(u, dtype=np.float64)
        if u.shape != self.grid_N:
            raise L.CFFTError(L.CFFT_ERR_SIZE, f"field shape {u.shape} does not match N = {self.grid_N}")
        if getattr(self, "p3c", None) is not None:
            self.p3c.forward(u, self.output_buffer)
            return self.output_buffer
        f0 = self.dev.full_to_sub(u)  # stride-L extraction on the device (k_full_to_sub)
        self.output_buffer[:] = self.dev.forward(f0)
        return self.output_buffer

    def _attach_p3c(self, kind, ops_shifted):
        """The family's own P3c tables and device engine when the ops allow
        them (cubic grid, the body-diagonal 3-fold with an even translation,
        an FFT length N/4 the device supports); otherwise the device pipeline
        stays in charge and `p3c` is None."""
        from . import p3c
        self.p3c = None
        self.p3c_tables = None
        n = self.grid_N[0]
        if not (self.grid_N == (n, n, n) and n % 4 == 0 and p3c_applies(ops_shifted, n)):
            return
        if n // 4 not in _P3C_LENS:
            return
        cls, stages, info = p3c.family_tables(kind, self.spec, ops_shifted)
        self.p3c = p3c.P3cEngine(n, cls, stages)
        self.p3c_tables = dict(info, representatives=len(cls[3]), nnz=self.p3c.nnz)


class RecursiveKRFFTPlan(_ForwardFamilyPlan):
    """src/recursive_blocks.jl:35-56: G0 at every M-grid point from the
    sub-sub-grid FFTs, then the A8 sum (`cfft_hip.p3c`)."""

    def __init__(self, spec, ops_shifted):
        super().__init__(spec, ops_shifted)
        self._attach_p3c("recursive", ops_shifted)


def plan_krfft_recursive(spec: SpectralIndexing, ops_shifted) -> RecursiveKRFFTPlan:
    """src/recursive_blocks.jl:64-144."""
    _check_p3c(spec, "Recursive KRFFT")
    return RecursiveKRFFTPlan(spec, ops_shifted)


def execute_recursive_krfft(plan: RecursiveKRFFTPlan, spec: SpectralIndexing, u):
    """src/recursive_blocks.jl:323-330 -> plan.output_buffer (n_spec)."""
    return plan._forward(u)


class SparseKRFFTPlan(_ForwardFamilyPlan):
    """src/recursive_blocks.jl:447-512: one CSR table from the sub-sub-grid
    FFT outputs straight to the spectral ASU (`cfft_hip.p3c`).  `centering`
    is recorded; the merged table keeps the entries the reference drops as
    reading identically-zero FFT outputs (they multiply zeros)."""

    def __init__(self, spec, ops_shifted, centering):
        super().__init__(spec, ops_shifted)
        self.centering = centering
        self._attach_p3c("sparse", ops_shifted)


def plan_krfft_sparse(spec: SpectralIndexing, ops_shifted, centering: str = "auto") -> SparseKRFFTPlan:
    """src/recursive_blocks.jl:514-629; centering :auto / :P / :I / :F."""
    N = _check_p3c(spec, "Sparse KRFFT")
    if centering == "auto":
        from .plans import detect_centering
        centering = detect_centering(ops_shifted, N).name
    return SparseKRFFTPlan(spec, ops_shifted, centering)


def execute_sparse_krfft(plan: SparseKRFFTPlan, spec: SpectralIndexing, u):
    """src/recursive_blocks.jl:631-670 -> plan.output_buffer."""
    return plan._forward(u)


class SelectiveG0Plan(_ForwardFamilyPlan):
    """src/recursive_blocks.jl:725-754: G0 only at the positions the A8 sum
    reads (`cfft_hip.p3c`)."""

    def __init__(self, spec, ops_shifted):
        super().__init__(spec, ops_shifted)
        self._attach_p3c("selective", ops_shifted)


def plan_krfft_selective(spec: SpectralIndexing, ops_shifted) -> SelectiveG0Plan:
    """src/recursive_blocks.jl:756-888."""
    _check_p3c(spec, "Selective G0 KRFFT")
    return SelectiveG0Plan(spec, ops_shifted)


def execute_selective_krfft(plan: SelectiveG0Plan, spec: SpectralIndexing, u):
    """src/recursive_blocks.jl:890-972 -> plan.output_buffer."""
    return plan._forward(u)


class G0ASUPlan(_ForwardFamilyPlan):
    """src/recursive_blocks.jl:1002-1032: G0 at the orbit representatives of
    the selective positions under the ops with even translations, the
    symmetry phases merged into the A8 weights (`cfft_hip.p3c`)."""

    def __init__(self, spec, ops_shifted):
        super().__init__(spec, ops_shifted)
        self._attach_p3c("g0asu", ops_shifted)


class GeneralG0ASUPlan(_ForwardFamilyPlan):
    """src/recursive_blocks.jl:1343-1376."""


def plan_krfft_g0asu_general(spec: SpectralIndexing, ops_shifted):
    """src/recursive_blocks.jl:1378-1552: stride-L G0-ASU for any group; with
    no subgrid reduction (L = 1) the reference warns and builds the sparse plan."""
    N = tuple(int(n) for n in spec.N)
    if len(N) != 3:
        raise L.CFFTError(L.CFFT_ERR_INVALID, "General G0 ASU currently supports 3D only")
    Lv = tuple(int(v) for v in auto_L(ops_shifted))
    if int(np.prod(Lv)) <= 1:
        warnings.warn("No sub-grid reduction possible (P1). Falling back to plan_krfft_sparse.")
        return plan_krfft_sparse(spec, ops_shifted)
    if any(n % l for n, l in zip(N, Lv)):
        raise L.CFFTError(L.CFFT_ERR_NOT_DIVISIBLE, f"Grid size N={N} not divisible by L={Lv}")
    return GeneralG0ASUPlan(spec, ops_shifted)


def plan_krfft_g0asu(spec: SpectralIndexing, ops_shifted):
    """src/recursive_blocks.jl:1034-1244: cubic groups (body-diagonal 3-fold
    axis) take the stride-4 G0-ASU plan, all others the general stride-L plan."""
    if len(spec.N) != 3:
        raise L.CFFTError(L.CFFT_ERR_INVALID, "G0 ASU KRFFT currently supports 3D only")
    if not has_cubic_p3c_symmetry(ops_shifted):
        return plan_krfft_g0asu_general(spec, ops_shifted)
    _check_p3c(spec, "G0 ASU KRFFT")
    return G0ASUPlan(spec, ops_shifted)


def execute_g0asu_krfft(plan, spec: SpectralIndexing, u):
    """src/recursive_blocks.jl:1246-1323 -> plan.output_buffer."""
    return plan._forward(u)


def execute_general_g0asu_krfft(plan, spec: SpectralIndexing, u):
    """src/recursive_blocks.jl:1554-1588 -> plan.output_buffer."""
    return plan._forward(u)


class G0ASUBackwardPlan:
    """src/recursive_blocks_backward.jl:95-124: F_spec -> u on the full grid."""

    def __init__(self, spec, ops_shifted):
        self.spec = spec
        self.grid_N = tuple(int(n) for n in spec.N)
        self.dev = DevicePlan(self.grid_N, ops_shifted, 0.0, None, L.PATH_AUTO)
        self.subgrid_dims = tuple(self.dev.M)
        self.sub_sub_dims = tuple(m // 2 for m in self.dev.M)
        self.n_spec = self.dev.n_spec


def plan_krfft_g0asu_backward(spec: SpectralIndexing, ops_shifted) -> G0ASUBackwardPlan:
    """src/recursive_blocks_backward.jl:135-458."""
    if len(spec.N) != 3:
        raise L.CFFTError(L.CFFT_ERR_INVALID, "G0 ASU backward currently supports 3D only")
    if not has_cubic_p3c_symmetry(ops_shifted):
        raise L.CFFTError(L.CFFT_ERR_UNSUPPORTED,
                          "G0 ASU backward requires cubic symmetry (P3c body-diagonal 3-fold rotation). "
                          "Non-cubic groups (tetragonal, orthorhombic, etc.) are not supported.")
    _check_p3c(spec, "G0 ASU backward")
    return G0ASUBackwardPlan(spec, ops_shifted)


def execute_g0asu_ikrfft(plan: G0ASUBackwardPlan, spec: SpectralIndexing, F_spec, u_out):
    """src/recursive_blocks_backward.jl:460-548: u_out (N^3, real) from the
    spectral-ASU values; the subgrid comes from the device backward transform
    and the rest of the grid from the symmetry fill map (k_sub_to_full)."""
    F_spec = np.asarray(F_spec, dtype=np.complex128)
    if F_spec.shape != (plan.n_spec,):
        raise L.CFFTError(L.CFFT_ERR_SIZE, f"F_spec has {F_spec.size} entries, the plan {plan.n_spec}")
    if tuple(np.shape(u_out)) != plan.grid_N:
        raise L.CFFTError(L.CFFT_ERR_SIZE, f"u_out shape {np.shape(u_out)} does not match N = {plan.grid_N}")
    f0 = np.real(plan.dev.backward(F_spec))
    u_out[...] = plan.dev.sub_to_full(f0)
    return u_out


def shift_ops_half_grid(ops, N):
    """src/fractal_krfft.jl:265-287: the ops of the grid sampled half a cell
    off the origin, t' = t + (I - R) b mod N with b = 1/2 on every axis.  The
    fractal plans take the group's own ops and expect a field symmetric under
    these (a mirror then maps even-parity sectors onto odd ones).  Signed
    permutation matrices give integer t'; a fractional one (some hexagonal
    settings) is refused."""
    from .symmetry_ops import SymOp
    Nv = np.asarray(N, dtype=np.int64)
    out = []
    for op in ops:
        R = np.asarray(op.Rm, dtype=np.int64)
        t2 = 2 * np.asarray(op.tv, dtype=np.int64) + (np.eye(len(Nv), dtype=np.int64) - R).sum(axis=1)
        if np.any(t2 % 2):
            raise L.CFFTError(L.CFFT_ERR_UNSUPPORTED,
                              "the half-grid shift leaves a fractional translation for this group")
        out.append(SymOp.make(R, np.mod(t2 // 2, Nv)))
    return out


class FractalCFFTPlan(_ForwardFamilyPlan):
    """src/fractal_krfft.jl:670-676 (also the v2 and v3 plans), planned from
    the group's own ops, for a field symmetric under `shift_ops_half_grid(ops)`
    (the reference's b = 1/2 sampling, src/fractal_krfft.jl:300-304, and its
    tests' `make_symmetric`).  On a cubic grid the decimation tree of
    `cfft_hip.p3c.fractal_tables` over those shifted ops — leaves at
    stride 2^depth, one per orbit of the offsets under the group, split by
    parity while that still reduces the leaves, and only the values the
    spectral ASU needs formed at each level (the v2 plan's needed sets,
    src/fractal_krfft.jl:1172-1208) — run on the device (cfft_p3c_*);
    `fractal` then summarises the tree.  With no level that reduces, or a
    leaf length the device does not transform, the device pipeline."""

    def __init__(self, spec, ops, depth=None):
        ops_half = shift_ops_half_grid(ops, tuple(int(n) for n in spec.N))
        super().__init__(spec, ops_half)
        from . import p3c
        self.ops_half = ops_half
        # the values are returned at the caller's spectral ASU (that of `ops`)
        self.n_spec = len(spec.points)
        self.output_buffer = np.zeros(self.n_spec, dtype=np.complex128)
        self.p3c = None
        self.fractal = None
        n = self.grid_N[0]
        if self.grid_N == (n, n, n):
            if depth is None:
                depth = p3c.fractal_depth(ops_half, n, lens=_P3C_LENS)
            if depth >= 1:
                stride, offs, stages, info = p3c.fractal_tables(spec, ops_half, depth)
                self.p3c = p3c.P3cEngine(n, None, stages, stride=stride, offsets=offs)
                self.fractal = dict(info, stride=stride, nnz=self.p3c.nnz)
                return
        self._map_to_device_asu(spec, ops_half)

    def _map_to_device_asu(self, spec, ops_half):
        """Without the tree the device pipeline runs on the spectral ASU of
        the shifted ops; each requested h is read there through an op,
        F(h) = exp(2 pi i (R^-T h).t / N) F(R^-T h mod N), or is 0 when no op
        reaches that ASU (h extinct for the shifted ops)."""
        N = np.asarray(self.grid_N, dtype=np.int64)
        dev_idx = self.dev.tables()[0][:, :len(N)]  # the points the device forward returns, in its order
        lin = lambda h: np.ravel_multi_index(tuple(np.asarray(h).T), tuple(int(v) for v in N), order="F")
        pos = {int(k): i for i, k in enumerate(np.atleast_1d(lin(dev_idx)))}
        H = np.asarray(spec.idx, dtype=np.int64)
        idx = np.full(len(H), -1, dtype=np.int64)
        phase = np.zeros(len(H), dtype=np.complex128)
        for op in ops_half:
            Rit = np.rint(np.linalg.inv(op.Rm).T).astype(np.int64)
            g = np.mod(H @ Rit.T, N)
            for i in np.nonzero(idx < 0)[0]:
                j = pos.get(int(lin(g[i])))
                if j is not None:
                    idx[i] = j
                    phase[i] = np.exp(2j * np.pi * np.sum(g[i] * op.tv / N))
        self._map_idx, self._map_phase = idx, phase

    def _forward(self, u):
        if self.p3c is not None:
            return super()._forward(u)
        u = np.asarray(u, dtype=np.float64)
        if u.shape != self.grid_N:
            raise L.CFFTError(L.CFFT_ERR_SIZE, f"field shape {u.shape} does not match N = {self.grid_N}")
        F = self.dev.forward(self.dev.full_to_sub(u))
        ok = self._map_idx >= 0
        self.output_buffer[:] = 0
        self.output_buffer[ok] = self._map_phase[ok] * F[self._map_idx[ok]]
        return self.output_buffer


def plan_fractal_krfft(spec: SpectralIndexing, ops, depth=None) -> FractalCFFTPlan:
    """src/fractal_krfft.jl:678-722: `ops` are the group's ops (not shifted;
    the plan applies the half-grid shift); `depth` overrides the tree depth."""
    return FractalCFFTPlan(spec, ops, depth)


def execute_fractal_krfft(plan: FractalCFFTPlan, u):
    """src/fractal_krfft.jl:869-901 -> plan.output_buffer."""
    return plan._forward(u)


plan_fractal_krfft_v2 = plan_fractal_krfft    # src/fractal_krfft.jl:1021-1170
execute_fractal_krfft_v2 = execute_fractal_krfft  # src/fractal_krfft.jl:1228-1279
plan_fractal_krfft_v3 = plan_fractal_krfft    # src/fractal_krfft_v3.jl:205-379
execute_fractal_krfft_v3 = execute_fractal_krfft  # src/fractal_krfft_v3.jl:381-452


class StagedPm3mPlan:
    """src/staged_krfft.jl:115-128: Pm-3m (#221, magic shift 1/2 pixel) on a
    cubic grid; F000 = FFT of f000(x) = u(2x) on the (N/2)^3 grid."""

    def __init__(self, N_full, use_yx=False):
        from .plans import find_optimal_shift
        from .symmetry_ops import get_ops
        self.N_full = tuple(int(n) for n in N_full)
        self.N_half = tuple(n // 2 for n in self.N_full)
        self.use_yx = bool(use_yx)
        ops = get_ops(221, 3, self.N_full)
        _, ops_s = find_optimal_shift(ops, self.N_full)
        self.dev = DevicePlan(self.N_full, ops_s, 0.0, None, L.PATH_M2)
        if tuple(self.dev.L) != (2, 2, 2):
            raise L.CFFTError(L.CFFT_ERR_UNSUPPORTED, "Pm-3m subgrid stride is not 2")
        self.a8_fft = np.zeros(self.N_half, dtype=np.complex128)
        # F000 from the P3c relation: 4 sub-sub-grid FFTs of (N/4)^3 and a
        # sparse stage of 8 entries per M-grid point, when N/4 is a length
        # the device transforms; otherwise the device pipeline's fold
        self.p3c = None
        n = self.N_full[0]
        if n % 4 == 0 and n // 4 in _P3C_LENS and p3c_applies(ops_s, n):
            from . import p3c
            cls, stages = p3c.g0_table(ops_s, n)
            self.p3c = p3c.P3cEngine(n, cls, stages)


def plan_staged_pm3m(N_full, use_yx: bool = False) -> StagedPm3mPlan:
    """src/staged_krfft.jl:115-128."""
    N_full = tuple(int(n) for n in N_full)
    if any(n % 2 for n in N_full):
        raise L.CFFTError(L.CFFT_ERR_NOT_DIVISIBLE, "Grid dimensions must be even")
    if not (len(N_full) == 3 and N_full[0] == N_full[1] == N_full[2]):
        raise L.CFFTError(L.CFFT_ERR_INVALID, "Must be cubic")
    return StagedPm3mPlan(N_full, use_yx)


def execute_staged(plan: StagedPm3mPlan, u):
    """src/staged_krfft.jl:562-567 -> F000, the FFT of u(2x) on the (N/2)^3
    grid: through the P3c engine (4 sub-sub-grid FFTs and one sparse stage)
    when N/4 is a supported length, else the 8-channel fold + channel FFTs +
    assembly of the device pipeline."""
    u = np.asarray(u, dtype=np.float64)
    if u.shape != plan.N_full:
        raise L.CFFTError(L.CFFT_ERR_SIZE, f"field shape {u.shape} does not match N = {plan.N_full}")
    if plan.p3c is not None:
        plan.a8_fft[...] = plan.p3c.forward(u).reshape(plan.N_half, order="F")
        return plan.a8_fft
    f0 = plan.dev.full_to_sub(u)
    plan.a8_fft[...] = plan.dev.forward_g0(f0)
    return plan.a8_fft


plan_staged_pm3m_full = plan_staged_pm3m  # src/staged_krfft.jl:863-875
execute_staged_full = execute_staged      # src/staged_krfft.jl:883-888


def lookup_full_spectrum(plan: StagedPm3mPlan, F000, hx, hy, hz):
    """src/staged_krfft.jl:585-617: F(h) of the full grid from F000 for the
    Pm-3m ops with the 1/2-pixel shift.  The 8 stride-2 subgrid spectra are F000
    at sign-flipped arguments times exp(2 pi i (flipped components)/N_half);
    F(h) = sum_eps exp(-2 pi i eps.h/N) F_eps(h mod N_half).  Accepts arrays of
    frequencies (vectorised)."""
    F000 = np.asarray(F000)
    Nh = plan.N_half[0]
    Nf = plan.N_full[0]
    h = [np.asarray(v, dtype=np.int64) for v in (hx, hy, hz)]
    i = [np.mod(v, Nh) for v in h]
    ph = [np.exp(-2j * np.pi * v / Nf) for v in h]
    out = 0.0
    for e0 in (0, 1):
        for e1 in (0, 1):
            for e2 in (0, 1):
                eps = (e0, e1, e2)
                idx = tuple(np.mod(-ii if e else ii, Nh) for ii, e in zip(i, eps))
                lin = sum(ii for ii, e in zip(i, eps) if e)
                sub = np.exp(2j * np.pi * lin / Nh) * F000[idx]
                w = 1.0
                for d in range(3):
                    if eps[d]:
                        w = w * ph[d]
                out = out + w * sub
    return out


def compute_spectral_asu(result, plan: StagedPm3mPlan, F000, spec: SpectralIndexing):
    """src/staged_krfft.jl:726-738: result[i] = F at the i-th spectral-ASU point."""
    n = len(spec)
    if len(result) < n:
        raise L.CFFTError(L.CFFT_ERR_SIZE, "result shorter than the spectral ASU")
    h = np.asarray(spec.idx, dtype=np.int64)
    result[:n] = lookup_full_spectrum(plan, F000, h[:, 0], h[:, 1], h[:, 2])
    return result
