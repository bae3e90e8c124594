"""Plan-time maps and device plans over the C ABI.

Python mirror of the reference's plan/execute API for the hot path.  Julia's
mutating ``f!`` functions keep their names without the ``!`` and mutate their
array arguments in place; Julia's 1-based point indices become 0-based.

Reference counterparts:
  find_optimal_shift          src/asu.jl:57-123
  calc_spectral_asu           src/spectral_indexing.jl:22-176
  get_k_vector                src/spectral_indexing.jl:195-238
  auto_L                      src/krfft.jl:225-279
  plan_krfft / fft_reconstruct / pack_stride / execute_krfft
                              src/krfft.jl:301-389, 554-621
  plan_m2_backward / execute_m2_backward
                              src/krfft.jl:1101-1277
  plan_m2_scft / execute_m2_scft / update_m2_kernel
                              src/q_fused_krfft.jl:948-1052
  plan_krfft_centered / execute_centered_krfft / fft_reconstruct_centered
                              src/centering_fold.jl:328-431,
                              src/centering_prefold.jl:162-257 (pre-fold)
  plan_centered_ikrfft / execute_centered_ikrfft / ifft_unrecon_centered
                              src/centering_fold.jl:630-775
  plan_centered_scft / execute_centered_scft / update_kernel
                              src/centering_fold.jl:868-953
  plan_m2_q / execute_m2_q    src/q_fused_krfft.jl:83-169, 553-576
  plan_m7_scft / execute_m7_scft
                              src/q_fused_krfft.jl:826-895
  subgrid_to_fullgrid / fullgrid_to_subgrid
                              src/q_fused_krfft.jl:681-719
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .symmetry_ops import (CenteringType, CentP, SymOp, get_ops, ops_arrays,
                           ops_from_arrays, pack_ops)


def _N3(N):
    N = tuple(int(n) for n in N)
    if len(N) == 2:
        return N + (1,), 2
    if len(N) == 3:
        return N, 3
    raise ValueError("N must have 2 or 3 entries")


def _pack_any(ops, dim):
    """Pack D-dimensional ops (D = 2 or 3) into the 3-D ABI layout."""
    R, t = ops_arrays(ops)
    n = R.shape[0]
    R3 = np.zeros((n, 3, 3), dtype=np.int64)
    t3 = np.zeros((n, 3), dtype=np.int64)
    R3[:, :dim, :dim] = R
    for d in range(dim, 3):
        R3[:, d, d] = 1
    t3[:, :dim] = t
    return np.ascontiguousarray(np.concatenate([R3.reshape(n, 9), t3], axis=1).astype(np.int32))


def _unpack_any(arr, dim):
    arr = np.asarray(arr, dtype=np.int64).reshape(-1, 12)
    R = arr[:, :9].reshape(-1, 3, 3)[:, :dim, :dim]
    t = arr[:, 9:9 + dim]
    return ops_from_arrays(R, t)


# ---------------------------------------------------------------------------
# plan-time maps
# ---------------------------------------------------------------------------

def find_optimal_shift(ops, N):
    """src/asu.jl:57-123 -> (shift b/N, shifted ops)."""
    N3, dim = _N3(N)
    packed = _pack_any(ops, dim)
    out = np.zeros_like(packed)
    shift = np.zeros(3)
    L.check(L.lib().cfft_find_optimal_shift(dim, L.i32p(np.array(N3, dtype=np.int32)),
                                            L.i32p(packed), len(ops),
                                            shift.ctypes.data_as(L.PD), L.i32p(out)))
    return shift[:dim].copy(), _unpack_any(out, dim)


def count_special_positions(N, ops) -> int:
    """src/asu.jl:148-214."""
    N3, dim = _N3(N)
    packed = _pack_any(ops, dim)
    c = C.c_int64(0)
    L.check(L.lib().cfft_count_special_positions(dim, L.i32p(np.array(N3, dtype=np.int32)),
                                                 L.i32p(packed), len(ops), C.byref(c)))
    return int(c.value)


@dataclass
class ASUPoint:
    """src/asu.jl:23-27."""
    idx: np.ndarray
    depth: np.ndarray
    multiplicity: int


@dataclass
class SpectralIndexing:
    """src/spectral_indexing.jl:9-13: points (0-based frequencies in [0,N),
    lexicographic order), the reciprocal ops and the grid."""
    idx: np.ndarray          # (n_spec, D)
    mult: np.ndarray         # (n_spec,)
    ops: list
    N: tuple

    @property
    def points(self):
        D = self.idx.shape[1]
        return [ASUPoint(self.idx[i].copy(), np.zeros(D, dtype=np.int64), int(self.mult[i]))
                for i in range(self.idx.shape[0])]

    def __len__(self):
        return int(self.idx.shape[0])


def calc_spectral_asu(ops, dim: int, N) -> SpectralIndexing:
    """src/spectral_indexing.jl:22-176 (also accepts (sg_num, dim, N))."""
    if isinstance(ops, (int, np.integer)):
        ops = get_ops(int(ops), dim, N)
    from .symmetry_ops import dual_ops
    N3, d = _N3(N)
    if d != dim:
        raise ValueError("dim does not match N")
    packed = _pack_any(ops, dim)
    n = C.c_int64(0)
    Np = np.array(N3, dtype=np.int32)
    L.check(L.lib().cfft_calc_spectral_asu(dim, L.i32p(Np), L.i32p(packed), len(ops), None, None,
                                           0, C.byref(n)))
    idx = np.zeros((n.value, 3), dtype=np.int32)
    mult = np.zeros(n.value, dtype=np.int32)
    L.check(L.lib().cfft_calc_spectral_asu(dim, L.i32p(Np), L.i32p(packed), len(ops), L.i32p(idx),
                                           L.i32p(mult), n.value, C.byref(n)))
    return SpectralIndexing(idx[:, :dim].astype(np.int64), mult.astype(np.int64),
                            dual_ops(ops), tuple(N))


def get_k_vector(spec: SpectralIndexing, i: int) -> np.ndarray:
    """src/spectral_indexing.jl:195-238 (0-based point index)."""
    n = len(spec)
    if i < 0 or i >= n:
        raise IndexError(f"Index {i} out of bounds for SpectralIndexing with {n} points")
    raw = spec.idx[i]
    return np.array([v - nn if v >= nn - nn // 2 else v for v, nn in zip(raw, spec.N)],
                    dtype=np.int64)


def auto_L(ops_shifted) -> np.ndarray:
    """src/krfft.jl:225-279."""
    dim = len(ops_shifted[0].t)
    packed = _pack_any(ops_shifted, dim)
    Lout = np.zeros(3, dtype=np.int32)
    L.check(L.lib().cfft_auto_L(dim, L.i32p(packed), len(ops_shifted), L.i32p(Lout)))
    return Lout[:dim].astype(np.int64)


def select_rep_ops(ops_shifted, Lv, N) -> np.ndarray:
    """src/krfft.jl:958-987 -> 0-based op indices, cosets first-dimension fastest."""
    N3, dim = _N3(N)
    packed = _pack_any(ops_shifted, dim)
    L3 = np.ones(3, dtype=np.int32)
    L3[:dim] = np.asarray(Lv)
    rep = np.zeros(int(np.prod(L3)), dtype=np.int32)
    L.check(L.lib().cfft_select_rep_ops(dim, L.i32p(np.array(N3, dtype=np.int32)), L.i32p(packed),
                                        len(ops_shifted), L.i32p(L3), L.i32p(rep)))
    return rep.astype(np.int64)


def detect_centering(ops, N) -> CenteringType:
    N3, dim = _N3(N)
    packed = _pack_any(ops, dim)
    c = C.c_int32(0)
    L.check(L.lib().cfft_detect_centering(dim, L.i32p(np.array(N3, dtype=np.int32)),
                                          L.i32p(packed), len(ops), C.byref(c)))
    return CenteringType(c.value)


def build_fill_map(ops_shifted, Lv, N) -> np.ndarray:
    """src/q_fused_krfft.jl:733-783 -> 0-based M-grid index per full-grid point
    (flat, column-major)."""
    N3, dim = _N3(N)
    packed = _pack_any(ops_shifted, dim)
    L3 = np.ones(3, dtype=np.int32)
    L3[:dim] = np.asarray(Lv)
    out = np.zeros(int(np.prod(N3)), dtype=np.int32)
    L.check(L.lib().cfft_fill_map(dim, L.i32p(np.array(N3, dtype=np.int32)), L.i32p(packed),
                                  len(ops_shifted), L.i32p(L3), L.i32p(out)))
    return out.astype(np.int64)


# ---------------------------------------------------------------------------
# device plans
# ---------------------------------------------------------------------------

_I = np.eye(3)


def qs_eligible(N, ops_shifted, lattice) -> bool:
    """Whether path QS (plane-pair kernel) can plan this case: cubic grid,
    L = (2,2,2), an orthogonal cell and 8 diagonal representative ops with
    distinct sign patterns (host-side check, no GPU)."""
    N = tuple(int(n) for n in N)
    if len(N) != 3 or len(set(N)) != 1:
        return False
    lat = np.asarray(lattice, dtype=np.float64).reshape(3, 3)
    if np.count_nonzero(lat - np.diag(np.diag(lat))):
        return False
    Lv = auto_L(ops_shifted)
    if tuple(int(v) for v in Lv) != (2, 2, 2) or N[0] % 4:
        return False
    patterns = set()
    for r in select_rep_ops(ops_shifted, Lv, N):
        R = np.asarray(ops_shifted[int(r)].Rm)
        if np.count_nonzero(R - np.diag(np.diag(R))) or not np.all(np.abs(np.diag(R)) == 1):
            return False
        patterns.add(tuple(int(v) for v in np.diag(R)))
    return len(patterns) == 8


class DevicePlan:
    """One C plan (device tables + buffers) for a 3-D grid, a shifted op list,
    a pipeline path, a precision and a batch size."""

    def __init__(self, N, ops_shifted, ds=0.0, lattice=None, path=L.PATH_AUTO,
                 precision=L.FP64, batch=1):
        self.N = tuple(int(n) for n in N)
        if len(self.N) != 3:
            raise ValueError("device plans are 3-D")
        self.ops = list(ops_shifted)
        self.lattice = np.asarray(_I if lattice is None else lattice, dtype=np.float64)
        lat = np.ascontiguousarray(self.lattice.reshape(3, 3).T.ravel())  # column-major
        packed = pack_ops(self.ops)
        h = C.c_void_p()
        L.check(L.lib().cfft_plan_create(L.i32p(np.array(self.N, dtype=np.int32)), L.i32p(packed),
                                         len(self.ops), float(ds), lat.ctypes.data_as(L.PD),
                                         int(path), int(precision), int(batch), C.byref(h)))
        self._h = h
        info = np.zeros(18, dtype=np.int32)
        L.check(L.lib().cfft_plan_info(self._h, L.i32p(info)))
        self.M = tuple(int(v) for v in info[0:3])
        self.L = tuple(int(v) for v in info[3:6])
        self.n_spec = int(info[6])
        self.n_ch = int(info[7])
        self.centering = CenteringType(int(info[8]))
        self.d = int(info[9])
        self.path = int(info[10])
        self.options = {}  # plan options set through set_option
        self.precision = int(info[11])
        self.batch = int(info[12])
        self.H = tuple(int(v) for v in info[13:16])
        self.n_rep = int(info[16])
        self.n_alive = int(info[17])
        self.ds = float(ds)

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.lib().cfft_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    def tables(self):
        idx = np.zeros((self.n_spec, 3), dtype=np.int32)
        mult = np.zeros(self.n_spec, dtype=np.int32)
        K = np.zeros(self.n_spec)
        rep = np.zeros(self.d, dtype=np.int32)
        L.check(L.lib().cfft_plan_tables(self._h, L.i32p(idx), L.i32p(mult), K.ctypes.data_as(L.PD),
                                         L.i32p(rep)))
        return idx.astype(np.int64), mult.astype(np.int64), K, rep.astype(np.int64)

    # -- shapes ------------------------------------------------------------
    def _f0_in(self, f0):
        a = np.asarray(f0, dtype=np.float64)
        want = self.M if self.batch == 1 else (self.batch,) + self.M
        if a.shape != want:
            raise ValueError(f"f0 size {a.shape} does not match plan subgrid dims {want}. "
                             f"Use L={self.L} stride subgrid extraction.")
        return self._to_abi(a)

    def _to_abi(self, a):
        """numpy (batch, M0, M1, M2) or (M0, M1, M2) -> flat column-major per field."""
        if a.ndim == 3:
            return np.asfortranarray(a).ravel(order="F").copy()
        return np.concatenate([np.asarray(a[b]).ravel(order="F") for b in range(a.shape[0])])

    def _from_abi(self, flat, shape3):
        n = int(np.prod(shape3))
        if self.batch == 1:
            return flat[:n].reshape(shape3, order="F")
        return np.stack([flat[b * n:(b + 1) * n].reshape(shape3, order="F") for b in range(self.batch)])

    # -- execute (host arrays) ----------------------------------------------
    def forward(self, f0):
        x = self._f0_in(f0)
        F = np.zeros(2 * self.n_spec * self.batch)
        L.check(L.lib().cfft_forward(self._h, L.f64p(x), L.f64p(F)))
        Fc = F.view(np.complex128)
        return Fc.copy() if self.batch == 1 else Fc.reshape(self.batch, self.n_spec).copy()

    def backward(self, F_spec):
        F = np.ascontiguousarray(np.asarray(F_spec, dtype=np.complex128).reshape(-1))
        if F.size != self.n_spec * self.batch:
            raise ValueError(f"F_spec has {F.size} entries, plan expects {self.n_spec * self.batch}")
        out = np.zeros(int(np.prod(self.M)) * self.batch)
        L.check(L.lib().cfft_backward(self._h, F.view(np.float64).ctypes.data_as(L.PD), L.f64p(out)))
        return self._from_abi(out, self.M)

    def forward_g0(self, f0):
        x = self._f0_in(f0)
        G = np.zeros(2 * int(np.prod(self.M)) * self.batch)
        L.check(L.lib().cfft_forward_g0(self._h, L.f64p(x), L.f64p(G)))
        return self._from_abi(G.view(np.complex128), self.M)

    def diffuse(self, f0):
        x = self._f0_in(f0)
        L.check(L.lib().cfft_diffuse(self._h, L.f64p(x)))
        return self._from_abi(x, self.M)

    def step(self, q, w, dt, nsteps=1):
        x = self._f0_in(q)
        wv = self._f0_in(w)
        L.check(L.lib().cfft_step(self._h, L.f64p(x), L.f64p(wv), float(dt), int(nsteps)))
        return self._from_abi(x, self.M)

    def update_kernel(self, ds, lattice=None):
        lat = None
        if lattice is not None:
            self.lattice = np.asarray(lattice, dtype=np.float64)
        lat = np.ascontiguousarray(self.lattice.reshape(3, 3).T.ravel())
        L.check(L.lib().cfft_update_kernel(self._h, float(ds), lat.ctypes.data_as(L.PD)))
        self.ds = float(ds)

    def sub_to_full(self, f0):
        x = np.asfortranarray(np.asarray(f0, dtype=np.float64)).ravel(order="F").copy()
        if x.size != int(np.prod(self.M)):
            raise ValueError("f0 size does not match the subgrid")
        u = np.zeros(int(np.prod(self.N)))
        L.check(L.lib().cfft_sub_to_full(self._h, L.f64p(x), L.f64p(u)))
        return u.reshape(self.N, order="F")

    def full_to_sub(self, u):
        x = np.asfortranarray(np.asarray(u, dtype=np.float64)).ravel(order="F").copy()
        if x.size != int(np.prod(self.N)):
            raise ValueError("full-grid field size does not match N")
        f0 = np.zeros(int(np.prod(self.M)))
        L.check(L.lib().cfft_full_to_sub(self._h, L.f64p(x), L.f64p(f0)))
        return f0.reshape(self.M, order="F")

    # -- execute (device pointers) ------------------------------------------
    def diffuse_dev(self, ptr, stream=None):
        L.check(L.lib().cfft_diffuse_dev(self._h, C.c_void_p(ptr), C.c_void_p(stream)))

    def forward_dev(self, f0_ptr, F_ptr, stream=None):
        L.check(L.lib().cfft_forward_dev(self._h, C.c_void_p(f0_ptr), C.c_void_p(F_ptr),
                                         C.c_void_p(stream)))

    def backward_dev(self, F_ptr, f0_ptr, stream=None):
        L.check(L.lib().cfft_backward_dev(self._h, C.c_void_p(F_ptr), C.c_void_p(f0_ptr),
                                          C.c_void_p(stream)))

    def step_dev(self, q_ptr, E_ptr, nsteps=1, stream=None):
        L.check(L.lib().cfft_step_dev(self._h, C.c_void_p(q_ptr), C.c_void_p(E_ptr), int(nsteps),
                                      C.c_void_p(stream)))

    def exp_dev(self, w_ptr, E_ptr, factor, n, stream=None):
        L.check(L.lib().cfft_exp_dev(self._h, C.c_void_p(w_ptr), C.c_void_p(E_ptr), float(factor),
                                     int(n), C.c_void_p(stream)))

    def sync(self):
        L.check(L.lib().cfft_sync(self._h))

    def set_option(self, key, value):
        L.check(L.lib().cfft_plan_set_option(self._h, int(key), int(value)))
        self.options[int(key)] = int(value)


# ---------------------------------------------------------------------------
# reference-shaped plans
# ---------------------------------------------------------------------------

def _sg_setup(N, sg_num, dim):
    ops = get_ops(sg_num, dim, N)
    shift, ops_s = find_optimal_shift(ops, N)
    return ops, shift, ops_s


class GeneralCFFTPlan:
    """plan_krfft(spec, ops_shifted), src/krfft.jl:301-389: forward KRFFT on
    the stride-L subgrid.  ``input_buffer`` holds f0 (M-grid), ``output_buffer``
    the spectral-ASU coefficients after ``fft_reconstruct``."""

    def __init__(self, spec, ops_shifted, path=L.PATH_M2, precision=L.FP64):
        self.spec = spec
        self.dev = DevicePlan(spec.N, ops_shifted, 0.0, None, path, precision)
        self.grid_N = tuple(spec.N)
        self.subgrid_dims = self.dev.M
        self.L_factors = [np.array(self.dev.L)]
        self.n_ops = self.dev.d
        self.input_buffer = np.zeros(self.dev.M, dtype=np.complex128, order="F")
        self.output_buffer = np.zeros(self.dev.n_spec, dtype=np.complex128)


def plan_krfft(*args) -> GeneralCFFTPlan:
    """``plan_krfft(spec, ops_shifted)`` (src/krfft.jl:301-389), or
    ``plan_krfft(real_asu, spec, ops)`` (src/krfft.jl:119-...) with the real
    ASU of ``pack_asu_interleaved``: its one block is the stride-L subgrid,
    and the plan is the same as the first form's when that L is the one
    ``auto_L`` finds for the ops (otherwise CFFT_ERR_UNSUPPORTED)."""
    if len(args) == 2:
        return GeneralCFFTPlan(args[0], args[1])
    if len(args) != 3:
        raise TypeError("plan_krfft(spec, ops_shifted) or plan_krfft(real_asu, spec, ops)")
    real_asu, spec, ops = args
    blocks = [b for d in sorted(real_asu.dim_blocks) for b in real_asu.dim_blocks[d]]
    N = tuple(int(n) for n in spec.N)
    if len(blocks) != 1:
        raise L.CFFTError(L.CFFT_ERR_UNSUPPORTED, "the real ASU must be the one stride-L subgrid block")
    dims = blocks[0].data.shape
    if any(N[d] % dims[d] for d in range(len(N))):
        raise L.CFFTError(L.CFFT_ERR_INVALID, "Block size must divide Global size for Cooley-Tukey decomposition")
    Lb = tuple(N[d] // dims[d] for d in range(len(N)))
    if Lb != tuple(int(v) for v in auto_L(ops))[:len(N)]:
        raise L.CFFTError(L.CFFT_ERR_UNSUPPORTED,
                          f"subgrid stride {Lb} differs from auto_L of the ops {tuple(auto_L(ops))}")
    return GeneralCFFTPlan(spec, ops)


def fft_reconstruct(plan: GeneralCFFTPlan):
    """src/krfft.jl:554-562 (the real part of input_buffer is the field)."""
    plan.output_buffer[:] = plan.dev.forward(np.real(plan.input_buffer))
    return plan.output_buffer


fast_reconstruct = fft_reconstruct


def pack_stride(plan: GeneralCFFTPlan, u):
    """src/krfft.jl:571-600: input_buffer = u[::2, ::2, ::2]."""
    u = np.asarray(u, dtype=np.float64)
    M = plan.subgrid_dims
    plan.input_buffer[...] = u[:2 * M[0]:2, :2 * M[1]:2, :2 * M[2]:2]


def execute_krfft(plan: GeneralCFFTPlan, u):
    """src/krfft.jl:608-621 with its documented semantics: pack, FFT, reconstruct."""
    pack_stride(plan, u)
    return fft_reconstruct(plan)


class M2BackwardPlan:
    """src/krfft.jl:928-948."""

    def __init__(self, spec, ops_shifted, path=L.PATH_M2):
        self.dev = DevicePlan(spec.N, ops_shifted, 0.0, None, path)
        self.n_spec = self.dev.n_spec
        self.d = self.dev.d
        self.L = self.dev.L
        self.subgrid_dims = self.dev.M
        self.grid_N = tuple(spec.N)
        self.f0_buf = np.zeros(self.dev.M, dtype=np.complex128, order="F")


def plan_m2_backward(spec, ops_shifted) -> M2BackwardPlan:
    return M2BackwardPlan(spec, ops_shifted)


def execute_m2_backward(bplan: M2BackwardPlan, F_spec):
    """src/krfft.jl:1269-1277 -> bplan.f0_buf (complex M-grid; the field is real)."""
    bplan.f0_buf[...] = bplan.dev.backward(F_spec)
    return bplan.f0_buf


class _SCFTPlanBase:
    def _run(self, f0):
        f0 = np.asarray(f0)
        if f0.shape != self.dev.M:
            raise ValueError(f"f0 size {f0.shape} does not match plan subgrid dims {self.dev.M}. "
                             f"Use L={self.dev.L} stride subgrid extraction.")
        out = self.dev.diffuse(f0)
        f0[...] = out
        return f0


class M2SCFTPlan(_SCFTPlanBase):
    """src/q_fused_krfft.jl:922-946."""

    def __init__(self, N, ops_shifted, ds, lattice, spec=None, precision=L.FP64):
        self.dev = DevicePlan(N, ops_shifted, ds, lattice, L.PATH_M2, precision)
        self.n_spec = self.dev.n_spec
        _, _, K, _ = self.dev.tables()
        self.K_spec = K


def plan_m2_scft(N, sg_num, dim, ds, lattice, precision=L.FP64) -> M2SCFTPlan:
    """src/q_fused_krfft.jl:948-986."""
    if len(N) != dim:
        raise ValueError("len(N) must equal dim")
    _, _, ops_s = _sg_setup(N, sg_num, dim)
    return M2SCFTPlan(N, ops_s, ds, lattice, precision=precision)


def execute_m2_scft(plan: M2SCFTPlan, f0):
    """src/q_fused_krfft.jl:989-1026 (in place)."""
    return plan._run(f0)


def update_m2_kernel(plan: M2SCFTPlan, N, sg_num, dim, ds_new, lattice):
    """src/q_fused_krfft.jl:1036-1052."""
    plan.dev.update_kernel(ds_new, lattice)
    plan.K_spec = plan.dev.tables()[2]


class CenteredKRFFTPlan:
    """src/centering_fold.jl:311-316: centred forward KRFFT (fold + channel
    FFTs + reconstruction); ``f0_buffer`` holds the real stride-2 subgrid."""

    def __init__(self, spec, ops_shifted, precision=L.FP64):
        self.spec = spec
        self.dev = DevicePlan(spec.N, ops_shifted, 0.0, None, L.PATH_M7, precision)
        self.f0_buffer = np.zeros(self.dev.M, order="F")
        self.output_buffer = np.zeros(self.dev.n_spec, dtype=np.complex128)
        self.centering = self.dev.centering
        self.n_channels = self.dev.n_ch
        self.subgrid_dims = self.dev.M


class CenteringPreFoldKRFFTPlan:
    """src/centering_prefold.jl:162-205: full-grid centring pre-fold into
    4 (I) / 2 (F, C, A) channels on the halved grid, one FFT per channel and
    a merge onto the spectral ASU of ``ops``.  (The reference names this
    struct CenteredKRFFTPlan as well, in a second file; here it has a name of
    its own so that both centred forward plans can coexist.)"""

    def __init__(self, ops, N):
        N3, dim = _N3(N)
        if dim != 3:
            raise L.CFFTError(L.CFFT_ERR_INVALID, "the centring pre-fold is 3-D")
        packed = _pack_any(ops, 3)
        h = C.c_void_p()
        L.check(L.lib().cfft_prefold_plan_create(L.i32p(np.array(N3, dtype=np.int32)),
                                                 L.i32p(packed), len(ops), C.byref(h)))
        self._h = h
        info = np.zeros(6, dtype=np.int32)
        L.check(L.lib().cfft_prefold_plan_info(h, L.i32p(info), None))
        n = int(info[0])
        idx = np.zeros((max(n, 1), 3), dtype=np.int32)
        L.check(L.lib().cfft_prefold_plan_info(h, L.i32p(info), L.i32p(idx)))
        self.N_full = N3
        self.N_sub = tuple(int(v) for v in info[3:6])
        self.n_channels = int(info[1])
        self.centering = CenteringType(int(info[2]))
        self.full_spec_asu = calc_spectral_asu(ops, 3, N3)
        if not np.array_equal(self.full_spec_asu.idx, idx[:n].astype(self.full_spec_asu.idx.dtype)):
            raise L.CFFTError(L.CFFT_ERR_INVALID, "spectral ASU mismatch")
        self.output_buffer = np.zeros(n, dtype=np.complex128)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            L.lib().cfft_prefold_plan_destroy(h)
            self._h = None

    def execute(self, u):
        u = np.asfortranarray(np.asarray(u, dtype=np.float64))
        if u.shape != self.N_full:
            raise L.CFFTError(L.CFFT_ERR_INVALID, f"grid shape {u.shape} != {self.N_full}")
        out = self.output_buffer.view(np.float64)
        L.check(L.lib().cfft_prefold_forward(self._h, L.f64p(u), L.f64p(out)))
        return self.output_buffer


def plan_krfft_centered(spec, ops_shifted, centering="auto"):
    """src/centering_fold.jl:328-367: falls back to plan_krfft when the lattice
    is primitive, L != (2,2,2) or M is odd.

    ``plan_krfft_centered(ops, N)`` and ``plan_krfft_centered(sg_num, N)``
    (src/centering_prefold.jl:175-184) build the full-grid pre-fold plan."""
    if not isinstance(spec, SpectralIndexing):
        N = tuple(int(n) for n in ops_shifted)
        ops = get_ops(int(spec), 3, N) if isinstance(spec, (int, np.integer)) else list(spec)
        return CenteringPreFoldKRFFTPlan(ops, N)
    N = tuple(spec.N)
    cent = detect_centering(ops_shifted, N) if centering == "auto" else CenteringType(centering)
    Lv = auto_L(ops_shifted)
    if cent == CentP or len(N) != 3 or not np.all(Lv == 2) or any((n // 2) % 2 for n in N):
        return plan_krfft(spec, ops_shifted)
    return CenteredKRFFTPlan(spec, ops_shifted)


def pack_stride_real(f0, u):
    """src/centering_fold.jl:375-380."""
    u = np.asarray(u, dtype=np.float64)
    M = f0.shape
    f0[...] = u[:2 * M[0]:2, :2 * M[1]:2, :2 * M[2]:2]


def fft_reconstruct_centered(plan: CenteredKRFFTPlan):
    """src/centering_fold.jl:421-431."""
    plan.output_buffer[:] = plan.dev.forward(plan.f0_buffer)
    return plan.output_buffer


def execute_centered_krfft(plan, u):
    """src/centering_fold.jl:392-413 (and src/centering_prefold.jl:236-244
    for a pre-fold plan)."""
    if isinstance(plan, CenteringPreFoldKRFFTPlan):
        return plan.execute(u)
    if isinstance(plan, GeneralCFFTPlan):
        return execute_krfft(plan, u)
    pack_stride_real(plan.f0_buffer, u)
    return fft_reconstruct_centered(plan)


class CenteredKRFFTBackwardPlan:
    """src/centering_fold.jl:599-620."""

    def __init__(self, spec, ops_shifted, fwd=None, precision=L.FP64):
        self.dev = DevicePlan(spec.N, ops_shifted, 0.0, None, L.PATH_M7, precision)
        self.f0_buffer = fwd.f0_buffer if fwd is not None else np.zeros(self.dev.M, order="F")
        self.subgrid_dims = self.dev.M
        self.n_spec = self.dev.n_spec


def plan_centered_ikrfft(spec, ops_shifted, fwd_plan) -> CenteredKRFFTBackwardPlan:
    return CenteredKRFFTBackwardPlan(spec, ops_shifted, fwd_plan)


def execute_centered_ikrfft(bplan: CenteredKRFFTBackwardPlan, F_spec, f0_out):
    """src/centering_fold.jl:745-764."""
    f0_out[...] = bplan.dev.backward(F_spec)
    return f0_out


def ifft_unrecon_centered(bplan: CenteredKRFFTBackwardPlan, F_spec):
    return execute_centered_ikrfft(bplan, F_spec, bplan.f0_buffer)


class CenteredSCFTPlan(_SCFTPlanBase):
    """src/centering_fold.jl:840-866."""

    def __init__(self, spec, ops_shifted, N, ds, lattice, precision=L.FP64):
        self.dev = DevicePlan(N, ops_shifted, ds, lattice, L.PATH_M7, precision)
        self.n_spec = self.dev.n_spec
        self.K_spec = self.dev.tables()[2]


def plan_centered_scft(spec, ops_shifted, N, ds, lattice, precision=L.FP64) -> CenteredSCFTPlan:
    """src/centering_fold.jl:868-894."""
    return CenteredSCFTPlan(spec, ops_shifted, N, ds, lattice, precision)


def execute_centered_scft(plan: CenteredSCFTPlan, f0):
    """src/centering_fold.jl:908-931 (in place)."""
    return plan._run(f0)


def update_kernel(plan: CenteredSCFTPlan, spec, N, ds_new, lattice):
    """src/centering_fold.jl:939-953."""
    plan.dev.update_kernel(ds_new, lattice)
    plan.K_spec = plan.dev.tables()[2]


class M2QPlan(_SCFTPlanBase):
    """src/q_fused_krfft.jl:45-61."""

    def __init__(self, N, ops_shifted, ds, lattice, path=L.PATH_M2Q, precision=L.FP64):
        self.dev = DevicePlan(N, ops_shifted, ds, lattice, path, precision)
        self.L = self.dev.L
        self.M = self.dev.M
        self.N = tuple(N)
        self.ops_shifted = list(ops_shifted)
        self._fill = None

    @property
    def fill_map(self):
        if self._fill is None:
            self._fill = build_fill_map(self.ops_shifted, self.L, self.N)
        return self._fill


def plan_m2_q(N, sg_num, dim, ds, lattice, precision=L.FP64) -> M2QPlan:
    """src/q_fused_krfft.jl:83-169."""
    if len(N) != dim:
        raise ValueError("len(N) must equal dim")
    _, _, ops_s = _sg_setup(N, sg_num, dim)
    return M2QPlan(N, ops_s, ds, lattice, L.PATH_M2Q, precision)


def execute_m2_q(plan: M2QPlan, f0):
    """src/q_fused_krfft.jl:553-576 (in place)."""
    return plan._run(f0)


class M7SCFTPlan(M2QPlan):
    """src/q_fused_krfft.jl:807-812."""


def plan_m7_scft(N, sg_num, dim, ds, lattice, precision=L.FP64) -> M7SCFTPlan:
    """src/q_fused_krfft.jl:826-855 (errors for P-centring)."""
    _, _, ops_s = _sg_setup(N, sg_num, dim)
    return M7SCFTPlan(N, ops_s, ds, lattice, L.PATH_M7Q, precision)


def execute_m7_scft(plan: M7SCFTPlan, f0):
    """src/q_fused_krfft.jl:870-895 (in place)."""
    return plan._run(f0)


def subgrid_to_fullgrid(f_full, f0, plan):
    """src/q_fused_krfft.jl:681-688 (device gather through the fill map)."""
    f_full[...] = plan.dev.sub_to_full(f0)
    return f_full


def fullgrid_to_subgrid(f0, f_full, plan):
    """src/q_fused_krfft.jl:698-719 (device strided extraction)."""
    f0[...] = plan.dev.full_to_sub(f_full)
    return f0
