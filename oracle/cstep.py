"""TEST INFRASTRUCTURE ONLY: CPU baseline of one Strang diffusion step.

Builds the index/phase tables of the centred / M2 KRFFT step from the NumPy
oracle (calc_spectral_asu, select_rep_ops, reverse_lookup, kernel_spec:
src/spectral_indexing.jl, src/krfft.jl:301-389 and 997-1311,
src/q_fused_krfft.jl:966-973) and executes the step in C (oracle/cstep.c,
OpenMP).  Only tests and the ``cpu_baseline`` leg of bench.py use it.
Pinned by tests/test_oracle_cstep.py against the NumPy oracle's Strang step.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from .krfft import (CENT_P, alive_offsets, auto_L, calc_spectral_asu, detect_centering_type,
                    get_k_vector, kernel_spec, reverse_lookup, select_rep_ops)

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cstep.c")
LIB = os.path.join(HERE, "libcstep.so")

_lib = None


def build(verbose=False):
    """gcc -O3 (x86-64-v3, OpenMP) -> oracle/libcstep.so."""
    tmp = f"{LIB}.tmp{os.getpid()}"
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-o", tmp, SRC, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/libcstep.so is not built (oracle.cstep.build())")
        L = C.CDLL(LIB)
        P, PI, PD, PU = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_uint8)
        L.cstep_create.restype = P
        L.cstep_create.argtypes = [PI, C.c_int32, PI, C.c_int32, C.c_int32, PI, PD, PD, PI, PD, PU,
                                   C.c_double, C.c_int32]
        L.cstep_step.restype = None
        L.cstep_step.argtypes = [P, PD, PD, PD]
        L.cstep_destroy.restype = None
        L.cstep_destroy.argtypes = [P]
        L.cstep_max_threads.restype = C.c_int
        L.cstep_set_threads.restype = None
        L.cstep_set_threads.argtypes = [P, C.c_int32]
        _lib = L
    return _lib


def _lin(coords, dims):
    out = np.zeros(coords.shape[0], dtype=np.int64)
    stride = 1
    for d in range(len(dims)):
        out += coords[:, d] * stride
        stride *= int(dims[d])
    return out


class CStep:
    """Tables + C plan for q <- E D (E q) on the stride-2 subgrid (L = 2)."""

    def __init__(self, R, t, N, ds, lattice, nthreads=1):
        N = tuple(int(n) for n in N)
        Lv = auto_L(t)
        if not np.all(Lv == 2):
            raise ValueError("the C baseline covers L = (2,2,2) plans")
        M = tuple(n // 2 for n in N)
        if any(m % 2 for m in M):
            raise ValueError("M must be even")
        H = tuple(m // 2 for m in M)
        if max(H) > 64:
            raise ValueError("H <= 64")
        self.M, self.H = M, H
        spec = calc_spectral_asu(R, t, N)
        self.spec = spec
        rep = select_rep_ops(R, t, Lv, N)
        ridx, rph, rcj = reverse_lookup(spec, R, t, N)
        cent = detect_centering_type(R, t, N)
        offs = ([(e & 1, (e >> 1) & 1, (e >> 2) & 1) for e in range(8)] if cent == CENT_P
                else alive_offsets(cent))
        chmap = -np.ones(8, dtype=np.int64)
        for c, o in enumerate(offs):
            chmap[o[0] | (o[1] << 1) | (o[2] << 2)] = c
        H3 = int(np.prod(H))
        n_spec = spec.idx.shape[0]
        d = len(rep)
        hc = get_k_vector(spec.idx, N)
        fidx = np.zeros((n_spec, d), dtype=np.int32)
        fph = np.zeros((n_spec, d), dtype=np.complex128)
        for j, a in enumerate(rep):
            rot = np.mod(hc @ R[a], np.asarray(M))
            code = (rot[:, 0] % 2) | ((rot[:, 1] % 2) << 1) | ((rot[:, 2] % 2) << 2)
            c = chmap[code]
            k = _lin(rot // 2, H)
            fidx[:, j] = np.where(c >= 0, c * H3 + k, -1)
            phase = sum(hc[:, q] * t[a, q] / N[q] for q in range(3))
            fph[:, j] = np.exp(-2j * np.pi * phase)
        n_alive = len(offs) * H3
        bidx = -np.ones((n_alive, 8), dtype=np.int32)
        bph = np.zeros((n_alive, 8), dtype=np.complex128)
        bcj = np.zeros((n_alive, 8), dtype=np.uint8)
        kk = np.arange(H3)
        kc = np.stack([kk % H[0], (kk // H[0]) % H[1], kk // (H[0] * H[1])], axis=1)
        # G0(q) = (1/d) sum_alpha F(q + M alpha), F read through the reverse lookup
        for c, o in enumerate(offs):
            qv = 2 * kc + np.asarray(o)[None, :]
            for a in range(8):
                al = np.array([a & 1, (a >> 1) & 1, (a >> 2) & 1])
                h = _lin(qv + al[None, :] * np.asarray(M)[None, :], N)
                sl = slice(c * H3, (c + 1) * H3)
                bidx[sl, a] = ridx[h]
                bph[sl, a] = np.where(ridx[h] >= 0, rph[h] / d, 0.0)
                bcj[sl, a] = rcj[h]
        self.K = np.ascontiguousarray(kernel_spec(spec, ds, lattice))
        self._keep = [np.ascontiguousarray(a) for a in (fidx, fph.view(np.float64), bidx,
                                                        bph.view(np.float64), bcj)]
        self.offs = np.ascontiguousarray(np.array(offs, dtype=np.int32).ravel())
        Lb = lib()
        PI, PD, PU = C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_uint8)
        fidx_, fph_, bidx_, bph_, bcj_ = self._keep
        self._h = Lb.cstep_create(np.array(M, dtype=np.int32).ctypes.data_as(PI), len(offs),
                                  self.offs.ctypes.data_as(PI), n_spec, d, fidx_.ctypes.data_as(PI),
                                  fph_.ctypes.data_as(PD), self.K.ctypes.data_as(PD),
                                  bidx_.ctypes.data_as(PI), bph_.ctypes.data_as(PD),
                                  bcj_.ctypes.data_as(PU), 1.0, int(nthreads))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            try:
                _lib.cstep_destroy(h)
            except Exception:
                pass
            self._h = None

    def set_threads(self, n):
        lib().cstep_set_threads(self._h, int(n))

    def step(self, q, E=None, out=None):
        """One step on a subgrid field (shape M, x fastest); E = exp(-w ds/2)
        or None.  Returns the new field."""
        qf = np.asfortranarray(q, dtype=np.float64)
        o = np.empty(self.M, order="F") if out is None else out
        PD = C.POINTER(C.c_double)
        Ep = None if E is None else np.asfortranarray(E, dtype=np.float64).ctypes.data_as(PD)
        lib().cstep_step(self._h, qf.ctypes.data_as(PD), o.ctypes.data_as(PD), Ep)
        return o


def first_iteration(cs, wA, wB, f, chiN, ns):
    """Q, F and the densities of one SCFT sweep with the C step: the
    computation of oracle.scft.sweep / free_energy (Simpson densities over the
    stored q, pairing q(j) with q+(ns - j)), with q+ advanced alongside so that
    only q is stored.  Returns dict(Q, F, phiA, phiB, q, qd) with q, qd the
    end points q(1), q+(1)."""
    from .scft import contour_weights, free_energy
    ds = 1.0 / ns
    EA = np.asfortranarray(np.exp(-wA * ds / 2.0))
    EB = np.asfortranarray(np.exp(-wB * ds / 2.0))
    nA = int(round(f * ns))
    q = [np.ones(cs.M, order="F")]
    for j in range(ns):
        q.append(cs.step(q[-1], EA if j < nA else EB))
    qd = np.ones(cs.M, order="F")
    wa, wb = contour_weights(nA), contour_weights(ns - nA)
    phiA = np.zeros(cs.M)
    phiB = np.zeros(cs.M)
    for j in range(ns + 1):
        # qd holds q+(j); the density pairs q(ns - j) with q+(j)
        k = ns - j
        if k <= nA:
            phiA += wa[k] * q[k] * qd
        if k >= nA:
            phiB += wb[k - nA] * q[k] * qd
        if j < ns:
            qd = cs.step(qd, EB if j < ns - nA else EA)
    Q = float(np.mean(q[ns]))
    phiA *= ds / Q
    phiB *= ds / Q
    F = free_energy(wA, wB, phiA, phiB, Q, chiN)
    return dict(Q=Q, F=F, phiA=phiA, phiB=phiB, q=q[ns], qd=qd)
