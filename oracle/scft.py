"""NumPy FP64 SCFT quantities for an AB diblock melt (test oracle).

The reference has no SCFT solver (no Q, density, free energy or mixing; it
stops at the diffusion step, src/diffusion_solver.jl:254-282).  These
functions are the textbook pseudo-spectral SCFT of an incompressible AB diblock
(Matsen & Schick, PRL 72, 2660 (1994); Rasmussen & Kalosakas, J. Polym. Sci. B
40, 1777 (2002) for the Strang step), restated on the stride-L subgrid:

  q(s + ds)  = e^{-w(s) ds/2} IFFT[ e^{-|k|^2 ds} FFT[ e^{-w(s) ds/2} q(s) ] ]
  Q          = <q(1)>            (grid mean)
  phi_A(r)   = (1/Q) int_0^f  q(r,s) q+(r,1-s) ds
  phi_B(r)   = (1/Q) int_f^1  q(r,s) q+(r,1-s) ds
  F / nkT    = -ln Q + < chiN phi_A phi_B - w_A phi_A - w_B phi_B >

For a field symmetric under the group, the full-grid mean equals the mean over
the stride-L subgrid (the full grid is the union of prod(L) images of the
subgrid), so every mean here is taken on the subgrid.

Parity vs the reference: unpinned (no counterpart).  Pinned instead by closed
forms (tests/test_oracle_scft.py): homogeneous fields give Q = exp(-(f w_A +
(1-f) w_B)), phi_A = f, phi_B = 1-f, and the disordered fixed point of the
field update has F = chiN f (1-f); and by the full-grid pseudo-spectral sweep
computed with numpy FFTs on the whole N^3 grid.
"""
from __future__ import annotations

import numpy as np

__all__ = ["contour_weights", "propagate", "sweep", "field_update", "free_energy",
           "disordered_free_energy", "homogeneous_Q", "AndersonMixer"]


def contour_weights(n):
    """Simpson weights on n+1 points (trapezoid on an odd remainder)."""
    w = np.zeros(n + 1)
    if n == 0:
        return w
    if n % 2 == 0:
        w[0] = w[n] = 1.0
        w[1:n:2] = 4.0
        w[2:n:2] = 2.0
        return w / 3.0
    # odd: Simpson on the first n-1 intervals, trapezoid on the last
    w[:n] = contour_weights(n - 1)
    w[n - 1] += 0.5
    w[n] += 0.5
    return w


def propagate(q0, w_of_step, ns, ds, diffuse):
    """Return q(s_j) for j = 0..ns (list of arrays)."""
    out = [q0.copy()]
    q = q0.copy()
    for j in range(ns):
        e = np.exp(-w_of_step(j) * ds / 2.0)
        q = e * diffuse(e * q)
        out.append(q)
    return out


def sweep(wA, wB, f, ns, diffuse):
    """One propagator sweep: returns dict(Q, phiA, phiB, F, q, qd).

    q starts at the A end, q+ at the B end; contour step ds = 1/ns; the A
    block spans steps j < round(f ns).
    """
    ds = 1.0 / ns
    nA = int(round(f * ns))
    q0 = np.ones_like(wA)
    q = propagate(q0, lambda j: wA if j < nA else wB, ns, ds, diffuse)
    qd = propagate(q0, lambda j: wB if j < ns - nA else wA, ns, ds, diffuse)
    Q = float(np.mean(q[ns]))
    phiA = np.zeros_like(wA)
    phiB = np.zeros_like(wA)
    wa = contour_weights(nA)
    wb = contour_weights(ns - nA)
    for j in range(nA + 1):
        phiA += wa[j] * q[j] * qd[ns - j]
    for j in range(nA, ns + 1):
        phiB += wb[j - nA] * q[j] * qd[ns - j]
    phiA *= ds / Q
    phiB *= ds / Q
    return dict(Q=Q, phiA=phiA, phiB=phiB, q=q, qd=qd)


def free_energy(wA, wB, phiA, phiB, Q, chiN):
    return float(-np.log(Q) + np.mean(chiN * phiA * phiB - wA * phiA - wB * phiB))


def field_update(wA, wB, phiA, phiB, chiN):
    """Target fields and residual for the incompressible AB melt.

    xi = (wA + wB - chiN)/2; wA* = chiN phiB + xi; wB* = chiN phiA + xi.
    Residual = sqrt(<(wA*-wA)^2 + (wB*-wB)^2> / <wA^2 + wB^2>) + |<phiA+phiB> - 1|.
    """
    xi = 0.5 * (wA + wB - chiN)
    tA = chiN * phiB + xi
    tB = chiN * phiA + xi
    num = np.mean((tA - wA) ** 2 + (tB - wB) ** 2)
    den = max(np.mean(wA ** 2 + wB ** 2), 1e-300)
    res = float(np.sqrt(num / den) + abs(np.mean(phiA + phiB) - 1.0))
    return tA, tB, res


def disordered_free_energy(f, chiN):
    return chiN * f * (1.0 - f)


def homogeneous_Q(wA, wB, f):
    return float(np.exp(-(f * wA + (1.0 - f) * wB)))


class AndersonMixer:
    """Anderson mixing of the field iteration (Matsen, Eur. Phys. J. E 30,
    361 (2009), eqs. 31-35), FP64.

    History of inputs w_k and deviations d_k = w_k^out - w_k^in (both fields
    stacked).  With U_ij = <(d_n - d_{n-i}).(d_n - d_{n-j})>, V_i =
    <(d_n - d_{n-i}).d_n>, U c = V (least squares):
      w' = w_n + sum_i c_i (w_{n-i} - w_n),  d' = d_n + sum_i c_i (d_{n-i} - d_n),
      w_{n+1} = w' + lam d'.
    Before `start_after` history entries exist (or while the residual is
    above `simple_above`) it falls back to simple mixing w + lam_s d.
    """

    def __init__(self, m=10, lam=1.0, lam_simple=0.1, simple_above=1e-2):
        self.m, self.lam, self.lam_simple, self.simple_above = m, lam, lam_simple, simple_above
        self.W, self.D = [], []

    def step(self, w, d, res):
        self.W.append(w.copy())
        self.D.append(d.copy())
        if len(self.W) > self.m + 1:
            self.W.pop(0)
            self.D.pop(0)
        k = len(self.W) - 1
        if res > self.simple_above or k == 0:
            return w + self.lam_simple * d
        dn, wn = self.D[-1], self.W[-1]
        dd = [dn - self.D[-1 - i] for i in range(1, k + 1)]
        U = np.array([[np.mean(a * b) for b in dd] for a in dd])
        V = np.array([np.mean(a * dn) for a in dd])
        c = np.linalg.lstsq(U, V, rcond=None)[0]
        wp = wn + sum(c[i] * (self.W[-2 - i] - wn) for i in range(k))
        dp = dn + sum(c[i] * (self.D[-2 - i] - dn) for i in range(k))
        return wp + self.lam * dp
