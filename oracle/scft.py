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
           "disordered_free_energy", "homogeneous_Q"]


def contour_weights(n):
    """Simpson weights on n+1 points (trapezoid on an odd remainder)."""
    w = np.zeros(n + 1)
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
