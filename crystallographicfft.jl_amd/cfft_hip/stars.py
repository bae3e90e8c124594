"""Symmetry-adapted (star) coefficients of a subgrid field.

The spectral ASU holds one representative h per orbit (star) of the
reciprocal grid; the KRFFT forward gives F(h) = sum_x u(x) exp(-2 pi i h.x/N)
of the full-grid field at each representative, and the backward rebuilds the
field from them through the reverse lookup (phases and conjugation of every
star member, src/krfft.jl:1183-1311).  Star coefficients are

    c_i = F(h_i) / (N0 N1 N2),   i ordered by |2 pi A^-T h_c|^2, then by index,

so that c_0 is the cell mean and u(x) = sum over all h of c(h) exp(2 pi i h.x/N).
No counterpart in the reference (it has no basis I/O); this is the format the
warm-start model reads and writes.
"""
from __future__ import annotations

import numpy as np

from .plans import DevicePlan


def star_order(plan: DevicePlan, lattice=None):
    """Spectral-ASU indices sorted by |k|^2 (ties by index) and the |k|^2."""
    idx, mult, K, rep = plan.tables()
    lat = plan.lattice if lattice is None else np.asarray(lattice, dtype=np.float64)
    N = np.asarray(plan.N)
    hc = np.where(idx >= (N // 2)[None, :], idx - N[None, :], idx)
    B = 2.0 * np.pi * np.linalg.inv(lat).T
    kv = hc @ B.T
    k2 = np.sum(kv * kv, axis=1)
    order = np.lexsort((np.arange(len(k2)), np.round(k2, 9)))
    return order, k2[order]


def encode(plan: DevicePlan, f0, n, order=None):
    """First n star coefficients of the subgrid field f0 (complex)."""
    if order is None:
        order, _ = star_order(plan)
    F = plan.forward(f0)
    return F[order[:n]] / float(np.prod(plan.N))


def decode(plan: DevicePlan, coeffs, order=None):
    """Subgrid field whose first len(coeffs) star coefficients are `coeffs`
    and all others zero."""
    if order is None:
        order, _ = star_order(plan)
    coeffs = np.asarray(coeffs, dtype=np.complex128)
    F = np.zeros(plan.n_spec, dtype=np.complex128)
    F[order[:coeffs.size]] = coeffs * float(np.prod(plan.N))
    return plan.backward(F)
