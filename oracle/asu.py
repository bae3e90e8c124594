"""Real-space ASU and block packing (test oracle).

Restates src/asu.jl:216-325 (calc_asu, recursive parity filter of Kudlicki et
al. 2007) and src/pack_asu.jl:6-102 (pack_asu, greedy rectangular blocks per
depth group), plus the per-block FFT of the block CFFT (src/cfft_plan.jl,
src/execution.jl).

Pinned by the reference's tests (test/test_asu.jl, test/test_packing.jl,
test/test_cfft.jl): p2mm 8x8 -> 16 points, p2mg 8x8 -> 16, Pmmm 8^3 -> 64;
orbits of the points partition the grid (coverage + disjointness); the packed
block sizes sum to the number of points; block FFT round trip = identity.
Block ORDER inside a dimension group follows Julia's Dict iteration over depth
groups in the reference and is not reproducible; here depth groups are taken
in sorted order (parity unpinned for the order; the block set is pinned).
"""
from __future__ import annotations

import itertools

import numpy as np

__all__ = ["calc_asu", "pack_asu", "block_fft", "block_ifft"]


def _apply(R, t, x, N):
    return tuple(int(v) for v in np.mod(R @ np.asarray(x) + t, np.asarray(N)))


def calc_asu(N, R, t):
    """src/asu.jl:216-325 -> list of (idx tuple, depth tuple, multiplicity),
    sorted lexicographically by idx."""
    D = len(N)
    ops0 = [(np.array(R[g], dtype=np.int64), np.array(t[g], dtype=np.int64)) for g in range(R.shape[0])]
    queue = [(tuple(N), ops0, (1,) * D, (0,) * D, (0,) * D, (False,) * D)]
    pts = []
    while queue:
        cN, cops, cs, co, cd, cgp = queue.pop()
        if any(n == 0 for n in cN):
            continue
        if all(cgp) or all(n <= 1 for n in cN):
            visited = set()
            for p in itertools.product(*[range(n) for n in reversed(cN)]):
                p = tuple(reversed(p))  # Iterators.product: first dimension fastest
                if p in visited:
                    continue
                orbit = {p}
                stack = [p]
                while stack:
                    c = stack.pop()
                    for Ro, to in cops:
                        y = _apply(Ro, to, c, cN)
                        if y not in orbit:
                            orbit.add(y)
                            stack.append(y)
                rep = min(orbit)
                pts.append((tuple(cs[d] * rep[d] + co[d] for d in range(D)), tuple(cd), len(orbit)))
                visited |= orbit
            continue
        eff = list(cgp)
        for d in range(D):
            if cgp[d]:
                continue
            pe = np.zeros(D, dtype=np.int64)
            po = np.zeros(D, dtype=np.int64)
            po[d] = 1
            ok = True
            for Ro, to in cops:
                if _apply(Ro, to, pe, cN)[d] % 2 != 0:
                    ok = False
                    break
                if _apply(Ro, to, po, cN)[d] % 2 != 1:
                    ok = False
                    break
            if not ok:
                eff[d] = True
        ranges = [range(1) if (eff[d] or cN[d] <= 1) else range(2) for d in range(D)]
        children = []
        for par in itertools.product(*reversed(ranges)):
            par = np.array(tuple(reversed(par)), dtype=np.int64)
            S = np.array([1 if eff[d] else 2 for d in range(D)], dtype=np.int64)
            new_ops = []
            valid = True
            for Ro, to in cops:
                Rv = Ro * S[None, :]
                tmp = to - par + Ro @ par
                if np.any(Rv % S[:, None] != 0) or np.any(tmp % S != 0):
                    valid = False
                    break
                new_ops.append((np.fix(Rv / S[:, None]).astype(np.int64), np.fix(tmp / S).astype(np.int64)))
            if not valid:
                continue
            nN = tuple(cN[d] if eff[d] else cN[d] // 2 for d in range(D))
            ngp = tuple(eff[d] or par[d] == 1 for d in range(D))
            ndep = tuple(cd[d] if eff[d] else cd[d] + (1 if par[d] == 0 else 0) for d in range(D))
            children.append((nN, new_ops, tuple(cs[d] * S[d] for d in range(D)),
                             tuple(cs[d] * par[d] + co[d] for d in range(D)), ndep, ngp))
        queue.extend(children)
    pts.sort(key=lambda p: p[0])
    return pts


def pack_asu(points, N):
    """src/pack_asu.jl:6-102 -> list of blocks (start tuple, shape tuple,
    depth), greedy boxes grown dimension by dimension from the smallest
    remaining point of each depth group; depth groups in sorted order."""
    D = len(N)
    groups = {}
    for idx, depth, _ in points:
        groups.setdefault(depth, []).append(idx)
    blocks = []
    for depth in sorted(groups):
        pool = sorted(groups[depth])
        pool_set = set(pool)
        while pool:
            start = pool[0]
            lo = list(start)
            hi = list(start)
            for d in range(D):
                steps = 0
                while True:
                    nxt = start[d] + steps + 1
                    face = [range(lo[k], hi[k] + 1) for k in range(D)]
                    face[d] = range(nxt, nxt + 1)
                    if any(tuple(pt) not in pool_set for pt in itertools.product(*face)):
                        break
                    steps += 1
                hi[d] = start[d] + steps
            used = set(itertools.product(*[range(lo[k], hi[k] + 1) for k in range(D)]))
            blocks.append((tuple(lo), tuple(hi[k] - lo[k] + 1 for k in range(D)), depth))
            pool_set -= used
            pool = [p for p in pool if p in pool_set]
    return blocks


def block_fft(data):
    """FFTW forward on each block (src/cfft_plan.jl:43-53)."""
    return [np.fft.fftn(b) for b in data]


def block_ifft(data):
    """FFTW inverse (1/n) on each block (src/cfft_plan.jl:55-66)."""
    return [np.fft.ifftn(b) for b in data]
