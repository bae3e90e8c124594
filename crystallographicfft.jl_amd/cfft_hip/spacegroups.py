"""Space-group operation tables (the role Crystalline.jl plays for the reference).

The reference obtains its symmetry operations from Crystalline.jl through
``operations(spacegroup(sg, dim))`` (src/symmetry_ops.jl:51-53): the full
conventional-cell list, centring translations included, translations as
fractions in [0, 1).  Crystalline is not available here, so the operations are
generated from Hall symbols (S. R. Hall, Acta Cryst. A37, 517 (1981); the
explicit-origin notation of the International Tables, Vol. B, Sect. 1.4):
the generators a Hall symbol encodes are closed into a group modulo lattice
translations and then multiplied by the centring vectors.

Exact rational arithmetic is kept by storing translations in units of 1/12
(every space-group translation is a multiple of 1/12).

The ORDER of the generated list is this module's own (identity first, then
breadth-first products of the generators, centring blocks last).  The
reference's order comes from Crystalline's tables and is not reproduced; no
transform result depends on it (see the op-order note in the design notes).
"""
from __future__ import annotations

import numpy as np

# Hall symbols for the ITA standard settings used here.  Groups with two origin
# choices use origin choice 2 (centrosymmetric origin), which is what
# Crystalline returns: e.g. the Fddd d-glide x+1/4,y+1/4,-z of
# src/mirror_table.jl:98 and the Fd-3m glide translation (1/4,3/4,1/2) of
# src/mirror_table.jl:255 only occur in origin choice 2.
HALL_3D = {
    1: "P 1",
    2: "-P 1",
    10: "-P 2y",
    12: "-C 2y",
    14: "-P 2ybc",
    25: "P 2 -2",
    35: "C 2 -2",
    47: "-P 2 2",
    62: "-P 2ac 2n",
    63: "-C 2c 2",
    65: "-C 2 2",
    69: "-F 2 2",
    70: "-F 2uv 2vw",
    71: "-I 2 2",
    72: "-I 2 2c",
    74: "-I 2b 2",
    75: "P 4",
    99: "P 4 -2",
    123: "-P 4 2",
    136: "-P 4n 2n",
    139: "-I 4 2",
    141: "-I 4bd 2",
    200: "-P 2 2 3",
    221: "-P 4 2 3",
    223: "-P 4n 2 3",
    224: "-P 4bc 2bc 3",
    225: "-F 4 2 3",
    227: "-F 4vw 2vw 3",
    229: "-I 4 2 3",
    230: "-I 4bd 2c 3",
}

# Group orders |G| (conventional cell, centring included) for the table above:
# point-group order x number of centring vectors.
GROUP_ORDER_3D = {
    1: 1, 2: 2, 10: 4, 12: 8, 14: 4, 25: 4, 35: 8, 47: 8,
    62: 8, 63: 16, 65: 16, 69: 32, 70: 32, 71: 16, 72: 16, 74: 16, 75: 4, 99: 8,
    123: 16, 136: 16, 139: 32, 141: 32, 200: 24, 221: 48, 223: 48, 224: 48,
    225: 192, 227: 192, 229: 96, 230: 96,
}

# Plane groups (2-D), numbered as in the International Tables / Crystalline's
# ``spacegroup(n, 2)``.  Entries: list of (R (2x2), t in twelfths).
_PLANE_GROUPS = {
    1: [((1, 0, 0, 1), (0, 0))],                                       # p1
    2: [((1, 0, 0, 1), (0, 0)), ((-1, 0, 0, -1), (0, 0))],             # p2
    3: [((1, 0, 0, 1), (0, 0)), ((-1, 0, 0, 1), (0, 0))],              # pm
    4: [((1, 0, 0, 1), (0, 0)), ((-1, 0, 0, 1), (0, 6))],              # pg
    5: [((1, 0, 0, 1), (0, 0)), ((-1, 0, 0, 1), (0, 0)),               # cm
        ((1, 0, 0, 1), (6, 6)), ((-1, 0, 0, 1), (6, 6))],
    6: [((1, 0, 0, 1), (0, 0)), ((-1, 0, 0, -1), (0, 0)),              # p2mm
        ((-1, 0, 0, 1), (0, 0)), ((1, 0, 0, -1), (0, 0))],
    7: [((1, 0, 0, 1), (0, 0)), ((-1, 0, 0, -1), (0, 0)),              # p2mg
        ((-1, 0, 0, 1), (6, 0)), ((1, 0, 0, -1), (6, 0))],
    8: [((1, 0, 0, 1), (0, 0)), ((-1, 0, 0, -1), (0, 0)),              # p2gg
        ((-1, 0, 0, 1), (6, 6)), ((1, 0, 0, -1), (6, 6))],
    11: [((1, 0, 0, 1), (0, 0)), ((-1, 0, 0, -1), (0, 0)),             # p4mm
         ((0, -1, 1, 0), (0, 0)), ((0, 1, -1, 0), (0, 0)),
         ((-1, 0, 0, 1), (0, 0)), ((1, 0, 0, -1), (0, 0)),
         ((0, 1, 1, 0), (0, 0)), ((0, -1, -1, 0), (0, 0))],
}

_LATTICE_TRANSLATIONS = {
    "P": [],
    "A": [(0, 6, 6)],
    "B": [(6, 0, 6)],
    "C": [(6, 6, 0)],
    "I": [(6, 6, 6)],
    "F": [(0, 6, 6), (6, 0, 6), (6, 6, 0)],
    "R": [(8, 4, 4), (4, 8, 8)],
}

_TRANSLATION_SYMBOLS = {
    "a": (6, 0, 0), "b": (0, 6, 0), "c": (0, 0, 6), "n": (6, 6, 6),
    "u": (3, 0, 0), "v": (0, 3, 0), "w": (0, 0, 3), "d": (3, 3, 3),
}


_ROT = {
    # Hall (1981), table of rotation matrices for principal axes.
    ("x", 1): [[1, 0, 0], [0, 1, 0], [0, 0, 1]],
    ("y", 1): [[1, 0, 0], [0, 1, 0], [0, 0, 1]],
    ("z", 1): [[1, 0, 0], [0, 1, 0], [0, 0, 1]],
    ("x", 2): [[1, 0, 0], [0, -1, 0], [0, 0, -1]],
    ("y", 2): [[-1, 0, 0], [0, 1, 0], [0, 0, -1]],
    ("z", 2): [[-1, 0, 0], [0, -1, 0], [0, 0, 1]],
    ("x", 3): [[1, 0, 0], [0, 0, -1], [0, 1, -1]],
    ("y", 3): [[-1, 0, 1], [0, 1, 0], [-1, 0, 0]],
    ("z", 3): [[0, -1, 0], [1, -1, 0], [0, 0, 1]],
    ("x", 4): [[1, 0, 0], [0, 0, -1], [0, 1, 0]],
    ("y", 4): [[0, 0, 1], [0, 1, 0], [-1, 0, 0]],
    ("z", 4): [[0, -1, 0], [1, 0, 0], [0, 0, 1]],
    ("x", 6): [[1, 0, 0], [0, 1, -1], [0, 1, 0]],
    ("y", 6): [[0, 0, 1], [0, 1, 0], [-1, 0, 1]],
    ("z", 6): [[1, -1, 0], [1, 0, 0], [0, 0, 1]],
    # face-diagonal two-folds (primary axis z) and the body-diagonal three-fold
    ("'", 2): [[0, -1, 0], [-1, 0, 0], [0, 0, -1]],
    ('"', 2): [[0, 1, 0], [1, 0, 0], [0, 0, -1]],
    ("*", 3): [[0, 0, 1], [1, 0, 0], [0, 1, 0]],
}


def _rot(axis: str, n: int) -> np.ndarray:
    """Proper rotation matrix of order n about a principal or special axis."""
    try:
        return np.array(_ROT[(axis, n)], dtype=np.int64)
    except KeyError:
        raise ValueError(f"no {n}-fold rotation about axis {axis!r}") from None


def _axis_vector(axis: str) -> np.ndarray:
    return {"x": np.array([12, 0, 0]), "y": np.array([0, 12, 0]),
            "z": np.array([0, 0, 12]), "*": np.array([12, 12, 12])}[axis]


def parse_hall(symbol: str):
    """Parse a Hall symbol into (generators, centring vectors).

    Generators are (R, t12) with R a 3x3 int matrix and t12 the translation in
    twelfths.  Only the notation needed for the conventional settings above is
    supported (no change-of-basis suffix other than an origin shift vector).
    """
    toks = symbol.split()
    lat = toks[0]
    centro = lat.startswith("-")
    lat = lat.lstrip("-")
    if lat not in _LATTICE_TRANSLATIONS:
        raise ValueError(f"bad lattice symbol in {symbol!r}")
    gens = []
    if centro:
        gens.append((-np.eye(3, dtype=np.int64), np.zeros(3, dtype=np.int64)))
    shift = np.zeros(3, dtype=np.int64)
    prev_n = None
    for k, tok in enumerate(toks[1:]):
        if tok.startswith("("):
            shift = np.array([int(v) for v in tok.strip("()").split()], dtype=np.int64)
            continue
        improper = tok.startswith("-")
        s = tok.lstrip("-")
        n = int(s[0])
        rest = s[1:]
        axis = None
        if rest and rest[0] in "xyz'\"*":
            axis = rest[0]
            rest = rest[1:]
        if axis is None:
            if k == 0:
                axis = "z"
            elif k == 1 and n == 2:
                axis = "x" if prev_n in (2, 4) else "'"
            elif k == 2 and n == 3:
                axis = "*"
            elif n == 1:
                axis = "z"
            else:
                raise ValueError(f"cannot infer axis for {tok!r} in {symbol!r}")
        R = _rot(axis, n)
        if improper:
            R = -R
        t = np.zeros(3, dtype=np.int64)
        for ch in rest:
            if ch in _TRANSLATION_SYMBOLS:
                t += np.array(_TRANSLATION_SYMBOLS[ch], dtype=np.int64)
            elif ch.isdigit():
                # screw component: (digit / n) along the rotation axis
                t += _axis_vector(axis) * int(ch) // n
            else:
                raise ValueError(f"bad translation symbol {ch!r} in {symbol!r}")
        gens.append((R, t % 12))
        prev_n = n
    if np.any(shift):
        # origin shift v (twelfths): t' = t + (I - R) v
        gens = [(R, (t + (np.eye(3, dtype=np.int64) - R) @ shift) % 12) for R, t in gens]
    cents = [np.array(c, dtype=np.int64) for c in _LATTICE_TRANSLATIONS[lat]]
    return gens, cents


def _close(gens, cents, dim=3):
    """Close generators into a group modulo the lattice (BFS).

    Two operations are identified when their translations differ by an integer
    vector or a centring vector; the first translation generated is kept as
    the coset representative.
    """
    lattice = [np.zeros(dim, dtype=np.int64)] + list(cents)

    def key(op):
        t = op[1]
        canon = min(tuple(int(v) for v in (t + c) % 12) for c in lattice)
        return (tuple(op[0].ravel()), canon)

    ident = (np.eye(dim, dtype=np.int64), np.zeros(dim, dtype=np.int64))
    ops = [ident]
    seen = {key(ident)}
    frontier = [ident]
    while frontier:
        nxt = []
        for a in frontier:
            for g in gens:
                R = g[0] @ a[0]
                t = (g[0] @ a[1] + g[1]) % 12
                op = (R, t)
                k = key(op)
                if k not in seen:
                    seen.add(k)
                    ops.append(op)
                    nxt.append(op)
        frontier = nxt
    return ops


def space_group_ops12(sg: int, dim: int = 3):
    """All operations (R, t in twelfths, t in [0,12)) of space group ``sg``."""
    if dim == 2:
        if sg not in _PLANE_GROUPS:
            raise ValueError(f"plane group {sg} not tabulated")
        return [(np.array(R, dtype=np.int64).reshape(2, 2), np.array(t, dtype=np.int64))
                for R, t in _PLANE_GROUPS[sg]]
    if dim != 3:
        raise ValueError("dim must be 2 or 3")
    if sg not in HALL_3D:
        raise ValueError(f"space group {sg} not tabulated (have {sorted(HALL_3D)})")
    gens, cents = parse_hall(HALL_3D[sg])
    base = _close(gens, cents)
    ops = list(base)
    for c in cents:
        ops.extend((R, (t + c) % 12) for R, t in base)
    return ops


def centering_vectors12(sg: int):
    """Centring translations (twelfths) of a 3-D space group."""
    return [tuple(int(v) for v in c) for c in parse_hall(HALL_3D[sg])[1]]
