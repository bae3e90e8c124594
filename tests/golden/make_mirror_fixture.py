"""Extract the rows of the reference's generated mirror/glide table
(src/mirror_table.jl, written by scripts/compute_mirror_table.jl from
Crystalline's operation tables) that cover the space groups tabulated here,
as plain data: centring symbol, centring vectors, mirror normals and the glide
translation of the first operation with each mirror rotation.

Run from the repository root:  python tests/golden/make_mirror_fixture.py
Reads /root/reference/src/mirror_table.jl; writes tests/golden/mirror_table_rows.json.
"""
import json
import os
import re
from fractions import Fraction

SRC = "/root/reference/src/mirror_table.jl"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mirror_table_rows.json")
GROUPS = [1, 2, 10, 12, 14, 25, 35, 47, 62, 63, 65, 69, 70, 71, 72, 74, 123, 136, 139, 141,
          200, 221, 223, 224, 225, 227, 229, 230]


def parse_vecs(s):
    vecs = re.findall(r"\[([^\[\]]*)\]", s)
    out = []
    for v in vecs:
        if not v.strip():
            continue
        out.append([str(Fraction(x.strip().replace("//", "/"))) for x in v.split(",")])
    return out


def split_top(s):
    """Split a Julia argument list at top-level commas."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch == "[":
            depth += 1
        elif ch == "]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def main():
    rows = {}
    pat = re.compile(r"^\s*(\d+) => MirrorAdaptationEntry\((.*)\),?\s*$")
    for line in open(SRC):
        m = pat.match(line)
        if not m:
            continue
        sg = int(m.group(1))
        if sg not in GROUPS:
            continue
        args = split_top(m.group(2))
        normals = [[int(x) for x in v] for v in parse_vecs(args[1])]
        trans = parse_vecs(args[2])
        rows[sg] = {"n_mirrors": int(args[0]), "normals": normals, "translations": trans,
                    "centering": args[3].lstrip(":"), "centering_vecs": parse_vecs(args[4])}
    with open(OUT, "w") as f:
        json.dump({"source": "src/mirror_table.jl", "rows": rows}, f, indent=1, sort_keys=True)
    print(f"wrote {len(rows)} rows to {OUT}")


if __name__ == "__main__":
    main()
