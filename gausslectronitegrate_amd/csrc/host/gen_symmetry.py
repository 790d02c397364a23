#!/usr/bin/env python3
"""Write oh_tables.inc: the O_h matrices and their Cayley table as C initialisers.

Each group element is a signed axis permutation, given as (col_r, sign_r) per
row r; the element order is the ABI order of the exported array
(symmetry.c:9-58 of the reference).  The multiplication table is derived here
from the matrices (entry [i][j] = index of M_i * M_j), not transcribed.
"""
import sys

# (column of the non-zero entry in rows 0,1,2), (its sign in rows 0,1,2)
CODES = [
    ((0, 1, 2), (1, 1, 1)), ((1, 0, 2), (-1, -1, -1)), ((2, 1, 0), (1, -1, 1)), ((2, 0, 1), (-1, 1, -1)),
    ((1, 2, 0), (1, -1, -1)), ((0, 2, 1), (-1, 1, 1)), ((0, 1, 2), (1, 1, -1)), ((1, 0, 2), (-1, -1, 1)),
    ((2, 1, 0), (1, -1, -1)), ((2, 0, 1), (-1, 1, 1)), ((1, 2, 0), (1, -1, 1)), ((0, 2, 1), (-1, 1, -1)),
    ((2, 1, 0), (-1, -1, -1)), ((2, 0, 1), (1, 1, 1)), ((0, 1, 2), (-1, 1, -1)), ((1, 0, 2), (1, -1, 1)),
    ((0, 2, 1), (1, 1, -1)), ((1, 2, 0), (-1, -1, 1)), ((2, 1, 0), (-1, -1, 1)), ((2, 0, 1), (1, 1, -1)),
    ((0, 1, 2), (-1, 1, 1)), ((1, 0, 2), (1, -1, -1)), ((0, 2, 1), (1, 1, 1)), ((1, 2, 0), (-1, -1, -1)),
    ((1, 2, 0), (1, 1, -1)), ((0, 2, 1), (-1, -1, 1)), ((1, 0, 2), (-1, 1, -1)), ((0, 1, 2), (1, -1, 1)),
    ((2, 0, 1), (-1, -1, -1)), ((2, 1, 0), (1, 1, 1)), ((1, 2, 0), (1, 1, 1)), ((0, 2, 1), (-1, -1, -1)),
    ((1, 0, 2), (-1, 1, 1)), ((0, 1, 2), (1, -1, -1)), ((2, 0, 1), (-1, -1, 1)), ((2, 1, 0), (1, 1, -1)),
    ((0, 2, 1), (1, -1, -1)), ((1, 2, 0), (-1, 1, 1)), ((2, 0, 1), (1, -1, 1)), ((2, 1, 0), (-1, 1, -1)),
    ((1, 0, 2), (1, 1, 1)), ((0, 1, 2), (-1, -1, -1)), ((0, 2, 1), (1, -1, 1)), ((1, 2, 0), (-1, 1, -1)),
    ((2, 0, 1), (1, -1, -1)), ((2, 1, 0), (-1, 1, 1)), ((1, 0, 2), (1, 1, -1)), ((0, 1, 2), (-1, -1, 1)),
]


def matrix(code):
    cols, signs = code
    return tuple(tuple(signs[r] if c == cols[r] else 0 for c in range(3)) for r in range(3))


def mul(a, b):
    return tuple(tuple(sum(a[i][k] * b[k][j] for k in range(3)) for j in range(3)) for i in range(3))


def main(out):
    mats = [matrix(c) for c in CODES]
    index = {m: e for e, m in enumerate(mats)}
    assert len(index) == 48
    lines = ["/* generated by gen_symmetry.py: do not edit */"]
    lines.append("#define GLI_OH_MATRICES \\")
    rows = []
    for m in mats:
        rows.append("\t{ .data = { " + ", ".join("{ %2d, %2d, %2d }" % r for r in m) + " } }")
    lines.append(", \\\n".join(rows))
    lines.append("")
    lines.append("#define GLI_OH_CAYLEY \\")
    rows = []
    for a in mats:
        rows.append("\t{ " + ", ".join("%2d" % index[mul(a, b)] for b in mats) + " }")
    lines.append(", \\\n".join(rows))
    lines.append("")
    lines.append("/* signed axis permutation codes: row r of element e reads axis GLI_OH_COL[e][r] with sign GLI_OH_SGN[e][r] */")
    lines.append("#define GLI_OH_COL { " + ", ".join("{%d,%d,%d}" % c[0] for c in CODES) + " }")
    lines.append("#define GLI_OH_SGN { " + ", ".join("{%d,%d,%d}" % c[1] for c in CODES) + " }")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main(sys.argv[1])
