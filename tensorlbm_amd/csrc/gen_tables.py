"""Generate ``tables.inc``: exact moment matrices and cumulant transforms.

Run ``python gen_tables.py`` in this directory; the output is committed.

* MRT bases — D2Q9 (reference src/tensorlbm/solver.py:20-30), D3Q19
  polynomial rows (solver3d.py:24-58), D3Q27 Gram-Schmidt rows
  (d3q27.py:545-633, orthogonalised here in exact rationals).
* Raw-monomial moment matrices and 1-D binomial shift groups for the
  cascaded / cumulant operators (cascaded_collision.py:60-187; the D2Q9
  moment set m00..m22 of cumulant.py:160-175).
* Geier cumulant <-> central-moment transforms for D3Q27, derived from the
  set-partition (moment-cumulant) formula rather than typed in
  (the reference's coefficients are cumulant.py:674-815).

Matrix inverses are exact (Gauss-Jordan over ``fractions.Fraction``) and
printed as the nearest binary64, so every kernel constant is the correctly
rounded value of the exact rational.
"""

from __future__ import annotations

from collections import Counter
from fractions import Fraction
from pathlib import Path

D2Q9_C = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (1, 1, 0), (-1, 1, 0), (-1, -1, 0), (1, -1, 0)]
D3Q19_C = [
    (0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
    (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1),
    (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1),
]
D3Q27_C = [
    (0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
    (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1),
    (-1, 0, -1), (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1), (1, 1, 1), (-1, 1, 1),
    (1, -1, 1), (-1, -1, 1), (1, 1, -1), (-1, 1, -1), (1, -1, -1), (-1, -1, -1),
]
DEG9 = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (2, 0, 0), (0, 2, 0), (1, 1, 0), (2, 1, 0), (1, 2, 0), (2, 2, 0)]
DEG19 = [
    (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0),
    (1, 0, 1), (0, 1, 1), (2, 1, 0), (2, 0, 1), (1, 2, 0), (0, 2, 1), (1, 0, 2), (0, 1, 2),
    (2, 2, 0), (2, 0, 2), (0, 2, 2),
]
DEG27 = [
    (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0),
    (1, 0, 1), (0, 1, 1), (2, 1, 0), (2, 0, 1), (1, 2, 0), (0, 2, 1), (1, 0, 2), (0, 1, 2),
    (1, 1, 1), (2, 2, 0), (2, 0, 2), (0, 2, 2), (2, 1, 1), (1, 2, 1), (1, 1, 2), (2, 2, 1),
    (2, 1, 2), (1, 2, 2), (2, 2, 2),
]


def inverse(m):
    n = len(m)
    a = [list(r) + [Fraction(int(i == j)) for j in range(n)] for i, r in enumerate(m)]
    for col in range(n):
        piv = next(r for r in range(col, n) if a[r][col] != 0)
        a[col], a[piv] = a[piv], a[col]
        p = a[col][col]
        a[col] = [v / p for v in a[col]]
        for r in range(n):
            if r != col and a[r][col] != 0:
                k = a[r][col]
                a[r] = [x - k * y for x, y in zip(a[r], a[col])]
    return [r[n:] for r in a]


def mrt_d2q9():
    rows = [
        [1, 1, 1, 1, 1, 1, 1, 1, 1], [-4, -1, -1, -1, -1, 2, 2, 2, 2], [4, -2, -2, -2, -2, 1, 1, 1, 1],
        [0, 1, 0, -1, 0, 1, -1, -1, 1], [0, -2, 0, 2, 0, 1, -1, -1, 1], [0, 0, 1, 0, -1, 1, 1, -1, -1],
        [0, 0, -2, 0, 2, 1, 1, -1, -1], [0, 1, -1, 1, -1, 0, 0, 0, 0], [0, 0, 0, 0, 0, 1, -1, 1, -1],
    ]
    return [[Fraction(v) for v in r] for r in rows]


def mrt_d3q19():
    cols = []
    for cx, cy, cz in D3Q19_C:
        e2 = cx * cx + cy * cy + cz * cz
        e4 = e2 * e2
        cols.append([
            Fraction(1), Fraction(19 * e2 - 30), Fraction(21 * e4 - 53 * e2 + 24, 2), Fraction(cx),
            Fraction(cx * e2 * (5 * e2 - 9), 2), Fraction(cy), Fraction(cy * e2 * (5 * e2 - 9), 2),
            Fraction(cz), Fraction(cz * e2 * (5 * e2 - 9), 2), Fraction(3 * cx * cx - e2),
            Fraction(cx * cx - cy * cy), Fraction(cx * cy), Fraction(cx * cz), Fraction(cy * cz),
            Fraction((3 * e2 - 5) * (3 * cx * cx - e2), 2), Fraction((3 * e2 - 5) * (cx * cx - cy * cy), 2),
            Fraction(cx * cx * cy), Fraction(cx * cx * cz), Fraction(cy * cy * cx),
        ])
    return [list(r) for r in zip(*cols)]


def mrt_d3q27():
    cols = []
    for cx, cy, cz in D3Q27_C:
        e2 = cx * cx + cy * cy + cz * cz
        cols.append([
            1, cx, cy, cz, e2, 3 * cx * cx - e2, cy * cy - cz * cz, cx * cy, cx * cz, cy * cz,
            cx * e2, cy * e2, cz * e2, cx * cx * cy, cx * cx * cz, cy * cy * cx, cy * cy * cz,
            cz * cz * cx, cz * cz * cy, e2 * e2, cx * cx * e2, cy * cy * e2, cz * cz * e2,
            cx * cx * cy * cy, cx * cx * cz * cz, cy * cy * cz * cz, cx * cy * cz,
        ])
    raw = [[Fraction(v) for v in r] for r in zip(*cols)]

    def dot(a, b):
        return sum(x * y for x, y in zip(a, b))

    out = []
    for r in raw:
        v = list(r)
        for p in out:
            k = dot(v, p) / dot(p, p)
            v = [x - k * y for x, y in zip(v, p)]
        if all(x == 0 for x in v):
            for i in range(27):
                u = [Fraction(int(j == i)) for j in range(27)]
                for p in out:
                    k = dot(u, p) / dot(p, p)
                    u = [x - k * y for x, y in zip(u, p)]
                if any(x != 0 for x in u):
                    v = u
                    break
        out.append(v)
    return out


def monomial(c, deg):
    return [[Fraction(cx**a * cy**b * cz**d) for cx, cy, cz in c] for a, b, d in deg]


def shift_groups(deg):
    res = []
    for dim in range(3):
        other = [d for d in range(3) if d != dim]
        bucket = {}
        for i, dg in enumerate(deg):
            bucket.setdefault((dg[other[0]], dg[other[1]]), {})[dg[dim]] = i
        g = []
        for key in sorted(bucket):
            mm = bucket[key]
            if 0 in mm and 1 in mm and 2 in mm:
                g.append((mm[0], mm[1], mm[2]))
        res.append(g)
    return res


def set_partitions(items):
    if not items:
        yield []
        return
    first, rest = items[0], items[1:]
    for part in set_partitions(rest):
        for i in range(len(part)):
            yield part[:i] + [[first] + part[i]] + part[i + 1:]
        yield [[first]] + part


def cumulant_terms(deg, inverse_dir):
    """Monomial expansion of C~_deg in kappa~ (or kappa~ in C~ when inverse_dir)."""
    items = [0] * deg[0] + [1] * deg[1] + [2] * deg[2]
    acc: Counter = Counter()
    for part in set_partitions(list(range(len(items)))):
        if any(len(b) == 1 for b in part):
            continue
        n = len(part)
        coef = 1 if inverse_dir else (-1) ** (n - 1) * fact(n - 1)
        key = []
        for b in part:
            d = [0, 0, 0]
            for i in b:
                d[items[i]] += 1
            key.append(DEG27.index(tuple(d)))
        acc[tuple(sorted(key))] += coef
    return acc


def fact(n):
    r = 1
    for i in range(2, n + 1):
        r *= i
    return r


def fmt(x: Fraction) -> str:
    return repr(float(x)) if x.denominator != 1 else f"{x.numerator}.0"


def matmul(a, b):
    return [[sum(a[i][k] * b[k][j] for k in range(len(b))) for j in range(len(b[0]))] for i in range(len(a))]


def emit_matrix(name, m):
    rows = ",\n    ".join("{" + ", ".join(fmt(v) for v in r) + "}" for r in m)
    n = len(m)
    return f"static constexpr double {name}[{n}][{n}] = {{\n    {rows}}};\n"


def emit_groups(name, groups):
    lines = []
    for ax, g in zip("XYZ", groups):
        if g:  # a 2-D lattice has no z groups
            flat = ", ".join(f"{{{a}, {b}, {c}}}" for a, b, c in g)
            lines.append(f"static constexpr int {name}_{ax}[{len(g)}][3] = {{{flat}}};")
        lines.append(f"static constexpr int {name}_N{ax} = {len(g)};")
    return "\n".join(lines) + "\n"


def emit_transform(fn, src, dst, inverse_dir):
    out = [f"template <typename T>\n__device__ __forceinline__ void {fn}(const T* {src}, T* {dst}) {{"]
    for i, dg in enumerate(DEG27):
        if sum(dg) < 2:
            out.append(f"  {dst}[{i}] = {src}[{i}];")
            continue
        terms = cumulant_terms(dg, inverse_dir)
        # the single-block partition is kappa~ (or C~) itself
        parts = []
        for key, coef in sorted(terms.items(), key=lambda kv: (len(kv[0]), kv[0])):
            if coef == 0:
                continue
            prod = " * ".join(f"{src}[{k}]" for k in key)
            parts.append(f"T({coef}) * {prod}" if coef != 1 else prod)
        out.append(f"  {dst}[{i}] = " + " + ".join(parts) + ";")
    out.append("}\n")
    return "\n".join(out)


def render() -> str:
    """The text of tables.inc."""
    lines = [
        "// Generated by gen_tables.py -- exact moment tables and cumulant transforms.",
        "#pragma once",
        "namespace tlbm { namespace tables {",
    ]
    for name, m in (("MRT9", mrt_d2q9()), ("MRT19", mrt_d3q19()), ("MRT27", mrt_d3q27())):
        lines.append(emit_matrix(name, m))
        lines.append(emit_matrix(name + "_INV", inverse(m)))
    for name, c, deg in (("CM9", D2Q9_C, DEG9), ("CM19", D3Q19_C, DEG19), ("CM27", D3Q27_C, DEG27)):
        m = monomial(c, deg)
        lines.append(emit_matrix(name, m))
        lines.append(emit_matrix(name + "_INV", inverse(m)))
        lines.append(emit_groups(name + "_SHIFT", shift_groups(deg)))
    # MRT moments as combinations of the monomial moments (M = T CM, M^-1 = CM^-1 T^-1 with
    # T^-1 = CM M^-1): the kernels reach the MRT basis through the axis-factored monomial
    # transforms plus these sparse recombinations
    for name, mrt, c, deg in (("MRT19", mrt_d3q19(), D3Q19_C, DEG19), ("MRT27", mrt_d3q27(), D3Q27_C, DEG27)):
        cm = monomial(c, deg)
        lines.append(emit_matrix(name + "_T", matmul(mrt, inverse(cm))))
        lines.append(emit_matrix(name + "_TI", matmul(cm, inverse(mrt))))
    lines.append(emit_transform("central_to_cumulant27", "kt", "ct", False))
    lines.append(emit_transform("cumulant_to_central27", "ct", "kt", True))
    lines.append("}}  // namespace tlbm::tables")
    return "\n".join(lines) + "\n"


def main():
    Path(__file__).with_name("tables.inc").write_text(render())


if __name__ == "__main__":
    main()
