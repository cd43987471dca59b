"""Regenerate the per-pixel cubature tables from the mathematics.

The seven rules are the ones Lensed offers (src/quadrature.c:21-30):
point, sub2, sub4, gm75, g3k7, g5k11, g7k15.  Nothing is copied: the
Gauss-Kronrod rules are recomputed here from first principles with exact
rational arithmetic plus mpmath root finding, and the Genz-Malik rule comes
from its published closed form (Genz & Malik, J. Comp. Appl. Math. 6 (1980)
295-302).  All rules live on the unit pixel [-1/2, 1/2]^2 with weights that
sum to one; the error weights sum to zero.

Tensor Gauss-Kronrod rules: with 1-D Kronrod weights K and Gauss weights G
(G = 0 on Kronrod-only nodes) the value weights are K x K and the error
weights are (K - G) x (K - G), so that sum(err * f) is the product of the
1-D (Kronrod - Gauss) differences.  Node order: x outer, y inner, ascending.

Run:  python gen_quad.py > quad_tables.h
"""
from fractions import Fraction
import mpmath

mpmath.mp.dps = 60


def legendre_coeffs(n):
    """Monomial coefficients (lowest first) of the Legendre polynomial P_n."""
    p0, p1 = [Fraction(1)], [Fraction(0), Fraction(1)]
    if n == 0:
        return p0
    for k in range(1, n):
        # (k+1) P_{k+1} = (2k+1) x P_k - k P_{k-1}
        xp = [Fraction(0)] + p1
        nxt = [Fraction(0)] * (k + 2)
        for i, c in enumerate(xp):
            nxt[i] += Fraction(2 * k + 1, k + 1) * c
        for i, c in enumerate(p0):
            nxt[i] -= Fraction(k, k + 1) * c
        p0, p1 = p1, nxt
    return p1


def moment(m):
    """Integral of x^m over [-1, 1]."""
    return Fraction(0) if m % 2 else Fraction(2, m + 1)


def poly_mul(a, b):
    out = [Fraction(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] += x * y
    return out


def solve_rational(a, b):
    """Gaussian elimination over the rationals."""
    n = len(b)
    m = [row[:] + [rhs] for row, rhs in zip(a, b)]
    for c in range(n):
        piv = next(r for r in range(c, n) if m[r][c] != 0)
        m[c], m[piv] = m[piv], m[c]
        for r in range(n):
            if r != c and m[r][c] != 0:
                f = m[r][c] / m[c][c]
                m[r] = [x - f * y for x, y in zip(m[r], m[c])]
    return [m[i][n] / m[i][i] for i in range(n)]


def real_roots(coeffs):
    """All (real, simple) roots of a polynomial, ascending."""
    roots = mpmath.polyroots([mpmath.mpf(c.numerator) / c.denominator
                              for c in reversed(coeffs)], maxsteps=500, extraprec=400)
    return sorted(mpmath.re(r) for r in roots)


def weights_for(nodes, exact_degree):
    """Interpolatory weights on [-1, 1] from the moment equations."""
    n = len(nodes)
    a = mpmath.matrix(n, n)
    b = mpmath.matrix(n, 1)
    for k in range(n):
        for i, x in enumerate(nodes):
            a[k, i] = x ** k
        mk = moment(k)
        b[k] = mpmath.mpf(mk.numerator) / mk.denominator
    w = mpmath.lu_solve(a, b)
    # check the claimed degree of exactness on the extra monomials
    for k in range(n, exact_degree + 1):
        mk = moment(k)
        got = sum(w[i] * nodes[i] ** k for i in range(n))
        assert abs(got - mpmath.mpf(mk.numerator) / mk.denominator) < mpmath.mpf(10) ** -40, (n, k)
    return [w[i] for i in range(n)]


def gauss_kronrod(n):
    """n-point Gauss and (2n+1)-point Kronrod rule on [-1, 1]."""
    pn = legendre_coeffs(n)
    gauss = real_roots(pn)
    # Stieltjes polynomial E_{n+1} = x^{n+1} + sum_{j<=n} c_j x^j with
    # int P_n E x^k dx = 0 for k = 0..n
    rows, rhs = [], []
    for k in range(n + 1):
        row = []
        for j in range(n + 1):
            row.append(sum(c * moment(i + j + k) for i, c in enumerate(pn)))
        rows.append(row)
        rhs.append(-sum(c * moment(i + n + 1 + k) for i, c in enumerate(pn)))
    c = solve_rational(rows, rhs)
    stieltjes = c + [Fraction(1)]
    kron_extra = real_roots(stieltjes)
    nodes = sorted(gauss + kron_extra)
    wk = weights_for(nodes, 3 * n + 1)
    wg_only = weights_for(gauss, 2 * n - 1)
    wg = []
    for x in nodes:
        hit = [w for g, w in zip(gauss, wg_only) if abs(g - x) < mpmath.mpf(10) ** -40]
        wg.append(hit[0] if hit else mpmath.mpf(0))
    return nodes, wk, wg


def tensor_gk(n):
    nodes, wk, wg = gauss_kronrod(n)
    half = mpmath.mpf(1) / 2
    xs = [x * half for x in nodes]
    k = [w * half for w in wk]
    e = [(a - b) * half for a, b in zip(wk, wg)]
    pts, wht, err = [], [], []
    for i in range(len(xs)):
        for j in range(len(xs)):
            pts.append((xs[i], xs[j]))
            wht.append(k[i] * k[j])
            err.append(e[i] * e[j])
    return pts, wht, err


def genz_malik_75():
    """Degree-7 Genz-Malik rule with embedded degree-5 rule, n = 2, unit square."""
    n = 2
    l2 = mpmath.sqrt(mpmath.mpf(9) / 70) / 2
    l3 = mpmath.sqrt(mpmath.mpf(9) / 10) / 2
    l5 = mpmath.sqrt(mpmath.mpf(9) / 19) / 2
    f = lambda a, b: mpmath.mpf(a) / b
    w7 = [f(12824 - 9120 * n + 400 * n * n, 19683), f(980, 6561), f(1820 - 400 * n, 19683),
          f(200, 19683), f(6859, 19683) / 2 ** n]
    w5 = [f(729 - 950 * n + 50 * n * n, 729), f(245, 486), f(265 - 100 * n, 1458), f(25, 729), 0]
    z = mpmath.mpf(0)
    groups = [
        [(z, z)],
        [(-l2, z), (z, -l2), (z, l2), (l2, z)],
        [(-l3, z), (z, -l3), (z, l3), (l3, z)],
        [(-l3, -l3), (-l3, l3), (l3, -l3), (l3, l3)],
        [(-l5, -l5), (-l5, l5), (l5, -l5), (l5, l5)],
    ]
    pts, wht, err = [], [], []
    for g, a, b in zip(groups, w7, w5):
        for p in g:
            pts.append(p)
            wht.append(a)
            err.append(a - b)
    return pts, wht, err


def subsample(m):
    xs = [(mpmath.mpf(2 * i + 1) / (2 * m)) - mpmath.mpf(1) / 2 for i in range(m)]
    pts = [(x, y) for x in xs for y in xs]
    w = mpmath.mpf(1) / (m * m)
    return pts, [w] * len(pts), [mpmath.mpf(0)] * len(pts)


RULES = [
    ("point", "single point sampling without error estimate", lambda: subsample(1)),
    ("sub2", "2x2 subsampling without error estimate", lambda: subsample(2)),
    ("sub4", "4x4 subsampling without error estimate", lambda: subsample(4)),
    ("gm75", "Genz-Malik (7, 5) fully symmetric rule", genz_malik_75),
    ("g3k7", "Gauss-Kronrod (7, 3) Cartesian rule", lambda: tensor_gk(3)),
    ("g5k11", "Gauss-Kronrod (11, 5) Cartesian rule", lambda: tensor_gk(5)),
    ("g7k15", "Gauss-Kronrod (15, 7) Cartesian rule", lambda: tensor_gk(7)),
]


def fmt(x):
    s = mpmath.nstr(x, 21, min_fixed=-4, max_fixed=4)
    if s in ("0.0", "-0.0"):
        return "0.0"
    return s


def main():
    out = ["// Generated by gen_quad.py from the mathematics of each rule; do not edit.",
           "#pragma once", "", "namespace lensed_quad {", "",
           "struct Rule { const char* name; const char* info; int size;",
           "              const double (*pts)[2]; const double* wht; const double* err; };", ""]
    for name, info, fn in RULES:
        pts, wht, err = fn()
        assert abs(sum(wht) - 1) < mpmath.mpf(10) ** -40
        assert abs(sum(err)) < mpmath.mpf(10) ** -40
        out.append(f"static const double {name.upper()}_PTS[][2] = {{")
        out += [f"    {{{fmt(p[0])}, {fmt(p[1])}}}," for p in pts]
        out.append("};")
        out.append(f"static const double {name.upper()}_WHT[] = {{")
        out += [f"    {fmt(w)}," for w in wht]
        out.append("};")
        out.append(f"static const double {name.upper()}_ERR[] = {{")
        out += [f"    {fmt(w)}," for w in err]
        out.append("};")
        out.append("")
    out.append("static const Rule RULES[] = {")
    for name, info, fn in RULES:
        u = name.upper()
        out.append(f'    {{"{name}", "{info}", (int)(sizeof({u}_WHT)/sizeof(double)), {u}_PTS, {u}_WHT, {u}_ERR}},')
    out.append("    {nullptr, nullptr, 0, nullptr, nullptr, nullptr}")
    out.append("};")
    out.append("")
    out.append("}  // namespace lensed_quad")
    print("\n".join(out))


if __name__ == "__main__":
    main()
