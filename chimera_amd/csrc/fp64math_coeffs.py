"""Polynomial coefficients of fp64math.h, computed in 60-digit decimal arithmetic.

    python fp64math_coeffs.py

P(z): interpolates sin(pi sqrt z) / sqrt z (so sin(pi u) = u P(u^2)) at the 9 Chebyshev
nodes of z in [0, 1/4] (|u| <= 1/2).  Q(r): interpolates e^r at the 12 Chebyshev nodes of
[-ln 2 / 2, ln 2 / 2].  Interpolation at Chebyshev nodes is within a small factor of the
minimax polynomial of the same degree (Chebyshev economisation); the script prints the
coefficients (monomial form, constant term first) rounded to double and the largest
interpolation error on a dense grid.
"""
from decimal import Decimal as D, getcontext

getcontext().prec = 60
PI = D("3.14159265358979323846264338327950288419716939937510582097446")


def sin_d(x):
    s, t, n = D(0), x, 1
    while abs(t) > D(10) ** -58:
        s += t
        t = -t * x * x / ((n + 1) * (n + 2))
        n += 2
    return s


def cos_d(x):
    return sin_d(PI / 2 - x)


def interpolate(f, a, b, deg):
    """Monomial coefficients of the degree-`deg` interpolant of f at Chebyshev nodes."""
    n = deg + 1
    xs = [(a + b) / 2 + (b - a) / 2 * cos_d(PI * (D(k) + D("0.5")) / n) for k in range(n)]
    rows = [[x ** j for j in range(n)] + [f(x)] for x in xs]
    for i in range(n):                       # Gauss-Jordan with partial pivoting
        piv = max(range(i, n), key=lambda r: abs(rows[r][i]))
        rows[i], rows[piv] = rows[piv], rows[i]
        for r in range(n):
            if r != i:
                m = rows[r][i] / rows[i][i]
                rows[r] = [u - m * v for u, v in zip(rows[r], rows[i])]
    return [rows[i][n] / rows[i][i] for i in range(n)]


def max_error(f, c, a, b, npts=400):
    err = D(0)
    for q in range(npts + 1):
        x = a + (b - a) * q / npts
        p = D(0)
        for cj in reversed(c):
            p = p * x + cj
        err = max(err, abs(p - f(x)))
    return err


def main():
    g = lambda z: sin_d(PI * z.sqrt()) / z.sqrt() if z > 0 else PI
    P = interpolate(g, D(0), D("0.25"), 8)
    print("P", [repr(float(c)) for c in P], "max |err| * 1/2:",
          float(max_error(g, P, D(0), D("0.25")) / 2))
    h = D(2).ln() / 2
    Q = interpolate(lambda r: r.exp(), -h, h, 11)
    print("Q", [repr(float(c)) for c in Q], "max rel err:",
          float(max_error(lambda r: r.exp(), Q, -h, h) / (-h).exp()))


if __name__ == "__main__":
    main()
