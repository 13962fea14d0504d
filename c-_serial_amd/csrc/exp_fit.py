"""Derives the polynomial physics.hpp's exp_r uses for e^r on
|r| <= ln(2)/2: p(r) = 1 + r + c2 r^2 + ... + c10 r^10, fitted in relative
error by least squares on Chebyshev nodes with weights raised where the error
peaks (the usual route towards the minimax fit), in 50-digit arithmetic.
Prints the relative error of p with its coefficients rounded to double, and
the coefficients c2 .. c11.

    python exp_fit.py          # degree 10
    python exp_fit.py 11       # another degree
"""
import sys

import mpmath as mp

mp.mp.dps = 50
H = mp.log(2) / 2


def fit(deg, nodes=300, sweeps=6):
    xs = [H * mp.cos(mp.pi * (i + 0.5) / nodes) for i in range(nodes)]

    def solve(w):
        rows = [[wi * x ** j for j in range(2, deg + 1)] for x, wi in zip(xs, w)]
        rhs = [wi * (mp.exp(x) - 1 - x) for x, wi in zip(xs, w)]
        sol = mp.qr_solve(mp.matrix(rows), mp.matrix(rhs))[0]
        return [sol[j] for j in range(deg - 1)]

    def rel(c, x):
        return (1 + x + sum(cj * x ** (j + 2) for j, cj in enumerate(c))) / mp.exp(x) - 1

    w = [1 / mp.exp(x) for x in xs]
    c = solve(w)
    for _ in range(sweeps):
        err = [rel(c, x) for x in xs]
        top = max(abs(e) for e in err)
        w = [wi * (1 + 50 * abs(e) / top) for wi, e in zip(w, err)]
        c = solve(w)
    cd = [mp.mpf(float(v)) for v in c]
    grid = [-H + 2 * H * i / 4000 for i in range(4001)]
    return c, max(abs(rel(cd, x)) for x in grid)


if __name__ == "__main__":
    deg = int(sys.argv[1]) if len(sys.argv) > 1 else 10
    c, err = fit(deg)
    print(f"degree {deg} on [-ln2/2, ln2/2]: max relative error {mp.nstr(err, 4)}")
    print("c2..:", [float(v) for v in c])
