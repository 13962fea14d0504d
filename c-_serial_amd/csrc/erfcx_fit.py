"""Derives the rational approximation of erfcx(t) = exp(t^2) erfc(t) on
[0, 6] that physics.hpp's erf_x uses: R = P/Q with deg P = deg Q = 8, fitted
in relative error by linearised least squares on Chebyshev nodes, reweighted
by 1/Q a few times (the usual route towards the minimax fit), in 50-digit
arithmetic.  Prints the relative error of R on a fine grid and the
coefficients rounded to double, lowest degree first.

    python erfcx_fit.py            # degrees (8, 8) on [0, 6]
    python erfcx_fit.py 7 7 6      # other degrees / interval
"""
import sys

import mpmath as mp

mp.mp.dps = 50


def fit(m, k, x1, nodes=400, sweeps=6):
    xs = [x1 * (1 - mp.cos(mp.pi * (i + 0.5) / nodes)) / 2 for i in range(nodes)]
    fs = [mp.erfc(x) * mp.exp(x * x) for x in xs]

    def solve(w):
        # unknowns p0..pm, q1..qk (q0 = 1): sum p_j x^j - f sum_{j>=1} q_j x^j = f
        rows, rhs = [], []
        for x, f, wi in zip(xs, fs, w):
            rows.append([wi * x ** j for j in range(m + 1)] +
                        [-wi * f * x ** j for j in range(1, k + 1)])
            rhs.append(wi * f)
        sol = mp.qr_solve(mp.matrix(rows), mp.matrix(rhs))[0]
        return [sol[j] for j in range(m + 1)], [mp.mpf(1)] + [sol[m + 1 + j] for j in range(k)]

    ev = lambda c, x: sum(cj * x ** j for j, cj in enumerate(c))
    p, q = solve([1 / f for f in fs])
    for _ in range(sweeps):
        p, q = solve([1 / (f * ev(q, x)) for x, f in zip(xs, fs)])
    grid = [x1 * i / 3000 for i in range(3001)]
    err = max(abs(ev(p, x) / ev(q, x) / (mp.erfc(x) * mp.exp(x * x)) - 1) for x in grid)
    return p, q, err


if __name__ == "__main__":
    m = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    k = int(sys.argv[2]) if len(sys.argv) > 2 else 8
    x1 = mp.mpf(sys.argv[3]) if len(sys.argv) > 3 else mp.mpf(6)
    p, q, err = fit(m, k, x1)
    print(f"degrees ({m}, {k}) on [0, {x1}]: max relative error of R {mp.nstr(err, 4)}")
    print("P:", [float(c) for c in p])
    print("Q:", [float(c) for c in q])
