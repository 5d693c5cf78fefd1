"""Generate the device Rys root/weight tables (rys_table.inc) from the mathematics.

Method (this project's own; the reference instead solves moments -> Jacobi
matrix -> QL -> Newton per primitive quartet, src/shared/math/rys_quadrature.cpp:455-513):

* For n roots and T < T_cut(n): piecewise Chebyshev interpolants of every
  root u_i(T) and weight w_i(T) on unit intervals [k, k+1), DEG+1 terms each,
  evaluated on the device by Clenshaw.  The sample values come from the
  exact quadrature: moments F_k(T) = 1F1(k+1/2; k+3/2; -T)/(2k+1) at 90
  digits, the Chebyshev (moment) algorithm at that precision, and the
  eigen-decomposition of the resulting Jacobi matrix (Golub-Welsch).
* For T >= T_cut(n): the large-T limit of the Rys quadrature, the positive
  half of the 2n-point Gauss-Hermite rule: u_i = x_i^2 / T,
  w_i = W_i / sqrt(T).  T_cut(n) is the smallest integer from which this
  limit matches the exact rule to 2e-16 relative.

The script prints the worst relative error of the interpolants against the
exact rule at random points, per n.

    python gen_rys_table.py  ->  rys_table.inc (next to this script)
"""
import os
import random
import sys
from multiprocessing import Pool

import mpmath as mp

NMAX = 9
DEG = 13
DPS = 90


def boys_list(nmax, T):
    T = mp.mpf(T)
    if T == 0:
        return [mp.mpf(1) / (2 * k + 1) for k in range(nmax + 1)]
    F = [None] * (nmax + 1)
    F[nmax] = mp.hyp1f1(nmax + mp.mpf(1) / 2, nmax + mp.mpf(3) / 2, -T) / (2 * nmax + 1)
    e = mp.exp(-T)
    for k in range(nmax, 0, -1):
        F[k - 1] = (2 * T * F[k] + e) / (2 * k - 1)
    return F


def rys_exact(n, T):
    mp.mp.dps = DPS
    mom = boys_list(2 * n - 1, T)
    a = [mp.mpf(0)] * n
    b = [mp.mpf(0)] * n
    prev = [mp.mpf(0)] * (2 * n)
    cur = list(mom)
    a[0] = mom[1] / mom[0]
    b[0] = mom[0]
    for l in range(1, n):
        nxt = [mp.mpf(0)] * (2 * n)
        for k in range(l, 2 * n - l):
            nxt[k] = cur[k + 1] - a[l - 1] * cur[k] - b[l - 1] * prev[k]
        a[l] = nxt[l + 1] / nxt[l] - cur[l] / cur[l - 1]
        b[l] = nxt[l] / cur[l - 1]
        prev, cur = cur, nxt
    J = mp.matrix(n, n)
    for i in range(n):
        J[i, i] = a[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(b[i + 1])
    E, Q = mp.eigsy(J)
    idx = sorted(range(n), key=lambda i: E[i])
    return [E[i] for i in idx], [b[0] * Q[0, i] ** 2 for i in idx]


def hermite_half(n):
    """Positive nodes x_i and weights W_i of the 2n-point Gauss-Hermite rule."""
    mp.mp.dps = DPS
    m = 2 * n
    J = mp.matrix(m, m)
    for i in range(m - 1):
        J[i, i + 1] = J[i + 1, i] = mp.sqrt(mp.mpf(i + 1) / 2)
    E, Q = mp.eigsy(J)
    pairs = sorted((E[i], mp.sqrt(mp.pi) * Q[0, i] ** 2) for i in range(m) if E[i] > 0)
    return [p[0] for p in pairs], [p[1] for p in pairs]


def asym(n, T, hx, hw):
    T = mp.mpf(T)
    return [x * x / T for x in hx], [w / mp.sqrt(T) for w in hw]


def rel_err(ra, wa, rb, wb):
    e = 0.0
    for x, y in zip(ra, rb):
        e = max(e, float(abs(x - y) / abs(y)))
    ws = max(abs(y) for y in wb)
    for x, y in zip(wa, wb):
        e = max(e, float(abs(x - y) / ws))
    return e


def find_tcut(n):
    hx, hw = hermite_half(n)
    T = 8
    while True:
        ok = True
        for t in (T, T + 0.5, T + 1.5, T + 4):
            r, w = rys_exact(n, t)
            ra, wa = asym(n, t, hx, hw)
            if rel_err(ra, wa, r, w) > 2e-16:
                ok = False
                break
        if ok:
            return T
        T += 1


def cheb_nodes(k):
    return [mp.mpf(k) + (1 + mp.cos(mp.pi * (j + mp.mpf(1) / 2) / (DEG + 1))) / 2 for j in range(DEG + 1)]


def fit_interval(args):
    n, k = args
    mp.mp.dps = DPS
    N = DEG + 1
    xs = cheb_nodes(k)
    vals = [rys_exact(n, x) for x in xs]
    coefs = []
    for f in range(2 * n):
        fv = [vals[j][0][f] if f < n else vals[j][1][f - n] for j in range(N)]
        cs = []
        for m in range(N):
            s = mp.fsum(fv[j] * mp.cos(mp.pi * m * (j + mp.mpf(1) / 2) / N) for j in range(N))
            cs.append(2 * s / N)
        cs[0] /= 2
        coefs.append(cs)
    return n, k, coefs


def clenshaw(cs, x):
    d1 = d2 = 0.0
    for c in reversed(cs[1:]):
        d1, d2 = 2 * x * d1 - d2 + c, d1
    return cs[0] + x * d1 - d2


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        tcut = pool.map(find_tcut, range(1, NMAX + 1))
        tasks = [(n, k) for n in range(1, NMAX + 1) for k in range(tcut[n - 1])]
        res = pool.map(fit_interval, tasks)
    table = {(n, k): c for n, k, c in res}
    # accuracy report against the exact rule
    rnd = random.Random(7)
    for n in range(1, NMAX + 1):
        worst = 0.0
        for _ in range(60):
            T = rnd.uniform(0, tcut[n - 1])
            k = int(T)
            x = 2 * (T - k) - 1
            cs = table[(n, k)]
            r = [clenshaw([float(c) for c in cs[i]], x) for i in range(n)]
            w = [clenshaw([float(c) for c in cs[n + i]], x) for i in range(n)]
            re, we = rys_exact(n, T)
            worst = max(worst, rel_err(r, w, re, we))
        print("n=%d T_cut=%d max rel err %.2e" % (n, tcut[n - 1], worst), file=sys.stderr)
    # emit
    offs = []
    flat = []
    for n in range(1, NMAX + 1):
        offs.append(len(flat))
        for k in range(tcut[n - 1]):
            for f in range(2 * n):
                flat.extend(table[(n, k)][f])
    hx_all, hw_all, hoff = [], [], []
    for n in range(1, NMAX + 1):
        hx, hw = hermite_half(n)
        hoff.append(len(hx_all))
        hx_all.extend(x * x for x in hx)
        hw_all.extend(hw)
    g = lambda v: mp.nstr(v, 20, min_fixed=1, max_fixed=0)
    lines = [
        "// Generated by gen_rys_table.py (exact Rys quadrature at %d digits); do not edit." % DPS,
        "// Layout: for n = 1..%d, for interval k = 0..T_cut(n)-1, for function f = 0..2n-1" % NMAX,
        "// (roots u_0..u_{n-1}, then weights w_0..w_{n-1}): %d Chebyshev terms," % (DEG + 1),
        "// constant term halved, in x = 2(T-k)-1.",
        "#pragma once",
        "#define LAI_RYS_NMAX %d" % NMAX,
        "#define LAI_RYS_NTERMS %d" % (DEG + 1),
        "inline constexpr int lai_rys_tcut[%d] = {%s};" % (NMAX + 1, ", ".join(["0"] + [str(t) for t in tcut])),
        "inline constexpr int lai_rys_off[%d] = {%s};" % (NMAX + 1, ", ".join(["0"] + [str(o) for o in offs])),
        "inline constexpr int lai_rys_hoff[%d] = {%s};" % (NMAX + 1, ", ".join(["0"] + [str(o) for o in hoff])),
        "// x_i^2 and W_i of the positive half of the 2n-point Gauss-Hermite rule",
        "inline constexpr double lai_rys_hx2[%d] = {%s};" % (len(hx_all), ", ".join(g(v) for v in hx_all)),
        "inline constexpr double lai_rys_hw[%d] = {%s};" % (len(hw_all), ", ".join(g(v) for v in hw_all)),
        "// device table = lai_rys_cheb_h followed by lai_rys_hx2 and lai_rys_hw",
        "#define LAI_RYS_CHEB_SIZE %d" % len(flat),
        "#define LAI_RYS_HX2_OFF %d" % len(flat),
        "#define LAI_RYS_HW_OFF %d" % (len(flat) + len(hx_all)),
        "#define LAI_RYS_TABLE_SIZE %d" % (len(flat) + 2 * len(hx_all)),
        "static const double lai_rys_cheb_h[%d] = {" % len(flat),
    ]
    for i in range(0, len(flat), 4):
        lines.append("  " + ", ".join(g(v) for v in flat[i:i + 4]) + ",")
    lines.append("};")
    with open(os.path.join(here, "rys_table.inc"), "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
