"""Exact Rys quadrature for test checks: moments F_k(T) = 1F1(k+1/2; k+3/2; -T)/(2k+1)
at 90 digits, the Chebyshev (moment) algorithm at that precision and the
eigen-decomposition of the Jacobi matrix (Golub-Welsch)."""
import mpmath as mp
import numpy as np


def boys_exact(nmax, T):
    mp.mp.dps = 90
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
    F = boys_exact(2 * n - 1, T)
    a, b = [0] * n, [0] * n
    prev, cur = [mp.mpf(0)] * (2 * n), list(F)
    a[0], b[0] = F[1] / F[0], F[0]
    for l in range(1, n):
        nxt = [mp.mpf(0)] * (2 * n)
        for k in range(l, 2 * n - l):
            nxt[k] = cur[k + 1] - a[l - 1] * cur[k] - b[l - 1] * prev[k]
        a[l] = nxt[l + 1] / nxt[l] - cur[l] / cur[l - 1]
        b[l] = nxt[l] / cur[l - 1]
        prev, cur = cur, nxt
    Jm = mp.matrix(n, n)
    for i in range(n):
        Jm[i, i] = a[i]
        if i + 1 < n:
            Jm[i, i + 1] = Jm[i + 1, i] = mp.sqrt(b[i + 1])
    E, Q = mp.eigsy(Jm)
    idx = sorted(range(n), key=lambda i: E[i])
    return (np.array([float(E[i]) for i in idx]), np.array([float(b[0] * Q[0, i] ** 2) for i in idx]))
