"""Pins the CPU oracle (the checker) against the reference's golden data, the
compiled reference sources, and the mathematics.  CPU only."""
import json
import math

import mpmath as mp
import numpy as np
import pytest

import oracle
from conftest import golden
from libaccint import systems

LIT = json.load(open(golden("reference_literals.json")))


# ---------------------------------------------------------------------- Boys
def test_boys_vs_mpmath_reference_values():
    data = json.load(open(golden("boys_reference.json")))
    worst = 0.0
    for pt in data["test_points"]:
        T = float(pt["T"])
        ref = np.array([float(v) for v in pt["values"]])
        got = oracle.boys(len(ref) - 1, T)
        worst = max(worst, float(np.max(np.abs(got - ref) / np.abs(ref))))
    assert worst < 5e-14


@pytest.mark.parametrize("T", [1e-3, 0.5, 3.7, 12.0, 29.5, 35.99, 36.0, 48.0])
def test_boys_erf_closed_form(T):
    # F_0 = sqrt(pi/T) erf(sqrt T)/2, F_1 = (F_0 - e^-T)/(2T)  (reference test_boys_function.cpp:29-37)
    f = oracle.boys(1, T)
    f0 = 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
    assert abs(f[0] - f0) <= 2e-15 * f0
    f1 = float(mp.quad(lambda t: t ** 2 * mp.exp(-T * t ** 2), [0, 1]))
    assert abs(f[1] - f1) <= 1e-13 * f1


def test_boys_zero():
    f = oracle.boys(30, 0.0)
    assert np.allclose(f, 1.0 / (2 * np.arange(31) + 1), rtol=0, atol=0)


# ----------------------------------------------------------------------- Rys
from exact import rys_exact as _rys_exact  # noqa: E402


@pytest.mark.parametrize("n,T", [(1, 0.7), (2, 3.3), (3, 0.0), (3, 11.0), (4, 25.0), (5, 2.0)])
def test_rys_vs_exact(n, T):
    r, w = oracle.rys(n, T)
    re, we = _rys_exact(n, T)
    assert np.max(np.abs(r - re)) < 5e-12
    assert np.max(np.abs(w - we)) < 5e-12 * np.max(we)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_rys_moment_identity(n):
    # sum_i w_i u_i^k = F_k(T) for k < 2n (reference test_rys_quadrature.cpp:263-295)
    for T in (0.1, 4.0, 22.0, 60.0):
        r, w = oracle.rys(n, T)
        F = oracle.boys(2 * n - 1, T)
        for k in range(2 * n):
            assert abs(np.sum(w * r ** k) - F[k]) <= 1e-11 * F[k]


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_rys_T0_is_gauss_legendre(n):
    # T = 0: even Legendre polynomials; u_i = x_i^2 of the positive 2n-point nodes
    x, wt = np.polynomial.legendre.leggauss(2 * n)
    u = np.sort(x[x > 0] ** 2)
    r, w = oracle.rys(n, 0.0)
    assert np.allclose(r, u, atol=1e-13)
    assert np.allclose(np.sort(w), np.sort(wt[x > 0]), atol=1e-13)


def test_rys_vs_reference_data():
    # the reference's own C++ outputs (tests/data/rys_reference.json): same
    # algorithm; agreement is bounded by the moment method's conditioning
    d = json.load(open(golden("rys_reference.json")))["rys_quadrature"]["data"]
    tol = {1: 1e-15, 2: 1e-14, 3: 1e-13, 4: 1e-12, 5: 1e-11, 6: 1e-10, 7: 3e-9, 8: 1e-7, 9: 1e-5}
    for n_s, pts in d.items():
        n = int(n_s)
        for T_s, rw in pts.items():
            r, w = oracle.rys(n, float(T_s))
            rr = np.array([float(v) for v in rw["roots"]])
            ww = np.array([float(v) for v in rw["weights"]])
            assert np.max(np.abs(r - rr)) <= tol[n], (n, T_s)
            assert np.max(np.abs(w - ww)) <= tol[n] * np.max(ww), (n, T_s)


# ----------------------------------------------------------------- integrals
def _h2o():
    rb = oracle.RawBasis("sto-3g", systems.h2o_sto3g_geometry())
    return rb, oracle.Oracle(rb)


def test_eri_pyscf_literals():
    rb, o = _h2o()
    E = o.eri_tensor()
    for i, j, k, l, v in LIT["h2o_eri"]:
        assert abs(E[i, j, k, l] - v) < 1e-12
    h2 = oracle.Oracle(oracle.RawBasis("sto-3g", LIT["h2_geometry"]))
    E2 = h2.eri_tensor()
    assert abs(E2[0, 0, 0, 0] - LIT["h2_eri"]["0000"]) < 1e-12
    assert abs(E2[0, 0, 1, 1] - LIT["h2_eri"]["0011"]) < 1e-12
    assert abs(E2[0, 1, 0, 1] - LIT["h2_eri"]["0101"]) < 1e-12


def test_eri_eightfold_and_schwarz():
    rb = oracle.RawBasis("cc-pvdz", systems.h2o_sto3g_geometry())
    o = oracle.Oracle(rb)
    rng = np.random.default_rng(3)
    for _ in range(25):
        i, j, k, l = rng.integers(0, rb.n_shells, 4)
        g = o.eri_quartet(i, j, k, l)
        assert np.allclose(g, o.eri_quartet(j, i, k, l).transpose(1, 0, 2, 3), atol=1e-13)
        assert np.allclose(g, o.eri_quartet(k, l, i, j).transpose(2, 3, 0, 1), atol=1e-13)
        assert np.allclose(g, o.eri_quartet(i, j, l, k).transpose(0, 1, 3, 2), atol=1e-13)
        dij = np.einsum("abab->ab", o.eri_quartet(i, j, i, j))
        dkl = np.einsum("abab->ab", o.eri_quartet(k, l, k, l))
        bound = np.sqrt(dij.reshape(-1)[:, None] * dkl.reshape(-1)[None, :])
        assert np.all(np.abs(g.reshape(len(dij.reshape(-1)), -1)) <= bound * (1 + 1e-10) + 1e-14)


def test_ssss_closed_form():
    # (ss|ss) = 2 pi^(5/2)/(z e sqrt(z+e)) K_AB K_CD F_0(T) for single primitives
    a, b, c, d = 1.3, 0.7, 0.45, 2.1
    A, B, C, D = map(np.array, ([0, 0, 0.0], [0.3, -0.4, 0.9], [1.1, 0.2, -0.5], [-0.6, 0.8, 0.3]))
    atoms = [(1, tuple(X)) for X in (A, B, C, D)]
    rb = oracle.RawBasis("sto-3g", atoms)
    rb.L[:] = 0
    rb.nprim[:] = 1
    rb.exps = np.array([a, b, c, d])
    rb.raw_coefs = np.ones(4)
    o = oracle.Oracle(rb)
    coef, _ = o.normalized()
    z, e = a + b, c + d
    P, Q = (a * A + b * B) / z, (c * C + d * D) / e
    T = z * e / (z + e) * np.sum((P - Q) ** 2)
    ref = (2 * math.pi ** 2.5 / (z * e * math.sqrt(z + e)) * math.exp(-a * b / z * np.sum((A - B) ** 2)) *
           math.exp(-c * d / e * np.sum((C - D) ** 2)) * 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T)))
    assert abs(o.eri_quartet(0, 1, 2, 3)[0, 0, 0, 0] - ref * np.prod(coef)) < 1e-14


def test_one_electron_pyscf_literals():
    rb, o = _h2o()
    S, T, V = o.overlap(), o.kinetic(), o.nuclear()
    assert np.max(np.abs(S.ravel() - LIT["h2o_S"])) < 1e-12
    assert np.max(np.abs(T.ravel() - LIT["h2o_T"])) < 1e-12
    assert np.max(np.abs(V.ravel() - LIT["h2o_V"])) < 1e-12
    assert np.allclose(np.diag(S), 1.0, atol=1e-14)


def test_water_aug_cc_pvtz_pyscf():
    # PySCF (Cartesian, LibAccInt ordering) S/T/V and sampled ERIs for s..f shells
    w = np.load(golden("water_aug_cc_pvtz.npz"))
    atoms = [(8, w["geometry_bohr"][0]), (1, w["geometry_bohr"][1]), (1, w["geometry_bohr"][2])]
    rb = oracle.RawBasis("aug-cc-pvtz", atoms)
    o = oracle.Oracle(rb)
    assert rb.nbf == w["S"].shape[0]
    assert np.max(np.abs(o.overlap() - w["S"])) < 1e-12
    assert np.max(np.abs(o.kinetic() - w["T"])) < 1e-11
    assert np.max(np.abs(o.nuclear() - w["V"])) < 1e-11
    _, fo = o.normalized()
    nf = [(l + 1) * (l + 2) // 2 for l in rb.L]
    sh_of = np.repeat(np.arange(rb.n_shells), nf)
    for (p, q, r, s), v in zip(w["eri_idx"], w["eri_val"]):
        blk = o.eri_quartet(sh_of[p], sh_of[q], sh_of[r], sh_of[s])
        got = blk[p - fo[sh_of[p]], q - fo[sh_of[q]], r - fo[sh_of[r]], s - fo[sh_of[s]]]
        assert abs(got - v) < 1e-11


# ----------------------------------------------------------------------- J/K
def test_jk_symmetric_equals_full(density7):
    rb, o = _h2o()
    J1, K1 = o.jk_canonical(density7)
    J2, K2 = o.jk_full(density7, n_threads=2)
    E = o.eri_tensor()
    assert np.allclose(J1, J2, atol=1e-13) and np.allclose(K1, K2, atol=1e-13)
    assert np.allclose(J1, np.einsum("ijkl,kl->ij", E, density7), atol=1e-13)
    assert np.allclose(K1, np.einsum("ijkl,jl->ik", E, density7), atol=1e-13)


def test_h2_j00_closed_form():
    o = oracle.Oracle(oracle.RawBasis("sto-3g", LIT["h2_geometry"]))
    D = np.eye(2)
    J, _ = o.jk_canonical(D)
    assert abs(J[0, 0] - (LIT["h2_eri"]["0000"] + LIT["h2_eri"]["0011"])) < 1e-12


def test_screened_list_counts():
    rb = oracle.RawBasis("cc-pvdz", systems.water_cluster(2, 8.0, seed=3))
    o = oracle.Oracle(rb)
    Q = o.schwarz()
    n_all, _ = o.screened_list(Q, 0.0)
    npair = rb.n_shells * (rb.n_shells + 1) // 2
    assert n_all == npair * (npair + 1) // 2
    n12, _ = o.screened_list(Q, 1e-12)
    n6, _ = o.screened_list(Q, 1e-6)
    assert n6 < n12 <= n_all


# --------------------------------------------------- fixtures of the reference
def test_oracle_vs_reference_run_h2o_sto3g(density7):
    ref = np.load(golden("ref_h2o_sto3g.npz"))
    rb, o = _h2o()
    assert np.max(np.abs(o.eri_tensor() - ref["eri"])) < 1e-14
    for name, M in (("S", o.overlap()), ("T", o.kinetic()), ("V", o.nuclear())):
        assert np.max(np.abs(M - ref[name])) < 1e-13
    J, K = o.jk_full(ref["D"], n_threads=1)
    assert np.max(np.abs(J - ref["J"])) < 1e-13 and np.max(np.abs(K - ref["K"])) < 1e-13


@pytest.mark.skipif(oracle.ref_lib() is None, reason="oracle/_ref not built (needs the reference sources)")
def test_oracle_vs_compiled_reference():
    rb = oracle.RawBasis("cc-pvtz", systems.h2o_sto3g_geometry())
    o, r = oracle.Oracle(rb), oracle.Oracle(rb, use_ref=True)
    c1, f1 = o.normalized()
    c2, f2 = r.normalized()
    assert np.max(np.abs(c1 - c2)) <= 1e-14 * np.max(np.abs(c2)) and np.all(f1 == f2)
    rng = np.random.default_rng(11)
    for _ in range(40):
        q = rng.integers(0, rb.n_shells, 4)
        L = rb.L[q].sum()
        tol = 1e-12 if L <= 8 else 1e-10  # n_roots >= 6: the moment method's own conditioning
        assert np.max(np.abs(o.eri_quartet(*q) - r.eri_quartet(*q))) < tol
    for T in (0.0, 0.3, 2.5, 17.0, 40.0):
        assert np.max(np.abs(oracle.boys(20, T) - oracle.boys(20, T, use_ref=True))) < 1e-15
