"""The density-fitting oracle and the host side of DFFockBuilder (CPU only).

What pins oracle.df_jk (oracle/__init__.py), each check chosen so that a
plausible mistake fails one of them:

* exact fit: when the fitting functions span every product of two orbital
  functions on one centre, the Coulomb-metric fit is exact, so DF-J/K equal
  the four-centre J/K of the C restatement (a wrong prefactor or index of the
  two- or three-centre integrals, a transposed B or a dropped term breaks it);
* the reference's three-centre kernel (kernels::compute_three_center_tensor,
  compiled into oracle/_ref) on the classes it computes correctly (s fitting
  functions, bra L_a >= L_b);
* the reference's own DFFockBuilder contractions: given the reference's metric
  and three-centre tensor and its B convention, df_jk_from_tensors reproduces
  DFFockBuilder::compute_coulomb / compute_exchange;
* the Coulomb-metric fit is variational: E_J(DF) <= E_J for a PSD density.

The reference's metric kernel (two_center_coulomb_kernel.cpp:96-180) and its
three-centre kernel for p and higher fitting functions do not compute the
Coulomb integrals (see test_reference_two_centre_kernel_is_not_the_coulomb_metric),
and compute_B_tensor multiplies by L^-1 from the other side, so the
reference's DF J/K are not the parity target; the documented fit
(df_fock_builder.hpp:9-13) is.
"""
import numpy as np
import pytest

import oracle

WATER = [(8, (0.0, 0.0, 0.0)), (1, (0.0, 1.43233673, -1.10866041)), (1, (0.0, -1.43233673, -1.10866041))]


def _density(n, nocc, seed):
    C = np.random.default_rng(seed).standard_normal((n, nocc)) / np.sqrt(n)
    return C @ C.T


def exact_fit_bases(center=(0.1, -0.2, 0.3)):
    """Orbital s(0.5), s(3.0), p(0.8) on one centre; fitting functions with
    the exponent sums of every product: s(1.0, 3.5, 6.0), p(1.3, 3.8),
    Cartesian d(1.6) (x_i x_j e^{-1.6 r^2}, the p x p products)."""
    orb = oracle.RawBasis.from_shells([(0, center, [0.5], [1.0]), (0, center, [3.0], [1.0]),
                                       (1, center, [0.8], [1.0])])
    aux = oracle.RawBasis.from_shells([(0, center, [e], [1.0]) for e in (1.0, 3.5, 6.0)] +
                                      [(1, center, [e], [1.0]) for e in (1.3, 3.8)] + [(2, center, [1.6], [1.0])])
    return orb, aux


def placeholder_fitting_basis(atoms):
    """One s shell of four primitives per atom, the set make_df_fock_builder
    builds without atom data (df_fock_builder.cpp:491-499)."""
    return oracle.RawBasis.from_shells([(0, xyz, [10.0, 2.0, 0.5, 0.1], [0.25] * 4) for _, xyz in atoms])


def test_exact_fit_equals_four_centre_jk():
    orb, aux = exact_fit_bases()
    D = _density(orb.nbf, 2, 5)
    J, K = oracle.df_jk(orb, aux, D)
    J4, K4 = oracle.Oracle(orb).jk_canonical(D)
    assert np.abs(J - J4).max() < 1e-9 * np.abs(J4).max()
    assert np.abs(K - K4).max() < 1e-9 * np.abs(K4).max()
    # the other side of B breaks it
    V, T = oracle.df_tensors(orb, aux)
    Jr, _ = oracle.df_jk_from_tensors(V, T, D, b_side="right")
    assert np.abs(Jr - J4).max() > 1e-3 * np.abs(J4).max()


def test_exact_fit_uhf():
    orb, aux = exact_fit_bases()
    Da, Db = _density(orb.nbf, 2, 6), _density(orb.nbf, 1, 7)
    J, K, Ka, Kb = oracle.df_jk(orb, aux, Da, D_beta=Db)
    O = oracle.Oracle(orb)
    J4a, K4a = O.jk_canonical(Da)
    J4b, K4b = O.jk_canonical(Db)
    assert np.allclose(J, J4a + J4b, atol=1e-9 * np.abs(J).max(), rtol=0)
    assert np.allclose(Ka, K4a, atol=1e-9 * np.abs(Ka).max(), rtol=0)
    assert np.allclose(Kb, K4b, atol=1e-9 * np.abs(Kb).max(), rtol=0)
    assert np.allclose(K, Ka + Kb, atol=0, rtol=0)


def test_coulomb_fit_is_variational_and_close():
    orb = oracle.RawBasis("cc-pvdz", WATER)
    aux = oracle.RawBasis("cc-pvdz-rifit", WATER)
    D = _density(orb.nbf, 5, 1)
    J, K = oracle.df_jk(orb, aux, D)
    J4, K4 = oracle.Oracle(orb).jk_canonical(D)
    EJ, EJ4 = 0.5 * np.sum(D * J), 0.5 * np.sum(D * J4)
    assert EJ <= EJ4
    assert (EJ4 - EJ) < 1e-3 * EJ4
    assert np.abs(K - K4).max() < 1e-2 * np.abs(K4).max()
    assert np.allclose(J, J.T, atol=1e-13) and np.allclose(K, K.T, atol=1e-13)


def test_fitted_density_is_the_diagonal_of_b_times_d():
    """gamma_P = sum_ab B_P,ab D_ab equals sum_a (B_P D)_aa for a D that is not
    symmetric, because each B_P is symmetric in (a, b): the device's J+K build
    takes gamma from the diagonal of X = B D.  A B_P with an antisymmetric
    part breaks the identity, so the check has teeth."""
    orb = oracle.RawBasis("cc-pvdz", WATER)
    aux = oracle.RawBasis("cc-pvdz-rifit", WATER)
    V, T = oracle.df_tensors(orb, aux)
    B = np.einsum("PQ,abQ->Pab", np.linalg.inv(np.linalg.cholesky(V)), T)
    assert np.abs(B - B.transpose(0, 2, 1)).max() < 1e-12 * np.abs(B).max()
    D = np.random.default_rng(3).standard_normal((orb.nbf, orb.nbf))
    gamma = np.einsum("Pab,ab->P", B, D)
    assert np.abs(np.einsum("Paa->P", B @ D) - gamma).max() < 1e-12 * np.abs(gamma).max()
    A = np.random.default_rng(4).standard_normal(B.shape)
    B2 = B + 1e-3 * (A - A.transpose(0, 2, 1))
    gap = np.einsum("Paa->P", B2 @ D) - np.einsum("Pab,ab->P", B2, D)
    assert np.abs(gap).max() > 1e-6 * np.abs(gamma).max()


def test_metric_is_symmetric_positive_definite():
    aux = oracle.RawBasis("cc-pvdz-rifit", WATER)
    orb = oracle.RawBasis("sto-3g", WATER)
    for metric in ("coulomb", "overlap"):
        V, _ = oracle.df_tensors(orb, aux, metric)
        assert np.allclose(V, V.T, atol=1e-14 * np.abs(V).max())
        assert np.linalg.eigvalsh(V).min() > 0


def test_two_centre_ss_closed_form():
    # (P|Q) of two normalised s primitives = 2 pi^(5/2) / (a b sqrt(a+b)) N_a N_b F0(ab/(a+b) R^2)
    import math
    a, b, R = 1.3, 0.7, (0.0, 0.4, -1.1)
    aux = oracle.RawBasis.from_shells([(0, (0, 0, 0), [a], [1.0]), (0, R, [b], [1.0])])
    orb = oracle.RawBasis.from_shells([(0, (0, 0, 0), [1.0], [1.0])])
    V, _ = oracle.df_tensors(orb, aux)
    R2 = sum(x * x for x in R)
    T = a * b / (a + b) * R2
    F0 = 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
    N = lambda e: (2 * e / math.pi) ** 0.75
    ref = 2 * math.pi ** 2.5 / (a * b * math.sqrt(a + b)) * N(a) * N(b) * F0
    assert abs(V[0, 1] - ref) < 1e-14 * ref


def _need_ref():
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")


def test_three_centre_matches_reference_for_s_fitting_functions():
    _need_ref()
    orb = oracle.RawBasis("cc-pvdz", WATER)
    aux = oracle.RawBasis("cc-pvdz-rifit", WATER)
    _, T = oracle.df_tensors(orb, aux)
    _, Tr = oracle.ref_df_tensors(orb, aux)
    nf = lambda L: [(l + 1) * (l + 2) // 2 for l in L]
    fo = np.concatenate([[0], np.cumsum(nf(orb.L))])
    fp = np.concatenate([[0], np.cumsum(nf(aux.L))])
    checked = 0
    tmax = np.abs(Tr).max()
    for i in range(orb.n_shells):
        for j in range(orb.n_shells):
            if orb.L[i] < orb.L[j]:
                continue
            for P in range(aux.n_shells):
                if aux.L[P] != 0:
                    continue
                blk = T[fo[i]:fo[i + 1], fo[j]:fo[j + 1], fp[P]:fp[P + 1]]
                ref = Tr[fo[i]:fo[i + 1], fo[j]:fo[j + 1], fp[P]:fp[P + 1]]
                # relative 1e-11 per block, absolute 1e-14 of the largest integral
                assert np.abs(blk - ref).max() <= 1e-11 * np.abs(ref).max() + 1e-14 * tmax
                checked += 1
    assert checked > 100


def test_contractions_match_reference_builder():
    _need_ref()
    orb = oracle.RawBasis("sto-3g", WATER)
    aux = placeholder_fitting_basis(WATER)
    D = _density(orb.nbf, 3, 3)
    Vr, Tr = oracle.ref_df_tensors(orb, aux)
    J, K = oracle.df_jk_from_tensors(Vr, Tr, D, b_side="right")
    Jr, Kr = oracle.ref_df_jk(orb, aux, D)
    assert np.abs(J - Jr).max() <= 1e-13 * np.abs(Jr).max()
    assert np.abs(K - Kr).max() <= 1e-13 * np.abs(Kr).max()
    Db = _density(orb.nbf, 2, 4)
    J, K, Ka, Kb = oracle.df_jk_from_tensors(Vr, Tr, D, Db, b_side="right")
    Jr, Kr, Kar, Kbr = oracle.ref_df_jk(orb, aux, D, D_beta=Db)
    for x, y in ((J, Jr), (K, Kr), (Ka, Kar), (Kb, Kbr)):
        assert np.abs(x - y).max() <= 1e-13 * np.abs(y).max()


def test_reference_two_centre_kernel_is_not_the_coulomb_metric():
    """Records why the reference's DF output is not the parity target: its
    metric differs from the Coulomb integrals (closed form above) and, for
    the cc-pVDZ-RI set, is not positive definite."""
    _need_ref()
    aux = oracle.RawBasis("cc-pvdz-rifit", WATER)
    orb = oracle.RawBasis("sto-3g", WATER)
    V, _ = oracle.df_tensors(orb, aux)
    Vr, _ = oracle.ref_df_tensors(orb, aux)
    assert np.abs(V - Vr).max() > 0.1 * np.abs(V).max()
    assert np.linalg.eigvalsh(Vr).min() < 0
    with pytest.raises(RuntimeError, match="not positive definite"):
        oracle.ref_df_jk(orb, aux, np.eye(orb.nbf))


# ---- host side of the product API (no device calls) ------------------------------

def test_auxiliary_basis_loading():
    import libaccint as lai
    xyz = [list(p) for _, p in WATER]
    aux = lai.load_auxiliary_basis("cc-pVDZ-RI", [8, 1, 1], xyz)
    raw = oracle.RawBasis("cc-pvdz-rifit", WATER)
    assert aux.n_functions() == raw.nbf == 96 and aux.n_shells() == raw.n_shells == 30
    assert aux.max_angular_momentum() == 3
    assert [s.angular_momentum() for s in aux.shells()] == list(raw.L)
    assert aux.shell_to_function(1) == 1 and aux.function_to_shell(95) == 29
    assert sorted(lai.available_auxiliary_bases()) == ["cc-pVDZ-RI", "def2-SVP-RI"]
    assert lai.is_auxiliary_basis_available("def2-svp-rifit", [1, 6, 7, 8])
    assert not lai.is_auxiliary_basis_available("def2-TZVP-JKFIT", [8])
    with pytest.raises(ValueError, match="Unknown auxiliary basis"):
        lai.load_auxiliary_basis("def2-TZVP-JKFIT", [8], [[0, 0, 0]])
    with pytest.raises(ValueError, match="Mismatched sizes"):
        lai.load_auxiliary_basis("cc-pVDZ-RI", [8, 1], [[0, 0, 0]])


def test_builder_config_and_argument_checks():
    import libaccint as lai
    c = lai.DFFockBuilderConfig()
    assert c.exchange_fraction == 1.0 and c.compute_coulomb and c.compute_exchange
    assert c.conditioning_threshold == 1e8 and c.conditioning_hard_limit == 1e14
    assert c.metric_type == lai.DFMetricType.Coulomb and c.memory_limit_mb == 4096
    atoms = [lai.Atom(z, list(p)) for z, p in WATER]
    orb = lai.load_basis_set("cc-pvdz", atoms)
    aux = lai.load_auxiliary_basis("cc-pVDZ-RI", [8, 1, 1], [list(p) for _, p in WATER])
    f = lai.DFFockBuilder(orb, aux)
    assert f.n_orb() == 25 and f.n_aux() == 96 and not f.is_initialized()
    assert f.b_tensor_memory() == 25 * 25 * 96 * 8
    with pytest.raises(ValueError):
        f.set_density(np.zeros((24, 24)))
    with pytest.raises(ValueError, match="without atom data"):
        lai.make_df_fock_builder(orb, "")
    g = lai.make_df_fock_builder(orb, "placeholder")
    assert g.n_aux() == 3
    h = lai.make_df_fock_builder(orb, [8, 1, 1], [list(p) for _, p in WATER])
    assert h.n_aux() == 96


def test_df_cabi_rejects_invalid_arguments():
    """The density-fitting C ABI checks its arguments before any device work."""
    import ctypes
    from libaccint import _cabi
    L = _cabi.lib()
    out = ctypes.c_void_p()
    cond = ctypes.c_double()
    assert L.lai_dev_df_create(None, 3, 2, 0, 1e14, ctypes.byref(cond), ctypes.byref(out)) == -1
    assert b"invalid arguments" in L.lai_dev_last_error()
    assert L.lai_dev_df_jk(None, None, None, None, 3, None) == -1
    assert b"lai_dev_df_jk" in L.lai_dev_last_error()
    assert L.lai_dev_df_jk_profile(None, None, None, None, 3, None, None) == -1
    n = ctypes.c_int()
    assert L.lai_dev_df_dims(None, ctypes.byref(n), ctypes.byref(n)) == -1
    assert L.lai_dev_df_free(None) == 0
