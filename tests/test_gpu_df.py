"""Density-fitted J/K on the device (DFFockBuilder -> lai_dev_df_create /
lai_dev_df_jk) against the oracle (oracle.df_jk, pinned in
tests/test_df_oracle.py).  Bar: 1e-10 relative (max |diff| / max |ref|) where
the oracle's Cholesky and triangular inverse run the same algebra in another
order; the metric's condition number (<= 1e7 here) bounds the spread."""
import numpy as np
import pytest

import oracle
from libaccint import systems
import libaccint as lai

pytestmark = pytest.mark.gpu

WATER = systems.h2o_sto3g_geometry()
TOL = 1e-10


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def lai_shells(raw):
    out = []
    o = 0
    for s in range(raw.n_shells):
        k = int(raw.nprim[s])
        out.append(lai.Shell(int(raw.L[s]), list(raw.centers[s]), list(raw.exps[o:o + k]), list(raw.raw_coefs[o:o + k])))
        o += k
    return out


def builder(orb_raw, aux_raw, **cfg):
    c = lai.DFFockBuilderConfig()
    for k, v in cfg.items():
        setattr(c, k, v)
    orb = lai.BasisSet(lai_shells(orb_raw))
    aux = lai.AuxiliaryBasisSet(lai_shells(aux_raw))
    return lai.DFFockBuilder(orb, aux, c)


def density(n, nocc, seed):
    C = np.random.default_rng(seed).standard_normal((n, nocc)) / np.sqrt(n)
    return C @ C.T


@pytest.mark.parametrize("metric", ["coulomb", "overlap"])
def test_water_ccpvdz_matches_oracle(metric):
    print("df water cc-pvdz / cc-pvdz-rifit", metric, flush=True)
    orb = oracle.RawBasis("cc-pvdz", WATER)
    aux = oracle.RawBasis("cc-pvdz-rifit", WATER)
    D = density(orb.nbf, 5, 1)
    J0, K0 = oracle.df_jk(orb, aux, D, metric=metric)
    f = builder(orb, aux, metric_type=lai.DFMetricType.Coulomb if metric == "coulomb" else lai.DFMetricType.Overlap)
    f.set_density(D)
    J = f.compute_coulomb()
    K = f.compute_exchange()
    assert rel(J, J0) < TOL and rel(K, K0) < TOL
    assert f.is_initialized() and f.metric_condition_estimate() > 1.0


def test_exact_fit_equals_four_centre_jk():
    print("df exact fit", flush=True)
    centre = (0.1, -0.2, 0.3)
    orb = oracle.RawBasis.from_shells([(0, centre, [0.5], [1.0]), (0, centre, [3.0], [1.0]), (1, centre, [0.8], [1.0])])
    aux = oracle.RawBasis.from_shells([(0, centre, [e], [1.0]) for e in (1.0, 3.5, 6.0)] +
                                      [(1, centre, [e], [1.0]) for e in (1.3, 3.8)] + [(2, centre, [1.6], [1.0])])
    D = density(orb.nbf, 2, 5)
    J4, K4 = oracle.Oracle(orb).jk_canonical(D)
    f = builder(orb, aux)
    f.set_density(D)
    F = f.compute()
    assert rel(f.coulomb_matrix(), J4) < 1e-9 and rel(f.exchange_matrix(), K4) < 1e-9
    assert np.allclose(F, f.coulomb_matrix() - f.exchange_matrix(), atol=0, rtol=0)


def test_two_waters_uhf_matches_oracle():
    print("df two waters uhf", flush=True)
    atoms = systems.water_cluster(2, 8.0, seed=3)
    orb = oracle.RawBasis("cc-pvdz", atoms)
    aux = oracle.RawBasis("cc-pvdz-rifit", atoms)
    Da, Db = density(orb.nbf, 10, 2), density(orb.nbf, 9, 3)
    J0, K0, Ka0, Kb0 = oracle.df_jk(orb, aux, Da, D_beta=Db)
    f = builder(orb, aux)
    f.set_density_unrestricted(Da, Db)
    f.compute()
    assert f.is_uhf()
    assert rel(f.coulomb_matrix(), J0) < TOL and rel(f.exchange_matrix(), K0) < TOL
    assert rel(f.exchange_matrix_alpha(), Ka0) < TOL and rel(f.exchange_matrix_beta(), Kb0) < TOL


def test_fock_assembly_and_flags():
    print("df fock assembly", flush=True)
    orb = oracle.RawBasis("cc-pvdz", WATER)
    aux = oracle.RawBasis("def2-svp-rifit", WATER)
    D = density(orb.nbf, 5, 4)
    J0, K0 = oracle.df_jk(orb, aux, D)
    f = builder(orb, aux, exchange_fraction=0.25, compute_exchange=False)
    f.set_density(D)
    F = f.compute()
    assert rel(F, J0) < TOL and np.all(f.exchange_matrix() == 0.0)
    g = builder(orb, aux, exchange_fraction=0.25)
    g.set_density(D)
    F = np.ones((orb.nbf, orb.nbf))
    g.compute_accumulate(F)
    assert rel(F - 1.0, J0 - 0.25 * K0) < TOL
    H = np.full((orb.nbf, orb.nbf), 0.5)
    assert rel(g.fock_matrix(H), H + J0 - 0.25 * K0) < TOL


def test_metric_errors():
    print("df metric errors", flush=True)
    orb = oracle.RawBasis("sto-3g", WATER)
    aux = oracle.RawBasis("cc-pvdz-rifit", WATER)
    f = builder(orb, aux, metric_type=lai.DFMetricType.AttenuatedCoulomb)
    f.set_density(np.eye(orb.nbf))
    with pytest.raises(ValueError, match="AttenuatedCoulomb"):
        f.compute()
    g = builder(orb, aux, conditioning_hard_limit=2.0)
    g.set_density(np.eye(orb.nbf))
    with pytest.raises(ValueError, match="ill-conditioned"):
        g.compute()
    h = builder(orb, aux, conditioning_threshold=2.0)
    h.set_density(np.eye(orb.nbf))
    h.compute()
    assert h.metric_ill_conditioned()


def test_sixteen_waters_against_four_centre_build():
    """Full size (400 orbital, 1536 fitting functions): the Coulomb fit is
    variational and close, J and K symmetric, and J, K from two builds of the
    same handle identical."""
    print("df sixteen waters", flush=True)
    atoms = systems.water_cluster(16, 14.0, seed=2)
    b = lai.basis_set("cc-pvdz", atoms)
    e = lai.Engine(b)
    D = systems.density(400, 80, 5, 1.0 / np.sqrt(384.0))
    J4, K4 = e.build_jk(D, lai.JKOptions(schwarz_threshold=1e-12))
    f = lai.make_df_fock_builder(b, [z for z, _ in atoms], [list(p) for _, p in atoms], "cc-pVDZ-RI")
    assert f.n_orb() == 400 and f.n_aux() == 1536
    f.set_density(D)
    f.compute()
    J, K = f.coulomb_matrix(), f.exchange_matrix()
    EJ, EJ4 = 0.5 * np.sum(D * J), 0.5 * np.sum(D * J4)
    assert EJ <= EJ4 and (EJ4 - EJ) < 1e-3 * EJ4
    # cc-pVDZ-RI is a correlation fitting set: exchange is fitted less well
    # than Coulomb (1.3 % of max |K| here)
    assert rel(K, K4) < 3e-2
    assert rel(J, J.T) < 1e-12 and rel(K, K.T) < 1e-12
    f.compute()
    assert np.array_equal(J, f.coulomb_matrix()) and np.array_equal(K, f.exchange_matrix())
    # J of the combined build (gamma from the diagonal of X = B D) equals J of
    # the separate Coulomb build (gamma from a pass over B)
    assert rel(J, np.array(f.compute_coulomb())) < 1e-13
    # the combined symmetric build (skipped lower tiles, mirrored) equals the
    # exchange product over all tiles: lai_dev_df_jk with K alone, no symmetry
    import ctypes
    import torch
    from libaccint import _cabi
    L = _cabi.lib()
    h = ctypes.c_void_p(f.device_handle())
    dD = torch.from_numpy(np.ascontiguousarray(D)).cuda()
    dJ = torch.zeros_like(dD)
    dK1 = torch.zeros_like(dD)
    dK2 = torch.zeros_like(dD)
    _cabi.check(L.lai_dev_df_jk(h, ctypes.c_void_p(dD.data_ptr()), ctypes.c_void_p(dJ.data_ptr()),
                                ctypes.c_void_p(dK1.data_ptr()), 1 | 2 | 4, None))
    _cabi.check(L.lai_dev_df_jk(h, ctypes.c_void_p(dD.data_ptr()), None, ctypes.c_void_p(dK2.data_ptr()), 2, None))
    torch.cuda.synchronize()
    K1, K2 = dK1.cpu().numpy(), dK2.cpu().numpy()
    assert rel(K1, K2) < 1e-12 and rel(K2, K) < 1e-12
    assert rel(dJ.cpu().numpy(), J) < 1e-13


@pytest.fixture(scope="module")
def six_waters_tensors():
    atoms = systems.water_cluster(6, 11.0, seed=12)
    orb = oracle.RawBasis("cc-pvdz", atoms)
    aux = oracle.RawBasis("cc-pvdz-rifit", atoms)
    return orb, aux, oracle.df_tensors(orb, aux)


@pytest.mark.parametrize("symmetric", [True, False])
def test_six_waters_multi_tile_matches_oracle(symmetric, six_waters_tensors):
    """150 orbital functions: several 64 x 80 tiles of the exchange product
    (XCD tile order, skipped lower tiles with a symmetric D) and several
    slices of its summation, against the oracle."""
    print("df six waters, symmetric D" if symmetric else "df six waters, general D", flush=True)
    orb, aux, (V, T) = six_waters_tensors
    assert orb.nbf == 150
    D = density(orb.nbf, 25, 13)
    if not symmetric:
        D = D + 0.05 * np.random.default_rng(14).standard_normal(D.shape) / orb.nbf
    J0, K0 = oracle.df_jk_from_tensors(V, T, D)
    f = builder(orb, aux)
    f.set_density(np.ascontiguousarray(D))
    f.compute()
    assert rel(f.coulomb_matrix(), J0) < TOL and rel(f.exchange_matrix(), K0) < TOL
    assert rel(np.array(f.compute_coulomb()), J0) < TOL


def _numpy_rhf(H, S, jk, n_occ, e_nuc, iters=200):
    """The same RHF recipe on the CPU (core guess, DIIS on FDS - SDF,
    symmetric orthogonalisation); jk(D) -> (J, K)."""
    s, U = np.linalg.eigh(S)
    X = (U[:, s > 1e-10] / np.sqrt(s[s > 1e-10])) @ U[:, s > 1e-10].T

    def solve(F):
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        return X @ Cp

    C = solve(H)
    D = 2.0 * C[:, :n_occ] @ C[:, :n_occ].T
    Fs, Es, E_old = [], [], np.inf
    for _ in range(iters):
        J, K = jk(D)
        F = H + J - 0.5 * K
        E = 0.5 * np.sum((H + F) * D) + e_nuc
        Fs.append(F); Es.append(F @ D @ S - S @ D @ F)
        Fs, Es = Fs[-6:], Es[-6:]
        m = len(Fs)
        B = -np.ones((m + 1, m + 1)); B[m, m] = 0.0
        for i in range(m):
            for j in range(m):
                B[i, j] = np.sum(Es[i] * Es[j])
        rhs = np.zeros(m + 1); rhs[m] = -1.0
        c = np.linalg.lstsq(B, rhs, rcond=None)[0]
        C = solve(sum(c[i] * Fs[i] for i in range(m)))
        D_new = 2.0 * C[:, :n_occ] @ C[:, :n_occ].T
        if abs(E - E_old) < 1e-12 and np.abs(D_new - D).max() < 1e-10:
            D = D_new
            break
        D, E_old = D_new, E
    J, K = jk(D)
    return 0.5 * np.sum((2 * H + J - 0.5 * K) * D) + e_nuc


def test_df_rhf_energy_matches_oracle_scf():
    """RHF with density-fitted J/K (scf.rhf(..., density_fitting=...)) against
    the same SCF on the oracle's DF integrals, and close to the exact RHF."""
    print("df rhf water", flush=True)
    import scipy.linalg
    from libaccint import scf
    res = scf.rhf(WATER, "cc-pvdz", 10, density_fitting="cc-pVDZ-RI")
    exact = scf.rhf(WATER, "cc-pvdz", 10)
    assert res.converged and exact.converged
    orb = oracle.RawBasis("cc-pvdz", WATER)
    aux = oracle.RawBasis("cc-pvdz-rifit", WATER)
    o = oracle.Oracle(orb)
    S = o.overlap()
    H = o.kinetic() + o.nuclear()
    V, T = oracle.df_tensors(orb, aux)
    Lc = np.linalg.cholesky(V)
    B = np.einsum("PQ,abQ->Pab", scipy.linalg.solve_triangular(Lc, np.eye(len(V)), lower=True), T)

    def jk(D):
        J = np.einsum("Pab,P->ab", B, np.einsum("Pab,ab->P", B, D))
        return J, np.einsum("Pab,Pcb->ac", B @ D, B)

    e_nuc = scf.nuclear_repulsion(WATER)
    E0 = _numpy_rhf(H, S, jk, 5, e_nuc)
    assert abs(res.energy - E0) < 1e-8
    assert abs(res.energy - exact.energy) < 2e-3


def test_symmetric_and_general_density_paths():
    """An exactly symmetric D takes the exchange path that computes K's upper
    part and mirrors it; a non-symmetric D the full product.  Both against
    the oracle."""
    print("df symmetric / general D", flush=True)
    atoms = systems.water_cluster(2, 8.0, seed=3)
    orb = oracle.RawBasis("cc-pvdz", atoms)
    aux = oracle.RawBasis("cc-pvdz-rifit", atoms)
    D = density(orb.nbf, 10, 8)
    Ds = 0.5 * (D + D.T)  # exactly symmetric
    Dg = D + 0.1 * np.random.default_rng(9).standard_normal(D.shape) / orb.nbf  # not symmetric
    f = builder(orb, aux)
    for Dm in (Ds, Dg):
        J0, K0 = oracle.df_jk(orb, aux, Dm)
        f.set_density(np.ascontiguousarray(Dm))
        J1 = np.array(f.compute_coulomb())  # J alone: gamma from a pass over B
        assert rel(J1, J0) < TOL and rel(f.compute_exchange(), K0) < TOL
        f.compute()  # J and K together: gamma from the diagonal of X = B D
        assert rel(f.coulomb_matrix(), J1) < 1e-13 and rel(f.exchange_matrix(), K0) < TOL
    K = f.exchange_matrix()
    assert not np.allclose(K, K.T)  # the general density's K is not symmetric
    f.set_density(Ds)
    K = f.compute_exchange()
    assert np.array_equal(K, K.T)  # mirrored exactly
