"""The benchmark workloads at their full sizes, on the device, against the
compiled reference's outputs (tests/golden/make_fixtures.py):

  * 16 waters cc-pVDZ: the screened list from the device's own Schwarz table
    equals the reference's list (count and checksum);
  * C40H82 def2-SVP (1010 functions): FP64 J/K against the reference's
    tr(DJ), tr(DK), diagonals, row sums and 2000 sampled elements (1e-10
    relative to max |J|, max |K|); list identity from the device's own
    table; mixed precision |dE| <= 1e-6 Eh with dJ, dK stated;
  * six molecules of the 64-molecule cc-pVTZ batch in one BatchEngine: each
    molecule's block against the reference's per-molecule summaries, FP64
    and mixed;
  * 32 waters def2-SVP: J/K summaries, and the analytic gradient against
    central differences of the reference's energy (step 1e-4 bohr) at 10
    seeded coordinates, |g - FD| <= 1e-6 Eh/bohr, plus translational
    invariance.

The reference's own Rys roots are accurate to ~6e-13 (5 roots), ~3e-11 (6)
and ~6e-10 (7) relative; cc-pVTZ blocks with f shells use up to 7 roots, so
the batch tolerance is 1e-9 relative (the reference's error), 1e-10 elsewhere.
"""
import math
import os

import numpy as np
import pytest

from conftest import golden
from libaccint import systems
import libaccint as lai

pytestmark = pytest.mark.gpu


def check_summary(J, K, D, f, tol, prefix=""):
    """J, K (numpy) against the fixture's summaries under key prefix."""
    g = lambda k: f[prefix + k]
    jm, km = float(g("J_absmax")), float(g("K_absmax"))
    n = D.shape[0]
    assert abs(float(np.sum(D * J)) - float(g("trDJ"))) <= tol * abs(float(g("trDJ")))
    assert abs(float(np.sum(D * K)) - float(g("trDK"))) <= tol * abs(float(g("trDK")))
    assert np.max(np.abs(np.diag(J) - g("J_diag"))) <= tol * jm
    assert np.max(np.abs(np.diag(K) - g("K_diag"))) <= tol * km
    assert np.max(np.abs(J.sum(axis=1) - g("J_rowsum"))) <= tol * jm * math.sqrt(n)
    assert np.max(np.abs(K.sum(axis=1) - g("K_rowsum"))) <= tol * km * math.sqrt(n)
    idx = g("sample_idx")
    assert np.max(np.abs(J[idx[:, 0], idx[:, 1]] - g("J_sample"))) <= tol * jm
    assert np.max(np.abs(K[idx[:, 0], idx[:, 1]] - g("K_sample"))) <= tol * km


def tie_band(Q, thr, delta):
    """Pair-of-pairs (ordered) whose product lies within a relative delta of
    thr: the quartets whose selection could flip when Q changes by ~delta/2."""
    q = np.sort(Q[Q > 0])
    lo = np.searchsorted(q, thr * (1 - delta) / q, side="left")
    hi = np.searchsorted(q, thr * (1 + delta) / q, side="right")
    return int(np.sum(hi - lo))


@pytest.mark.skipif(not os.path.exists(golden("ref_16water_ccpvdz.npz")), reason="16-water fixture not generated")
def test_sixteen_waters_list_from_device_schwarz():
    ref = np.load(golden("ref_16water_ccpvdz.npz"))
    atoms = systems.water_cluster(16, 14.0, seed=2)
    b = lai.basis_set("cc-pvdz", atoms)
    e = lai.Engine(b)
    info = e.prepare_worklist(lai.JKOptions(schwarz_threshold=1e-12))
    Q = e.schwarz_bounds().packed()
    band = tie_band(Q, 1e-12, 1e-10)
    print("16 waters: device-Q tie band (1e-10) holds %d ordered pair products" % band)
    assert info["n_quartets"] == int(ref["n_quartets"])
    assert info["checksum"] == int(ref["checksum"])


@pytest.fixture(scope="module")
def alkane():
    f = np.load(golden("ref_c40h82_def2svp.npz"))
    atoms = systems.alkane(40)
    b = lai.basis_set("def2-svp", atoms)
    n = b.n_basis_functions()
    assert n == 1010
    D = systems.density(n, 161, 11, 1.0 / math.sqrt(n))
    e = lai.Engine(b)
    J, K = e.build_jk(D, lai.JKOptions(schwarz_threshold=1e-12))
    return f, b, e, D, J, K


def test_c40h82_fp64_against_reference(alkane):
    f, b, e, D, J, K = alkane
    info = e.worklist_info()
    Q = e.schwarz_bounds().packed()
    print("C40H82: %d quartets; device-Q tie band (1e-10) %d" % (info["n_quartets"], tie_band(Q, 1e-12, 1e-10)))
    assert info["n_quartets"] == int(f["n_quartets"])
    assert info["checksum"] == int(f["checksum"])
    check_summary(J, K, D, f, 1e-10)


def test_c40h82_mixed_precision_error(alkane):
    f, b, e, D, J, K = alkane
    Jm, Km = e.build_jk(D, lai.JKOptions(schwarz_threshold=1e-12, mixed_precision=True, fp32_threshold=1e-7))
    info = e.worklist_info()
    E_ref = 0.5 * float(f["trDJ"]) - 0.25 * float(f["trDK"])
    dE = 0.5 * float(np.sum(D * Jm)) - 0.25 * float(np.sum(D * Km)) - E_ref
    dJ = float(np.max(np.abs(Jm - J)))
    dK = float(np.max(np.abs(Km - K)))
    print("C40H82 mixed: fp32 quartets %d of %d, dE %.2e Eh, max|dJ| %.2e, max|dK| %.2e" %
          (info["n_fp32"], info["n_quartets"], dE, dJ, dK))
    assert info["n_fp32"] > info["n_quartets"] // 4
    assert abs(dE) <= 1e-6
    assert dJ <= 1e-6 * float(f["J_absmax"]) and dK <= 1e-6 * float(f["K_absmax"])
    # the mixed build against the reference's summaries, at the FP32 error level
    check_summary(Jm, Km, D, f, 1e-6)


@pytest.mark.skipif(not os.path.exists(golden("ref_batch_ccpvtz.npz")), reason="batch fixture not generated")
def test_cc_pvtz_batch_subset_against_reference():
    f = np.load(golden("ref_batch_ccpvtz.npz"))
    rng = np.random.default_rng(4)
    mols = [systems.random_molecule(rng) for _ in range(64)]
    sel = [int(k) for k in f["molecules"]]
    bases = [lai.basis_set("cc-pvtz", mols[k]) for k in sel]
    be = lai.BatchEngine(bases)
    nbf = be.n_basis_functions()
    offs = be.matrix_offsets()
    Ds = [systems.density(n, max(2, n // 5), 200 + k, 1.0 / math.sqrt(n)) for k, n in zip(sel, nbf)]
    Db = np.concatenate([d.ravel() for d in Ds])
    J, K = be.build_jk(Db, lai.JKOptions(schwarz_threshold=1e-12))
    info = be.worklist_info()
    assert info["n_quartets"] == sum(int(f["m%d_n_quartets" % k]) for k in sel)
    assert info["checksum"] == sum(int(f["m%d_checksum" % k]) for k in sel) % (1 << 64)
    Jm_, Km_ = be.build_jk(Db, lai.JKOptions(schwarz_threshold=1e-12, mixed_precision=True, fp32_threshold=1e-7))
    for m, k in enumerate(sel):
        a, z, n = offs[m], offs[m + 1], nbf[m]
        Jb, Kb = J[a:z].reshape(n, n), K[a:z].reshape(n, n)
        check_summary(Jb, Kb, Ds[m], f, 1e-9, prefix="m%d_" % k)
        Jm, Km = Jm_[a:z].reshape(n, n), Km_[a:z].reshape(n, n)
        E_ref = 0.5 * float(f["m%d_trDJ" % k]) - 0.25 * float(f["m%d_trDK" % k])
        dE = 0.5 * float(np.sum(Ds[m] * Jm)) - 0.25 * float(np.sum(Ds[m] * Km)) - E_ref
        print("molecule %d (nbf %d): mixed dE %.2e Eh" % (k, n, dE))
        assert abs(dE) <= 1e-6


@pytest.mark.skipif(not os.path.exists(golden("ref_water32_fd.npz")), reason="32-water fixture not generated")
def test_water32_jk_and_gradient_against_reference_fd():
    f = np.load(golden("ref_water32_fd.npz"))
    atoms = systems.water_cluster(32, 18.0, seed=4)
    b = lai.basis_set("def2-svp", atoms)
    n = b.n_basis_functions()
    D = systems.density(n, 160, 6, 1.0 / math.sqrt(768.0))
    e = lai.Engine(b)
    J, K = e.build_jk(D, lai.JKOptions(schwarz_threshold=1e-12))
    info = e.worklist_info()
    assert info["n_quartets"] == int(f["n_quartets"]) and info["checksum"] == int(f["checksum"])
    check_summary(J, K, D, f, 1e-10)
    g = e.compute_jk_gradient(D, 1.0)
    assert g.shape == (96, 3)
    assert np.max(np.abs(g.sum(axis=0))) < 1e-8 * max(1.0, float(np.max(np.abs(g))))
    h = float(f["fd_step"])
    fd = (f["e_plus"] - f["e_minus"]) / (2 * h)
    got = g[f["fd_atoms"], f["fd_axes"]]
    print("32 waters: max |g - FD| = %.2e Eh/bohr over %d coordinates" % (np.max(np.abs(got - fd)), len(fd)))
    assert np.max(np.abs(got - fd)) <= 1e-6
