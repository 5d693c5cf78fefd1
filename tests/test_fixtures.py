"""CPU checks of the compiled-reference fixtures and of how they were made.

* The 16-water list identity does not hang on ties: no product Q_ij Q_kl of
  the reference's Schwarz table lies within a relative 1e-8 of the threshold
  1e-12, while the device table agrees with it to 1e-11 relative, so the two
  tables select the same quartets.
* The finite-difference fixture of the 32-water gradient takes, per
  coordinate, only the canonical quartets that touch the moved atom; the
  energy difference of that part equals the difference of the full energy
  (checked with the compiled reference on three waters).
"""
import math
import os

import numpy as np
import pytest

import oracle
from conftest import golden
from libaccint import systems


def tie_band(Q, thr, delta):
    q = np.sort(Q[Q > 0])
    lo = np.searchsorted(q, thr * (1 - delta) / q, side="left")
    hi = np.searchsorted(q, thr * (1 + delta) / q, side="right")
    return int(np.sum(hi - lo))


def test_sixteen_water_schwarz_table_has_no_ties_near_threshold():
    Q = np.load(golden("ref_16water_ccpvdz.npz"))["Q"]
    assert tie_band(Q, 1e-12, 1e-8) == 0
    assert tie_band(Q, 1e-12, 1e-4) > 0  # the band test itself finds products near thr


@pytest.mark.skipif(oracle.ref_lib() is None, reason="compiled reference (oracle/_ref) not built")
def test_partial_energy_difference_equals_full_difference():
    atoms = systems.water_cluster(3, 9.0, seed=4)
    rb = oracle.RawBasis("def2-svp", atoms)
    D = systems.density(rb.nbf, 10, 6, 1.0 / math.sqrt(rb.nbf))
    a, x, h = 4, 1, 1e-4
    mask = (rb.atom == a).astype(np.int32)

    def energy(J, K):
        return 0.5 * float(np.sum(D * J)) - 0.25 * float(np.sum(D * K))

    full, part = [], []
    for sgn in (1.0, -1.0):
        moved = [(z, tuple(p)) for z, p in atoms]
        p = list(moved[a][1])
        p[x] += sgn * h
        moved[a] = (moved[a][0], tuple(p))
        r = oracle.Oracle(oracle.RawBasis("def2-svp", moved), use_ref=True)
        Q = r.schwarz()
        full.append(energy(*r.jk_canonical(D, Q, 1e-12)))
        part.append(energy(*r.jk_canonical_touching(D, Q, 1e-12, mask)))
    g_full = (full[0] - full[1]) / (2 * h)
    g_part = (part[0] - part[1]) / (2 * h)
    assert abs(g_full) > 1e-3
    assert abs(g_full - g_part) < 1e-9


@pytest.mark.parametrize("name", ["ref_c40h82_def2svp.npz", "ref_water32_fd.npz", "ref_batch_ccpvtz.npz"])
def test_fixture_summaries_are_consistent(name):
    if not os.path.exists(golden(name)):
        pytest.skip("fixture not generated")
    f = np.load(golden(name))
    prefixes = [""] if "molecules" not in f.files else ["m%d_" % k for k in f["molecules"]]
    for p in prefixes:
        J_diag, K_diag = f[p + "J_diag"], f[p + "K_diag"]
        n = len(J_diag)
        idx = f[p + "sample_idx"]
        assert idx.min() >= 0 and idx.max() < n
        # the sampled diagonal entries agree with the stored diagonals
        d = idx[:, 0] == idx[:, 1]
        assert np.array_equal(f[p + "J_sample"][d], J_diag[idx[d, 0]])
        assert np.array_equal(f[p + "K_sample"][d], K_diag[idx[d, 0]])
        assert float(f[p + "trDJ"]) > 0 and float(f[p + "trDK"]) > 0
        assert int(f[p + "n_quartets"]) > 0
    if "fd_atoms" in f.files:
        assert len(f["fd_atoms"]) == 10 and np.all(np.isfinite(f["e_plus"] - f["e_minus"]))
