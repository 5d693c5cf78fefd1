"""Write the golden fixtures of the test suite (run in a container that holds
the reference tree; the outputs are committed, the reference is not needed at
test time).

    python tests/golden/make_fixtures.py [--sixteen-waters]

Writes into tests/golden/:
  reference_literals.json  PySCF values quoted as literals by the reference's
                           tests: H2 and H2O/STO-3G ERIs
                           (tests/integration/test_two_electron.cpp:57-156),
                           H2 and H2O/STO-3G S/T/V
                           (tests/integration/test_one_electron.cpp:76-130),
                           RHF energies and E_nuc (test_hf_energy.cpp:314-389)
  boys_reference.json      copy of the reference's tests/data/boys_reference.json
                           (mpmath, 50 digits)
  rys_reference.json       tests/data/rys_reference.json, n <= 9
  water_aug_cc_pvtz.npz    tests/data/reference/water_1_aug_cc_pVTZ.json
                           (PySCF, Cartesian, LibAccInt ordering): S, T, V, D,
                           J, K, sampled ERIs, geometry
  ref_h2o_sto3g.npz        outputs of the compiled reference CPU path
                           (oracle/_ref) for H2O/STO-3G: 7^4 ERIs, S, T, V, J, K
                           for the seeded density
  ref_16water_ccpvdz.npz   (--sixteen-waters) the compiled reference for the seeded
                           16-water cc-pVDZ cluster: Schwarz table, screened
                           list count/checksum at 1e-12, J, K
"""
import argparse
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "libaccint-public_amd"))

NUM = r"[-+]?\d+\.\d+(?:e[-+]?\d+)?"


def _array_literal(text, name, n):
    m = re.search(r"constexpr Real %s\[%d\] = \{(.*?)\};" % (name, n), text, re.S)
    body = re.sub(r"//[^\n]*", "", m.group(1))
    vals = [float(x) for x in re.findall(NUM, body)]
    assert len(vals) == n, (name, len(vals))
    return vals


def literals():
    two = open(os.path.join(REF, "tests/integration/test_two_electron.cpp")).read()
    one = open(os.path.join(REF, "tests/integration/test_one_electron.cpp")).read()
    hf = open(os.path.join(REF, "tests/integration/test_hf_energy.cpp")).read()
    h2o = []
    for m in re.finditer(r"compute_element\((\d+), (\d+), (\d+), (\d+)\), (%s), ERI_TOL" % NUM, two):
        h2o.append([int(m.group(i)) for i in range(1, 5)] + [float(m.group(5))])
    h2 = {
        "0000": float(re.search(r"\(0 0 \| 0 0\)\n\s+compute_eri\(s0, s0, s0, s0, buffer\);\n\s+EXPECT_NEAR\(buffer\(0, 0, 0, 0\), (%s)" % NUM, two).group(1)),
        "0011": float(re.search(r"compute_eri\(s0, s0, s1, s1, buffer\);\n\s+EXPECT_NEAR\(buffer\(0, 0, 0, 0\), (%s)" % NUM, two).group(1)),
        "0101": float(re.search(r"compute_eri\(s0, s1, s0, s1, buffer\);\n\s+EXPECT_NEAR\(buffer\(0, 0, 0, 0\), (%s)" % NUM, two).group(1)),
    }
    out = {
        "source": "PySCF literals from the reference's integration tests",
        "h2_geometry": [[1, [0.0, 0.0, 0.0]], [1, [0.0, 0.0, 1.39839733]]],
        "h2o_geometry": [[8, [0.0, 0.0, 0.0]], [1, [0.0, 1.43233673, -1.10866041]],
                         [1, [0.0, -1.43233673, -1.10866041]]],
        "h2_eri": h2,
        "h2o_eri": h2o,
        "h2_S": _array_literal(one, "H2_S_REF", 4),
        "h2_T": _array_literal(one, "H2_T_REF", 4),
        "h2_V": _array_literal(one, "H2_V_REF", 4),
        "h2o_S": _array_literal(one, "H2O_S_REF", 49),
        "h2o_T": _array_literal(one, "H2O_T_REF", 49),
        "h2o_V": _array_literal(one, "H2O_V_REF", 49),
        "rhf_h2o_sto3g": float(re.search(r"REF_ENERGY = (%s);" % NUM, hf).group(1)),
        # H2O, H2, CH4 in the order of test_hf_energy.cpp
        "rhf_sto3g_energies": [float(x) for x in re.findall(r"REF_ENERGY = (%s);" % NUM, hf)],
        "ch4_b": float(re.search(r"constexpr Real b = (%s);" % NUM, hf).group(1)),
        "enuc_h2o": float(re.search(r"EXPECT_NEAR\(E_nuc, (%s)" % NUM, hf).group(1)),
    }
    with open(os.path.join(HERE, "reference_literals.json"), "w") as f:
        json.dump(out, f, indent=1)


def data_files():
    b = json.load(open(os.path.join(REF, "tests/data/boys_reference.json")))
    with open(os.path.join(HERE, "boys_reference.json"), "w") as f:
        json.dump(b, f)
    r = json.load(open(os.path.join(REF, "tests/data/rys_reference.json")))
    rq = r["rys_quadrature"]
    rq["data"] = {k: v for k, v in rq["data"].items() if int(k) <= 9}
    rq["n_roots_range"] = [1, 9]
    with open(os.path.join(HERE, "rys_reference.json"), "w") as f:
        json.dump(r, f)
    w = json.load(open(os.path.join(REF, "tests/data/reference/water_1_aug_cc_pVTZ.json")))
    n = w["n_basis"]
    np.savez_compressed(
        os.path.join(HERE, "water_aug_cc_pvtz.npz"),
        S=np.array(w["integrals"]["overlap"]["matrix"]).reshape(n, n),
        T=np.array(w["integrals"]["kinetic"]["matrix"]).reshape(n, n),
        V=np.array(w["integrals"]["nuclear_attraction"]["matrix"]).reshape(n, n),
        D=np.array(w["density_matrix"]["matrix"]).reshape(n, n),
        J=np.array(w["fock_from_density"]["coulomb_matrix"]).reshape(n, n),
        K=np.array(w["fock_from_density"]["exchange_matrix"]).reshape(n, n),
        eri_idx=np.array([e["indices"] for e in w["sampled_eri"]], dtype=np.int32),
        eri_val=np.array([e["value"] for e in w["sampled_eri"]]),
        geometry_bohr=np.array(w["molecule"]["geometry_bohr"], dtype=float),
        rhf_energy=w["rhf"]["total_energy"],
    )


def ref_h2o_sto3g():
    import oracle
    from libaccint import systems
    rb = oracle.RawBasis("sto-3g", systems.h2o_sto3g_geometry())
    r = oracle.Oracle(rb, use_ref=True)
    rng = np.random.default_rng(1)
    C = rng.standard_normal((7, 5)) / np.sqrt(7)
    D = C @ C.T
    J, K = r.jk_full(D, n_threads=1)
    np.savez_compressed(os.path.join(HERE, "ref_h2o_sto3g.npz"), eri=r.eri_tensor(), S=r.overlap(), T=r.kinetic(),
                        V=r.nuclear(), D=D, J=J, K=K)


def ref_16water_ccpvdz():
    import oracle
    from libaccint import systems
    atoms = systems.water_cluster(16, 14.0, seed=2)
    rb = oracle.RawBasis("cc-pvdz", atoms)
    r = oracle.Oracle(rb, use_ref=True)
    D = systems.density(rb.nbf, 80, 5, 1.0 / np.sqrt(384.0))
    Q = r.schwarz()
    n, cs = oracle.Oracle(rb).screened_list(Q, 1e-12)
    J, K = r.jk_canonical(D, Q, 1e-12, n_threads=0)
    np.savez_compressed(os.path.join(HERE, "ref_16water_ccpvdz.npz"), Q=Q, n_quartets=n, checksum=np.uint64(cs), J=J, K=K,
                        atoms_xyz=np.array([p for _, p in atoms]), atoms_Z=np.array([z for z, _ in atoms]))


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--sixteen-waters", action="store_true")
    a = ap.parse_args()
    if a.sixteen_waters:
        ref_16water_ccpvdz()
    else:
        literals()
        data_files()
        ref_h2o_sto3g()
