"""Restricted Hartree-Fock on the device J/K build.

The SCF the reference runs on top of this path (tests/integration/test_hf_energy.cpp:1-300,
examples/mini-apps/python-hf/hf.py): core-Hamiltonian guess; F = H + J - K/2
with D = 2 C_occ C_occ^T; E = 1/2 tr[(H + F) D] + E_nuc; Pulay DIIS on the
commutator F D S - S D F; symmetric orthogonalisation X = U s^{-1/2} U^T with
eigenvalues of S at or below 1e-10 dropped; converged when the energy moves by
less than 1e-12 and the density by less than 1e-10 (max abs).

D, J, K, F, S and X stay in HBM as torch tensors on one stream; J and K come
from Engine.build_jk_device (the screened, fused gfx950 J/K kernels) or, with
density_fitting set to a fitting-basis name, from the device DFFockBuilder
(lai_dev_df_jk on the same stream), the eigenproblems from torch.linalg.eigh
on the device.  Only the DIIS coefficients (a system of at most 7 equations)
are solved on the host.
"""
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _cabi, _core
from .convenience import atoms_from_list, basis_set


def nuclear_repulsion(atoms):
    """sum_{A<B} Z_A Z_B / |R_A - R_B| for [(Z, (x, y, z)), ...] in bohr."""
    e = 0.0
    for a in range(len(atoms)):
        za, ra = atoms[a][0], np.asarray(atoms[a][1], dtype=float)
        for b in range(a):
            zb, rb = atoms[b][0], np.asarray(atoms[b][1], dtype=float)
            e += za * zb / float(np.linalg.norm(ra - rb))
    return e


@dataclass
class RHFResult:
    energy: float
    iterations: int
    converged: bool
    orbital_energies: np.ndarray
    C: np.ndarray
    D: np.ndarray


class _DIIS:
    def __init__(self, size=6):
        self.size = size
        self.F, self.E = [], []

    def push(self, F, err):
        self.F.append(F.clone())
        self.E.append(err.clone())
        if len(self.F) > self.size:
            self.F.pop(0)
            self.E.pop(0)

    def extrapolate(self):
        n = len(self.F)
        if n < 2:
            return self.F[-1]
        B = -np.ones((n + 1, n + 1))
        B[n, n] = 0.0
        for i in range(n):
            for j in range(i + 1):
                B[i, j] = B[j, i] = float(torch.sum(self.E[i] * self.E[j]))
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        c = np.linalg.lstsq(B, rhs, rcond=None)[0]
        out = torch.zeros_like(self.F[0])
        for i in range(n):
            out += float(c[i]) * self.F[i]
        return out


def rhf(atoms, basis_name, n_electrons, max_iter=100, e_tol=1e-12, d_tol=1e-10, options=None,
        device="cuda", density_fitting=None):
    """RHF energy of [(Z, (x, y, z)), ...] (bohr) in basis_name with n_electrons (even).
    density_fitting: a fitting-basis name ("cc-pVDZ-RI", "def2-SVP-RI") for
    density-fitted J and K, else the exact four-centre build."""
    if n_electrons % 2:
        raise ValueError("rhf: closed shells only (even electron count)")
    n_occ = n_electrons // 2
    dev = torch.device(device)
    basis = basis_set(basis_name, atoms)
    engine = _core.Engine(basis)
    n = basis.n_basis_functions()
    A = atoms_from_list(atoms)
    t = lambda M: torch.from_numpy(np.ascontiguousarray(M)).to(dev)
    S = t(engine.compute_overlap_matrix())
    H = t(engine.compute_core_hamiltonian(A))
    e_nuc = nuclear_repulsion(atoms)
    opts = options if options is not None else _core.JKOptions()

    s, U = torch.linalg.eigh(S)
    keep = s > 1e-10
    X = (U[:, keep] * torch.rsqrt(s[keep])) @ U[:, keep].T

    def solve(F):
        eps, Cp = torch.linalg.eigh(X.T @ F @ X)
        return eps, X @ Cp

    def density(C):
        Co = C[:, :n_occ]
        return 2.0 * Co @ Co.T

    stream = torch.cuda.current_stream(dev)
    J = torch.zeros((n, n), dtype=torch.float64, device=dev)
    K = torch.zeros_like(J)

    dfb = None
    if density_fitting:
        dfb = _core.make_df_fock_builder(basis, [int(z) for z, _ in atoms], [list(map(float, p)) for _, p in atoms],
                                         density_fitting)
        dfb.initialize()
        dfh = dfb.device_handle()

    def fock(D):
        Dc = D.contiguous()
        if dfb is not None:  # J and K overwritten
            # D = 2 C_occ C_occ^T is symmetric: LAI_DF_SYMMETRIC (4) computes K's upper part
            _cabi.check(_cabi.lib().lai_dev_df_jk(dfh, Dc.data_ptr(), J.data_ptr(), K.data_ptr(), 3 | 4,
                                                  stream.cuda_stream))
        else:
            J.zero_()
            K.zero_()
            engine.build_jk_device(Dc.data_ptr(), J.data_ptr(), K.data_ptr(), opts, stream.cuda_stream, None)
        return H + J - 0.5 * K

    eps, C = solve(H)
    D = density(C)
    diis = _DIIS(6)
    e_old = math.inf
    converged = False
    it = 0
    for it in range(1, max_iter + 1):
        F = fock(D)
        E = float(0.5 * torch.sum((H + F) * D)) + e_nuc
        diis.push(F, F @ D @ S - S @ D @ F)
        eps, C = solve(diis.extrapolate())
        D_new = density(C)
        dD = float(torch.max(torch.abs(D_new - D)))
        dE = abs(E - e_old)
        e_old = E
        D = D_new
        if dE < e_tol and dD < d_tol:
            converged = True
            break
    # energy of the final density
    F = fock(D)
    E = float(0.5 * torch.sum((H + F) * D)) + e_nuc
    return RHFResult(E, it, converged, eps.cpu().numpy(), C.cpu().numpy(), D.cpu().numpy())
