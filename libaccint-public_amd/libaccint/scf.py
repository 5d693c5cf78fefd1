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

With graph=True every iteration after the first is one replay of a captured
hipGraph: the J/K build, the energy, DIIS and the new density, with no host
synchronisation inside.  torch.linalg.eigh checks its status on the host, so
inside the graph the density comes from the occupied-space projector by
second-order spectral projection (SP2) and DIIS solves its bordered system on
the device; the host reads the energy and density changes after each replay.
The final orbitals come from one eigh outside the graph.
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


def _solve_small(A, b):
    """x with A x = b for a small dense A (k x k) on the device: Gauss-Jordan
    elimination with partial pivoting, in tensor operations only (no host
    synchronisation)."""
    k = A.shape[0]
    M = torch.cat([A, b.view(k, 1)], dim=1)
    rows = torch.arange(k, device=A.device)
    for j in range(k):
        col = torch.abs(M[:, j]).masked_fill(rows < j, -1.0)
        p = torch.argmax(col).view(1)
        rj = M[j:j + 1].clone()
        rp = torch.index_select(M, 0, p)
        M = M.index_copy(0, p, rj)
        M = M.index_copy(0, rows[j:j + 1], rp)
        f = M[:, j] / M[j, j]
        f = f.masked_fill(rows == j, 0.0)
        M = M - f.view(k, 1) * M[j].view(1, -1)
    return M[:, k] / torch.diagonal(M[:, :k])


class _DeviceDIIS:
    """Pulay DIIS with a fixed-size history on the device (a ring of `size`
    Fock and error matrices): the same extrapolation as _DIIS, capturable."""

    def __init__(self, n, size, dev):
        self.size = size
        self.F = torch.zeros((size, n, n), dtype=torch.float64, device=dev)
        self.E = torch.zeros_like(self.F)
        self.B = torch.zeros((size, size), dtype=torch.float64, device=dev)
        self.count = torch.zeros((), dtype=torch.int64, device=dev)
        self.ar = torch.arange(size, device=dev)
        self.rhs = torch.zeros(size + 1, dtype=torch.float64, device=dev)
        self.rhs[size] = -1.0

    def step(self, F, err):
        m = self.size
        slot = torch.remainder(self.count, m).view(1)
        self.F.index_copy_(0, slot, F.unsqueeze(0))
        self.E.index_copy_(0, slot, err.unsqueeze(0))
        dots = self.E.reshape(m, -1) @ err.reshape(-1)
        self.B.index_copy_(0, slot, dots.view(1, m))
        self.B.index_copy_(1, slot, dots.view(m, 1))
        self.count += 1
        v = (self.ar < self.count).to(F.dtype)
        # the coefficients do not change when B is scaled: keep its entries O(1)
        scale = torch.clamp_min(torch.max(torch.diagonal(self.B) * v), 1e-300)
        A = torch.zeros((m + 1, m + 1), dtype=F.dtype, device=F.device)
        A[:m, :m] = self.B / scale * (v.view(m, 1) * v.view(1, m)) + torch.diag(1.0 - v)
        A[:m, m] = -v
        A[m, :m] = -v
        c = _solve_small(A, self.rhs)[:m]
        return (c @ self.F.reshape(m, -1)).reshape(F.shape)


def sp2_projector(Fp, n_occ, n_iter=80):
    """Projector onto the n_occ lowest eigenvectors of the symmetric matrix Fp
    by second-order spectral projection (A. M. N. Niklasson, Phys. Rev. B 66,
    155115 (2002)): X = (hi I - Fp) / (hi - lo) with Gershgorin bounds lo, hi
    of the spectrum, then n_iter steps X <- X^2 or 2X - X^2, whichever trace
    lies nearer n_occ (a projector is a fixed point of both).  Needs a gap
    between the n_occ-th and the next eigenvalue; device operations only."""
    d = torch.diagonal(Fp)
    r = torch.sum(torch.abs(Fp), dim=1) - torch.abs(d)
    lo, hi = torch.min(d - r), torch.max(d + r)
    X = (hi * torch.eye(Fp.shape[0], dtype=Fp.dtype, device=Fp.device) - Fp) / (hi - lo)
    for _ in range(n_iter):
        X2 = X @ X
        t1, t2 = torch.trace(X), torch.trace(X2)
        c = (torch.abs(t2 - n_occ) < torch.abs(2.0 * t1 - t2 - n_occ)).to(X.dtype)
        X = c * X2 + (1.0 - c) * (2.0 * X - X2)
    return X


def rhf(atoms, basis_name, n_electrons, max_iter=100, e_tol=1e-12, d_tol=1e-10, options=None,
        device="cuda", density_fitting=None, graph=False):
    """RHF energy of [(Z, (x, y, z)), ...] (bohr) in basis_name with n_electrons (even).
    density_fitting: a fitting-basis name ("cc-pVDZ-RI", "def2-SVP-RI") for
    density-fitted J and K, else the exact four-centre build.  graph: run the
    iterations as replays of one captured hipGraph (see the module text)."""
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
        st = torch.cuda.current_stream(dev).cuda_stream  # the capture stream while capturing
        if dfb is not None:  # J and K overwritten
            # D = 2 C_occ C_occ^T is symmetric: LAI_DF_SYMMETRIC (4) computes K's upper part
            _cabi.check(_cabi.lib().lai_dev_df_jk(dfh, Dc.data_ptr(), J.data_ptr(), K.data_ptr(), 3 | 4, st))
        else:
            J.zero_()
            K.zero_()
            engine.build_jk_device(Dc.data_ptr(), J.data_ptr(), K.data_ptr(), opts, st, None)
        return H + J - 0.5 * K

    if graph:
        return _rhf_graph(fock, H, S, X, n_occ, e_nuc, max_iter, e_tol, d_tol, dev, density, solve)

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


def _rhf_graph(fock, H, S, X, n_occ, e_nuc, max_iter, e_tol, d_tol, dev, density, solve):
    """The SCF loop of rhf() with every iteration after the first one replay of
    a captured graph."""
    eps, C = solve(H)
    D = density(C).contiguous()
    n = D.shape[0]
    diis = _DeviceDIIS(n, 6, dev)
    E_prev = torch.full((), math.inf, dtype=torch.float64, device=dev)
    out = torch.zeros(3, dtype=torch.float64, device=dev)  # E, |dE|, max |dD|

    def iteration():
        F = fock(D)
        E = 0.5 * torch.sum((H + F) * D) + e_nuc
        P = sp2_projector(X.T @ diis.step(F, F @ D @ S - S @ D @ F) @ X, n_occ)
        D_new = 2.0 * X @ P @ X.T
        out[0] = E
        out[1] = torch.abs(E - E_prev)
        out[2] = torch.max(torch.abs(D_new - D))
        E_prev.copy_(E)
        D.copy_(D_new)

    # the first iteration runs eagerly on a side stream (it also creates the
    # J/K build's worklist and streams), then one iteration is captured
    side = torch.cuda.Stream(device=dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(side):
        iteration()
    torch.cuda.current_stream(dev).wait_stream(side)
    torch.cuda.synchronize(dev)
    it = 1
    converged = False
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        iteration()
    # capturing does not run the work: replays continue from the eager iteration
    while it < max_iter:
        g.replay()
        it += 1
        E, dE, dD = out.tolist()
        if dE < e_tol and dD < d_tol:
            converged = True
            break
    F = fock(D)
    E = float(0.5 * torch.sum((H + F) * D)) + e_nuc
    eps, C = solve(F)
    return RHFResult(E, it, converged, eps.cpu().numpy(), C.cpu().numpy(), D.cpu().numpy())
