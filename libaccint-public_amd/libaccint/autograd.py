"""torch.autograd.Function for the Coulomb/exchange energy of a fixed AO density
as a function of the atomic coordinates.

    E(R) = 1/2 tr(D J[D]) - x/4 tr(D K[D])

with J and K built on the GPU, and dE/dR from the gfx950 adjoint gradient
kernels (derivative integrals contracted with D (x) D on the fly).  The basis
follows the atoms (basis functions move with their nuclei); D is held fixed,
as in a finite-difference check of the same energy.
"""
import numpy as np
import torch

from . import _core


def _engine_for(coords, Z, basis_name):
    xyz = coords.detach().to("cpu", torch.float64).numpy()
    atoms = [_core.Atom(int(z), [float(c) for c in p]) for z, p in zip(Z, xyz)]
    basis = _core.load_basis_set(basis_name, atoms)
    return basis, _core.Engine(basis)


class JKEnergy(torch.autograd.Function):
    """E = 1/2 tr(DJ) - x/4 tr(DK); forward(coords[n_atoms,3], Z, basis_name, D, x)."""

    @staticmethod
    def forward(ctx, coords, Z, basis_name, D, exchange_scale=1.0):
        basis, engine = _engine_for(coords, Z, basis_name)
        Dn = np.ascontiguousarray(D.detach().to("cpu", torch.float64).numpy())
        J, K = engine.build_jk(Dn)
        E = 0.5 * float(np.sum(Dn * J)) - 0.25 * exchange_scale * float(np.sum(Dn * K))
        ctx.state = (basis, engine, Dn, float(exchange_scale), len(Z), coords.device)
        return torch.tensor(E, dtype=torch.float64, device=coords.device)

    @staticmethod
    def backward(ctx, grad_out):
        basis, engine, Dn, x, n_atoms, dev = ctx.state
        g = engine.compute_jk_gradient(Dn, x, n_atoms)
        g = torch.from_numpy(np.ascontiguousarray(g)).to(dev)
        return grad_out * g, None, None, None, None


def jk_energy(coords, Z, basis_name, D, exchange_scale=1.0):
    return JKEnergy.apply(coords, list(Z), basis_name, D, exchange_scale)
