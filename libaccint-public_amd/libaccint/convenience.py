"""Convenience functions with the reference's names
(python/libaccint/convenience.py:25-279)."""
import numpy as np

from . import _core


def atoms_from_list(atoms):
    """[(Z, (x, y, z)), ...] in bohr -> [Atom, ...]."""
    return [_core.Atom(int(z), [float(c) for c in xyz]) for z, xyz in atoms]


def basis_set(name, atoms):
    """Built-in STO-3G or a Basis Set Exchange JSON from the data directory."""
    if atoms and not isinstance(atoms[0], _core.Atom):
        atoms = atoms_from_list(atoms)
    return _core.load_basis_set(name, atoms)


def build_fock(engine, D, H_core=None, exchange_fraction=1.0, screening=None):
    """Fock matrix F = H + J - x K for density D (J, K built on the GPU)."""
    D = np.ascontiguousarray(D, dtype=np.float64)
    n = D.shape[0]
    fb = _core.FockBuilder(n)
    fb.set_density(D)
    if screening is None:
        engine.compute_and_consume(_core.Operator.coulomb(), fb)
    else:
        engine.compute_and_consume_screened(_core.Operator.coulomb(), fb, screening)
    if H_core is None:
        H_core = np.zeros((n, n))
    return fb.get_fock_matrix(np.ascontiguousarray(H_core, dtype=np.float64), exchange_fraction)
