"""LibAccInt on MI355X -- Rys-quadrature ERIs and the Fock (J/K) build on gfx950.

Python surface of the C++ host layer (module ``_core``), with the names of the
reference package (python/libaccint/__init__.py, python/src/*_bindings.cpp):

>>> import libaccint
>>> atoms = [libaccint.Atom(8, [0.0, 0.0, 0.0]),
...          libaccint.Atom(1, [0.0, 1.43, -1.11]),
...          libaccint.Atom(1, [0.0, -1.43, -1.11])]
>>> basis = libaccint.basis_set("sto-3g", atoms)
>>> engine = libaccint.Engine(basis)
>>> S = engine.compute_overlap_matrix()

Every integral runs on the GPU; importing fails when the native libraries
are not built, and compute calls raise BackendError without a HIP device.
"""
__version__ = "0.1.0a2"

# torch and this package must share one HIP runtime (both load the SONAME
# libamdhip64.so.7); when torch is present it is loaded first so that its
# runtime is the one in the process and device pointers and streams from
# torch are valid here.
try:  # pragma: no cover - depends on the environment
    import torch as _torch  # noqa: F401
except ImportError:
    _torch = None

from . import _core as _core  # noqa: E402  (fails loudly when not built)

for _name in dir(_core):
    if not _name.startswith("_"):
        globals()[_name] = getattr(_core, _name)

from .convenience import basis_set, build_fock, atoms_from_list  # noqa: E402,F401
