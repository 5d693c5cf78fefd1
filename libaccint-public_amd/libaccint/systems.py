"""Molecular systems and densities of the benchmark workloads (seeded).

All coordinates in bohr.  Each builder returns a list of (Z, (x, y, z)).

* ``h2o_sto3g_geometry``: O at the origin, H at (0, +-1.43233673, -1.10866041).
* ``water_cluster(n, cube, seed)``: rigid waters (O-H 1.81 bohr, HOH 104.5 deg),
  uniformly random orientations, O positions drawn uniformly in a cube of the
  given edge and rejected unless every O-O distance is >= 5.0 bohr.
* ``alkane(n_c)``: all-trans zigzag CnH(2n+2), C-C 2.91 bohr, C-H 2.06 bohr,
  tetrahedral angles, backbone in the xy plane.
* ``random_molecule(rng)``: 3-60 atoms of H/C/N/O drawn with weights
  0.5/0.3/0.1/0.1, grown by random-walk bonding at 2.0-2.9 bohr.
* ``density(nbf, nocc, seed, scale)``: D = C C^T with C ~ N(0, scale^2).
"""
import math

import numpy as np

TETRA = math.acos(-1.0 / 3.0)  # 109.47 deg


def h2o_sto3g_geometry():
    return [(8, (0.0, 0.0, 0.0)), (1, (0.0, 1.43233673, -1.10866041)),
            (1, (0.0, -1.43233673, -1.10866041))]


def _random_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster(n, cube=14.0, seed=0, r_oh=1.81, hoh_deg=104.5, min_oo=5.0):
    rng = np.random.default_rng(seed)
    half = math.radians(hoh_deg) / 2
    body = np.array([[0.0, 0.0, 0.0],
                     [0.0, r_oh * math.sin(half), -r_oh * math.cos(half)],
                     [0.0, -r_oh * math.sin(half), -r_oh * math.cos(half)]])
    oxy = []
    atoms = []
    tries = 0
    while len(oxy) < n:
        tries += 1
        if tries > 200000:
            raise RuntimeError("water_cluster: could not place %d waters" % n)
        p = rng.uniform(0.0, cube, 3)
        if any(np.linalg.norm(p - o) < min_oo for o in oxy):
            continue
        R = _random_rotation(rng)
        xyz = body @ R.T + p
        oxy.append(p)
        atoms.append((8, tuple(xyz[0])))
        atoms.append((1, tuple(xyz[1])))
        atoms.append((1, tuple(xyz[2])))
    return atoms


def alkane(n_c=40, r_cc=2.91, r_ch=2.06):
    half = TETRA / 2
    dx = r_cc * math.sin(half)
    dy = r_cc * math.cos(half)
    C = np.array([[i * dx, (i % 2) * dy, 0.0] for i in range(n_c)])
    atoms = [(6, tuple(c)) for c in C]

    def h_pair(i, prev, nxt):
        u = -(_unit(prev - C[i]) + _unit(nxt - C[i]))
        u = _unit(u)
        a = TETRA / 2
        return [C[i] + r_ch * (math.cos(a) * u + s * math.sin(a) * np.array([0.0, 0.0, 1.0]))
                for s in (1.0, -1.0)]

    for i in range(n_c):
        prev = C[i - 1] if i > 0 else np.array([-dx, dy if (1 % 2) else 0.0, 0.0])
        nxt = C[i + 1] if i + 1 < n_c else np.array([(i + 1) * dx, ((i + 1) % 2) * dy, 0.0])
        if i == 0:
            virt = np.array([-dx, dy, 0.0])  # zigzag continuation
            atoms.append((1, tuple(C[0] + r_ch * _unit(virt - C[0]))))
            prev = virt
        if i == n_c - 1:
            virt = np.array([(i + 1) * dx, ((i + 1) % 2) * dy, 0.0])
            atoms.append((1, tuple(C[i] + r_ch * _unit(virt - C[i]))))
            nxt = virt
        for h in h_pair(i, prev, nxt):
            atoms.append((1, tuple(h)))
    return atoms


def _unit(v):
    return v / np.linalg.norm(v)


def random_molecule(rng, n_min=3, n_max=60):
    n = int(rng.integers(n_min, n_max + 1))
    Zs = rng.choice([1, 6, 7, 8], size=n, p=[0.5, 0.3, 0.1, 0.1])
    pos = [np.zeros(3)]
    for _ in range(1, n):
        for _attempt in range(1000):
            base = pos[int(rng.integers(len(pos)))]
            d = rng.standard_normal(3)
            d /= np.linalg.norm(d)
            p = base + rng.uniform(2.0, 2.9) * d
            if all(np.linalg.norm(p - q) >= 1.8 for q in pos):
                pos.append(p)
                break
        else:
            raise RuntimeError("random_molecule: placement failed")
    return [(int(z), tuple(p)) for z, p in zip(Zs, pos)]


def density(nbf, nocc, seed, scale):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((nbf, nocc)) * scale
    return C @ C.T
