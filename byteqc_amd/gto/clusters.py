"""Seeded water clusters used by the J/K workloads.

Water monomer: O-H 0.957 Angstrom, H-O-H 104.5 degrees.  Each molecule gets a uniform
random rotation (normalised Gaussian quaternion) and an O position uniform in [0, L)^3,
redrawn while any O-O distance is below `min_oo` (2.6 Angstrom).  All randomness comes
from ``numpy.random.default_rng(seed)`` so a (n, L, seed) triple names a geometry.
"""
from __future__ import annotations

import numpy as np

from .mole import Mole

R_OH = 0.957
THETA_HOH = np.deg2rad(104.5)


def water_monomer():
    """(3, 3) coordinates of O, H, H in Angstrom, O at the origin."""
    h = THETA_HOH / 2
    return np.array([[0., 0., 0.],
                     [R_OH * np.sin(h), 0., R_OH * np.cos(h)],
                     [-R_OH * np.sin(h), 0., R_OH * np.cos(h)]])


def _rotation(q):
    q = q / np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster_coords(n, box, seed, min_oo=2.6, max_tries=100000):
    rng = np.random.default_rng(seed)
    mono = water_monomer()
    oxy = []
    coords = []
    tries = 0
    while len(oxy) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f'could not place {n} waters in a {box} A box')
        rot = _rotation(rng.normal(size=4))
        o = rng.uniform(0., box, size=3)
        if oxy and np.min(np.linalg.norm(np.asarray(oxy) - o, axis=1)) < min_oo:
            continue
        oxy.append(o)
        coords.append(mono @ rot.T + o)
    return np.concatenate(coords, axis=0)


def water_cluster(n, box, seed, basis='cc-pvdz', **kw):
    xyz = water_cluster_coords(n, box, seed)
    atoms = [(('O', 'H', 'H')[i % 3], tuple(xyz[i])) for i in range(len(xyz))]
    return Mole(atoms, basis=basis, **kw)
