"""Named workloads (seeded water clusters) used by the benchmark and the tests."""
from __future__ import annotations

from ..gto import water_cluster

# name: (n_waters, box_angstrom, seed, basis)
WORKLOADS = {
    'h2o8-ccpvdz': (8, 9.0, 1, 'cc-pvdz'),        # 192 AO
    'h2o64-def2tzvp': (64, 16.0, 1, 'def2-tzvp'),  # 2752 AO, 3072 Cartesian
    'h2o256-ccpvdz': (256, 25.0, 1, 'cc-pvdz'),    # 6144 AO
}


def make_mol(name):
    n, box, seed, basis = WORKLOADS[name]
    return water_cluster(n, box, seed, basis=basis)
