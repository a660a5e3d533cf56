"""SCF checkpoint files: PySCF's chkfile layout in a NumPy .npz container.

PySCF's `chkfile.dump_scf(chkfile, mol, e_tot, mo_energy, mo_coeff, mo_occ)` writes an HDF5
file with the molecule (`mol`, a JSON string) and an `scf` group holding e_tot, mo_energy,
mo_occ and mo_coeff; `load_scf(chkfile)` returns (mol, scf_dict), and the reference's
callers (embyte/Framework/SIE.py:251-272) reload orbitals from it.  h5py is not available
here, so the same keys ('mol', 'scf/e_tot', 'scf/mo_energy', 'scf/mo_occ', 'scf/mo_coeff')
are stored in an uncompressed .npz, read back with allow_pickle=False.
"""
from __future__ import annotations

import json
import os

import numpy as np

from ..gto.mole import Mole

SCF_KEYS = ('e_tot', 'mo_energy', 'mo_occ', 'mo_coeff')


def _path(chkfile):
    return str(chkfile) if str(chkfile).endswith('.npz') else f'{chkfile}.npz'


def mol_dumps(mol):
    """The molecule as JSON: atoms in the unit they were given, basis name or dictionary."""
    return json.dumps({'atom': [[s, list(xyz)] for s, xyz in mol.atom], 'unit': mol.unit,
                       'basis': mol.basis, 'charge': mol.charge, 'spin': mol.spin,
                       'cart': bool(mol.cart)})


def mol_loads(text):
    d = json.loads(text)
    return Mole([(s, tuple(xyz)) for s, xyz in d['atom']], basis=d['basis'], unit=d['unit'],
                cart=d['cart'], charge=d['charge'], spin=d['spin'])


def dump_scf(chkfile, mol, e_tot, mo_energy, mo_coeff, mo_occ, overwrite_mol=True):
    """Write (or update) the checkpoint; other keys already in the file are kept."""
    path = _path(chkfile)
    data = {}
    if os.path.exists(path):
        with np.load(path, allow_pickle=False) as f:
            data = {k: f[k] for k in f.files}
    if overwrite_mol or 'mol' not in data:
        data['mol'] = np.asarray(mol_dumps(mol))
    for k, v in zip(SCF_KEYS, (e_tot, mo_energy, mo_occ, mo_coeff)):
        data[f'scf/{k}'] = np.asarray(v.cpu().numpy() if hasattr(v, 'cpu') else v, dtype=np.float64)
    tmp = path + '.tmp.npz'
    np.savez(tmp, **data)
    os.replace(tmp, path)
    return path


def load(chkfile, key):
    """One entry: 'mol' returns a Mole, 'scf' a dict, 'scf/<name>' an array or float."""
    with np.load(_path(chkfile), allow_pickle=False) as f:
        if key == 'mol':
            return mol_loads(str(f['mol']))
        if key == 'scf':
            return {k: _value(f[f'scf/{k}']) for k in SCF_KEYS if f'scf/{k}' in f.files}
        return _value(f[key])


def load_scf(chkfile):
    return load(chkfile, 'mol'), load(chkfile, 'scf')


def _value(a):
    return float(a) if a.ndim == 0 else a
