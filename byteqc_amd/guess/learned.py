"""Learned initial density: pair features -> perceptron -> local-frame blocks -> D.

A trained model file (`weights_<basis>.npz`, written by `byteqc_amd.guess.train`) holds
the perceptron weights, the input scaling, the per-pair-type output mean and spread, the
basis layout it was trained for and its baseline.  `learned_guess(mol, s)` predicts every
block of pairs closer than R_PAIR (zero beyond) as a correction to the superposition of
fragment densities (`fragments.py`), rotates the blocks into the molecule's AO basis,
symmetrises and scales D so that tr(D S) equals the electron count.  The SCF then judges
the guess by energy and iteration count; SAD stays the default and the fallback for
elements or bases the model was not trained on.
"""
from __future__ import annotations

import os

import numpy as np

from . import frames as fr

_HERE = os.path.dirname(os.path.abspath(__file__))


def model_path(basis):
    key = str(basis).lower().replace('*', 's').replace('_', '')
    return os.path.join(_HERE, f'weights_{key}.npz')


class Model:
    def __init__(self, data):
        self.data = {k: data[k] for k in data.files} if hasattr(data, 'files') else dict(data)
        d = self.data
        self.nmax = int(d['nmax'])
        self.dims = tuple(int(v) for v in d['dims'])
        self.basis = str(d['basis'])
        self.elements = [str(e) for e in d['elements']]
        self.x_mean, self.x_std = d['x_mean'], d['x_std']
        self.y_mean, self.y_std = d['y_mean'], d['y_std']
        # 'fragments': the network predicts D - (superposed fragment densities)
        self.baseline = str(d['baseline']) if 'baseline' in d else 'none'
        self._mlp = None

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as f:
            return cls(f)

    def save(self, path):
        np.savez(path, **self.data)

    def params(self):
        return {k: self.data[k] for k in ('W1', 'W2', 'W3', 'b1', 'b2', 'b3')}

    def mlp(self):
        if self._mlp is None:
            from .mlp import MLP
            self._mlp = MLP(self.dims, params=self.params())
        return self._mlp

    def scale_inputs(self, feats):
        return ((feats - self.x_mean) / self.x_std).astype(np.float32)

    def blocks(self, feats, types):
        """Local-frame blocks (P, nmax * nmax) predicted on the GPU."""
        y = self.mlp().predict(self.scale_inputs(feats)).cpu().numpy()
        return self.y_mean[types] + self.y_std[types] * y

    def supports(self, mol):
        syms = {mol.atom_symbol(i) for i in range(mol.natm)}
        return syms <= set(self.elements) and str(mol.basis).lower() == self.basis


def learned_guess(mol, s, model=None, via=None):
    """Initial density (nao, nao) from the trained model for mol's basis, or, with via=<basis>,
    from the model of that basis projected into mol's basis (project.py)."""
    if via is not None and str(via).lower() != str(mol.basis).lower():
        from ..gto.mole import Mole
        from .project import project_density
        small = Mole(mol.atom, basis=via)
        from ..scf.hf import basis_seg_contraction
        from ..scf.rhf import int1e
        from types import SimpleNamespace
        bas, env, coeff = basis_seg_contraction(small)
        s_small, _, _ = int1e(SimpleNamespace(atm=np.ascontiguousarray(small._atm, dtype=np.int32), bas=bas,
                                              env=env, coeff=coeff))
        dm = project_density(small, learned_guess(small, s_small, model), mol)
        ne = float(np.sum(dm * s))
        if ne > 0:
            dm *= mol.nelectron / ne
        return dm
    if model is None:
        path = model_path(mol.basis)
        if not os.path.exists(path):
            raise FileNotFoundError(f'no learned-guess model for basis {mol.basis!r} ({path})')
        model = Model.load(path)
    if not model.supports(mol):
        raise ValueError(f'model trained for {model.basis}/{model.elements} does not cover {mol!r}')
    sym, xyz = fr.mol_geometry(mol)
    feats, frames, types, pairs = fr.features(sym, xyz)
    blocks = model.blocks(feats, types)
    dm = fr.assemble(mol, blocks, pairs, frames, model.nmax)
    if model.baseline == 'fragments':
        from .fragments import fragment_density
        dm = dm + fragment_density(mol)
    ne = float(np.sum(dm * s))
    if ne > 0:
        dm *= mol.nelectron / ne
    return dm
