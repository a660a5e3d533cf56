"""Molecule and Gaussian basis model in the PySCF `_atm/_bas/_env` layout.

The reference takes a PySCF ``Mole`` (cuobc/scf/hf.py:53 `get_jk(mol, ...)`) and only
reads its integer/float tables (`_atm`, `_bas`, `_env`, hf.py:296, 464-468) plus a few
helpers (`ao_loc_nr`, `cart2sph_coeff`, `aoslice_by_atom`).  PySCF is not available in
this environment, so this module provides a small, self-contained molecule object with
exactly those tables and helpers.  The slot layout and the normalisation convention
(radially normalised primitives, each contracted column normalised to one, Cartesian
s/p carrying the 1/(2 sqrt(pi)) and sqrt(3/(4 pi)) angular factors) follow PySCF so a
real PySCF ``Mole`` can be handed to the same driver.
"""
from __future__ import annotations

import json
import os
from functools import lru_cache

import numpy as np

from .cart2sph import cart2sph_block

BOHR = 0.52917721092  # Angstrom per bohr (the value PySCF uses)

# _atm slots
CHARGE_OF, PTR_COORD, NUC_MOD_OF, PTR_ZETA, PTR_FRAC_CHARGE, ATM_SLOTS = 0, 1, 2, 3, 4, 6
# _bas slots
ATOM_OF, ANG_OF, NPRIM_OF, NCTR_OF, KAPPA_OF, PTR_EXP, PTR_COEFF, BAS_SLOTS = 0, 1, 2, 3, 4, 5, 6, 8
PTR_ENV_START = 20

ELEMENTS = {'H': 1, 'He': 2, 'C': 6, 'N': 7, 'O': 8}
ANGULAR = 'spdfghi'

_BASIS_DIR = os.path.join(os.path.dirname(__file__), 'basis')
_BASIS_ALIASES = {
    'ccpvdz': 'cc-pvdz', 'cc-pvdz': 'cc-pvdz',
    'def2tzvp': 'def2-tzvp', 'def2-tzvp': 'def2-tzvp',
    '631gs': '6-31gs', '6-31gs': '6-31gs', '6-31g*': '6-31gs', '631g*': '6-31gs',
}


def nf_cart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def gaussian_int(n, alpha):
    """int_0^inf x^n exp(-alpha x^2) dx."""
    n1 = (n + 1) * .5
    from scipy.special import gamma
    return gamma(n1) / (2. * alpha ** n1)


def gto_norm(l, expnt):
    """Radial normalisation 1/sqrt(int r^(2l+2) exp(-2 a r^2) dr)."""
    return 1. / np.sqrt(gaussian_int(l * 2 + 2, 2 * np.asarray(expnt, dtype=float)))


@lru_cache(maxsize=None)
def load_basis(name: str) -> dict:
    key = _BASIS_ALIASES.get(name.lower().replace('_', ''), name.lower())
    path = os.path.join(_BASIS_DIR, key + '.json')
    if not os.path.exists(path):
        raise KeyError(f'basis {name!r} is not available (known: {sorted(set(_BASIS_ALIASES.values()))})')
    with open(path) as f:
        data = json.load(f)
    data.pop('_comment', None)
    return data


def _normalized_contraction(l, exps, coefs):
    """Coefficients with gto_norm folded in and each column normalised (PySCF make_bas_env)."""
    exps = np.asarray(exps, dtype=float)
    cs = np.asarray(coefs, dtype=float).reshape(len(exps), -1)
    cs = cs * gto_norm(l, exps)[:, None]
    ee = gaussian_int(l * 2 + 2, exps[:, None] + exps[None, :])
    s1 = 1. / np.sqrt(np.einsum('pi,pq,qi->i', cs, ee, cs))
    return cs * s1[None, :]


class Mole:
    """Minimal molecule with PySCF-layout integral tables.

    atoms: list of (symbol, (x, y, z)) in `unit` (Angstrom by default).
    basis: basis name (one of the JSON files shipped in ``gto/basis``) or a dict
           {symbol: [[l, [exp, c1, c2, ...], ...], ...]}.
    """

    def __init__(self, atoms, basis='cc-pvdz', unit='Angstrom', cart=False, charge=0, spin=0):
        self.unit = unit
        self.cart = cart
        self.charge = charge
        self.spin = spin
        self.verbose = 0
        self.basis = basis
        self.atom = [(sym, tuple(float(c) for c in xyz)) for sym, xyz in atoms]
        bas_dict = load_basis(basis) if isinstance(basis, str) else basis
        fac = 1. / BOHR if unit.lower().startswith('a') else 1.
        env = [0.] * PTR_ENV_START
        atm, bas = [], []
        templates = {}
        for ia, (sym, xyz) in enumerate(self.atom):
            z = ELEMENTS[sym]
            ptr = len(env)
            env.extend([xyz[0] * fac, xyz[1] * fac, xyz[2] * fac, 0.])
            atm.append([z, ptr, 1, ptr + 3, 0, 0])
            if sym not in templates:
                shells = []
                for shell in bas_dict[sym]:
                    l = int(shell[0])
                    prim = np.asarray(shell[1:], dtype=float)
                    exps = prim[:, 0]
                    cs = _normalized_contraction(l, exps, prim[:, 1:])
                    shells.append((l, exps, cs))
                templates[sym] = shells
            for l, exps, cs in templates[sym]:
                pexp = len(env)
                env.extend(exps.tolist())
                pcoeff = len(env)
                env.extend(cs.T.ravel().tolist())  # [nctr][nprim], as PySCF stores it
                bas.append([ia, l, len(exps), cs.shape[1], 0, pexp, pcoeff, 0])
        self._atm = np.asarray(atm, dtype=np.int32).reshape(-1, ATM_SLOTS)
        self._bas = np.asarray(bas, dtype=np.int32).reshape(-1, BAS_SLOTS)
        self._env = np.asarray(env, dtype=np.float64)

    # ---- basic properties -------------------------------------------------
    @property
    def natm(self):
        return len(self._atm)

    @property
    def nbas(self):
        return len(self._bas)

    @property
    def nelectron(self):
        return int(self._atm[:, CHARGE_OF].sum()) - self.charge

    @property
    def nelec(self):
        ne = self.nelectron
        return ((ne + self.spin) // 2, (ne - self.spin) // 2)

    def atom_charges(self):
        return self._atm[:, CHARGE_OF].astype(float)

    def atom_coords(self):
        ptr = self._atm[:, PTR_COORD]
        return np.stack([self._env[ptr], self._env[ptr + 1], self._env[ptr + 2]], axis=1)

    def atom_symbol(self, ia):
        return self.atom[ia][0]

    def bas_atom(self, ib):
        return int(self._bas[ib, ATOM_OF])

    def bas_angular(self, ib):
        return int(self._bas[ib, ANG_OF])

    def bas_nprim(self, ib):
        return int(self._bas[ib, NPRIM_OF])

    def bas_nctr(self, ib):
        return int(self._bas[ib, NCTR_OF])

    def bas_exp(self, ib):
        p, n = self._bas[ib, PTR_EXP], self._bas[ib, NPRIM_OF]
        return self._env[p:p + n].copy()

    def bas_ctr_coeff(self, ib):
        """(nprim, nctr) contraction coefficients as stored in _env (norm folded in)."""
        p, n, c = self._bas[ib, PTR_COEFF], self._bas[ib, NPRIM_OF], self._bas[ib, NCTR_OF]
        return self._env[p:p + n * c].reshape(c, n).T.copy()

    def bas_coord(self, ib):
        ia = self._bas[ib, ATOM_OF]
        p = self._atm[ia, PTR_COORD]
        return self._env[p:p + 3].copy()

    def ao_loc_nr(self, cart=None):
        cart = self.cart if cart is None else cart
        l = self._bas[:, ANG_OF]
        nctr = self._bas[:, NCTR_OF]
        dims = ((l + 1) * (l + 2) // 2 if cart else 2 * l + 1) * nctr
        return np.append(0, np.cumsum(dims)).astype(np.int64)

    @property
    def ao_loc(self):
        return self.ao_loc_nr()

    @property
    def nao(self):
        return int(self.ao_loc_nr()[-1])

    def nao_nr(self, cart=None):
        return int(self.ao_loc_nr(cart)[-1])

    def aoslice_by_atom(self):
        """(natm, 4): shell start, shell stop, AO start, AO stop."""
        ao_loc = self.ao_loc_nr()
        out = np.zeros((self.natm, 4), dtype=np.int64)
        atoms = self._bas[:, ATOM_OF]
        for ia in range(self.natm):
            idx = np.where(atoms == ia)[0]
            if len(idx):
                out[ia] = (idx[0], idx[-1] + 1, ao_loc[idx[0]], ao_loc[idx[-1] + 1])
            elif ia:
                out[ia] = (out[ia - 1, 1], out[ia - 1, 1], out[ia - 1, 3], out[ia - 1, 3])
        return out

    def ao_labels(self):
        labels = []
        for ib in range(self.nbas):
            ia, l, nctr = self._bas[ib, ATOM_OF], self._bas[ib, ANG_OF], self._bas[ib, NCTR_OF]
            n = nf_cart(l) if self.cart else 2 * l + 1
            for c in range(nctr):
                for m in range(n):
                    labels.append((int(ia), self.atom[ia][0], ANGULAR[l], c, m))
        return labels

    def cart2sph_coeff(self):
        """(nao_cart, nao_sph) block-diagonal transform, PySCF AO order."""
        blocks = []
        for ib in range(self.nbas):
            l, nctr = int(self._bas[ib, ANG_OF]), int(self._bas[ib, NCTR_OF])
            for _ in range(nctr):
                blocks.append(cart2sph_block(l))
        nc = sum(b.shape[0] for b in blocks)
        ns = sum(b.shape[1] for b in blocks)
        c = np.zeros((nc, ns))
        i = j = 0
        for b in blocks:
            c[i:i + b.shape[0], j:j + b.shape[1]] = b
            i += b.shape[0]
            j += b.shape[1]
        return c

    def energy_nuc(self):
        z = self.atom_charges()
        r = self.atom_coords()
        e = 0.
        for i in range(self.natm):
            d = np.linalg.norm(r[:i] - r[i], axis=1)
            e += float(np.sum(z[i] * z[:i] / d))
        return e

    def copy(self):
        import copy
        new = copy.copy(self)
        new._atm = self._atm.copy()
        new._bas = self._bas.copy()
        new._env = self._env.copy()
        return new

    def __repr__(self):
        return f'Mole(natm={self.natm}, nbas={self.nbas}, nao={self.nao}, basis={self.basis!r})'


def M(atom, basis='cc-pvdz', **kw):
    """Build a Mole from PySCF-style atom text ('O 0 0 0; H 0 1 0') or a list."""
    if isinstance(atom, str):
        atoms = []
        for item in atom.replace('\n', ';').split(';'):
            item = item.strip()
            if not item:
                continue
            s = item.split()
            atoms.append((s[0], tuple(float(v) for v in s[1:4])))
        atom = atoms
    return Mole(atom, basis=basis, **kw)
