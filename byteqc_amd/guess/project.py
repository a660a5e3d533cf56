"""Projection of a density between basis sets on the same atoms.

D_to = P D_from P^T with P = S_tt^{-1} S_tf, the least-squares image of the "from" orbitals in
the "to" basis (the projection PySCF applies when it reads a density from another basis).
The overlaps come from one combined basis (every atom carries the shells of both sets), whose
one-electron integrals the library computes on the host.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from ..gto.mole import Mole, load_basis


def _combined(mol_from, mol_to):
    syms = sorted({mol_from.atom_symbol(i) for i in range(mol_from.natm)})
    bf, bt = load_basis(mol_from.basis), load_basis(mol_to.basis)
    basis = {s: list(bf[s]) + list(bt[s]) for s in syms}
    return Mole(mol_from.atom, basis=basis)


def cross_overlap(mol_from, mol_to):
    """(S_tt, S_tf): overlaps in the 'to' basis and between 'to' and 'from' (spherical AOs)."""
    from ..scf.hf import basis_seg_contraction
    from ..scf.rhf import int1e
    comb = _combined(mol_from, mol_to)
    bas, env, coeff = basis_seg_contraction(comb)
    s, _, _ = int1e(SimpleNamespace(atm=np.ascontiguousarray(comb._atm, dtype=np.int32),
                                    bas=np.ascontiguousarray(bas), env=np.ascontiguousarray(env),
                                    coeff=coeff))
    # AO indices of each basis inside the combined one: per atom, the 'from' shells come first
    sf, st = mol_from.aoslice_by_atom(), mol_to.aoslice_by_atom()
    sc = comb.aoslice_by_atom()
    idx_f, idx_t = [], []
    for ia in range(comb.natm):
        nf = sf[ia, 3] - sf[ia, 2]
        nt = st[ia, 3] - st[ia, 2]
        idx_f.extend(range(sc[ia, 2], sc[ia, 2] + nf))
        idx_t.extend(range(sc[ia, 2] + nf, sc[ia, 2] + nf + nt))
    idx_f, idx_t = np.asarray(idx_f), np.asarray(idx_t)
    return s[np.ix_(idx_t, idx_t)], s[np.ix_(idx_t, idx_f)]


def project_density(mol_from, dm_from, mol_to):
    s_tt, s_tf = cross_overlap(mol_from, mol_to)
    p = np.linalg.solve(s_tt, s_tf)
    return p @ dm_from @ p.T
