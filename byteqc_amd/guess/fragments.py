"""Superposition of fragment densities: the baseline the learned guess corrects.

Atoms are split into covalently bonded fragments (distance below 1.3 x the sum of covalent
radii).  Fragments that are rigid copies of one another (same elements, same internal
distances to 1e-3 Angstrom) share one reference density: this package's RHF on the isolated
fragment, computed once per (fragment, basis) and cached.  Each copy gets the reference
density rotated into its orientation, D_copy = M(R) P D_ref P^T M(R)^T, where R and the atom
permutation P come from a Kabsch fit (every permutation of equal elements is tried for small
fragments) and M(R) is the block-diagonal real-solid-harmonic representation of R on the
fragment's shells.  Blocks between fragments are zero.  For water clusters of rigid
monomers this baseline is exact for every isolated molecule, so what is left to learn is the
intermolecular polarisation and charge transfer.
"""
from __future__ import annotations

import itertools

import numpy as np

from . import frames as fr

COVALENT = {'H': 0.31, 'C': 0.76, 'N': 0.71, 'O': 0.66}
MAX_PERM_ATOMS = 8
_REF_CACHE = {}


def find_fragments(symbols, xyz):
    """Connected components of the covalent-bond graph -> list of atom index lists."""
    n = len(symbols)
    rad = np.array([COVALENT.get(s, 0.8) for s in symbols])
    d = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1)
    bonded = d < 1.3 * (rad[:, None] + rad[None])
    seen = np.zeros(n, dtype=bool)
    out = []
    for a in range(n):
        if seen[a]:
            continue
        stack, comp = [a], []
        seen[a] = True
        while stack:
            u = stack.pop()
            comp.append(u)
            for v in np.nonzero(bonded[u])[0]:
                if not seen[v]:
                    seen[v] = True
                    stack.append(int(v))
        out.append(sorted(comp))
    return out


def _kabsch(x, y):
    """Rotation R (det +1) minimising |y - x R^T| for centred x, y; returns (R, rmsd)."""
    h = x.T @ y
    u, _, vt = np.linalg.svd(h)
    dd = np.sign(np.linalg.det(vt.T @ u.T))
    r = vt.T @ np.diag([1.0, 1.0, dd]) @ u.T
    rmsd = float(np.sqrt(np.mean(np.sum((x @ r.T - y) ** 2, axis=1))))
    return r, rmsd


def _key(symbols):
    return tuple(sorted(symbols))


def align(ref_sym, ref_xyz, sym, xyz, tol=1e-3):
    """(R, perm) with xyz[perm[k]] - c = R (ref_xyz[k] - c_ref), or None if not a rigid copy."""
    if _key(ref_sym) != _key(sym) or len(sym) > MAX_PERM_ATOMS:
        return None
    x = ref_xyz - ref_xyz.mean(0)
    y0 = xyz - xyz.mean(0)
    groups = {}
    for k, s in enumerate(ref_sym):
        groups.setdefault(s, []).append(k)
    cand = {s: [i for i, t in enumerate(sym) if t == s] for s in groups}
    best = None
    for choice in itertools.product(*[itertools.permutations(cand[s]) for s in groups]):
        perm = np.empty(len(sym), dtype=int)
        for s, pc in zip(groups, choice):
            perm[groups[s]] = pc
        r, rmsd = _kabsch(x, y0[perm])
        if best is None or rmsd < best[2]:
            best = (r, perm, rmsd)
    if best is None or best[2] > tol:
        return None
    return best[0], best[1]


def reference_density(sym, xyz, basis):
    """Converged RHF density of an isolated fragment (spherical AOs, atoms in this order)."""
    from ..gto.mole import Mole
    from ..scf.rhf import RHF
    key = (tuple(sym), tuple(np.round(np.asarray(xyz).ravel(), 6)), str(basis).lower())
    if key not in _REF_CACHE:
        mol = Mole(list(zip(sym, [tuple(p) for p in xyz])), basis=basis)
        mf = RHF(mol, conv_tol=1e-10)
        mf.kernel()
        if not mf.converged:
            raise RuntimeError(f'fragment SCF did not converge ({sym})')
        _REF_CACHE[key] = mf.dm.cpu().numpy()
    return _REF_CACHE[key]


def fragment_density(mol):
    """Block-diagonal superposition of fragment densities in mol's AO basis."""
    sym, xyz = fr.mol_geometry(mol)
    layout = fr.element_layout(mol)
    sl = mol.aoslice_by_atom()
    dm = np.zeros((mol.nao, mol.nao))
    refs = []   # (symbols, coordinates) of the fragment types met so far
    for frag in find_fragments(sym, xyz):
        fs = [sym[a] for a in frag]
        fx = xyz[frag]
        hit = None
        for rs, rx in refs:
            al = align(rs, rx, fs, fx)
            if al is not None:
                hit = (rs, rx, al)
                break
        if hit is None:
            refs.append((fs, fx))
            hit = (fs, fx, (np.eye(3), np.arange(len(fs))))
        rs, rx, (rot, perm) = hit
        dref = reference_density(rs, rx, mol.basis)
        # AO offsets of the reference fragment's atoms (its own AO order)
        sizes = [layout[s][1] for s in rs]
        roff = np.append(0, np.cumsum(sizes))
        m = {s: fr.atom_rotations(layout[s][0], rot[None])[0] for s in set(rs)}
        for a in range(len(rs)):
            ia = frag[perm[a]]
            for b in range(len(rs)):
                ib = frag[perm[b]]
                blk = dref[roff[a]:roff[a + 1], roff[b]:roff[b + 1]]
                dm[sl[ia, 2]:sl[ia, 3], sl[ib, 2]:sl[ib, 3]] = m[rs[a]] @ blk @ m[rs[b]].T
    return dm
