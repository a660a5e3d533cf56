"""Atom-pair local frames, features and density-block rotations for the learned guess.

Every atom pair (A, B) closer than `R_PAIR` (and every atom with itself) gets a frame:
  origin  the midpoint of A and B (A itself for a diagonal block);
  z       along B - A (for a diagonal block: the distance-weighted mean direction of the
          neighbours);
  x       the principal axis of the neighbours' weighted perpendicular scatter, its sign
          fixed by the weighted third moment along it;
  y       z cross x.
The frame turns with the molecule, so a block predicted in it and rotated back is
rotation-covariant: D_AB = M_A(F)^T D_loc M_B(F), where M_l(F) maps the real solid
harmonics of degree l, Y_l(F r) = M_l(F) Y_l(r), and M_A is block-diagonal over A's shells.
Features are invariant: the pair type, the distance, and the K nearest neighbours of the
origin in frame coordinates with a smooth cutoff.
"""
from __future__ import annotations

from functools import lru_cache

import numpy as np

from ..gto.cart2sph import cart2sph_block, cart_components
from ..gto.mole import BOHR

R_PAIR = 4.5     # Angstrom: pairs further apart get a zero block
R_ENV = 6.0      # Angstrom: neighbour cutoff of the features
SIGMA = 2.0      # Angstrom: frame weights exp(-(r / SIGMA)^2)
K_NEIGH = 16
ELEMS = ('O', 'H')
# pair types: diagonal O, diagonal H, then ordered element pairs with Z_A >= Z_B
TYPES = ('OO0', 'HH0', 'OO', 'OH', 'HH')
N_FEAT_RAW = len(TYPES) + 2 + 5 * K_NEIGH
N_IN = 96        # padded input width (a multiple of 32)


def pair_type(sa, sb, diag):
    if diag:
        return TYPES.index(sa + sa + '0')
    return TYPES.index(sa + sb)


# ---- rotations of real solid harmonics ------------------------------------------------
@lru_cache(maxsize=None)
def _sample(l):
    rng = np.random.default_rng(1234 + l)
    pts = rng.normal(size=(3, 4 * (2 * l + 1)))
    return pts, np.linalg.pinv(_harmonics(l, pts))


def _harmonics(l, pts):
    """Y_l at points (3, S) -> (2l+1, S) in the AO order of cart2sph_block."""
    comps = cart_components(l)
    mono = np.stack([pts[..., 0, :] ** a * pts[..., 1, :] ** b * pts[..., 2, :] ** c for a, b, c in comps], axis=-2)
    return np.swapaxes(cart2sph_block(l), 0, 1) @ mono


def harmonic_rotations(l, frames):
    """M_l(F) for frames (P, 3, 3) (rows: local x, y, z in global coordinates) -> (P, 2l+1, 2l+1)."""
    if l == 0:
        return np.ones((len(frames), 1, 1))
    if l == 1:
        return frames.copy()      # p functions are (x, y, z): Y_1(F r) = F r
    pts, pinv = _sample(l)
    rotated = frames @ pts        # (P, 3, S)
    return _harmonics(l, rotated) @ pinv


def atom_rotations(shells, frames):
    """Block-diagonal M over an atom's shells [(l, nctr), ...] -> (P, n, n)."""
    n = sum((2 * l + 1) * c for l, c in shells)
    out = np.zeros((len(frames), n, n))
    cache = {}
    i = 0
    for l, c in shells:
        if l not in cache:
            cache[l] = harmonic_rotations(l, frames)
        d = 2 * l + 1
        for _ in range(c):
            out[:, i:i + d, i:i + d] = cache[l]
            i += d
    return out


# ---- frames and features ---------------------------------------------------------------
def _frame(origin, zdir, others, wts):
    z = zdir / np.linalg.norm(zdir)
    x = None
    if len(others):
        u = others - origin
        u = u - np.outer(u @ z, z)
        t = (u * wts[:, None]).T @ u
        ev, evec = np.linalg.eigh(t)
        if ev[-1] > 1e-8:
            x = evec[:, -1]
            if np.sum(wts * (u @ x) ** 3) < 0:
                x = -x
    if x is None:
        a = np.eye(3)[np.argmin(np.abs(z))]
        x = a - (a @ z) * z
    x = x / np.linalg.norm(x)
    y = np.cross(z, x)
    return np.stack([x, y, z])


def pair_list(symbols, xyz):
    """Pairs (a, b, diag) with Z_a >= Z_b, |r_ab| < R_PAIR, plus every (a, a)."""
    n = len(symbols)
    d = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1)
    out = [(a, a, True) for a in range(n)]
    for a in range(n):
        for b in range(a + 1, n):
            if d[a, b] < R_PAIR:
                if symbols[a] == 'H' and symbols[b] == 'O':
                    out.append((b, a, False))
                else:
                    out.append((a, b, False))
    return out


def features(symbols, xyz_ang, pairs=None):
    """(features (P, N_IN) float32, frames (P, 3, 3), types (P,), pairs)."""
    xyz = np.asarray(xyz_ang, dtype=float)
    if pairs is None:
        pairs = pair_list(symbols, xyz)
    isO = np.array([s == 'O' for s in symbols], dtype=float)
    feats = np.zeros((len(pairs), N_IN), dtype=np.float32)
    frames = np.zeros((len(pairs), 3, 3))
    types = np.zeros(len(pairs), dtype=np.int64)
    idx = np.arange(len(symbols))
    for p, (a, b, diag) in enumerate(pairs):
        origin = xyz[a] if diag else 0.5 * (xyz[a] + xyz[b])
        mask = (idx != a) & (idx != b)
        others = xyz[mask]
        r = np.linalg.norm(others - origin, axis=1)
        wts = np.exp(-(r / SIGMA) ** 2)
        if diag:
            zdir = (wts[:, None] * (others - origin) / np.maximum(r, 1e-12)[:, None]).sum(0) if len(others) else np.array([0., 0., 1.])
            if np.linalg.norm(zdir) < 1e-8:
                zdir = np.array([0., 0., 1.])
        else:
            zdir = xyz[b] - xyz[a]
        f = _frame(origin, zdir, others, wts)
        frames[p] = f
        t = pair_type(symbols[a], symbols[b], diag)
        types[p] = t
        row = feats[p]
        row[t] = 1.0
        rab = 0.0 if diag else float(np.linalg.norm(xyz[b] - xyz[a]))
        row[len(TYPES)] = rab
        row[len(TYPES) + 1] = 1.0 / rab if rab > 0 else 0.0
        loc = (others - origin) @ f.T
        # nearest first; equal distances ordered by frame coordinates (rotation-invariant)
        order = np.lexsort((loc[:, 2], loc[:, 1], loc[:, 0], np.round(r, 6)))[:K_NEIGH]
        base = len(TYPES) + 2
        oth_idx = idx[mask]
        for k, j in enumerate(order):
            if r[j] >= R_ENV:
                break
            fc = 0.5 * (np.cos(np.pi * r[j] / R_ENV) + 1.0)
            row[base + 5 * k:base + 5 * k + 3] = loc[j] * fc
            row[base + 5 * k + 3] = fc * isO[oth_idx[j]]
            row[base + 5 * k + 4] = fc * (1.0 - isO[oth_idx[j]])
    return feats, frames, types, pairs


# ---- basis layout ------------------------------------------------------------------------
def atom_shells(mol, ia):
    """[(l, nctr), ...] of atom ia in AO order (spherical)."""
    bas = mol._bas
    out = []
    for ib in range(mol.nbas):
        if bas[ib, 0] == ia:
            out.append((int(bas[ib, 1]), int(bas[ib, 3])))
    return out


def element_layout(mol):
    """{symbol: (shells, nao)} for every element of mol."""
    out = {}
    for ia in range(mol.natm):
        s = mol.atom_symbol(ia)
        if s not in out:
            sh = atom_shells(mol, ia)
            out[s] = (sh, sum((2 * l + 1) * c for l, c in sh))
    return out


def mol_geometry(mol):
    symbols = [mol.atom_symbol(i) for i in range(mol.natm)]
    return symbols, mol.atom_coords() * BOHR


def extract_blocks(mol, dm, pairs, frames, nmax):
    """Local-frame blocks M_A D_AB M_B^T of a density (nao, nao) -> (P, nmax * nmax)."""
    layout = element_layout(mol)
    sl = mol.aoslice_by_atom()
    symbols = [mol.atom_symbol(i) for i in range(mol.natm)]
    out = np.zeros((len(pairs), nmax * nmax), dtype=np.float32)
    rot = {s: atom_rotations(layout[s][0], frames) for s in layout}
    for p, (a, b, diag) in enumerate(pairs):
        blk = dm[sl[a, 2]:sl[a, 3], sl[b, 2]:sl[b, 3]]
        loc = rot[symbols[a]][p] @ blk @ rot[symbols[b]][p].T
        na, nb = loc.shape
        tmp = np.zeros((nmax, nmax))
        tmp[:na, :nb] = loc
        out[p] = tmp.ravel()
    return out


def assemble(mol, blocks, pairs, frames, nmax):
    """Inverse of extract_blocks: a symmetric (nao, nao) density from local-frame blocks."""
    layout = element_layout(mol)
    sl = mol.aoslice_by_atom()
    symbols = [mol.atom_symbol(i) for i in range(mol.natm)]
    rot = {s: atom_rotations(layout[s][0], frames) for s in layout}
    dm = np.zeros((mol.nao, mol.nao))
    for p, (a, b, diag) in enumerate(pairs):
        na, nb = layout[symbols[a]][1], layout[symbols[b]][1]
        loc = np.asarray(blocks[p], dtype=float).reshape(nmax, nmax)[:na, :nb]
        blk = rot[symbols[a]][p].T @ loc @ rot[symbols[b]][p]
        if diag:
            dm[sl[a, 2]:sl[a, 3], sl[a, 2]:sl[a, 3]] = 0.5 * (blk + blk.T)
        else:
            dm[sl[a, 2]:sl[a, 3], sl[b, 2]:sl[b, 3]] = blk
            dm[sl[b, 2]:sl[b, 3], sl[a, 2]:sl[a, 3]] = blk.T
    return dm


def block_masks(mol, nmax):
    """(len(TYPES), nmax * nmax) float32: which entries of a padded block are real."""
    layout = element_layout(mol)
    m = np.zeros((len(TYPES), nmax, nmax), dtype=np.float32)
    for t, name in enumerate(TYPES):
        sa, sb = name[0], name[1]
        if sa in layout and sb in layout:
            m[t, :layout[sa][1], :layout[sb][1]] = 1.0
    return m.reshape(len(TYPES), -1)
