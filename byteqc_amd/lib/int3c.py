"""Three- and two-centre Coulomb integrals and density-fitted J/K on the device.

The counterpart of cuobc/lib/int3c.py:177-373 (`get_int3c`, which fills (ij|P) blocks with
the 4-centre machinery).  Here the same is done without new kernels: the basis-product
cache is built over [orbital shells, one unit s shell, auxiliary shells], with pair classes
holding either orbital pairs (i, j) or auxiliary pairs (P, unit).  The unit shell has
exponent 0 and coefficient 2 sqrt(pi), the inverse of the s-shell angular factor the pair
data carries, so the pair (P, unit) is the auxiliary function P itself, and
`mjk_fill_int2e` of an auxiliary class against an orbital class writes (P 1|i j) = (ij|P).
Two auxiliary classes give (P|Q).

`Int3c(mol, auxmol)` holds the plan; `int3c()` returns (ij|P) as a (nao, nao, naux) device
tensor in the caller's spherical AOs, `int2c()` returns (P|Q), and `df_jk(dm)` builds J and
K through the Cholesky factor of (P|Q): B_P = L^-1 (P|ij), J = sum_P B_P (B_P . D),
K = sum_P B_P D B_P.  The (ij|P) tensor is nao^2 naux doubles: a tool for small and medium
systems and the golden data of density fitting, not a replacement for the direct J/K builder.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import check, load, ptr

UNIT_EXP = 0.0
UNIT_COEFF = 2.0 * np.sqrt(np.pi)  # 1 / (angular factor of an s shell, 1 / (2 sqrt(pi)))

# PySCF _bas column layout
ATOM_OF, ANG_OF, NPRIM_OF, NCTR_OF, PTR_EXP, PTR_COEFF = 0, 1, 2, 3, 5, 6


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('byteqc_amd.lib.int3c needs a ROCm GPU (torch.cuda is unavailable)')
    return torch


def _nf(l):
    return (l + 1) * (l + 2) // 2


def even_tempered_basis(mol, ratio=2.5, lmax=3, start=(0.12, 0.2, 0.3, 0.4), count=(10, 7, 5, 3),
                        light_count=(7, 4, 2, 0)):
    """An even-tempered auxiliary basis {symbol: [[l, [exp, 1.0]], ...]}: exponents
    start[l] * ratio**k, k < count[l] (light_count[l] for H and He), l <= lmax."""
    out = {}
    for sym, _ in mol.atom:
        if sym in out:
            continue
        n = light_count if sym in ('H', 'He') else count
        shells = []
        for l in range(lmax + 1):
            for k in range(n[l]):
                shells.append([l, [start[l] * ratio ** k, 1.0]])
        out[sym] = shells
    return out


class Int3c:
    """Plan for (ij|P) and (P|Q) over `mol` and the auxiliary basis `auxmol` (same atoms)."""

    def __init__(self, mol, auxmol, cutoff=1e-13):
        from ..scf.hf import _VHFOpt, basis_seg_contraction
        a1 = np.asarray([xyz for _, xyz in mol.atom])
        a2 = np.asarray([xyz for _, xyz in auxmol.atom])
        if a1.shape != a2.shape or not np.allclose(a1, a2):
            raise ValueError('the auxiliary basis must sit on the same atoms as the orbital basis')
        self.mol, self.auxmol = mol, auxmol
        vo = _VHFOpt(mol).build(cutoff, diag_block_with_triu=True)
        vo.clear()  # only its sorted shells, pair lists and coeff are used
        self.vo = vo
        abas, aenv, acoeff = basis_seg_contraction(auxmol)
        # auxiliary shells grouped by (l, nprim), stable inside a group
        order = np.lexsort((np.arange(len(abas)), abas[:, NPRIM_OF], abas[:, ANG_OF]))
        nfa = np.array([_nf(int(l)) for l in abas[:, ANG_OF]])
        aloc = np.append(0, np.cumsum(nfa))
        self.aux_coeff = np.ascontiguousarray(
            acoeff[np.concatenate([np.arange(aloc[i], aloc[i + 1]) for i in order])])
        abas = abas[order].copy()
        # combined tables: orbital env, then auxiliary env, then the unit shell
        off = len(vo.env)
        abas[:, PTR_EXP] += off
        abas[:, PTR_COEFF] += off
        env = np.concatenate([vo.env, aenv, [UNIT_EXP, UNIT_COEFF]])
        unit = np.array([[0, 0, 1, 1, 0, len(env) - 2, len(env) - 1, 0]], dtype=np.int32)
        nb = vo.nbas
        self.bas = np.ascontiguousarray(np.vstack([vo.bas, unit, abas]).astype(np.int32))
        self.env = np.ascontiguousarray(env)
        self.atm = vo.atm
        self.nao_c = vo.nao
        self.iunit = nb
        self.naux_c = int(nfa.sum())
        self.aux_off = self.nao_c + 1  # AO offset of the auxiliary block in the combined basis
        # pair classes: the orbital plan's, then one (P, unit) class per auxiliary group
        pi = [vo.bas_pair2shls[0]]
        pj = [vo.bas_pair2shls[1]]
        locs = list(vo.bas_pairs_locs)
        q = [vo.pair_q]
        keys = abas[:, [ANG_OF, NPRIM_OF]]
        starts = np.flatnonzero(np.r_[True, np.any(keys[1:] != keys[:-1], axis=1)])
        bounds = np.append(starts, len(abas))
        self.aux_classes = []
        for g0, g1 in zip(bounds[:-1], bounds[1:]):
            idx = np.arange(g0, g1, dtype=np.int32) + nb + 1
            pi.append(idx)
            pj.append(np.full(len(idx), nb, dtype=np.int32))
            q.append(np.ones(len(idx)))
            self.aux_classes.append(len(locs) - 1)
            locs.append(locs[-1] + len(idx))
        self.pair2shls = np.ascontiguousarray(np.vstack([np.concatenate(pi), np.concatenate(pj)]).astype(np.int32))
        self.locs = np.asarray(locs, dtype=np.int32)
        self.pair_q = np.ascontiguousarray(np.concatenate(q))
        self.orb_classes = list(range(vo.ncptype))
        self.bp = ctypes.c_void_p()
        check(load().mjk_init_basis_prod(ctypes.byref(self.bp), 1.0, ptr(self.pair2shls), ptr(self.locs),
                                         len(self.locs) - 1, ptr(self.pair_q), ptr(self.atm), len(self.atm),
                                         ptr(self.bas), len(self.bas), ptr(self.env)), 'mjk_init_basis_prod')
        self.nao_tot = int(load().mjk_nao(self.bp))
        assert self.nao_tot == self.nao_c + 1 + self.naux_c

    def close(self):
        if self.bp is not None and self.bp.value:
            load().mjk_del_basis_prod(ctypes.byref(self.bp))
        self.bp = None
        self.vo.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _fill(self, out, strides, offsets, ca, cb, stream):
        check(load().mjk_fill_int2e(self.bp, ptr(out), self.nao_tot, ptr(np.asarray(strides, dtype=np.int64)),
                                    ptr(np.asarray(offsets, dtype=np.int32)), ca, cb, ctypes.c_void_p(stream)),
              'mjk_fill_int2e')

    def int3c_cart(self):
        """(ij|P) over the plan's Cartesian orbital and auxiliary bases, (nao_c, nao_c, naux_c)."""
        torch = _torch()
        n, na = self.nao_c, self.naux_c
        out = torch.zeros((n, n, na), dtype=torch.float64, device='cuda')
        st = torch.cuda.current_stream().cuda_stream
        # element (P, unit, i, j) -> out[i, j, P]
        strides = [1, 0, n * na, na]
        offsets = [self.aux_off, self.nao_c, 0, 0]
        for ca in self.aux_classes:
            for cb in self.orb_classes:
                self._fill(out, strides, offsets, ca, cb, st)
        # orbital classes of two different groups hold (i, j) with i in the higher group only
        vo = self.vo
        l = vo.bas[:, ANG_OF]
        ao = np.append(0, np.cumsum((l + 1) * (l + 2) // 2))
        offs = vo.l_ctr_offsets
        g = len(offs) - 1
        for p in range(g):
            for qg in range(p):
                sp = slice(int(ao[offs[p]]), int(ao[offs[p + 1]]))
                sq = slice(int(ao[offs[qg]]), int(ao[offs[qg + 1]]))
                out[sq, sp] = out[sp, sq].transpose(0, 1)
        return out

    def int2c_cart(self):
        """(P|Q) over the plan's Cartesian auxiliary basis, (naux_c, naux_c)."""
        torch = _torch()
        na = self.naux_c
        out = torch.zeros((na, na), dtype=torch.float64, device='cuda')
        st = torch.cuda.current_stream().cuda_stream
        # element (P, unit, Q, unit) -> out[P, Q]
        for ca in self.aux_classes:
            for cb in self.aux_classes:
                self._fill(out, [na, 0, 1, 0], [self.aux_off, self.nao_c, self.aux_off, self.nao_c], ca, cb, st)
        return out

    def int3c(self):
        """(ij|P) in the caller's spherical orbital and auxiliary functions, (nao, nao, naux)."""
        torch = _torch()
        c, _ = self.vo.coeff_dev()
        ca = torch.as_tensor(self.aux_coeff, device='cuda')
        x = self.int3c_cart()
        x = torch.tensordot(x, ca, dims=([2], [0]))                     # (i, j, P)
        x = torch.tensordot(c, x, dims=([0], [0]))                      # (a, j, P)
        x = torch.tensordot(c, x, dims=([0], [1]))                      # (b, a, P)
        return x.transpose(0, 1).contiguous()

    def int2c(self):
        torch = _torch()
        ca = torch.as_tensor(self.aux_coeff, device='cuda')
        return (ca.T @ self.int2c_cart() @ ca).contiguous()

    def df_jk(self, dm, with_j=True, with_k=True, int3c=None, int2c=None):
        """Density-fitted J and K (Coulomb metric) of a symmetric spherical density."""
        torch = _torch()
        b3 = self.int3c() if int3c is None else int3c
        j2 = self.int2c() if int2c is None else int2c
        n = b3.shape[0]
        d = torch.as_tensor(np.asarray(dm) if not hasattr(dm, 'device') else dm, dtype=torch.float64,
                            device='cuda').reshape(n, n)
        low = torch.linalg.cholesky(j2)
        bmat = torch.linalg.solve_triangular(low, b3.reshape(n * n, -1).T, upper=False)  # (naux, n*n)
        vj = vk = None
        if with_j:
            vj = (bmat.T @ (bmat @ d.reshape(-1))).reshape(n, n)
        if with_k:
            bp = bmat.reshape(-1, n, n)
            vk = torch.matmul(torch.matmul(bp, d), bp).sum(0)
        return vj, vk
