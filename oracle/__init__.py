"""CPU oracle for the direct-SCF J/K path -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check in __graft_entry__.py and the CPU-baseline leg of bench.py
use this package, and only as the checker.  The product (byteqc_amd/) never imports it.

Parts and how each is pinned (tests/test_oracle_*.py):
  boys / rys_roots      C code in boys.c / rys_roots.c.  Pinned by mpmath quadrature of
                        the Boys integrals, the moment identities sum_i w_i x_i^k = F_k(T)
                        (k < 2N), the T = 0 Gauss-Legendre limit and the large-T
                        Gauss-Hermite limit.
  md_eri                McMurchie-Davidson ERIs (md_eri.c).  Pinned by the (ss|ss) closed
                        form, the 8-fold permutation symmetry, the Schwarz inequality,
                        translation invariance and agreement with the Rys quartets.
  rys_jk                Rys-quadrature restatement of the reference J/K (rys_jk.c,
                        cuobc/lib/gvhf/*).  Pinned against exact_jk (einsum over the full
                        McMurchie-Davidson tensor) on small molecules.
  seg_basis             Restatement of cuobc/scf/hf.py:481-564 (basis_seg_contraction):
                        Cartesian segmented basis and its (nao_cart, nao_sph) transform.
                        Pinned by C^T S_cart C = S_sph having a unit diagonal.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ATM_SLOTS, BAS_SLOTS = 6, 8
ANG_OF, NPRIM_OF, NCTR_OF, PTR_EXP, PTR_COEFF = 1, 2, 3, 5, 6


def build():
    subprocess.run(['make', '-s', '-C', _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'liboracle.so')
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
        d, i = ctypes.c_double, ctypes.c_int
        p = ctypes.c_void_p
        _LIB.oracle_boys.argtypes = [i, d, p]
        _LIB.oracle_rys_roots.argtypes = [i, d, p, p]
        _LIB.oracle_rys_roots_direct.argtypes = [i, d, p, p]
        _LIB.oracle_md_eri.argtypes = [p, i, p, i, p, p]
        _LIB.oracle_md_shell_quartet.argtypes = [p, p, p, i, i, i, i, p]
        _LIB.oracle_rys_shell_quartet.argtypes = [p, p, p, i, i, i, i, p]
        _LIB.oracle_q_cond.argtypes = [p, i, p, i, p, p, i]
        _LIB.oracle_md_int1e.argtypes = [p, i, p, i, p, p, p, p]
        _LIB.oracle_rys_jk.argtypes = [p, i, p, i, p, p, d, p, i, i, p, p, i]
        _LIB.oracle_rys_jk_rows.argtypes = [p, i, p, i, p, p, d, p, i, i, p, p, i, i, i]
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def boys(nmax, T):
    f = np.zeros(nmax + 1)
    lib().oracle_boys(nmax, float(T), _ptr(f))
    return f


def rys_roots(n, T, direct=False):
    x = np.zeros(n)
    w = np.zeros(n)
    fn = lib().oracle_rys_roots_direct if direct else lib().oracle_rys_roots
    fn(n, float(T), _ptr(x), _ptr(w))
    return x, w


class SegBasis:
    """Cartesian, segmented (nctr == 1) basis tables plus the (nao_cart, nao_sph) map."""

    def __init__(self, atm, bas, env, coeff):
        self.atm = np.ascontiguousarray(atm, dtype=np.int32)
        self.bas = np.ascontiguousarray(bas, dtype=np.int32)
        self.env = np.ascontiguousarray(env, dtype=np.float64)
        self.coeff = coeff

    @property
    def nbas(self):
        return len(self.bas)

    @property
    def nao(self):
        l = self.bas[:, ANG_OF]
        return int(((l + 1) * (l + 2) // 2).sum())


def _gto_norm(l, e):
    from scipy.special import gamma
    return 1. / np.sqrt(gamma(l + 1.5) / (2. * (2 * e) ** (l + 1.5)))


def seg_basis(mol, replica=True):
    """Segmented Cartesian basis for `mol` (restates cuobc/scf/hf.py:481-564).

    replica=True:  a generally contracted shell becomes nctr shells, each with all the
                   primitives and one contraction column (hf.py:514-522);
    replica=False: it becomes nprim one-primitive shells and the contraction moves into
                   the transform (hf.py:523-540), the layout the reference kernels use.
    """
    from byteqc_amd.gto.cart2sph import cart2sph_block  # plain data conventions

    env = list(mol._env)
    bas_out, blocks = [], []
    for sh in mol._bas:
        l, nprim, nctr = int(sh[ANG_OF]), int(sh[NPRIM_OF]), int(sh[NCTR_OF])
        nf = (l + 1) * (l + 2) // 2
        c2s = cart2sph_block(l)
        pexp, pcoef = int(sh[PTR_EXP]), int(sh[PTR_COEFF])
        cs = np.asarray(mol._env[pcoef:pcoef + nprim * nctr]).reshape(nctr, nprim)
        if nctr == 1:
            bas_out.append(sh.copy())
            blocks.append(c2s)
        elif replica:
            for c in range(nctr):
                b = sh.copy()
                b[NCTR_OF] = 1
                b[PTR_COEFF] = pcoef + c * nprim
                bas_out.append(b)
                blocks.append(c2s)
        else:
            exps = np.asarray(mol._env[pexp:pexp + nprim])
            norm = _gto_norm(l, exps)
            pc = len(env)
            env.extend(norm.tolist())
            for p in range(nprim):
                b = sh.copy()
                b[NPRIM_OF] = 1
                b[NCTR_OF] = 1
                b[PTR_EXP] = pexp + p
                b[PTR_COEFF] = pc + p
                bas_out.append(b)
            # contracted AO (c, m) = sum_p cs[c, p] / norm[p] * primitive AO (p, m)
            blk = np.zeros((nprim * nf, nctr * c2s.shape[1]))
            for c in range(nctr):
                for p in range(nprim):
                    blk[p * nf:(p + 1) * nf, c * c2s.shape[1]:(c + 1) * c2s.shape[1]] = \
                        cs[c, p] / norm[p] * c2s
            blocks.append(blk)
    nc = sum(b.shape[0] for b in blocks)
    ns = sum(b.shape[1] for b in blocks)
    coeff = np.zeros((nc, ns))
    i = j = 0
    for b in blocks:
        coeff[i:i + b.shape[0], j:j + b.shape[1]] = b
        i += b.shape[0]
        j += b.shape[1]
    return SegBasis(mol._atm, np.asarray(bas_out), np.asarray(env), coeff)


def md_eri_cart(sb: SegBasis):
    n = sb.nao
    eri = np.zeros((n, n, n, n))
    err = lib().oracle_md_eri(_ptr(sb.atm), len(sb.atm), _ptr(sb.bas), sb.nbas, _ptr(sb.env), _ptr(eri))
    if err:
        raise RuntimeError('oracle_md_eri failed')
    return eri


def md_int1e_cart(sb: SegBasis):
    n = sb.nao
    s, t, v = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    err = lib().oracle_md_int1e(_ptr(sb.atm), len(sb.atm), _ptr(sb.bas), sb.nbas, _ptr(sb.env),
                                _ptr(s), _ptr(t), _ptr(v))
    if err:
        raise RuntimeError('oracle_md_int1e failed')
    return s, t, v


def md_int1e_sph(mol):
    sb = seg_basis(mol, replica=True)
    c = sb.coeff
    return tuple(c.T @ m @ c for m in md_int1e_cart(sb))


def md_eri_sph(mol):
    sb = seg_basis(mol, replica=True)
    eri = md_eri_cart(sb)
    c = sb.coeff
    return np.einsum('pi,qj,rk,sl,pqrs->ijkl', c, c, c, c, eri, optimize=True)


def shell_quartet(sb: SegBasis, i, j, k, l, method='md'):
    ls = sb.bas[[i, j, k, l], ANG_OF]
    nf = [(x + 1) * (x + 2) // 2 for x in ls]
    out = np.zeros(nf)
    fn = lib().oracle_md_shell_quartet if method == 'md' else lib().oracle_rys_shell_quartet
    err = fn(_ptr(sb.atm), _ptr(sb.bas), _ptr(sb.env), i, j, k, l, _ptr(out))
    if err:
        raise RuntimeError('shell quartet failed')
    return out


def q_cond(sb: SegBasis, nthreads=0):
    q = np.zeros((sb.nbas, sb.nbas))
    lib().oracle_q_cond(_ptr(sb.atm), len(sb.atm), _ptr(sb.bas), sb.nbas, _ptr(sb.env), _ptr(q), nthreads)
    return q


def rys_jk_cart(sb: SegBasis, dm_cart, with_j=True, with_k=True, cutoff=1e-13, nthreads=0,
                q=None, stride=1, offset=0):
    """J/K in the segmented Cartesian basis.  cutoff=None: no screening.
    stride > 1 computes only bra pair rows pij % stride == offset (timing samples)."""
    dms = np.ascontiguousarray(dm_cart, dtype=np.float64).reshape(-1, sb.nao, sb.nao)
    n_dm = dms.shape[0]
    vj = np.zeros_like(dms) if with_j else None
    vk = np.zeros_like(dms) if with_k else None
    if cutoff is not None and q is None:
        q = q_cond(sb, nthreads)
    err = lib().oracle_rys_jk_rows(_ptr(sb.atm), len(sb.atm), _ptr(sb.bas), sb.nbas, _ptr(sb.env),
                                   _ptr(q) if cutoff is not None else None,
                                   float(cutoff or 0.), _ptr(dms), n_dm, sb.nao, _ptr(vj), _ptr(vk),
                                   nthreads, int(stride), int(offset))
    if err:
        raise RuntimeError(f'oracle_rys_jk failed ({err})')
    return vj, vk


def get_jk(mol, dm, with_j=True, with_k=True, cutoff=1e-13, nthreads=0, replica=False):
    """Oracle J/K in the spherical AO basis for a symmetric density (hermi=1)."""
    sb = seg_basis(mol, replica=replica)
    c = sb.coeff
    dm = np.asarray(dm, dtype=np.float64)
    dms = dm.reshape(-1, dm.shape[-1], dm.shape[-1])
    dmc = np.einsum('pi,nij,qj->npq', c, dms, c)
    vj, vk = rys_jk_cart(sb, dmc, with_j, with_k, cutoff, nthreads)
    out = []
    for v in (vj, vk):
        if v is None:
            out.append(None)
        else:
            v = np.einsum('pi,npq,qj->nij', c, v, c)
            out.append(v.reshape(dm.shape))
    return tuple(out)


def exact_jk(mol, dm):
    """J and K by einsum over the full McMurchie-Davidson ERI tensor (small mol only)."""
    eri = md_eri_sph(mol)
    dm = np.asarray(dm)
    dms = dm.reshape(-1, dm.shape[-1], dm.shape[-1])
    vj = np.einsum('ijkl,nkl->nij', eri, dms)
    vk = np.einsum('ikjl,nkl->nij', eri, dms)
    return vj.reshape(dm.shape), vk.reshape(dm.shape)
