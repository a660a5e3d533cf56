"""Direct-SCF J/K on one MI355X: the drop-in for ByteQC's ``cuobc.scf.hf``.

Mirrors the reference's Python surface:
  get_jk(mol, dm, hermi=1, vhfopt=None, with_j=True, with_k=True, omega=None,
         verbose=None, group_size=None, gpus=None, prescreen=True, nbins=None)
                                            cuobc/scf/hf.py:53-227
  _get_jk(mf, mol, dm, hermi, with_j, with_k, omega)      hf.py:230-250
  _VHFOpt(mol, intor, ...).build(cutoff, group_size, diag_block_with_triu, gpus)
                                            hf.py:277-448
  basis_seg_contraction(mol)                hf.py:481-564

The integrals run in libmi355jk.so (hand-written gfx950 HIP); torch is used only to
own device memory, run the cart<->sph GEMMs and give the current stream.  The with_j /
with_k flags work independently (the reference fails unless both are set: hf.py:79-89
define vj_ptr / vk_ptr conditionally but hf.py:147-148 use both).
"""
from __future__ import annotations

import time

import numpy as np

from .. import lib as mjk
from ..gto.cart2sph import cart2sph_block
from ..gto.mole import ANG_OF, NPRIM_OF, NCTR_OF, PTR_EXP, PTR_COEFF, gto_norm

LMAX_ON_GPU = mjk.MJK_LMAX
PAIR_BUCKETS_PER_DECADE = 1


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('byteqc_amd.scf.hf needs a ROCm GPU (torch.cuda is unavailable)')
    return torch


def basis_seg_contraction(mol):
    """Decontract generally contracted shells (nctr > 1) into one-primitive shells.

    Returns (bas, env, coeff): the Cartesian segmented `_bas`/`_env` tables (atoms are
    mol._atm) and coeff (nao_cart, nao_sph) with AO_sph = sum_c coeff[c, s] AO_cart[c],
    so D_cart = coeff D_sph coeff^T and V_sph = coeff^T V_cart coeff.
    Same rule as cuobc/scf/hf.py:481-564: segmented shells are kept, a shell with nctr
    columns and nprim primitives becomes nprim radially-normalised primitive shells and
    the contraction moves into coeff.
    """
    env = list(mol._env)
    bas_out, blocks = [], []
    for sh in mol._bas:
        l, nprim, nctr = int(sh[ANG_OF]), int(sh[NPRIM_OF]), int(sh[NCTR_OF])
        nf = (l + 1) * (l + 2) // 2
        c2s = cart2sph_block(l) if not getattr(mol, 'cart', False) else np.eye(nf)
        if nctr == 1:
            bas_out.append(sh.copy())
            blocks.append(c2s)
            continue
        pexp, pcoef = int(sh[PTR_EXP]), int(sh[PTR_COEFF])
        exps = np.asarray(mol._env[pexp:pexp + nprim])
        norm = gto_norm(l, exps)
        cs = np.asarray(mol._env[pcoef:pcoef + nprim * nctr]).reshape(nctr, nprim)
        pnorm = len(env)
        env.extend(norm.tolist())
        for p in range(nprim):
            b = sh.copy()
            b[NPRIM_OF] = 1
            b[NCTR_OF] = 1
            b[PTR_EXP] = pexp + p
            b[PTR_COEFF] = pnorm + p
            bas_out.append(b)
        ns = c2s.shape[1]
        blk = np.zeros((nprim * nf, nctr * ns))
        for c in range(nctr):
            for p in range(nprim):
                blk[p * nf:(p + 1) * nf, c * ns:(c + 1) * ns] = cs[c, p] / norm[p] * c2s
        blocks.append(blk)
    nc = sum(b.shape[0] for b in blocks)
    ns = sum(b.shape[1] for b in blocks)
    coeff = np.zeros((nc, ns))
    i = j = 0
    for b in blocks:
        coeff[i:i + b.shape[0], j:j + b.shape[1]] = b
        i += b.shape[0]
        j += b.shape[1]
    return (np.asarray(bas_out, dtype=np.int32).reshape(-1, 8),
            np.asarray(env, dtype=np.float64), coeff)


class _VHFOpt:
    """Shell sorting, Schwarz pair lists and the device basis-product cache.

    build() follows cuobc/scf/hf.py:288-408: shells are grouped by (l, nprim); for every
    group pair (p >= q) the shell pairs with Schwarz factor above `cutoff` are kept,
    lower triangle only on diagonal groups; the pair data goes to the device with e12
    halved on same-shell pairs.  Where the reference sorts pairs by descending factor
    for its log-q bins, here pairs are ordered by decade of the factor and then by (i, j),
    with shells inside a group ordered along a space-filling curve: tiles of pairs from
    one decade are screened as a whole, and neighbouring pairs share shells, which the
    kernels use to cut exchange atomics.
    """

    def __init__(self, mol, intor='int2e', prescreen='CVHFnoscreen',
                 qcondname='CVHFsetnr_direct_scf', dmcondname=None):
        if intor not in ('int2e', 'int2e_cart', 'int2e_sph'):
            raise NotImplementedError(f'intor {intor!r}')
        self.mol = mol
        self._intor = intor
        self.bas, self.env, self.coeff = basis_seg_contraction(mol)
        self.atm = np.ascontiguousarray(mol._atm, dtype=np.int32)
        self.bp = None
        self.direct_scf_tol = 1e-13
        self._coeff_dev = None
        self.timings = {}

    # ---- setup ------------------------------------------------------------------
    def build(self, cutoff=1e-13, group_size=None, diag_block_with_triu=False, gpus=None):
        if gpus not in (None, 1, 0, [0], (0,)):
            raise ValueError('byteqc_amd runs on one GPU: gpus must be None or a single device')
        t0 = time.perf_counter()
        bas = self.bas
        l_ctrs = bas[:, [ANG_OF, NPRIM_OF]]
        uniq, inv, counts = np.unique(l_ctrs, axis=0, return_inverse=True, return_counts=True)
        inv = np.asarray(inv).ravel()
        if uniq[:, 0].max() > LMAX_ON_GPU:
            raise NotImplementedError(f'angular momentum above {"spdf"[LMAX_ON_GPU]} is not supported')
        if group_size is not None:
            uniq, counts = _split_l_ctr_groups(uniq, counts, group_size)
        # within each (l, nprim) group, shells follow a Morton curve through space so
        # that neighbouring pair indices are neighbouring in space
        order = np.lexsort((_morton_keys(self.atm, self.env, bas), inv))
        nf = (bas[:, ANG_OF] + 1) * (bas[:, ANG_OF] + 2) // 2
        ao_loc = np.append(0, np.cumsum(nf))
        ao_idx = np.concatenate([np.arange(ao_loc[i], ao_loc[i + 1]) for i in order])
        self.coeff = np.ascontiguousarray(self.coeff[ao_idx])
        self.bas = bas = np.ascontiguousarray(bas[order])
        self.nbas = len(bas)
        self.uniq_l_ctr = uniq
        self.l_ctr_offsets = np.append(0, np.cumsum(counts))
        self.direct_scf_tol = cutoff
        self.diag_block_with_triu = bool(diag_block_with_triu)

        lib = mjk.load()
        q = np.zeros((self.nbas, self.nbas))
        mjk.check(lib.mjk_q_cond(mjk.ptr(self.atm), len(self.atm), mjk.ptr(bas), self.nbas,
                                 mjk.ptr(self.env), mjk.ptr(q)), 'mjk_q_cond')
        self.q_cond = q
        t1 = time.perf_counter()

        pair_i, pair_j, pair_q, locs = [], [], [], [0]
        offs = self.l_ctr_offsets
        for p in range(len(uniq)):
            p0, p1 = offs[p], offs[p + 1]
            for qg in range(p + 1):
                q0, q1 = offs[qg], offs[qg + 1]
                sub = q[p0:p1, q0:q1]
                ii, jj = np.meshgrid(np.arange(p0, p1), np.arange(q0, q1), indexing='ij')
                ii, jj, vals = ii.ravel(), jj.ravel(), sub.ravel()
                mask = vals > cutoff
                if qg == p and not diag_block_with_triu:
                    mask &= ii >= jj
                ii, jj, vals = ii[mask], jj[mask], vals[mask]
                # decades of the Schwarz factor, strongest first, then (i, j): a tile of
                # pairs from one decade is screened as a whole, and inside a decade
                # consecutive pairs share shell i, which the kernels use to keep exchange
                # partial sums in registers and to reduce across lanes
                dec = np.floor(np.log10(np.maximum(vals, 1e-300)) * PAIR_BUCKETS_PER_DECADE)
                idx = np.lexsort((jj, ii, -dec))
                pair_i.append(ii[idx])
                pair_j.append(jj[idx])
                pair_q.append(vals[idx])
                locs.append(locs[-1] + len(idx))
        self.bas_pair2shls = np.ascontiguousarray(
            np.vstack([np.concatenate(pair_i), np.concatenate(pair_j)]).astype(np.int32))
        self.bas_pairs_locs = np.asarray(locs, dtype=np.int32)
        self.pair_q = np.ascontiguousarray(np.concatenate(pair_q))
        self.ncptype = len(locs) - 1
        diag_fac = 1.0 if diag_block_with_triu else 0.5
        self.clear()
        bp = __import__('ctypes').c_void_p()
        mjk.check(lib.mjk_init_basis_prod(
            __import__('ctypes').byref(bp), diag_fac, mjk.ptr(self.bas_pair2shls),
            mjk.ptr(self.bas_pairs_locs), self.ncptype, mjk.ptr(self.pair_q),
            mjk.ptr(self.atm), len(self.atm), mjk.ptr(bas), self.nbas, mjk.ptr(self.env)),
            'mjk_init_basis_prod')
        self.bp = bp
        self.nao = int(lib.mjk_nao(bp))
        self._coeff_dev = None
        self.timings = {'q_cond': t1 - t0, 'pairs+cache': time.perf_counter() - t1}
        return self

    def coeff_dev(self):
        if self._coeff_dev is None:
            torch = _torch()
            c = torch.as_tensor(self.coeff, dtype=torch.float64, device='cuda')
            self._coeff_dev = (c, c.T.contiguous())
        return self._coeff_dev

    def clear(self):
        if self.bp is not None and self.bp.value:
            import ctypes
            mjk.load().mjk_del_basis_prod(ctypes.byref(self.bp))
        self.bp = None
        return self

    def __del__(self):
        try:
            self.clear()
        except Exception:
            pass

    @property
    def npairs(self):
        return int(self.bas_pairs_locs[-1])

    def set_streams(self, lanes):
        """Streams a build spreads its class pairs over (1 = the caller's stream only)."""
        mjk.check(mjk.load().mjk_set_streams(self.bp, int(lanes)), 'mjk_set_streams')
        return self

    def set_ket_chunk(self, cap):
        """Longest ket chunk per workgroup of the register and group kernels (0: default)."""
        mjk.check(mjk.load().mjk_set_ket_chunk(self.bp, int(cap)), 'mjk_set_ket_chunk')
        return self

    def stats(self):
        out = np.zeros(4)
        mjk.check(mjk.load().mjk_last_stats(self.bp, mjk.ptr(out)), 'mjk_last_stats')
        return {'quartets_fp64': int(out[0]), 'quartets_fp32': int(out[1]), 'launches': int(out[2]),
                'slots': int(out[3])}


def _morton_keys(atm, env, bas, cell=2.0):
    """Z-order key of each shell's centre on a grid of `cell` bohr."""
    ptr = atm[bas[:, 0], 1]
    xyz = np.stack([env[ptr], env[ptr + 1], env[ptr + 2]], axis=1)
    g = np.floor((xyz - xyz.min(axis=0)) / cell).astype(np.int64)
    key = np.zeros(len(g), dtype=np.int64)
    for bit in range(16):
        for d in range(3):
            key |= ((g[:, d] >> bit) & 1) << (3 * bit + d)
    return key


def _split_l_ctr_groups(uniq_l_ctr, l_ctr_counts, group_size):
    """Split (l, nprim) groups so none holds more than `group_size` AOs (hf.py:581-606)."""
    ls, cs = [], []
    for l_ctr, n in zip(uniq_l_ctr, l_ctr_counts):
        nf = (l_ctr[0] + 1) * (l_ctr[0] + 2) // 2
        max_shells = max(group_size // nf, 2)
        while n > max_shells:
            ls.append(l_ctr)
            cs.append(max_shells)
            n -= max_shells
        if n > 0:
            ls.append(l_ctr)
            cs.append(n)
    return np.vstack(ls), np.asarray(cs)


def jk_device(vhfopt, dm_cart, with_j=True, with_k=True, precision_cutoff=0.0,
              density_screen=True, stream=None, direct_scf_tol=None):
    """J/K in the Cartesian basis from a device density (n_dm, nao, nao) torch tensor.

    Returns device tensors (vj, vk) already symmetrised: J = 2 (V + V^T), K = V + V^T.
    """
    torch = _torch()
    lib = mjk.load()
    nao = vhfopt.nao
    dm_cart = dm_cart.reshape(-1, nao, nao).contiguous()
    n_dm = dm_cart.shape[0]
    vj = torch.empty_like(dm_cart) if with_j else None
    vk = torch.empty_like(dm_cart) if with_k else None
    tol = vhfopt.direct_scf_tol if direct_scf_tol is None else max(float(direct_scf_tol), vhfopt.direct_scf_tol)
    sc = mjk.Screen(tol, float(precision_cutoff), int(bool(density_screen)))
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    mjk.check(lib.mjk_build_jk_all(vhfopt.bp, mjk.ptr(vj), mjk.ptr(vk), mjk.ptr(dm_cart), nao,
                                   n_dm, __import__('ctypes').byref(sc),
                                   __import__('ctypes').c_void_p(stream)), 'mjk_build_jk_all')
    return vj, vk


def get_jk(mol, dm, hermi=1, vhfopt=None, with_j=True, with_k=True, omega=None, verbose=None,
           group_size=None, gpus=None, prescreen=True, nbins=None, *, precision_cutoff=0.0,
           density_screen=None, out_device=False, direct_scf_tol=None):
    """Coulomb and exchange matrices for symmetric density matrices (hermi=1).

    dm: (nao, nao) or (n_dm, nao, nao) in the spherical (or mol.cart) AO basis, numpy or
    a torch tensor.  Returns (vj, vk) with dm's shape, numpy (or device tensors when
    out_device=True); an entry is None when its flag is off.
    precision_cutoff > 0 runs quartets whose bound q_ij q_kl max|D| is below it in FP32.
    density_screen (default: prescreen) weights the Schwarz bound by max |D| over the six
    shell blocks a quartet touches; nbins is accepted for compatibility and unused.
    direct_scf_tol (default: the plan's cutoff) screens quartets at a looser tolerance for
    this call only; it cannot go below the cutoff the pair lists were built with.
    """
    if hermi != 1:
        raise NotImplementedError('JK-builder only supports hermitian density matrix')
    if omega is not None:
        raise NotImplementedError('range-separated J/K (omega) is not supported')
    is_tensor = hasattr(dm, 'data_ptr') and not isinstance(dm, np.ndarray)
    # a NaN or Inf in D would pass the screening (max |D| ignores NaN) and spread through J/K
    if not is_tensor and not np.isfinite(np.asarray(dm)).all():
        raise ValueError('get_jk: dm contains NaN or Inf')
    torch = _torch()
    if vhfopt is None:
        vhfopt = _VHFOpt(mol, 'int2e').build(group_size=group_size, gpus=gpus)
    if density_screen is None:
        density_screen = prescreen
    if not (with_j or with_k):
        return None, None
    c, ct = vhfopt.coeff_dev()
    nao0 = c.shape[1]
    dms = (dm if is_tensor else torch.as_tensor(np.asarray(dm), dtype=torch.float64))
    dms = dms.to(device='cuda', dtype=torch.float64)
    if is_tensor and not bool(torch.isfinite(dms).all()):
        raise ValueError('get_jk: dm contains NaN or Inf')
    shape = dms.shape
    dms = dms.reshape(-1, nao0, nao0)
    dmc = torch.matmul(torch.matmul(c, dms), ct)
    vj, vk = jk_device(vhfopt, dmc, with_j, with_k, precision_cutoff, density_screen,
                       direct_scf_tol=direct_scf_tol)
    out = []
    for v in (vj, vk):
        if v is None:
            out.append(None)
            continue
        v = torch.matmul(torch.matmul(ct, v), c).reshape(shape)
        out.append(v if out_device else v.cpu().numpy())
    return tuple(out)


def _get_jk(mf, mol=None, dm=None, hermi=1, with_j=True, with_k=True, omega=None):
    """RHF hook: the basis-product cache lives in mf._opt_gpu (hf.py:230-250)."""
    if omega is not None:
        raise NotImplementedError('omega')
    if mol is None:
        mol = mf.mol
    if dm is None:
        dm = mf.make_rdm1()
    if getattr(mf, '_opt_gpu', None) is None:
        mf._opt_gpu = _VHFOpt(mol, 'int2e').build(mf.direct_scf_tol)
    return get_jk(mol, dm, hermi, mf._opt_gpu, with_j, with_k, omega,
                  precision_cutoff=getattr(mf, 'precision_cutoff', 0.0))


from .rhf import RHF, eig  # noqa: E402,F401  (the reference exposes hf.RHF, hf.eig)
