"""Dense ERI tensor on the device: the restatement of cuobc/lib/int4c.py:46-127.

`get_int4c(mol)` returns (ij|kl) over the plan's Cartesian, segmented, group-sorted basis
(the basis the reference's `vhfopt.mol` holds), as a (nao, nao, nao, nao) device tensor:
every class-pair block is filled by `mjk_fill_int2e` and copied to its eight permutations,
as the reference's task loop does.  `to_spherical(eri, vhfopt)` applies the plan's `coeff`
to reach the caller's AO basis.  For small systems and tests: the tensor is nao^4 doubles.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import check, load, ptr


def _torch():
    import torch
    return torch


def _ao_loc(bas):
    l = bas[:, 1]
    return np.append(0, np.cumsum((l + 1) * (l + 2) // 2)).astype(np.int64)


def ind2pair(c):
    """Class index -> (p, q) with p >= q, c = p (p + 1) / 2 + q (cuobc/lib/int3c.py ind2pair)."""
    p = int((np.sqrt(8 * c + 1) - 1) // 2)
    while p * (p + 1) // 2 > c:
        p -= 1
    while (p + 1) * (p + 2) // 2 <= c:
        p += 1
    return p, c - p * (p + 1) // 2


def get_int4c_slice(cp_ij, cp_kl, vhfopt, stream=None):
    """(si, sj, sk, sl, block) of one class pair, as the reference's get_int4c_slice."""
    torch = _torch()
    ao_loc = _ao_loc(vhfopt.bas)
    offs = vhfopt.l_ctr_offsets
    ranges = []
    for g in (*ind2pair(cp_ij), *ind2pair(cp_kl)):
        ranges.append((int(ao_loc[offs[g]]), int(ao_loc[offs[g + 1]])))
    (i0, i1), (j0, j1), (k0, k1), (l0, l1) = ranges
    ni, nj, nk, nl = i1 - i0, j1 - j0, k1 - k0, l1 - l0
    blk = torch.zeros((ni, nj, nk, nl), dtype=torch.float64, device='cuda')
    strides = np.array([nj * nk * nl, nk * nl, nl, 1], dtype=np.int64)
    ao_offsets = np.array([i0, j0, k0, l0], dtype=np.int32)
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    check(load().mjk_fill_int2e(vhfopt.bp, ptr(blk), vhfopt.nao, ptr(strides), ptr(ao_offsets),
                                cp_ij, cp_kl, ctypes.c_void_p(stream)), 'mjk_fill_int2e')
    return slice(i0, i1), slice(j0, j1), slice(k0, k1), slice(l0, l1), blk


def get_int4c(mol, blksize=None, vhfopt=None, batch=1, buf_size=0, nbins=1):
    """Cartesian ERI tensor over the plan's basis (reference signature; batch, buf_size and
    nbins are accepted and unused).  The plan must hold full pair lists on diagonal groups
    (`build(diag_block_with_triu=True)`), as the reference's test builds it."""
    from ..scf.hf import _VHFOpt
    torch = _torch()
    if vhfopt is None:
        vhfopt = _VHFOpt(mol, 'int2e').build(group_size=blksize, diag_block_with_triu=True)
    if not getattr(vhfopt, 'diag_block_with_triu', False):
        raise ValueError('get_int4c needs a plan built with diag_block_with_triu=True')
    nao = vhfopt.nao
    eri = torch.zeros((nao, nao, nao, nao), dtype=torch.float64, device='cuda')
    for cp_ij, cp_kl in zip(*np.tril_indices(vhfopt.ncptype)):
        si, sj, sk, sl, b = get_int4c_slice(int(cp_ij), int(cp_kl), vhfopt)
        eri[si, sj, sk, sl] = b
        eri[sj, si, sk, sl] = b.permute(1, 0, 2, 3)
        eri[si, sj, sl, sk] = b.permute(0, 1, 3, 2)
        eri[sj, si, sl, sk] = b.permute(1, 0, 3, 2)
        eri[sk, sl, si, sj] = b.permute(2, 3, 0, 1)
        eri[sk, sl, sj, si] = b.permute(2, 3, 1, 0)
        eri[sl, sk, si, sj] = b.permute(3, 2, 0, 1)
        eri[sl, sk, sj, si] = b.permute(3, 2, 1, 0)
    return eri


def to_spherical(eri, vhfopt):
    """(ij|kl) in the caller's AO basis: coeff^T on every index (each contraction over the
    leading index appends the new one last, so four of them restore the order)."""
    torch = _torch()
    c, _ = vhfopt.coeff_dev()
    x = eri
    for _ in range(4):
        x = torch.tensordot(x, c, dims=([0], [0]))
    return x.contiguous()
