"""Explicit ERI tensors (mjk_fill_int2e, byteqc_amd.lib.int4c) against the oracle's
McMurchie-Davidson tensor on the same sorted, segmented Cartesian basis, the way the
reference's test compares get_int4c with libcint (cuobc/test/int4c.py)."""
import numpy as np
import pytest

import oracle
from oracle import SegBasis

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('basis,n', [('6-31gs', 2), ('cc-pvdz', 1), ('def2-tzvp', 1)])
def test_int4c_matches_oracle(basis, n):
    from byteqc_amd.gto.clusters import water_cluster
    from byteqc_amd.lib import int4c
    from byteqc_amd.scf.hf import _VHFOpt
    print(f'[int4c] {basis} {n} water(s)', flush=True)
    mol = water_cluster(n, 6.0, 40 + n, basis=basis)
    opt = _VHFOpt(mol).build(diag_block_with_triu=True)
    eri = int4c.get_int4c(mol, vhfopt=opt).cpu().numpy()
    ref = oracle.md_eri_cart(SegBasis(opt.atm, opt.bas, opt.env, opt.coeff))
    assert eri.shape == ref.shape
    assert np.abs(eri - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())
    # (ij|kl) = (kl|ij): exact across class pairs (copies), to rounding inside one class pair,
    # whose quartets (p, q) and (q, p) are computed separately
    assert np.abs(eri - eri.transpose(2, 3, 0, 1)).max() < 1e-14 * np.abs(ref).max()


def test_int4c_spherical_and_jk_agree():
    """The spherical tensor contracted with a density gives the J/K builder's J and K."""
    import torch

    from byteqc_amd.gto.clusters import water_cluster
    from byteqc_amd.lib import int4c
    from byteqc_amd.scf import hf
    from byteqc_amd.scf.hf import _VHFOpt
    mol = water_cluster(2, 6.0, 7, basis='cc-pvdz')
    opt = _VHFOpt(mol).build(diag_block_with_triu=True)
    eri = int4c.to_spherical(int4c.get_int4c(mol, vhfopt=opt), opt)
    rng = np.random.default_rng(3)
    d = rng.random((mol.nao, mol.nao))
    d = d + d.T
    dt = torch.as_tensor(d, device='cuda')
    j = torch.einsum('ijkl,lk->ij', eri, dt).cpu().numpy()
    k = torch.einsum('ijkl,jk->il', eri, dt).cpu().numpy()
    vj, vk = hf.get_jk(mol, d)
    assert np.abs(j - vj).max() < 1e-10 * np.abs(vj).max()
    assert np.abs(k - vk).max() < 1e-10 * np.abs(vk).max()


def test_fill_rejects_bad_arguments():
    import ctypes

    from byteqc_amd import lib as mjk
    lib = mjk.load()
    assert lib.mjk_fill_int2e(None, None, 1, None, None, 0, 0, None) == mjk.MJK_ERR_ARG
