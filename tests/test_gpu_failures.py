"""Failure handling on the device: injected allocation failures, fully screened and empty
class pairs, non-finite densities."""
import numpy as np
import pytest

from byteqc_amd import lib as mjk
from byteqc_amd.gto import M
from byteqc_amd.scf import hf

pytestmark = pytest.mark.gpu

WATER = 'O 0 0 0; H 0.757 0 0.587; H -0.757 0 0.587'


def _dm(n, seed=2):
    rng = np.random.default_rng(seed)
    d = rng.random((n, n))
    return d + d.T


def test_injected_allocation_failure_is_clean():
    """Every device allocation of the plan failing in turn gives MJK_ERR_OOM (MJKError), frees
    what was already allocated and leaves the library usable."""
    import torch
    lib = mjk.load()
    mol = M(WATER, basis='cc-pvdz')
    dm = _dm(mol.nao)
    rj, rk = hf.get_jk(mol, dm)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    failures = 0
    try:
        for n in range(64):
            lib.mjk_fail_alloc_after(n)
            try:
                opt = hf._VHFOpt(mol).build()
            except mjk.MJKError as e:
                assert 'out of memory' in str(e)
                failures += 1
                continue
            lib.mjk_fail_alloc_after(-1)
            opt.clear()
            break
    finally:
        lib.mjk_fail_alloc_after(-1)
    print(f'{failures} injected failures before the plan could be built', flush=True)
    assert failures >= 10
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] >= free0 - (4 << 20)
    vj, vk = hf.get_jk(mol, dm)
    # the same plan again: equal up to the order of the atomic additions
    assert np.abs(vj - rj).max() < 1e-12 and np.abs(vk - rk).max() < 1e-12


def test_zero_density_is_fully_screened():
    mol = M(WATER, basis='cc-pvdz')
    opt = hf._VHFOpt(mol).build()
    vj, vk = hf.get_jk(mol, np.zeros((mol.nao, mol.nao)), vhfopt=opt)
    assert not vj.any() and not vk.any()
    st = opt.stats()
    assert st['quartets_fp64'] == 0 and st['quartets_fp32'] == 0
    opt.clear()


def test_empty_classes_are_skipped():
    """A Schwarz cutoff of 1 empties the (p|s), (d|p) and (d|s) pair classes of water/cc-pVDZ
    (their largest factors are 0.91-0.97; the oracle's q_cond gives the same); the build skips
    them and J, K stay symmetric and reproducible."""
    mol = M(WATER, basis='cc-pvdz')
    opt = hf._VHFOpt(mol).build(cutoff=1.0)
    locs = opt.bas_pairs_locs
    assert (np.diff(locs) == 0).sum() == 3
    dm = _dm(mol.nao, 5)
    vj, vk = hf.get_jk(mol, dm, vhfopt=opt)
    vj2, vk2 = hf.get_jk(mol, dm, vhfopt=opt)
    assert np.abs(vj - vj.T).max() < 1e-12 and np.abs(vk - vk.T).max() < 1e-12
    assert np.abs(vj - vj2).max() < 1e-12 and np.abs(vk - vk2).max() < 1e-12
    opt.clear()


def test_non_finite_device_density_is_rejected():
    import torch
    mol = M(WATER, basis='cc-pvdz')
    d = torch.zeros((mol.nao, mol.nao), dtype=torch.float64, device='cuda')
    d[1, 2] = float('nan')
    with pytest.raises(ValueError, match='NaN'):
        hf.get_jk(mol, d, out_device=True)
