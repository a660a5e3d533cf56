"""GPU J/K parity against the CPU oracle (runs through libmi355jk.so)."""
import numpy as np
import pytest

import oracle
from byteqc_amd.gto import M, water_cluster
from byteqc_amd.scf import hf

pytestmark = pytest.mark.gpu

WATER = 'O 0 0 0; H 0.757 0 0.587; H -0.757 0 0.587'


def _rand_dm(n, seed=2341):
    rng = np.random.default_rng(seed)
    d = rng.random((n, n))
    return d + d.T


@pytest.mark.parametrize('basis', ['cc-pvdz', 'def2-tzvp', '6-31gs'])
def test_water_exact(basis):
    """(H2O)1 J/K vs einsum over the full McMurchie-Davidson ERI tensor."""
    mol = M(WATER, basis=basis)
    dm = _rand_dm(mol.nao)
    print(f'case: water {basis} nao={mol.nao}', flush=True)
    vj, vk = hf.get_jk(mol, dm, prescreen=False)
    rj, rk = oracle.exact_jk(mol, dm)
    assert np.abs(vj - rj).max() < 1e-11
    assert np.abs(vk - rk).max() < 1e-11


def test_water_dimer_exact():
    mol = M(WATER + '; O 2.9 0.1 0.2; H 3.6 0.2 0.8; H 2.7 -0.8 0.5', basis='cc-pvdz')
    dm = np.stack([_rand_dm(mol.nao, 1), _rand_dm(mol.nao, 2)])
    print(f'case: water dimer cc-pvdz nao={mol.nao}, n_dm=2', flush=True)
    vj, vk = hf.get_jk(mol, dm, prescreen=False)
    rj, rk = oracle.exact_jk(mol, dm)
    assert vj.shape == dm.shape
    assert np.abs(vj - rj).max() < 1e-11
    assert np.abs(vk - rk).max() < 1e-11


@pytest.mark.parametrize('basis,n', [('cc-pvdz', 8), ('def2-tzvp', 2), ('6-31gs', 4)])
def test_cluster_vs_rys_oracle(basis, n):
    """Water clusters vs the oracle's Rys J/K without screening (1e-10)."""
    mol = water_cluster(n, 9.0, seed=7, basis=basis)
    dm = _rand_dm(mol.nao)
    print(f'case: (H2O){n} {basis} nao={mol.nao}', flush=True)
    vj, vk = hf.get_jk(mol, dm, prescreen=False)
    rj, rk = oracle.get_jk(mol, dm, cutoff=None)
    assert np.abs(vj - rj).max() < 1e-10
    assert np.abs(vk - rk).max() < 1e-10


def test_j_only_k_only():
    mol = M(WATER, basis='cc-pvdz')
    dm = _rand_dm(mol.nao)
    vj, vk = hf.get_jk(mol, dm)
    j_only, none_k = hf.get_jk(mol, dm, with_k=False)
    none_j, k_only = hf.get_jk(mol, dm, with_j=False)
    assert none_k is None and none_j is None
    assert np.abs(j_only - vj).max() < 1e-13
    assert np.abs(k_only - vk).max() < 1e-13


@pytest.fixture(scope='module')
def tzvp4():
    mol = water_cluster(4, 9.0, seed=11, basis='def2-tzvp')
    dm = _rand_dm(mol.nao, 5)
    return mol, dm, oracle.get_jk(mol, dm, cutoff=None)


@pytest.mark.parametrize('lanes', [1, 2, 4])
def test_streams_and_ket_chunks_agree_with_oracle(tzvp4, lanes):
    """Class pairs spread over 1, 2 or 4 streams, with default and short ket chunks, give the
    oracle's J/K (1e-10), also when the builder is reused on the same cache."""
    mol, dm, (rj, rk) = tzvp4
    print(f'case: (H2O)4 def2-tzvp nao={mol.nao}, streams {lanes}', flush=True)
    opt = hf._VHFOpt(mol).build()
    try:
        opt.set_streams(lanes)
        for cap in (0, 4):
            opt.set_ket_chunk(cap)
            for _ in range(2):
                vj, vk = hf.get_jk(mol, dm, vhfopt=opt, prescreen=False)
                assert np.abs(vj - rj).max() < 1e-10
                assert np.abs(vk - rk).max() < 1e-10
    finally:
        opt.clear()
