"""Failure handling that the host forces before any device work (no GPU needed)."""
import numpy as np
import pytest

from byteqc_amd import lib as mjk
from byteqc_amd.gto import M
from byteqc_amd.scf import hf

WATER = 'O 0 0 0; H 0.757 0 0.587; H -0.757 0 0.587'


def test_get_jk_rejects_unsupported_requests_before_device_work():
    mol = M(WATER, basis='cc-pvdz')
    dm = np.eye(mol.nao)
    with pytest.raises(NotImplementedError):
        hf.get_jk(mol, dm, hermi=0)
    with pytest.raises(NotImplementedError):
        hf.get_jk(mol, dm, omega=0.3)
    bad = dm.copy()
    bad[3, 4] = np.nan
    with pytest.raises(ValueError, match='NaN'):
        hf.get_jk(mol, bad)
    bad[3, 4] = np.inf
    with pytest.raises(ValueError, match='NaN'):
        hf.get_jk(mol, bad)


def test_angular_momentum_above_f_is_refused():
    """A g shell: the plan raises NotImplementedError and the C ABI returns
    MJK_ERR_UNSUPPORTED while it checks the shells, before touching the device."""
    mol = M('H 0 0 0; H 0 0 1.4', basis={'H': [[0, [1.0, 1.0]], [4, [0.9, 1.0]]]}, unit='Bohr')
    with pytest.raises(NotImplementedError):
        hf._VHFOpt(mol).build()
    lib = mjk.load()
    q = np.zeros((mol.nbas, mol.nbas))
    rc = lib.mjk_q_cond(mjk.ptr(mol._atm), mol.natm, mjk.ptr(mol._bas), mol.nbas, mjk.ptr(mol._env), mjk.ptr(q))
    assert rc == mjk.MJK_ERR_UNSUPPORTED
    assert b'MJK_LMAX' in lib.mjk_last_error()


def test_allocation_failure_hook_switches_off():
    lib = mjk.load()
    assert lib.mjk_fail_alloc_after(3) == mjk.MJK_OK
    assert lib.mjk_fail_alloc_after(-1) == mjk.MJK_OK
