"""The C ABI: every declared symbol is exported and argument errors come back as codes.
No GPU compute is launched here."""
import ctypes
import os
import re

import numpy as np

from byteqc_amd import lib as mjk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for fn in os.listdir(os.path.join(ROOT, 'include')):
        if fn.endswith('.h'):
            text = open(os.path.join(ROOT, 'include', fn)).read()
            names += re.findall(r'^(?:int|const char \*|void)\s*\*?\s*(mjk_\w+)\s*\(', text, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = mjk.load()
    names = _declared()
    assert len(names) >= 9
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) <= set(mjk.EXPORTS)


def test_version_and_error_paths():
    lib = mjk.load()
    assert lib.mjk_version() >= 1
    bp = ctypes.c_void_p()
    rc = lib.mjk_init_basis_prod(ctypes.byref(bp), 0.5, None, None, 0, None, None, 0, None, 0, None)
    assert rc == mjk.MJK_ERR_ARG
    assert b'bad argument' in lib.mjk_last_error()
    assert lib.mjk_del_basis_prod(ctypes.byref(ctypes.c_void_p())) == 0
    assert lib.mjk_del_basis_prod(None) == 0
    # generally contracted shells are rejected before any device work
    from byteqc_amd.gto import M
    mol = M('O 0 0 0', basis='cc-pvdz')
    q = np.zeros((mol.nbas, mol.nbas))
    rc = lib.mjk_q_cond(mjk.ptr(mol._atm), mol.natm, mjk.ptr(mol._bas), mol.nbas,
                        mjk.ptr(mol._env), mjk.ptr(q))
    assert rc == mjk.MJK_ERR_ARG and b'segmented' in lib.mjk_last_error()
    sc = mjk.Screen(1e-13, 0.0, 1)
    rc = lib.mjk_build_jk_all(None, None, None, None, 10, 1, ctypes.byref(sc), None)
    assert rc == mjk.MJK_ERR_ARG
    assert lib.mjk_set_streams(None, 4) == mjk.MJK_ERR_ARG
    assert lib.mjk_set_ket_chunk(None, 32) == mjk.MJK_ERR_ARG


def test_check_raises():
    import pytest
    with pytest.raises(mjk.MJKError):
        mjk.check(mjk.MJK_ERR_ARG, 'x')
    with pytest.raises(NotImplementedError):
        mjk.check(mjk.MJK_ERR_UNSUPPORTED, 'x')


def test_perceptron_entry_points_reject_bad_arguments():
    """Argument checks of the learned-guess kernels return codes before any launch."""
    lib = mjk.load()
    assert lib.mjk_gemm_bf16(3, 4, 4, 4, 8, 4, 8, 4, 8, 4, 0.0, None) == mjk.MJK_ERR_ARG
    assert b'layout' in lib.mjk_last_error()
    assert lib.mjk_gemm_bf16(0, 4, 4, 4, 8, 2, 8, 4, 8, 4, 0.0, None) == mjk.MJK_ERR_ARG
    assert b'leading dimension' in lib.mjk_last_error()
    assert lib.mjk_gemm_bf16(0, 4, 4, 4, None, 4, None, 4, 8, 4, 0.0, None) == mjk.MJK_ERR_ARG
    assert lib.mjk_bias_act(4, 4, None, None, None, 1, None) == mjk.MJK_ERR_ARG
    assert lib.mjk_act_backward(4, 4, 8, None, 8, None, 1, None) == mjk.MJK_ERR_ARG
    assert lib.mjk_adam(4, 8, 8, 8, 8, None, 1e-3, 0.9, 0.999, 1e-8, 0, None) == mjk.MJK_ERR_ARG
    assert lib.mjk_mse_grad(4, None, 8, None, 1.0, 8, None, None) == mjk.MJK_ERR_ARG
    assert lib.mjk_to_bf16(4, None, 8, None) == mjk.MJK_ERR_ARG
