"""ctypes binding of libmi355jk.so (the C ABI in include/mi355jk.h).

The reference loads its libraries with ``numpy.ctypeslib.load_library``
(cuobc/lib/__init__.py:41-42); this module does the same for the single MI355X library
and fails loudly when it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libmi355jk.so')

MJK_OK, MJK_ERR_ARG, MJK_ERR_UNSUPPORTED, MJK_ERR_HIP, MJK_ERR_OOM = range(5)
MJK_LMAX = 3
_ERRNAMES = {MJK_ERR_ARG: 'bad argument', MJK_ERR_UNSUPPORTED: 'unsupported',
             MJK_ERR_HIP: 'HIP error', MJK_ERR_OOM: 'out of memory'}


class Screen(ctypes.Structure):
    """mjk_screen."""
    _fields_ = [('direct_scf_tol', ctypes.c_double),
                ('precision_cutoff', ctypes.c_double),
                ('density_screen', ctypes.c_int),
                ('reserved', ctypes.c_int * 3)]


EXPORTS = ('mjk_version', 'mjk_last_error', 'mjk_q_cond', 'mjk_init_basis_prod',
           'mjk_del_basis_prod', 'mjk_nao', 'mjk_build_jk', 'mjk_build_jk_all',
           'mjk_last_stats', 'mjk_int1e')

_lib = None


class MJKError(RuntimeError):
    pass


def load():
    """Load libmi355jk.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f'{LIB_PATH} is missing: build it with `python -c "import '
                          f'__graft_entry__ as g; g.build()"` (hipcc, gfx950)')
    lib = ctypes.CDLL(LIB_PATH)
    p, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.mjk_version.restype = i
    lib.mjk_last_error.restype = ctypes.c_char_p
    lib.mjk_q_cond.argtypes = [p, i, p, i, p, p]
    lib.mjk_init_basis_prod.argtypes = [ctypes.POINTER(p), d, p, p, i, p, p, i, p, i, p]
    lib.mjk_del_basis_prod.argtypes = [ctypes.POINTER(p)]
    lib.mjk_nao.argtypes = [p]
    lib.mjk_build_jk.argtypes = [p, p, p, p, i, i, i, i, ctypes.POINTER(Screen), p]
    lib.mjk_build_jk_all.argtypes = [p, p, p, p, i, i, ctypes.POINTER(Screen), p]
    lib.mjk_last_stats.argtypes = [p, p]
    lib.mjk_int1e.argtypes = [p, i, p, i, p, p, p, p, i]
    _lib = lib
    return lib


def check(rc, what=''):
    if rc != MJK_OK:
        msg = load().mjk_last_error().decode(errors='replace')
        exc = NotImplementedError if rc == MJK_ERR_UNSUPPORTED else MJKError
        raise exc(f'{what}: {_ERRNAMES.get(rc, rc)}: {msg}')


def ptr(a):
    """Address of a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(ctypes.c_void_p)
    return ctypes.c_void_p(a.data_ptr())
