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
           'mjk_last_stats', 'mjk_set_streams', 'mjk_set_ket_chunk', 'mjk_fail_alloc_after', 'mjk_int1e', 'mjk_gemm_bf16', 'mjk_bias_act', 'mjk_act_backward',
           'mjk_mse_grad', 'mjk_adam', 'mjk_to_bf16', 'mjk_batch_create', 'mjk_batch_destroy',
           'mjk_batch_build', 'mjk_batch_num_class_pairs', 'mjk_batch_tiles', 'mjk_batch_set_kc',
           'mjk_batch_time_class_pair', 'mjk_batch_stats', 'mjk_fill_int2e')

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
    try:
        # let torch create its HIP context before the library makes its first HIP call: the
        # Python host hands torch device memory and streams to the library
        import torch
        torch.cuda.is_available()
    except ImportError:
        pass
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
    lib.mjk_set_streams.argtypes = [p, ctypes.c_int]
    lib.mjk_set_ket_chunk.argtypes = [p, ctypes.c_int]
    lib.mjk_fail_alloc_after.argtypes = [ctypes.c_int]
    lib.mjk_int1e.argtypes = [p, i, p, i, p, p, p, p, i]
    f, sz = ctypes.c_float, ctypes.c_size_t
    lib.mjk_gemm_bf16.argtypes = [i, i, i, i, p, i, p, i, p, i, f, p]
    lib.mjk_bias_act.argtypes = [i, i, p, p, p, i, p]
    lib.mjk_act_backward.argtypes = [i, i, p, p, p, p, i, p]
    lib.mjk_mse_grad.argtypes = [sz, p, p, p, f, p, p, p]
    lib.mjk_adam.argtypes = [sz, p, p, p, p, p, f, f, f, f, i, p]
    lib.mjk_to_bf16.argtypes = [sz, p, p, p]
    pp = ctypes.POINTER(p)
    lib.mjk_batch_create.argtypes = [pp, pp, i, p, p, p, p]
    lib.mjk_batch_destroy.argtypes = [pp]
    lib.mjk_batch_build.argtypes = [p, p, p, p, ctypes.POINTER(Screen), i, p]
    lib.mjk_batch_num_class_pairs.argtypes = [p]
    lib.mjk_batch_tiles.argtypes = [p, p, i]
    lib.mjk_batch_set_kc.argtypes = [p, p, i]
    lib.mjk_batch_time_class_pair.argtypes = [p, i, p, p, p, ctypes.POINTER(Screen), i, p,
                                              ctypes.POINTER(ctypes.c_float)]
    lib.mjk_batch_stats.argtypes = [p, p]
    lib.mjk_fill_int2e.argtypes = [p, p, i, p, p, i, i, p]
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
