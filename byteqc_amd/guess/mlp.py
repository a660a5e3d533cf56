"""Three-layer perceptron on the library's HIP kernels (bf16 MFMA GEMMs, fp32 accumulate).

y = W3^T silu(W2^T silu(W1^T x + b1) + b2) + b3 with weights stored [in][out].  fp32
master weights live on the device next to their bf16 copies (the GEMM operands); Adam
updates both in one kernel.  Torch only allocates device memory and supplies the current
stream; every arithmetic step is a call into libmi355jk.so (include/mi355jk.h), so the
module fails loudly without the built library.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import lib as mjk

SILU, IDENTITY = 1, 0


def _torch():
    import torch
    return torch


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _lib():
    return mjk.load()


def to_bf16(x):
    """fp32 device tensor -> bf16 bit pattern (int16 tensor of the same shape)."""
    torch = _torch()
    if x.dtype != torch.float32:
        raise TypeError(f'to_bf16 takes float32, got {x.dtype}')
    x = x.contiguous()
    out = torch.empty(x.shape, dtype=torch.int16, device=x.device)
    mjk.check(_lib().mjk_to_bf16(x.numel(), _p(x), _p(out), _stream()), 'mjk_to_bf16')
    return out


def gemm(layout, a, b, m, n, k, lda, ldb, out=None, beta=0.0):
    """c[m][n] = op(a) op(b) (+ beta c); a, b bf16 (int16) device tensors."""
    torch = _torch()
    if out is None:
        out = torch.empty((m, n), dtype=torch.float32, device=a.device)
    mjk.check(_lib().mjk_gemm_bf16(layout, m, n, k, _p(a), lda, _p(b), ldb, _p(out), n, beta, _stream()),
              'mjk_gemm_bf16')
    return out


class MLP:
    """dims = (n_in, n_hidden, n_hidden, n_out)."""

    def __init__(self, dims, seed=0, params=None):
        torch = _torch()
        self.dims = tuple(int(d) for d in dims)
        if params is None:
            rng = np.random.default_rng(seed)
            params = {}
            for i in range(3):
                fan_in = self.dims[i]
                params[f'W{i + 1}'] = (rng.normal(size=(self.dims[i], self.dims[i + 1])) / np.sqrt(fan_in)).astype(np.float32)
                params[f'b{i + 1}'] = np.zeros(self.dims[i + 1], dtype=np.float32)
        self.w = [torch.as_tensor(np.asarray(params[f'W{i + 1}'], dtype=np.float32), device='cuda').contiguous() for i in range(3)]
        self.b = [torch.as_tensor(np.asarray(params[f'b{i + 1}'], dtype=np.float32), device='cuda').contiguous() for i in range(3)]
        self.wb = [to_bf16(w) for w in self.w]
        self.m1 = [torch.zeros_like(t) for t in self.w + self.b]
        self.m2 = [torch.zeros_like(t) for t in self.w + self.b]
        self.step = 0

    def params(self):
        out = {}
        for i in range(3):
            out[f'W{i + 1}'] = self.w[i].cpu().numpy()
            out[f'b{i + 1}'] = self.b[i].cpu().numpy()
        return out

    # ---- passes ------------------------------------------------------------------------
    def forward(self, xb, m):
        """xb: bf16 [m][n_in] -> (y fp32 [m][n_out], cache)."""
        torch = _torch()
        lib = _lib()
        d = self.dims
        acts = [xb]
        zs = []
        for i in range(3):
            z = gemm(0, acts[-1], self.wb[i], m, d[i + 1], d[i], d[i], d[i + 1])
            if i < 2:
                a = torch.empty((m, d[i + 1]), dtype=torch.int16, device='cuda')
                mjk.check(lib.mjk_bias_act(m, d[i + 1], _p(z), _p(self.b[i]), _p(a), SILU, _stream()), 'mjk_bias_act')
                acts.append(a)
            else:
                mjk.check(lib.mjk_bias_act(m, d[i + 1], _p(z), _p(self.b[i]), None, IDENTITY, _stream()), 'mjk_bias_act')
            zs.append(z)
        return zs[2], (acts, zs)

    def backward(self, cache, dy, m):
        """Gradients [dW1, dW2, dW3, db1, db2, db3] of sum(dy * y)."""
        torch = _torch()
        lib = _lib()
        acts, zs = cache
        d = self.dims
        dw = [None] * 3
        db = [None] * 3
        da = dy
        for i in (2, 1, 0):
            dz = torch.empty((m, d[i + 1]), dtype=torch.int16, device='cuda')
            db[i] = torch.empty(d[i + 1], dtype=torch.float32, device='cuda')
            act = IDENTITY if i == 2 else SILU
            mjk.check(lib.mjk_act_backward(m, d[i + 1], _p(da), _p(zs[i]), _p(dz), _p(db[i]), act, _stream()),
                      'mjk_act_backward')
            # dW_i[in][out] = a_i^T dz  (TN: a_i is [m][in], dz is [m][out])
            dw[i] = gemm(2, acts[i], dz, d[i], d[i + 1], m, d[i], d[i + 1])
            if i > 0:
                # da_{i}[m][in] = dz W_i^T  (NT: W_i is [in][out] = [n][k])
                da = gemm(1, dz, self.wb[i], m, d[i], d[i + 1], d[i + 1], d[i + 1])
        return dw + db

    def adam(self, grads, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        lib = _lib()
        self.step += 1
        tensors = self.w + self.b
        for j, (t, g) in enumerate(zip(tensors, grads)):
            wb = self.wb[j] if j < 3 else None
            mjk.check(lib.mjk_adam(t.numel(), _p(t), _p(g), _p(self.m1[j]), _p(self.m2[j]), _p(wb), lr, beta1, beta2,
                                   eps, self.step, _stream()), 'mjk_adam')

    def loss_grad(self, y, target, mask, scale, loss):
        """dy = scale * mask * (y - target); loss (device double) += sum mask (y - target)^2."""
        torch = _torch()
        dy = torch.empty_like(y)
        mjk.check(_lib().mjk_mse_grad(y.numel(), _p(y), _p(target), _p(mask), scale, _p(dy), _p(loss), _stream()),
                  'mjk_mse_grad')
        return dy

    def predict(self, x, batch=65536):
        """fp32 host or device features [m][n_in] -> fp32 device outputs [m][n_out]."""
        torch = _torch()
        x = torch.as_tensor(np.asarray(x, dtype=np.float32) if isinstance(x, np.ndarray) else x,
                            device='cuda').contiguous()
        out = []
        for s in range(0, x.shape[0], batch):
            xs = x[s:s + batch].contiguous()
            y, _ = self.forward(to_bf16(xs), xs.shape[0])
            out.append(y)
        return torch.cat(out) if out else torch.empty((0, self.dims[-1]), device='cuda')
