"""Generate piecewise-Chebyshev fits of Rys roots and weights (FP64 and FP32 tables).

Rys quadrature replaces the Boys-function kernel exp(-T t^2) on t in [0, 1] by an
N-point Gauss rule in x = t^2:

    F_k(T) = int_0^1 t^(2k) exp(-T t^2) dt = sum_i w_i x_i^k,   k = 0 .. 2N-1

(the u = t^2/(1-t^2) variable of the reference, cuobc/lib/gint/rys_roots.cu:1318-1331,
is a reparametrisation of the same nodes; this project works in x = t^2 directly).

Method, at `mp.dps` decimal digits:
  1. moments F_0..F_{2N-1}(T) from 1F1(n+1/2; n+3/2; -T)/(2n+1) at the top order and
     the stable downward recursion F_n = (2T F_{n+1} + e^-T)/(2n+1);
  2. recurrence coefficients of the monic orthogonal polynomials by Chebyshev's
     algorithm (W. Gautschi, "Orthogonal Polynomials: Computation and Approximation",
     Oxford 2004, Alg. 2.1);
  3. nodes and weights from the Jacobi matrix (Golub & Welsch, Math. Comp. 23 (1969)).

For T >= TMAX(N) the nodes are those of the half-range Hermite rule:
x_i = s_i^2 / T and w_i = W_i / sqrt(T), with s_i > 0 the roots of H_2N and W_i the
Gauss-Hermite weights; the neglected tail is int_1^inf, below 1e-17 relative there.
Below TMAX the interval [0, TMAX) is split into pieces of width H and each root and
weight gets its own Chebyshev series in s = 2 (T - T0)/H - 1.

Run:  python -m byteqc_amd.rys.gen_tables   (writes byteqc_amd/csrc/rys_tables.inc)
"""
from __future__ import annotations

import math
import os
import sys
from multiprocessing import Pool

import numpy as np

NMAX = 7            # up to (ff|ff): floor(12/2) + 1
H64, DEG64 = 1.0, 11  # interval width and Chebyshev degree of the FP64 table
H32, DEG32 = 2.0, 8   # FP32 table
DPS = 50


def tmax(n: int) -> float:
    return 40.0 + 5.0 * n


def _mp():
    import mpmath
    mpmath.mp.dps = DPS
    return mpmath


def boys_mp(T, kmax):
    """[F_0 .. F_kmax](T) as mpmath numbers."""
    mp = _mp()
    T = mp.mpf(T)
    top = mp.hyp1f1(kmax + mp.mpf(1) / 2, kmax + mp.mpf(3) / 2, -T) / (2 * kmax + 1)
    out = [mp.mpf(0)] * (kmax + 1)
    out[kmax] = top
    e = mp.exp(-T)
    for n in range(kmax - 1, -1, -1):
        out[n] = (2 * T * out[n + 1] + e) / (2 * n + 1)
    return out


def rys_mp(n, T):
    """(x[n], w[n]) mpmath nodes (x = t^2, ascending) and weights for exp(-T t^2)."""
    mp = _mp()
    mu = boys_mp(T, 2 * n - 1)
    # Chebyshev algorithm on ordinary moments
    alpha = [mp.mpf(0)] * n
    beta = [mp.mpf(0)] * n
    sig_prev = [mp.mpf(0)] * (2 * n)
    sig = list(mu)
    alpha[0] = mu[1] / mu[0]
    beta[0] = mu[0]
    for k in range(1, n):
        sig_new = [mp.mpf(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            sig_new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l]
        alpha[k] = sig_new[k + 1] / sig_new[k] - sig[k] / sig[k - 1]
        beta[k] = sig_new[k] / sig[k - 1]
        sig_prev, sig = sig, sig_new
    J = mp.matrix(n, n)
    for i in range(n):
        J[i, i] = alpha[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(beta[i + 1])
    E, Q = mp.eigsy(J)
    pairs = sorted((E[i], beta[0] * Q[0, i] ** 2) for i in range(n))
    return [p[0] for p in pairs], [p[1] for p in pairs]


def hermite_half(n):
    """Positive roots s_i (ascending) and weights of the 2n-point Gauss-Hermite rule."""
    s, w = np.polynomial.hermite.hermgauss(2 * n)
    keep = s > 0
    order = np.argsort(s[keep])
    return s[keep][order], w[keep][order]


def _fit_interval(args):
    n, t0, h, deg = args
    mp = _mp()
    m = deg + 1
    nodes = [math.cos(math.pi * (k + 0.5) / m) for k in range(m)]
    vals = np.zeros((2 * n, m))
    for k, s in enumerate(nodes):
        T = t0 + (s + 1) * h / 2
        x, w = rys_mp(n, T)
        vals[:n, k] = [float(v) for v in x]
        vals[n:, k] = [float(v) for v in w]
    # Chebyshev coefficients by the discrete cosine sums at Chebyshev nodes
    coef = np.zeros((2 * n, m))
    for j in range(m):
        cj = np.array([math.cos(math.pi * j * (k + 0.5) / m) for k in range(m)])
        coef[:, j] = (2.0 / m) * vals @ cj
    coef[:, 0] *= 0.5
    return coef


def clenshaw(c, s):
    b1 = b2 = 0.0
    for a in c[:0:-1]:
        b1, b2 = 2 * s * b1 - b2 + a, b1
    return s * b1 - b2 + c[0]


def build_table(n, h, deg, pool=None):
    nint = int(math.ceil(tmax(n) / h))
    jobs = [(n, i * h, h, deg) for i in range(nint)]
    res = pool.map(_fit_interval, jobs) if pool else list(map(_fit_interval, jobs))
    return np.stack(res)  # (nint, 2n, deg+1)


def check_table(n, h, tab, npts=24, seed=0):
    """Max relative errors of nodes and weights against mpmath at random T."""
    rng = np.random.default_rng(seed + n)
    Ts = np.concatenate([rng.uniform(0, tmax(n), npts), [0.0, 1e-12, tmax(n) - 1e-9]])
    ex = ew = 0.0
    for T in Ts:
        x, w = rys_mp(n, T)
        i = min(int(T / h), tab.shape[0] - 1)
        s = 2 * (T - i * h) / h - 1
        for r in range(n):
            xr = clenshaw(tab[i, r], s)
            wr = clenshaw(tab[i, n + r], s)
            ex = max(ex, abs(xr - float(x[r])) / float(x[r]))
            ew = max(ew, abs(wr - float(w[r])) / float(w[r]))
    return ex, ew


def check_asymptote(n):
    s, W = hermite_half(n)
    T = tmax(n)
    x, w = rys_mp(n, T)
    ex = max(abs(s[r] ** 2 / T - float(x[r])) / float(x[r]) for r in range(n))
    ew = max(abs(W[r] / math.sqrt(T) - float(w[r])) / float(w[r]) for r in range(n))
    return ex, ew


def _fmt(v):
    return repr(float(v))


BOYS_H, BOYS_TMAX, BOYS_K = 0.1, 40.0, 10  # grid step, range, highest order stored


def boys_grid():
    """F_0..F_BOYS_K at T = 0, h, 2h, ... < BOYS_TMAX (mpmath), shape (npts, BOYS_K+1)."""
    n = int(round(BOYS_TMAX / BOYS_H))
    return np.array([[float(v) for v in boys_mp(i * BOYS_H, BOYS_K)] for i in range(n + 1)])


def emit(path, t64, t32, grid):
    lines = [
        '// Generated by byteqc_amd/rys/gen_tables.py -- do not edit by hand.',
        '// Piecewise Chebyshev fits of Rys nodes x = t^2 and weights w for exp(-T t^2)',
        '// on t in [0,1], N = 1..%d roots.  Layout per N: [interval][2N][DEG+1], the first' % NMAX,
        '// N rows are nodes (ascending), the next N rows the matching weights.',
        '#pragma once',
        '// RYS_STORAGE lets a HIP translation unit place the tables in device memory.',
        '#ifndef RYS_STORAGE',
        '#define RYS_STORAGE static const',
        '#endif',
        '#define RYS_NMAX %d' % NMAX,
        '#define RYS_H64 %r' % H64,
        '#define RYS_DEG64 %d' % DEG64,
        '#define RYS_H32 %rf' % H32,
        '#define RYS_DEG32 %d' % DEG32,
    ]
    off64 = [0]
    off32 = [0]
    for n in range(1, NMAX + 1):
        off64.append(off64[-1] + t64[n].size)
        off32.append(off32[-1] + t32[n].size)
    lines.append('RYS_STORAGE int rys_nint64[RYS_NMAX + 1] = {0, %s};' %
                 ', '.join(str(t64[n].shape[0]) for n in range(1, NMAX + 1)))
    lines.append('RYS_STORAGE int rys_nint32[RYS_NMAX + 1] = {0, %s};' %
                 ', '.join(str(t32[n].shape[0]) for n in range(1, NMAX + 1)))
    lines.append('RYS_STORAGE int rys_off64[RYS_NMAX + 1] = {0, %s};' % ', '.join(map(str, off64[:-1])))
    lines.append('RYS_STORAGE int rys_off32[RYS_NMAX + 1] = {0, %s};' % ', '.join(map(str, off32[:-1])))
    # offset of the N-point Hermite data inside rys_herm_*: N (N - 1) / 2
    lines.append('RYS_STORAGE double rys_tmax[RYS_NMAX + 1] = {0.0, %s};' %
                 ', '.join(_fmt(tmax(n)) for n in range(1, NMAX + 1)))
    hs, hw = [], []
    for n in range(1, NMAX + 1):
        s, W = hermite_half(n)
        hs.extend(s ** 2)
        hw.extend(W)
    lines.append('// half-range Gauss-Hermite: s_i^2 and W_i, N = 1..NMAX concatenated')
    lines.append('RYS_STORAGE double rys_herm_s2[%d] = {%s};' % (len(hs), ', '.join(map(_fmt, hs))))
    lines.append('RYS_STORAGE double rys_herm_w[%d] = {%s};' % (len(hw), ', '.join(map(_fmt, hw))))
    lines.append('// Boys F_k(T_i), T_i = i * RYS_BOYS_H, k = 0..RYS_BOYS_K: the one-root rule uses')
    lines.append('// F_n(T_i - d) = sum_k F_(n+k)(T_i) d^k / k! with |d| <= h/2')
    lines.append('#define RYS_BOYS_H %r' % BOYS_H)
    lines.append('#define RYS_BOYS_TMAX %r' % BOYS_TMAX)
    lines.append('#define RYS_BOYS_K %d' % BOYS_K)
    lines.append('RYS_STORAGE double rys_boys_grid[%d] = {' % grid.size)
    for row in grid:
        lines.append('  ' + ', '.join(map(_fmt, row)) + ',')
    lines.append('};')
    lines.append('RYS_STORAGE double rys_tab64[%d] = {' % off64[-1])
    for n in range(1, NMAX + 1):
        lines.append('  // N = %d' % n)
        for row in t64[n].reshape(-1, DEG64 + 1):
            lines.append('  ' + ', '.join(map(_fmt, row)) + ',')
    lines.append('};')
    lines.append('RYS_STORAGE float rys_tab32[%d] = {' % off32[-1])
    for n in range(1, NMAX + 1):
        lines.append('  // N = %d' % n)
        for row in t32[n].reshape(-1, DEG32 + 1):
            lines.append('  ' + ', '.join('%.9ef' % v for v in row) + ',')
    lines.append('};')
    with open(path, 'w') as f:
        f.write('\n'.join(lines) + '\n')


def main(argv=None):
    out = os.path.join(os.path.dirname(__file__), '..', 'csrc', 'rys_tables.inc')
    out = os.path.abspath(out)
    t64, t32 = {}, {}
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        for n in range(1, NMAX + 1):
            t64[n] = build_table(n, H64, DEG64, pool)
            t32[n] = build_table(n, H32, DEG32, pool)
            e64 = check_table(n, H64, t64[n])
            e32 = check_table(n, H32, t32[n])
            ea = check_asymptote(n)
            print(f'N={n}: fp64 fit rel err x={e64[0]:.2e} w={e64[1]:.2e}; '
                  f'fp32 fit x={e32[0]:.2e} w={e32[1]:.2e}; asymptote at T={tmax(n)}: '
                  f'x={ea[0]:.2e} w={ea[1]:.2e}', flush=True)
    grid = boys_grid()
    emit(out, t64, t32, grid)
    np.savez_compressed(os.path.join(os.path.dirname(out), 'rys_tables.npz'),
                        boys_grid=grid,
                        **{f'tab64_{n}': t64[n] for n in t64}, **{f'tab32_{n}': t32[n] for n in t32})
    print('wrote', out)


if __name__ == '__main__':
    main(sys.argv[1:])
