"""The product's generated Rys tables (byteqc_amd/csrc/rys_tables.*) against mathematics:
moment identities with mpmath quadrature Boys values, the T = 0 Gauss-Legendre limit and
the large-T Gauss-Hermite limit (CPU evaluation of the same Chebyshev data the kernels use)."""
import os

import numpy as np
import pytest

from byteqc_amd.rys import gen_tables as g

mpmath = pytest.importorskip('mpmath')
HERE = os.path.dirname(os.path.abspath(__file__))
TAB = np.load(os.path.join(HERE, '..', 'byteqc_amd', 'csrc', 'rys_tables.npz'))


def clenshaw(c, s):
    b1 = b2 = 0.0
    for a in c[:0:-1]:
        b1, b2 = 2 * s * b1 - b2 + a, b1
    return s * b1 - b2 + c[0]


def roots(n, T, kind=64):
    h = g.H64 if kind == 64 else g.H32
    if T >= g.tmax(n):
        s, W = g.hermite_half(n)
        return s ** 2 / T, W / np.sqrt(T)
    tab = TAB[f'tab{kind}_{n}']
    i = int(T / h)
    s = 2 * (T - i * h) / h - 1
    x = np.array([clenshaw(tab[i, r], s) for r in range(n)])
    w = np.array([clenshaw(tab[i, n + r], s) for r in range(n)])
    return x, w


def boys_quad(k, T):
    mpmath.mp.dps = 40
    return float(mpmath.quad(lambda t: t ** (2 * k) * mpmath.exp(-T * t * t), mpmath.linspace(0, 1, 21)))


@pytest.mark.parametrize('n', range(1, g.NMAX + 1))
def test_moments_fp64(n):
    rng = np.random.default_rng(100 + n)
    for T in np.concatenate([[0.0, 0.5, g.tmax(n) - 1e-7, g.tmax(n) + 1e-7, 300.0],
                             rng.uniform(0, g.tmax(n), 6)]):
        x, w = roots(n, T)
        for k in range(2 * n):
            ref = boys_quad(k, T)
            assert abs(np.sum(w * x ** k) - ref) <= 6e-14 * ref, (n, T, k)


@pytest.mark.parametrize('n', [1, 3, 5, 7])
def test_moments_fp32_table(n):
    rng = np.random.default_rng(200 + n)
    for T in rng.uniform(0, g.tmax(n), 8):
        x, w = roots(n, T, kind=32)
        for k in range(2 * n):
            ref = boys_quad(k, T)
            assert abs(np.sum(w * x ** k) - ref) <= 2e-7 * ref, (n, T, k)


@pytest.mark.parametrize('n', [1, 4, 7])
def test_limits(n):
    x, w = roots(n, 0.0)
    t, W = np.polynomial.legendre.leggauss(2 * n)
    order = np.argsort(t[t > 0])
    assert np.allclose(np.sqrt(x), t[t > 0][order], rtol=1e-13)
    assert np.allclose(w, W[t > 0][order], rtol=1e-13)
    T = g.tmax(n) - 1e-6  # last fitted interval meets the Hermite branch continuously
    x1, w1 = roots(n, T)
    x2, w2 = roots(n, T + 2e-6)
    assert np.allclose(x1, x2, rtol=1e-12) and np.allclose(w1, w2, rtol=1e-12)


def test_header_matches_npz():
    text = open(os.path.join(HERE, '..', 'byteqc_amd', 'csrc', 'rys_tables.inc')).read()
    assert f'#define RYS_DEG64 {g.DEG64}' in text and f'#define RYS_NMAX {g.NMAX}' in text
    n64 = sum(TAB[f'tab64_{n}'].size for n in range(1, g.NMAX + 1))
    assert f'rys_tab64[{n64}]' in text


def root1_grid(T):
    """CPU evaluation of the kernels' one-root rule: Taylor series on the Boys grid below
    RYS_BOYS_TMAX, the asymptote above it."""
    if T >= g.BOYS_TMAX:
        return 0.5 / T, 0.5 * np.sqrt(np.pi / T)
    grid = TAB['boys_grid'].reshape(-1, g.BOYS_K + 1)
    i = int(T / g.BOYS_H + 0.5)
    d = i * g.BOYS_H - T
    fact = np.cumprod([1.0] + list(range(1, g.BOYS_K)))
    f0 = sum(grid[i, k] * d ** k / fact[k] for k in range(g.BOYS_K))
    f1 = sum(grid[i, k + 1] * d ** k / fact[k] for k in range(g.BOYS_K))
    return f1 / f0, f0


def test_boys_grid_root1():
    rng = np.random.default_rng(7)
    for T in np.concatenate([[0.0, 1e-9, 0.05, 0.0500001, g.BOYS_TMAX - 1e-9, g.BOYS_TMAX, 77.0],
                             rng.uniform(0, g.BOYS_TMAX + 5, 12)]):
        x, w = root1_grid(T)
        f0, f1 = boys_quad(0, T), boys_quad(1, T)
        assert abs(w - f0) <= 4e-16 * f0 + 1e-300, T
        assert abs(x - f1 / f0) <= 1e-15 * (f1 / f0), T
