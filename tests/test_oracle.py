"""Pins the CPU oracle against mathematics it cannot fake (no GPU needed)."""
import numpy as np
import pytest

import oracle
from byteqc_amd.gto import M

mpmath = pytest.importorskip('mpmath')

WATER = 'O 0 0 0; H 0.757 0 0.587; H -0.757 0 0.587'


def boys_quad(n, T):
    mpmath.mp.dps = 40
    f = lambda t: t ** (2 * n) * mpmath.exp(-T * t * t)
    return float(mpmath.quad(f, mpmath.linspace(0, 1, 21)))


@pytest.mark.parametrize('T', [0.0, 1e-9, 0.37, 4.2, 17.0, 29.99, 30.01, 55.0, 120.0])
def test_boys_vs_quadrature(T):
    f = oracle.boys(10, T)
    for n in range(11):
        ref = boys_quad(n, T)
        assert abs(f[n] - ref) <= 2e-15 * ref + 1e-300, (n, T, f[n], ref)


def test_boys_closed_forms():
    assert np.allclose(oracle.boys(6, 0.0), [1 / (2 * n + 1) for n in range(7)], rtol=1e-15)
    from math import erf, pi, sqrt
    for T in (0.5, 3.0, 40.0):
        assert abs(oracle.boys(0, T)[0] - 0.5 * sqrt(pi / T) * erf(sqrt(T))) < 1e-15


@pytest.mark.parametrize('n', [1, 2, 3, 5, 7])
@pytest.mark.parametrize('T', [0.0, 0.2, 3.3, 11.7, 39.5, 62.0, 90.0, 400.0])
def test_rys_moments(n, T):
    """sum_i w_i x_i^k = F_k(T), k < 2N; F_k from mpmath quadrature."""
    x, w = oracle.rys_roots(n, T)
    for k in range(2 * n):
        ref = boys_quad(k, T)
        assert abs(np.sum(w * x ** k) - ref) <= 5e-14 * ref, (n, T, k)


def test_rys_limits():
    for n in (2, 4, 7):
        x, w = oracle.rys_roots(n, 0.0)
        t, W = np.polynomial.legendre.leggauss(2 * n)
        assert np.allclose(np.sqrt(x), np.sort(t[t > 0]), rtol=1e-13)
        assert np.allclose(w, W[t > 0][np.argsort(t[t > 0])], rtol=1e-13)
        T = 5000.0
        x, w = oracle.rys_roots(n, T)
        s, W = np.polynomial.hermite.hermgauss(2 * n)
        pos = np.sort(s[s > 0])
        assert np.allclose(x * T, pos ** 2, rtol=1e-12)
        assert np.allclose(w * np.sqrt(T), W[s > 0][np.argsort(s[s > 0])], rtol=1e-12)


def _sb(basis='cc-pvdz', atom=WATER):
    return oracle.seg_basis(M(atom, basis=basis), replica=False)


def test_ssss_closed_form():
    """(ss|ss) of unnormalised-by-hand primitives against 2 pi^(5/2)/(pq sqrt(p+q)) K F0."""
    from math import pi, sqrt, exp, erf
    mol = M('H 0 0 0; H 0.3 0.4 -1.1', basis='cc-pvdz')
    sb = _sb(atom='H 0 0 0; H 0.3 0.4 -1.1')
    s_shells = [i for i in range(sb.nbas) if sb.bas[i, 1] == 0]
    i, j, k, l = s_shells[0], s_shells[4], s_shells[2], s_shells[5]
    got = oracle.shell_quartet(sb, i, j, k, l)[0, 0, 0, 0]
    a = [sb.env[sb.bas[x, 5]] for x in (i, j, k, l)]
    c = [sb.env[sb.bas[x, 6]] * 0.282094791773878143 for x in (i, j, k, l)]
    r = [sb.env[sb.atm[sb.bas[x, 0], 1]:sb.atm[sb.bas[x, 0], 1] + 3] for x in (i, j, k, l)]
    p, q = a[0] + a[1], a[2] + a[3]
    P = (a[0] * r[0] + a[1] * r[1]) / p
    Q = (a[2] * r[2] + a[3] * r[3]) / q
    kab = exp(-a[0] * a[1] / p * np.sum((r[0] - r[1]) ** 2))
    kcd = exp(-a[2] * a[3] / q * np.sum((r[2] - r[3]) ** 2))
    T = p * q / (p + q) * np.sum((P - Q) ** 2)
    f0 = 0.5 * sqrt(pi / T) * erf(sqrt(T)) if T > 0 else 1.0
    ref = 2 * pi ** 2.5 / (p * q * sqrt(p + q)) * kab * kcd * f0 * np.prod(c)
    assert abs(got - ref) < 1e-14 * abs(ref)


def test_eri_symmetry_schwarz_translation():
    mol = M(WATER, basis='def2-tzvp')
    eri = oracle.md_eri_sph(mol)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.abs(eri - eri.transpose(perm)).max() < 1e-13
    n = eri.shape[0]
    diag = np.einsum('ijij->ij', eri)
    assert np.all(np.abs(eri) <= np.sqrt(np.abs(diag))[:, :, None, None]
                  * np.sqrt(np.abs(diag))[None, None] * (1 + 1e-10) + 1e-14)
    shifted = M('O 1.1 -2 0.3; H 1.857 -2 0.887; H 0.343 -2 0.887', basis='def2-tzvp')
    assert np.abs(oracle.md_eri_sph(shifted) - eri).max() < 1e-12


@pytest.mark.parametrize('basis', ['cc-pvdz', 'def2-tzvp', '6-31gs'])
def test_rys_quartets_match_md(basis):
    sb = _sb(basis, WATER + '; O 2.9 0.1 0.2')
    rng = np.random.default_rng(3)
    for _ in range(150):
        i, j, k, l = rng.integers(0, sb.nbas, 4)
        a = oracle.shell_quartet(sb, i, j, k, l, 'md')
        b = oracle.shell_quartet(sb, i, j, k, l, 'rys')
        assert np.abs(a - b).max() < 1e-13 * max(1.0, np.abs(a).max())


@pytest.mark.parametrize('basis', ['cc-pvdz', 'def2-tzvp'])
def test_rys_jk_vs_exact(basis):
    mol = M(WATER, basis=basis)
    rng = np.random.default_rng(5)
    dm = rng.random((mol.nao, mol.nao))
    dm = dm + dm.T
    ej, ek = oracle.exact_jk(mol, dm)
    for replica in (True, False):
        rj, rk = oracle.get_jk(mol, dm, cutoff=None, replica=replica)
        assert np.abs(rj - ej).max() < 1e-12
        assert np.abs(rk - ek).max() < 1e-12
    # Schwarz screening at 1e-13 changes nothing visible
    sj, sk = oracle.get_jk(mol, dm, cutoff=1e-13)
    assert np.abs(sj - ej).max() < 1e-11 and np.abs(sk - ek).max() < 1e-11


def test_int1e_normalisation_and_hermiticity():
    for basis in ('cc-pvdz', 'def2-tzvp', '6-31gs'):
        s, t, v = oracle.md_int1e_sph(M(WATER, basis=basis))
        assert np.abs(np.diag(s) - 1).max() < 1e-13
        for m in (s, t, v):
            assert np.abs(m - m.T).max() < 1e-12
        assert np.linalg.eigvalsh(t).min() > 0
