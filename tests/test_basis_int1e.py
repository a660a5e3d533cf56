"""Basis data, AO conventions and the library's one-electron integrals (CPU)."""

import numpy as np
import pytest

import oracle
from byteqc_amd import lib as mjk
from byteqc_amd.gto import M, water_cluster, water_cluster_coords, cart_components
from byteqc_amd.gto.cart2sph import cart2sph_block
from byteqc_amd.scf.hf import basis_seg_contraction

WATER = 'O 0 0 0; H 0.757 0 0.587; H -0.757 0 0.587'


@pytest.mark.parametrize('basis,nao,ncart', [('cc-pvdz', 24, 25), ('def2-tzvp', 43, 48),
                                             ('6-31gs', 18, 19)])
def test_sizes(basis, nao, ncart):
    mol = M(WATER, basis=basis)
    assert mol.nao == nao and mol.nao_nr(cart=True) == ncart
    assert mol.nelectron == 10


def test_config_sizes():
    assert water_cluster(8, 9.0, 1, 'cc-pvdz').nao == 192
    m = water_cluster(64, 16.0, 1, 'def2-tzvp')
    assert m.nao == 2752 and m.nao_nr(cart=True) == 3072 and m.nbas == 1216


def test_cluster_geometry():
    xyz = water_cluster_coords(32, 12.0, 11)
    o = xyz[0::3]
    d = np.linalg.norm(o[:, None] - o[None], axis=-1) + np.eye(len(o)) * 10
    assert d.min() >= 2.6
    oh = np.linalg.norm(xyz[1::3] - o, axis=1)
    assert np.allclose(oh, 0.957)
    ang = np.degrees(np.arccos(np.einsum('ij,ij->i', xyz[1::3] - o, xyz[2::3] - o) / 0.957 ** 2))
    assert np.allclose(ang, 104.5)
    assert np.array_equal(xyz, water_cluster_coords(32, 12.0, 11))


def _sphere_moment(a, b, c):
    """int x^a y^b z^c dOmega over the unit sphere."""
    from math import gamma
    if a % 2 or b % 2 or c % 2:
        return 0.0
    return 2 * gamma((a + 1) / 2) * gamma((b + 1) / 2) * gamma((c + 1) / 2) / gamma((a + b + c + 3) / 2)


@pytest.mark.parametrize('l', [2, 3, 4])
def test_solid_harmonics_orthonormal(l):
    c = cart2sph_block(l)
    comps = cart_components(l)
    g = np.array([[_sphere_moment(*(np.add(p, q))) for q in comps] for p in comps])
    assert np.abs(c.T @ g @ c - np.eye(2 * l + 1)).max() < 1e-13


def test_d_f_harmonics_known_forms():
    d = cart2sph_block(2)  # xx xy xz yy yz zz ; m = -2..2
    assert np.isclose(d[1, 0], np.sqrt(15 / (4 * np.pi)))          # d_xy
    assert np.isclose(d[5, 2], 2 * np.sqrt(5 / (16 * np.pi)))      # d_z2 zz
    assert np.isclose(d[0, 4], np.sqrt(15 / (16 * np.pi)))         # d_x2-y2
    f = cart2sph_block(3)  # xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz
    assert np.isclose(f[4, 1], np.sqrt(105 / (4 * np.pi)))         # f_xyz
    assert np.isclose(f[6, 0], -np.sqrt(35 / (32 * np.pi)))        # -y^3 in f_-3


def _lib_int1e(mol):
    bas, env, c = basis_seg_contraction(mol)
    n = c.shape[0]
    s, t, v = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    rc = mjk.load().mjk_int1e(mjk.ptr(mol._atm), mol.natm, mjk.ptr(bas), len(bas), mjk.ptr(env),
                              mjk.ptr(s), mjk.ptr(t), mjk.ptr(v), 0)
    assert rc == 0
    return c.T @ s @ c, c.T @ t @ c, c.T @ v @ c


@pytest.mark.parametrize('basis', ['cc-pvdz', 'def2-tzvp', '6-31gs'])
def test_int1e_vs_oracle(basis):
    mol = M(WATER + '; O 2.9 0.1 0.2; H 3.6 0.2 0.8; H 2.7 -0.8 0.5', basis=basis)
    got = _lib_int1e(mol)
    ref = oracle.md_int1e_sph(mol)
    for g, r in zip(got, ref):
        assert np.abs(g - r).max() < 1e-12
    assert np.abs(np.diag(got[0]) - 1).max() < 1e-13


def test_decontraction_preserves_span():
    """C^T S_cart C must equal the contracted overlap (unit diagonal)."""
    mol = M(WATER, basis='cc-pvdz')
    s, _, _ = _lib_int1e(mol)
    assert np.abs(np.diag(s) - 1).max() < 1e-13
    bas, env, c = basis_seg_contraction(mol)
    assert c.shape == (41, 24) and len(bas) == 24  # O 9s 4p 1d, H 4s 1p (primitive shells)
