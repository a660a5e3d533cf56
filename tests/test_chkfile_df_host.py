"""Checkpoint files and the density-fitting set-up, on the CPU (no device work)."""
import numpy as np
import pytest
from scipy.special import erf

import oracle
from byteqc_amd.gto import M
from byteqc_amd.lib.int3c import UNIT_COEFF, even_tempered_basis
from byteqc_amd.scf import chkfile

WATER = 'O 0 0 0; H 0.757 0 0.587; H -0.757 0 0.587'


def test_chkfile_round_trip(tmp_path):
    mol = M(WATER, basis='cc-pvdz')
    rng = np.random.default_rng(1)
    c = rng.random((mol.nao, mol.nao))
    e = np.sort(rng.random(mol.nao))
    occ = np.r_[np.full(5, 2.0), np.zeros(mol.nao - 5)]
    path = chkfile.dump_scf(tmp_path / 'h2o', mol, -76.02, e, c, occ)
    mol2, scf = chkfile.load_scf(path)
    assert scf['e_tot'] == -76.02
    assert np.array_equal(scf['mo_coeff'], c) and np.array_equal(scf['mo_energy'], e)
    assert np.array_equal(scf['mo_occ'], occ)
    assert np.array_equal(mol2._env, mol._env) and np.array_equal(mol2._bas, mol._bas)
    assert chkfile.load(path, 'scf/e_tot') == -76.02
    # updating keeps the file consistent and replaces the scf entries
    chkfile.dump_scf(path, mol, -76.03, e, c, occ)
    assert chkfile.load(path, 'scf/e_tot') == -76.03


def test_chkfile_keeps_basis_dictionaries(tmp_path):
    basis = {'H': [[0, [0.8, 1.0]], [1, [1.1, 1.0]]]}
    mol = M('H 0 0 0; H 0 0 0.74', basis=basis)
    path = chkfile.dump_scf(tmp_path / 'h2.npz', mol, -1.0, np.zeros(8), np.eye(8), np.zeros(8))
    mol2 = chkfile.load(path, 'mol')
    assert mol2.nao == mol.nao and np.array_equal(mol2._env, mol._env)


def test_even_tempered_basis_shape():
    mol = M(WATER, basis='cc-pvdz')
    aux = even_tempered_basis(mol)
    assert set(aux) == {'O', 'H'}
    ls = [s[0] for s in aux['O']]
    assert ls.count(0) == 10 and ls.count(3) == 3 and max(s[0] for s in aux['H']) == 2
    e = [s[1][0] for s in aux['O'] if s[0] == 0]
    assert np.allclose(np.diff(np.log(e)), np.log(2.5))


def _f0(t):
    return 1.0 if t < 1e-14 else 0.5 * np.sqrt(np.pi / t) * erf(np.sqrt(t))


def test_unit_shell_gives_three_and_two_centre_integrals():
    """A shell of exponent 0 and coefficient 2 sqrt(pi) paired with P is P itself: the oracle's
    four-centre (ij|P unit) and (P unit|Q unit) equal the closed forms of (ss|s) and (s|s)."""
    geom = 'H 0 0 0; H 0.3 0.5 -0.7'
    mol = M(geom, basis={'H': [[0, [0.8, 1.0]]]}, unit='Bohr')
    aux = M(geom, basis={'H': [[0, [1.7, 1.0]]]}, unit='Bohr')
    env = np.concatenate([mol._env, aux._env, [0.0, UNIT_COEFF]])
    ab = aux._bas.copy()
    ab[:, 5] += len(mol._env)
    ab[:, 6] += len(mol._env)
    unit = np.array([[0, 0, 1, 1, 0, len(env) - 2, len(env) - 1, 0]])
    sb = oracle.SegBasis(mol._atm, np.vstack([mol._bas, unit, ab]), env, None)
    nrm = lambda a: (2 * a / np.pi) ** 0.75
    R = [mol._env[mol._atm[i, 1]:mol._atm[i, 1] + 3] for i in range(2)]
    a, c = 0.8, 1.7
    for i, j, k in ((0, 1, 3), (0, 1, 4), (0, 0, 4), (1, 1, 3)):
        A, B, C = R[i], R[j], R[k - 3]
        p = 2 * a
        P = (A + B) / 2
        ref = (nrm(a) ** 2 * nrm(c) * 2 * np.pi ** 2.5 / (p * c * np.sqrt(p + c))
               * np.exp(-a * a / p * np.sum((A - B) ** 2)) * _f0(p * c / (p + c) * np.sum((P - C) ** 2)))
        got = oracle.shell_quartet(sb, i, j, k, 2)[0, 0, 0, 0]
        assert abs(got - ref) < 1e-13 * abs(ref)
    ref2 = nrm(c) ** 2 * 2 * np.pi ** 2.5 / (c * c * np.sqrt(2 * c)) * _f0(c / 2 * np.sum((R[0] - R[1]) ** 2))
    assert abs(oracle.shell_quartet(sb, 3, 2, 4, 2)[0, 0, 0, 0] - ref2) < 1e-13 * ref2
