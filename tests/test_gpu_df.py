"""Density fitting on the device (byteqc_amd.lib.int3c): (ij|P) and (P|Q) against the oracle's
four-centre integrals with the unit shell, and the properties of Coulomb-metric fitting."""
import numpy as np
import pytest

import oracle
from byteqc_amd.gto import M

pytestmark = pytest.mark.gpu

WATER = 'O 0 0 0; H 0.757 0 0.587; H -0.757 0 0.587'


def _ao_loc(bas):
    l = bas[:, 1]
    return np.append(0, np.cumsum((l + 1) * (l + 2) // 2))


def test_int3c_int2c_match_oracle():
    from byteqc_amd.lib.int3c import Int3c, even_tempered_basis
    mol = M(WATER, basis='cc-pvdz')
    auxb = even_tempered_basis(mol, lmax=2, count=(6, 4, 2, 0), light_count=(4, 2, 1, 0))
    plan = Int3c(mol, M(WATER, basis=auxb))
    print(f'case: water cc-pvdz, {plan.naux_c} Cartesian auxiliary functions', flush=True)
    g3 = plan.int3c_cart().cpu().numpy()
    g2 = plan.int2c_cart().cpu().numpy()
    sb = oracle.SegBasis(plan.atm, plan.bas, plan.env, None)
    ao = _ao_loc(plan.bas)
    u, nb = plan.iunit, plan.vo.nbas
    aux = range(u + 1, len(plan.bas))
    ref3 = np.zeros_like(g3)
    for p in aux:
        sp = slice(ao[p] - plan.aux_off, ao[p + 1] - plan.aux_off)
        for i in range(nb):
            for j in range(nb):
                blk = oracle.shell_quartet(sb, p, u, i, j)[:, 0]          # (P, i, j)
                ref3[ao[i]:ao[i + 1], ao[j]:ao[j + 1], sp] = blk.transpose(1, 2, 0)
    ref2 = np.zeros_like(g2)
    for p in aux:
        for q in aux:
            ref2[ao[p] - plan.aux_off:ao[p + 1] - plan.aux_off, ao[q] - plan.aux_off:ao[q + 1] - plan.aux_off] = \
                oracle.shell_quartet(sb, p, u, q, u)[:, 0, :, 0]
    assert np.abs(g3 - ref3).max() < 1e-11 * np.abs(ref3).max()
    assert np.abs(g2 - ref2).max() < 1e-11 * np.abs(ref2).max()
    # (ij|P) = (ji|P): copied across shell groups, computed twice (to rounding) inside one
    assert np.abs(g3 - g3.transpose(1, 0, 2)).max() < 1e-14 * np.abs(ref3).max()
    plan.close()


def test_df_j_is_exact_when_the_density_is_in_the_auxiliary_span():
    """One s primitive of exponent a: its square is an s Gaussian of exponent 2a, so with that
    single auxiliary function the fit is exact and J_DF = (ss|ss) D."""
    from byteqc_amd.lib.int3c import Int3c
    from byteqc_amd.scf import hf
    mol = M('H 0 0 0', basis={'H': [[0, [0.5, 1.0]]]}, spin=1)
    aux = M('H 0 0 0', basis={'H': [[0, [1.0, 1.0]]]}, spin=1)
    plan = Int3c(mol, aux)
    d = np.array([[0.7]])
    vj, vk = plan.df_jk(d)
    rj, rk = hf.get_jk(mol, d)
    assert abs(float(vj.cpu()[0, 0]) - rj[0, 0]) < 1e-12 * abs(rj[0, 0])
    assert abs(float(vk.cpu()[0, 0]) - rk[0, 0]) < 1e-12 * abs(rk[0, 0])
    plan.close()


def test_df_coulomb_energy_is_bounded_and_converges():
    """Coulomb-metric fitting never overestimates the Coulomb energy of a PSD density, and a
    larger auxiliary space (containing the smaller one) gets closer: E_small <= E_large <= E."""
    from byteqc_amd.lib.int3c import Int3c, even_tempered_basis
    from byteqc_amd.scf import hf
    mol = M(WATER, basis='cc-pvdz')
    rng = np.random.default_rng(4)
    c = rng.standard_normal((mol.nao, 5)) * 0.3
    d = c @ c.T
    rj, rk = hf.get_jk(mol, d)
    e = float(np.sum(d * rj))
    ek = float(np.sum(d * rk))
    small = even_tempered_basis(mol, lmax=1, count=(6, 3, 0, 0), light_count=(4, 1, 0, 0))
    large = even_tempered_basis(mol)
    es = []
    for auxb in (small, large):
        plan = Int3c(mol, M(WATER, basis=auxb))
        vj, vk = plan.df_jk(d)
        vj, vk = vj.cpu().numpy(), vk.cpu().numpy()
        assert np.abs(vj - vj.T).max() < 1e-12 * np.abs(vj).max()
        es.append((float(np.sum(d * vj)), float(np.sum(d * vk))))
        plan.close()
    print('E_J exact / small / large:', e, es[0][0], es[1][0], flush=True)
    assert es[0][0] <= es[1][0] + 1e-10 and es[1][0] <= e + 1e-10
    assert (e - es[1][0]) < 2e-3 * e
    assert abs(ek - es[1][1]) < 2e-2 * ek
