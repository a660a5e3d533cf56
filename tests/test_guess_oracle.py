"""CPU checks of the learned guess: the perceptron oracle against finite differences, the
bf16 rounding against exact cases, and the local frames' invariance and covariance."""
import numpy as np
import pytest

from oracle import mlp as om
from byteqc_amd.guess import frames as fr
from byteqc_amd.gto.clusters import water_cluster, _rotation
from byteqc_amd.gto.mole import Mole


def _params(rng, dims):
    p = {}
    for i in range(3):
        p[f'W{i + 1}'] = rng.normal(size=(dims[i], dims[i + 1])) / np.sqrt(dims[i])
        p[f'b{i + 1}'] = 0.1 * rng.normal(size=dims[i + 1])
    return p


def test_oracle_backward_matches_finite_differences():
    """Pins the oracle's backward formulas: float64, no rounding, central differences."""
    rng = np.random.default_rng(0)
    dims = (7, 6, 5, 4)
    p = _params(rng, dims)
    x = rng.normal(size=(9, dims[0]))
    t = rng.normal(size=(9, dims[3]))
    mask = (rng.uniform(size=(9, dims[3])) > 0.3).astype(float)

    def loss(pp):
        y, _ = om.forward(pp, x, emulate=False)
        return om.mse(y, t, mask, 1.0)[0]

    y, cache = om.forward(p, x, emulate=False)
    _, dy = om.mse(y, t, mask, 2.0)
    g = om.backward(cache, dy, emulate=False)
    h = 1e-6
    for name in ('W1', 'W2', 'W3', 'b1', 'b2', 'b3'):
        num = np.zeros_like(p[name])
        it = np.nditer(p[name], flags=['multi_index'])
        for _ in it:
            i = it.multi_index
            pp = {k: v.copy() for k, v in p.items()}
            pp[name][i] += h
            lp = loss(pp)
            pp[name][i] -= 2 * h
            lm = loss(pp)
            num[i] = (lp - lm) / (2 * h)
        assert np.allclose(g[name], num, rtol=1e-6, atol=1e-8), name


def test_bf16_rounding_exact_cases():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0, 1e-30],
                 dtype=np.float32)
    r = om.bf16(v)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa wins (1); 1 + 3 2^-8 -> 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2 ** -6)
    assert r[3] == 1.0 and r[4] == -2.5 and r[5] == 0.0
    assert abs(r[6] - 1e-30) <= 2 ** -8 * 1e-30


def test_adam_oracle_first_step_is_signed_lr():
    w, g = np.array([1.0, -2.0, 3.0]), np.array([0.5, -0.1, 0.0])
    w1, m1, m2 = om.adam(w, g, np.zeros(3), np.zeros(3), 0.01, 0.9, 0.999, 1e-12, 1)
    assert np.allclose(w1 - w, -0.01 * np.sign(g))


def test_harmonic_rotations_are_an_orthogonal_representation():
    rng = np.random.default_rng(1)
    F1, F2 = _rotation(rng.normal(size=4)), _rotation(rng.normal(size=4))
    for l in range(4):
        M1 = fr.harmonic_rotations(l, F1[None])[0]
        M2 = fr.harmonic_rotations(l, F2[None])[0]
        M12 = fr.harmonic_rotations(l, (F1 @ F2)[None])[0]
        assert np.allclose(M1 @ M1.T, np.eye(2 * l + 1), atol=1e-13)
        assert np.allclose(M12, M1 @ M2, atol=1e-13)


@pytest.mark.parametrize('basis', ['cc-pvdz', 'def2-tzvp'])
def test_features_invariant_blocks_covariant(basis):
    """Rotating and translating the cluster leaves the features and the local-frame blocks of
    the correspondingly rotated density unchanged."""
    rng = np.random.default_rng(2)
    mol = water_cluster(5, 7.5, 11, basis=basis)
    sym, xyz = fr.mol_geometry(mol)
    f, F, ty, pairs = fr.features(sym, xyz)
    Q = _rotation(rng.normal(size=4))
    mol2 = Mole([(s, tuple(Q @ x + np.array([1.0, -2.0, 0.5]))) for s, x in zip(sym, xyz)], basis=basis)
    f2, F2, ty2, pairs2 = fr.features(*fr.mol_geometry(mol2))
    assert pairs == pairs2 and np.array_equal(ty, ty2)
    assert np.abs(f - f2).max() < 1e-10
    n = mol.nao
    D = rng.normal(size=(n, n))
    D = D + D.T
    lay = fr.element_layout(mol)
    sl = mol.aoslice_by_atom()
    R = np.zeros((n, n))
    for ia in range(mol.natm):
        R[sl[ia, 2]:sl[ia, 3], sl[ia, 2]:sl[ia, 3]] = fr.atom_rotations(lay[sym[ia]][0], Q[None])[0]
    nmax = max(v[1] for v in lay.values())
    b1 = fr.extract_blocks(mol, D, pairs, F, nmax)
    b2 = fr.extract_blocks(mol2, R @ D @ R.T, pairs2, F2, nmax)
    assert np.abs(b1 - b2).max() < 1e-5 * np.abs(b1).max()
    # assemble inverts extract on the blocks it covers
    Dr = fr.assemble(mol, b1, pairs, F, nmax)
    covered = Dr != 0
    assert np.abs((Dr - D)[covered]).max() < 1e-5 * np.abs(D).max()


def test_fragments_and_alignment():
    """Water clusters split into their molecules; every molecule aligns onto the first with
    rmsd below 1e-3 A, including a copy whose hydrogens are listed in the other order."""
    from byteqc_amd.guess import fragments as fg
    mol = water_cluster(6, 8.0, 21, basis='cc-pvdz')
    sym, xyz = fr.mol_geometry(mol)
    frags = fg.find_fragments(sym, xyz)
    assert sorted(len(f) for f in frags) == [3] * 6
    ref = frags[0]
    for f in frags[1:]:
        r, perm = fg.align([sym[a] for a in ref], xyz[ref], [sym[a] for a in f], xyz[f])
        assert np.allclose(r @ r.T, np.eye(3)) and abs(np.linalg.det(r) - 1) < 1e-12
        moved = (xyz[ref] - xyz[ref].mean(0)) @ r.T + xyz[f].mean(0)
        assert np.abs(moved - xyz[f][perm]).max() < 1e-3
    swapped = [f[0], f[2], f[1]]
    r, perm = fg.align([sym[a] for a in ref], xyz[ref], [sym[a] for a in swapped], xyz[swapped])
    assert np.abs((xyz[ref] - xyz[ref].mean(0)) @ r.T + xyz[swapped].mean(0) - xyz[swapped][perm]).max() < 1e-3
    # a distorted copy is not a rigid copy
    bent = xyz[f].copy()
    bent[1] += 0.05
    assert fg.align([sym[a] for a in ref], xyz[ref], [sym[a] for a in f], bent) is None


def test_projection_between_bases():
    """Self-projection is the identity; the 'to' block of the combined overlap equals the
    'to' basis' own overlap; projecting the cc-pVDZ overlap inverse (tr(D S) = AO count) into
    def2-TZVP loses weight but little."""
    from byteqc_amd.guess.project import cross_overlap, project_density
    from byteqc_amd.gto.mole import Mole
    from byteqc_amd.scf.hf import basis_seg_contraction
    from byteqc_amd.scf.rhf import int1e
    from types import SimpleNamespace

    def overlap(m):
        bas, env, coeff = basis_seg_contraction(m)
        return int1e(SimpleNamespace(atm=np.ascontiguousarray(m._atm, dtype=np.int32), bas=bas, env=env,
                                     coeff=coeff))[0]

    mol = water_cluster(2, 6.0, 3, basis='cc-pvdz')
    tz = Mole(mol.atom, basis='def2-tzvp')
    rng = np.random.default_rng(9)
    d = rng.normal(size=(mol.nao, mol.nao))
    d = d + d.T
    assert np.allclose(project_density(mol, d, mol), d, atol=1e-9)
    s_tt, s_tf = cross_overlap(mol, tz)
    assert np.allclose(s_tt, overlap(tz), atol=1e-12)
    s_ff = overlap(mol)
    dz = np.linalg.inv(s_ff)            # tr(dz s_ff) = nao_dz
    dt = project_density(mol, dz, tz)
    # a projection can only lose weight: what def2-TZVP cannot represent of arbitrary
    # cc-pVDZ functions (about 3% here)
    tr = np.sum(dt * s_tt)
    assert 0.9 * mol.nao < tr <= mol.nao + 1e-9
