"""Batched J/K (many molecules per launch, hipGraph replay, autotuned ket chunks) against
the single-molecule path, which the parity tests pin to the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _clusters(sizes, basis='6-31gs', seed0=500):
    from byteqc_amd.gto.clusters import water_cluster
    from byteqc_amd.guess.train import box_for
    return [water_cluster(n, box_for(n), seed0 + i, basis=basis) for i, n in enumerate(sizes)]


def _psd_density(mol, rng):
    """D = 2 Q Q^T, Q orthonormal (nao x n_occ): PSD with trace = electron count."""
    n, nocc = mol.nao, mol.nelectron // 2
    q, _ = np.linalg.qr(rng.normal(size=(n, nocc)))
    return 2.0 * q @ q.T


def _rel(a, b):
    a = a.cpu().numpy() if hasattr(a, 'cpu') else a
    b = b.cpu().numpy() if hasattr(b, 'cpu') else b
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_batch_matches_single_molecule_builds():
    from byteqc_amd.scf import hf
    from byteqc_amd.scf.batch import BatchJK
    print('[batch] 6-31G* clusters of 1..5 waters', flush=True)
    mols = _clusters([3, 1, 5, 2, 4])
    rng = np.random.default_rng(1)
    dms = [_psd_density(m, rng) for m in mols]
    b = BatchJK(mols, use_graph=False)
    vj, vk = b.get_jk(dms)
    for m, d, j, k in zip(mols, dms, vj, vk):
        rj, rk = hf.get_jk(m, d, density_screen=False)
        assert _rel(j, rj) < 1e-12 and _rel(k, rk) < 1e-12
    st = b.stats()
    assert st['quartets_fp64'] > 0 and st['quartets_fp32'] == 0


def test_graph_replay_follows_the_density():
    """The graph is recorded once and replayed; D is rewritten in place between replays with
    eager work and synchronisations in between, and every replay matches an eager build."""
    import torch

    from byteqc_amd.scf.batch import BatchJK
    print('[batch] graph replays', flush=True)
    mols = _clusters([2, 4, 1])
    rng = np.random.default_rng(2)
    b = BatchJK(mols, use_graph=True)
    dm = b.to_cart([_psd_density(m, rng) for m in mols])
    vj = torch.empty_like(dm)
    vk = torch.empty_like(dm)
    for rep in range(4):
        new = b.to_cart([_psd_density(m, rng) for m in mols])
        dm.copy_(new)
        b.build_cart(dm, vj, vk, use_graph=True)
        torch.cuda.synchronize()
        ej, ek = b.build_cart(new.clone(), use_graph=False)
        torch.cuda.synchronize()
        assert _rel(vj, ej) < 1e-12 and _rel(vk, ek) < 1e-12, rep
        _ = (new * 2.0).sum().item()   # unrelated eager work between replays


def test_autotuned_chunks_keep_results_and_round_trip(tmp_path):
    from byteqc_amd.scf.batch import BatchJK
    print('[batch] autotune', flush=True)
    mols = _clusters([6, 3, 8])
    rng = np.random.default_rng(3)
    dms = [_psd_density(m, rng) for m in mols]
    b = BatchJK(mols, use_graph=False)
    flat = b.to_cart(dms)
    ref_j, ref_k = b.build_cart(flat)
    table = b.tune(flat, reps=1)
    assert len(table) == b.ncp
    vj, vk = b.build_cart(flat)
    assert _rel(vj, ref_j) < 1e-12 and _rel(vk, ref_k) < 1e-12
    path = tmp_path / 'kc.json'
    b.save_tuning(str(path))
    b2 = BatchJK(mols, use_graph=False)
    assert b2.load_tuning(str(path))
    assert np.array_equal(b2.tiles()[:, 6], b.tiles()[:, 6])


def test_single_molecule_batch_covers_f_shells():
    """One def2-TZVP water: the (ff|fd) and (ff|ff) classes use the block-LDS kernel, which
    a one-molecule batch runs on computed tiles."""
    from byteqc_amd.scf import hf
    from byteqc_amd.scf.batch import BatchJK
    mols = _clusters([2], basis='def2-tzvp')
    rng = np.random.default_rng(4)
    d = _psd_density(mols[0], rng)
    b = BatchJK(mols, use_graph=False)
    vj, vk = b.get_jk([d])
    rj, rk = hf.get_jk(mols[0], d, density_screen=False)
    assert _rel(vj[0], rj) < 1e-12 and _rel(vk[0], rk) < 1e-12


def test_multi_molecule_batch_rejects_block_lds_classes():
    from byteqc_amd.scf.batch import BatchJK
    with pytest.raises(NotImplementedError):
        BatchJK(_clusters([1, 1], basis='def2-tzvp'))
