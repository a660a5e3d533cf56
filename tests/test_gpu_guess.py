"""GPU tests of the learned guess: the bf16 MFMA GEMM in its three layouts, the perceptron's
forward and backward passes and Adam against the CPU oracle (oracle/mlp.py), training on
the device, and the trained model's gain over SAD on held-out clusters."""
import os

import numpy as np
import pytest

from oracle import mlp as om

pytestmark = pytest.mark.gpu


def _bits(x):
    """float values -> bf16 bit patterns (int16), rounded as the oracle rounds."""
    return (om.bf16(x).view(np.uint32) >> 16).astype(np.uint16).view(np.int16)


def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device='cuda')


@pytest.mark.parametrize('m,n,k', [(1, 1, 1), (70, 45, 37), (256, 196, 96), (130, 64, 300)])
@pytest.mark.parametrize('layout', [0, 1, 2])
def test_gemm_layouts(layout, m, n, k):
    from byteqc_amd.guess.mlp import gemm
    print(f'[gemm] layout {layout} m={m} n={n} k={k}', flush=True)
    rng = np.random.default_rng(layout * 1000 + m + n + k)
    A = om.bf16(rng.normal(size=(m, k)).astype(np.float32))
    B = om.bf16(rng.normal(size=(k, n)).astype(np.float32))
    a_store = A.T if layout == 2 else A
    b_store = B.T if layout == 1 else B
    lda = a_store.shape[1]
    ldb = b_store.shape[1]
    ref = A.astype(np.float64) @ B.astype(np.float64)
    c = gemm(layout, _dev(_bits(a_store)), _dev(_bits(b_store)), m, n, k, lda, ldb)
    tol = 1e-6 * k * np.abs(A).max() * np.abs(B).max()
    assert np.abs(c.cpu().numpy() - ref).max() <= tol
    c2 = gemm(layout, _dev(_bits(a_store)), _dev(_bits(b_store)), m, n, k, lda, ldb, out=c, beta=1.0)
    assert np.abs(c2.cpu().numpy() - 2 * ref).max() <= 2 * tol


def test_to_bf16_rounding_is_bit_exact():
    import torch

    from byteqc_amd.guess.mlp import to_bf16
    rng = np.random.default_rng(3)
    x = (rng.normal(size=4097) * 10.0 ** rng.integers(-20, 20, 4097)).astype(np.float32)
    ties = np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -(1.0 + 2 ** -8)], dtype=np.float32)
    x = np.concatenate([x, ties])
    assert x.dtype == np.float32
    got = to_bf16(torch.as_tensor(x, device='cuda')).cpu().numpy()
    assert np.array_equal(got, _bits(x))


def _params(rng, dims):
    p = {}
    for i in range(3):
        p[f'W{i + 1}'] = (rng.normal(size=(dims[i], dims[i + 1])) / np.sqrt(dims[i])).astype(np.float32)
        p[f'b{i + 1}'] = (0.1 * rng.normal(size=dims[i + 1])).astype(np.float32)
    return p


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize('dims,batch', [((96, 256, 256, 196), 300), ((96, 64, 64, 25), 77)])
def test_forward_backward_match_oracle(dims, batch):
    """HIP forward/backward vs the oracle with the same bf16 rounding points; the tolerance
    covers fp32 summation order and the rare bf16 rounding flip it causes."""
    import torch

    from byteqc_amd.guess.mlp import MLP, to_bf16
    print(f'[mlp] dims {dims} batch {batch}', flush=True)
    rng = np.random.default_rng(5)
    p = _params(rng, dims)
    x = rng.normal(size=(batch, dims[0])).astype(np.float32)
    t = rng.normal(size=(batch, dims[3])).astype(np.float32)
    mask = (rng.uniform(size=(batch, dims[3])) > 0.2).astype(np.float32)
    scale = 2.0 / mask.sum()
    net = MLP(dims, params=p)
    y, cache = net.forward(to_bf16(_dev(x)), batch)
    loss = torch.zeros(1, dtype=torch.float64, device='cuda')
    dy = net.loss_grad(y, _dev(t), _dev(mask), scale, loss)
    grads = net.backward(cache, dy, batch)
    y_ref, c_ref = om.forward(p, x)
    l_ref, dy_ref = om.mse(y_ref, t, mask, scale)
    g_ref = om.backward(c_ref, dy_ref.astype(np.float32))
    assert _rel(y.cpu().numpy(), y_ref) < 2e-3
    assert abs(loss.item() - l_ref) < 1e-3 * l_ref
    for j, name in enumerate(('W1', 'W2', 'W3', 'b1', 'b2', 'b3')):
        assert _rel(grads[j].cpu().numpy(), g_ref[name]) < 5e-3, name


def test_adam_matches_oracle():
    import torch

    from byteqc_amd.guess.mlp import MLP
    rng = np.random.default_rng(8)
    dims = (96, 64, 64, 25)
    p = _params(rng, dims)
    net = MLP(dims, params=p)
    w = p['W1'].astype(np.float64)
    m1 = np.zeros_like(w)
    m2 = np.zeros_like(w)
    for step in range(1, 4):
        g = rng.normal(size=w.shape).astype(np.float32)
        gs = [_dev(g)] + [torch.zeros_like(t) for t in net.w[1:] + net.b]
        net.adam(gs, 1e-2)
        w, m1, m2 = om.adam(w, g.astype(np.float64), m1, m2, 1e-2, 0.9, 0.999, 1e-8, step)
    got = net.w[0].cpu().numpy()
    assert np.abs(got - w).max() < 1e-5
    # the bf16 copy is the rounded master weight
    assert np.array_equal(net.wb[0].cpu().numpy(), _bits(got))


def test_training_on_device_reduces_loss():
    """Fit a random teacher network with the device kernels: the loss falls tenfold."""
    import torch

    from byteqc_amd.guess.mlp import MLP, to_bf16
    rng = np.random.default_rng(11)
    dims = (96, 128, 128, 32)
    teacher = _params(rng, dims)
    x = rng.normal(size=(8192, dims[0])).astype(np.float32)
    t, _ = om.forward(teacher, x, emulate=False)
    t = t.astype(np.float32)
    net = MLP(dims, seed=1)
    xb, td = to_bf16(_dev(x)), _dev(t)
    ones = torch.ones_like(td)
    loss = torch.zeros(1, dtype=torch.float64, device='cuda')
    first = None
    for step in range(300):
        loss.zero_()
        y, cache = net.forward(xb, x.shape[0])
        dy = net.loss_grad(y, td, ones, 2.0 / t.size, loss)
        net.adam(net.backward(cache, dy, x.shape[0]), 3e-3)
        if first is None:
            first = loss.item()
    assert loss.item() < 0.1 * first


def test_learned_guess_beats_sad_on_held_out_clusters():
    """The committed cc-pVDZ model: fewer SCF cycles than SAD to the same energy on held-out
    seeds (training seeds start at 10000, these at 900000)."""
    from byteqc_amd.guess import train
    from byteqc_amd.guess.learned import Model, model_path
    path = model_path('cc-pvdz')
    assert os.path.exists(path), 'trained model missing'
    model = Model.load(path)
    rows = train.evaluate(model, 'cc-pvdz', 6, [train.EVAL_SEED0 + 100, train.EVAL_SEED0 + 101], log=print)
    for r in rows:
        assert r['sad']['converged'] and r['learned']['converged']
        assert abs(r['delta_e']) < 1e-7
        assert r['learned']['cycles'] < r['sad']['cycles']
        e = r['sad']['e_tot']
        assert abs(r['learned']['e_first'] - e) < abs(r['sad']['e_first'] - e)


def test_fragment_baseline_is_exact_for_distant_copies():
    """Two rigid waters 20 A apart in different orientations: the converged density equals the
    superposition of the isolated water's density rotated into each orientation (what remains
    is their interaction, far below the 1e-4 tolerance; a wrong rotation convention gives
    errors of order 0.1)."""
    from byteqc_amd.gto.clusters import _rotation, water_monomer
    from byteqc_amd.gto.mole import Mole
    from byteqc_amd.guess.fragments import fragment_density
    from byteqc_amd.scf.rhf import RHF
    rng = np.random.default_rng(12)
    mono = water_monomer()
    xyz = np.concatenate([mono @ _rotation(rng.normal(size=4)).T,
                          mono @ _rotation(rng.normal(size=4)).T + np.array([20.0, 0.0, 0.0])])
    mol = Mole([(('O', 'H', 'H')[i % 3], tuple(p)) for i, p in enumerate(xyz)], basis='cc-pvdz')
    mf = RHF(mol, conv_tol=1e-10)
    mf.kernel()
    assert mf.converged
    d = mf.dm.cpu().numpy()
    assert np.abs(fragment_density(mol) - d).max() < 1e-4
