"""Train the learned guess on converged densities from this package's own SCF.

    python -m byteqc_amd.guess.train --basis cc-pvdz --clusters 2000 --min-waters 2 \
        --max-waters 16 --out byteqc_amd/guess/weights_cc-pvdz.npz

Data: seeded water clusters (O-H 0.957 A, 104.5 deg; sizes uniform in [min, max]; box
edge (n / 0.0156)^(1/3) A, at least 5 A, i.e. liquid-water density) converged by
`byteqc_amd.scf.rhf.RHF` from SAD; the target is D minus the superposition of fragment
densities, and every pair block within R_PAIR of it, in its local frame, is one sample.  Training: the perceptron kernels of libmi355jk.so (bf16 MFMA forward and
backward, Adam), minibatches drawn on the device, loss = mean squared block error in
density units over the real (unpadded) entries.  Training seeds start at TRAIN_SEED0;
held-out evaluation seeds at EVAL_SEED0, so the two sets never share a geometry.
"""
from __future__ import annotations

import argparse
import math
import sys
import time

import numpy as np

from . import frames as fr
from .fragments import fragment_density
from .learned import Model, learned_guess

TRAIN_SEED0 = 10_000
EVAL_SEED0 = 900_000
DENSITY = 0.0156   # waters per cubic Angstrom


def box_for(n):
    return max(5.0, (n / DENSITY) ** (1.0 / 3.0))


def _log(msg):
    print(msg, flush=True)


def converged(mol, conv_tol=1e-8, init_guess='sad', dm0=None, precision_cutoff=0.0, early_tol=None):
    from ..scf.rhf import RHF
    mf = RHF(mol, conv_tol=conv_tol, precision_cutoff=precision_cutoff, early_tol=early_tol,
             early_until=1e-4)
    mf.kernel(dm0=dm0, init_guess=init_guess)
    return mf


def make_dataset(basis, n_clusters, min_w, max_w, seed0=TRAIN_SEED0, conv_tol=1e-7, max_seconds=None, early_tol=None,
                 log=_log):
    from ..gto.clusters import water_cluster
    rng = np.random.default_rng(seed0)
    sizes = rng.integers(min_w, max_w + 1, n_clusters)
    X, Y, T, C = [], [], [], []
    nmax = None
    t0 = time.perf_counter()
    for c in range(n_clusters):
        if max_seconds is not None and time.perf_counter() - t0 > max_seconds:
            log(f'[train] time limit: stopped after {c} clusters')
            break
        n = int(sizes[c])
        mol = water_cluster(n, box_for(n), seed0 + 1 + c, basis=basis)
        if nmax is None:
            nmax = max(v[1] for v in fr.element_layout(mol).values())
        mf = converged(mol, conv_tol, early_tol=early_tol)
        if not mf.converged:
            log(f'[train] cluster {c} ({n} waters) did not converge; skipped')
            continue
        dm = mf.dm.cpu().numpy() - fragment_density(mol)
        sym, xyz = fr.mol_geometry(mol)
        f, frames, types, pairs = fr.features(sym, xyz)
        X.append(f)
        Y.append(fr.extract_blocks(mol, dm, pairs, frames, nmax))
        T.append(types)
        C.append(np.full(len(pairs), c))
        if c % 25 == 0 or c == n_clusters - 1:
            log(f'[train] data {c + 1}/{n_clusters}: {n} waters, {mf.cycles} cycles, '
                f'{len(pairs)} pairs, {time.perf_counter() - t0:.0f} s')
    return np.concatenate(X), np.concatenate(Y), np.concatenate(T), np.concatenate(C), nmax, mol


def fit(X, Y, types, clusters, masks, basis, elements, nmax, hidden=256, epochs=80, batch=4096, lr=2e-3,
        seed=0, val_frac=0.05, log=_log):
    """Train on the GPU; returns a Model."""
    import torch

    from .mlp import MLP, to_bf16
    ntype, nout = masks.shape
    x_mean = X.mean(0)
    x_std = X.std(0)
    x_std[x_std < 1e-6] = 1.0
    y_mean = np.zeros((ntype, nout), dtype=np.float32)
    y_std = np.ones((ntype, nout), dtype=np.float32)
    for t in range(ntype):
        sel = types == t
        if sel.any():
            y_mean[t] = Y[sel].mean(0) * masks[t]
            s = Y[sel].std(0)
            y_std[t] = np.where(masks[t] > 0, np.maximum(s, 1e-5), 1.0)
    # loss weights: the squared spread, so the loss is the block error in density units
    w = masks * y_std ** 2
    w = w / w[masks > 0].mean()
    ncl = int(clusters.max()) + 1
    val = clusters >= ncl - max(1, int(round(val_frac * ncl)))
    tr = ~val
    Xs = ((X - x_mean) / x_std).astype(np.float32)
    Ys = ((Y - y_mean[types]) / y_std[types]).astype(np.float32)

    def dev(a):
        return torch.as_tensor(a, device='cuda').contiguous()

    xb_tr, y_tr, t_tr = to_bf16(dev(Xs[tr])), dev(Ys[tr]), dev(types[tr])
    xb_va, y_va, t_va = to_bf16(dev(Xs[val])), dev(Ys[val]), dev(types[val])
    wd = dev(w.astype(np.float32))
    dims = (X.shape[1], hidden, hidden, nout)
    mlp = MLP(dims, seed=seed)
    ntr = int(tr.sum())
    steps_per_epoch = max(1, ntr // batch)
    total = steps_per_epoch * epochs
    loss = torch.zeros(1, dtype=torch.float64, device='cuda')
    gen = torch.Generator(device='cuda')
    gen.manual_seed(seed)
    valid_per_row = float(masks.sum(1).mean())
    t0 = time.perf_counter()
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(ntr, device='cuda', generator=gen)
        loss.zero_()
        for s in range(steps_per_epoch):
            step = ep * steps_per_epoch + s
            idx = perm[s * batch:(s + 1) * batch]
            m = idx.numel()
            xb = xb_tr.index_select(0, idx)
            yt = y_tr.index_select(0, idx)
            mk = wd.index_select(0, t_tr.index_select(0, idx))
            y, cache = mlp.forward(xb, m)
            dy = mlp.loss_grad(y, yt, mk, 2.0 / (m * valid_per_row), loss)
            grads = mlp.backward(cache, dy, m)
            lr_t = lr * 0.5 * (1.0 + math.cos(math.pi * step / total))
            mlp.adam(grads, lr_t)
        train_mse = float(loss.item()) / (steps_per_epoch * batch * valid_per_row)
        if ep % 10 == 0 or ep == epochs - 1:
            loss.zero_()
            for s in range(0, xb_va.shape[0], 65536):
                xs = xb_va[s:s + 65536]
                y, _ = mlp.forward(xs, xs.shape[0])
                mlp.loss_grad(y, y_va[s:s + 65536], wd.index_select(0, t_va[s:s + 65536]), 0.0, loss)
            val_mse = float(loss.item()) / (max(1, xb_va.shape[0]) * valid_per_row)
            hist.append((ep, train_mse, val_mse))
            log(f'[train] epoch {ep}: train mse {train_mse:.3e}  val mse {val_mse:.3e}  '
                f'{time.perf_counter() - t0:.1f} s')
    torch.cuda.synchronize()
    data = dict(mlp.params())
    data.update(dims=np.asarray(dims), nmax=np.asarray(nmax), basis=np.asarray(str(basis).lower()),
                elements=np.asarray(elements), x_mean=x_mean.astype(np.float32), x_std=x_std.astype(np.float32),
                y_mean=y_mean, y_std=y_std, loss_weights=w.astype(np.float32), baseline=np.asarray('fragments'),
                history=np.asarray(hist, dtype=np.float64),
                train_info=np.asarray([len(X), ntr, epochs, batch, lr, time.perf_counter() - t0]))
    return Model(data)


def evaluate(model, basis, n_waters, seeds, precision_cutoff=0.0, log=_log):
    """SCF from SAD and from the learned guess on held-out clusters."""
    from ..gto.clusters import water_cluster
    from ..scf.rhf import RHF
    out = []
    for sd in seeds:
        mol = water_cluster(n_waters, box_for(n_waters), sd, basis=basis)
        row = {'seed': int(sd), 'n_waters': n_waters}
        for key in ('sad', 'learned'):
            mf = RHF(mol, precision_cutoff=precision_cutoff)
            mf.build()
            t0 = time.perf_counter()
            dm0 = learned_guess(mol, mf.get_ovlp(), model) if key == 'learned' else None
            mf.kernel(dm0=dm0, init_guess='sad')
            row[key] = {'cycles': mf.cycles, 'e_tot': mf.e_tot, 'converged': mf.converged,
                        'time_s': time.perf_counter() - t0, 'e_first': mf.history[0]['e_tot']}
        row['delta_e'] = row['learned']['e_tot'] - row['sad']['e_tot']
        log(f'[eval] seed {sd}: SAD {row["sad"]["cycles"]} cycles {row["sad"]["time_s"]:.2f} s, '
            f'learned {row["learned"]["cycles"]} cycles {row["learned"]["time_s"]:.2f} s, dE {row["delta_e"]:.1e}')
        out.append(row)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split('\n\n')[0])
    ap.add_argument('--basis', default='cc-pvdz')
    ap.add_argument('--clusters', type=int, default=2000)
    ap.add_argument('--min-waters', type=int, default=2)
    ap.add_argument('--max-waters', type=int, default=16)
    ap.add_argument('--epochs', type=int, default=80)
    ap.add_argument('--batch', type=int, default=4096)
    ap.add_argument('--lr', type=float, default=2e-3)
    ap.add_argument('--hidden', type=int, default=256)
    ap.add_argument('--conv-tol', type=float, default=1e-7)
    ap.add_argument('--max-seconds', type=float, default=None, help='stop generating data after this long')
    ap.add_argument('--early-tol', type=float, default=None,
                    help='looser quartet screening in the data SCFs until max|[F,D]| < 1e-4')
    ap.add_argument('--out', required=True)
    ap.add_argument('--eval-waters', type=int, default=32)
    ap.add_argument('--eval-seeds', type=int, default=2)
    a = ap.parse_args(argv)
    X, Y, T, C, nmax, mol = make_dataset(a.basis, a.clusters, a.min_waters, a.max_waters, conv_tol=a.conv_tol,
                                         max_seconds=a.max_seconds, early_tol=a.early_tol)
    _log(f'[train] {len(X)} pair samples, nmax {nmax}')
    masks = fr.block_masks(mol, nmax)
    elements = sorted({mol.atom_symbol(i) for i in range(mol.natm)})
    model = fit(X, Y, T, C, masks, a.basis, elements, nmax, hidden=a.hidden, epochs=a.epochs, batch=a.batch, lr=a.lr)
    model.data['train_clusters'] = np.asarray([int(C.max()) + 1, a.min_waters, a.max_waters, TRAIN_SEED0])
    model.save(a.out)
    _log(f'[train] wrote {a.out}')
    if a.eval_seeds > 0:
        evaluate(model, a.basis, a.eval_waters, [EVAL_SEED0 + i for i in range(a.eval_seeds)])
    return 0


if __name__ == '__main__':
    sys.exit(main())
