"""Batched J/K for many molecules of uneven size: one launch per quartet class for all of
them, recorded once into a hipGraph and replayed every SCF iteration, with the ket chunk
of every class pair picked by timing (the autotuner).

The reference computes one molecule per `get_jk` call (cuobc/scf/hf.py:53-227).  Here
each molecule gets its own `_VHFOpt` pair lists (same screening, same ordering); the
batch concatenates their shells, atoms and pair lists class by class (molecule 0's pairs
first, then molecule 1's, ...), so no pair ever spans two molecules, and hands the merged
cache to `mjk_batch_create`.  D, J and K of all molecules live in one flat device array,
molecule m's nao_m x nao_m Cartesian block at offset dm_off[m].
"""
from __future__ import annotations

import ctypes
import json
import os
import time

import numpy as np

from .. import lib as mjk
from ..gto.mole import ATOM_OF, PTR_COEFF, PTR_EXP
from . import hf

KC_CANDIDATES = {0: (16, 32, 64, 128, 256), 1: (4, 8, 16, 32, 64)}


def _torch():
    import torch
    return torch


def _label(gb, gk):
    """'(l nprim, l nprim | l nprim, l nprim)' of a bra and a ket pair class, e.g. 's6s3|p3s1'."""
    def one(g):
        return f'{"spdf"[g[0][0]]}{g[0][1]}{"spdf"[g[1][0]]}{g[1][1]}'
    return f'{one(gb)}|{one(gk)}'


class BatchJK:
    """J/K of many molecules per call.  mols: list of Mole; all share one device."""

    def __init__(self, mols, cutoff=1e-13, use_graph=True):
        torch = _torch()
        self.mols = list(mols)
        self.cutoff = cutoff
        self.use_graph = use_graph
        t0 = time.perf_counter()
        self.opts = [hf._VHFOpt(m).build(cutoff) for m in self.mols]
        for o in self.opts:
            o.clear()  # per-molecule device caches are not needed here
        self._merge()
        self.timings = {'setup': time.perf_counter() - t0}
        self.stream = torch.cuda.Stream()
        self.tuned = None

    # ---- construction --------------------------------------------------------------------
    def _merge(self):
        opts = self.opts
        groups = sorted({tuple(int(v) for v in g) for o in opts for g in o.uniq_l_ctr})
        ng = len(groups)
        atm, bas, env = [], [], []
        shell_off, ao_off = [], []
        nsh = nat = nenv = nao = 0
        for o in opts:
            a = o.atm.copy()
            b = o.bas.copy()
            a[:, 1] += nenv           # PTR_COORD
            a[:, 3] += nenv           # PTR_ZETA
            b[:, ATOM_OF] += nat
            b[:, PTR_EXP] += nenv
            b[:, PTR_COEFF] += nenv
            atm.append(a)
            bas.append(b)
            env.append(o.env)
            shell_off.append(nsh)
            ao_off.append(nao)
            nsh += len(b)
            nat += len(a)
            nenv += len(o.env)
            nao += o.nao
        self.atm = np.ascontiguousarray(np.vstack(atm), dtype=np.int32)
        self.bas = np.ascontiguousarray(np.vstack(bas), dtype=np.int32)
        self.env = np.ascontiguousarray(np.concatenate(env))
        # class (p, q) of the merged group list, p >= q, in the same p-major order
        pi, pj, pq, locs = [], [], [], [0]
        n_mol = len(opts)
        mol_off = []
        for p in range(ng):
            for q in range(p + 1):
                offs = [0]
                for m, o in enumerate(opts):
                    mg = {tuple(int(v) for v in g): i for i, g in enumerate(o.uniq_l_ctr)}
                    if groups[p] in mg and groups[q] in mg:
                        gp, gq = mg[groups[p]], mg[groups[q]]
                        c = gp * (gp + 1) // 2 + gq
                        s0, s1 = o.bas_pairs_locs[c], o.bas_pairs_locs[c + 1]
                        pi.append(o.bas_pair2shls[0, s0:s1] + shell_off[m])
                        pj.append(o.bas_pair2shls[1, s0:s1] + shell_off[m])
                        pq.append(o.pair_q[s0:s1])
                        offs.append(offs[-1] + int(s1 - s0))
                    else:
                        offs.append(offs[-1])
                locs.append(locs[-1] + offs[-1])
                mol_off.append(offs)
        self.groups = groups
        self.bas_pair2shls = np.ascontiguousarray(np.vstack([np.concatenate(pi), np.concatenate(pj)]).astype(np.int32))
        self.bas_pairs_locs = np.asarray(locs, dtype=np.int32)
        self.pair_q = np.ascontiguousarray(np.concatenate(pq))
        self.ncptype = len(locs) - 1
        self.mol_pair_off = np.ascontiguousarray(np.asarray(mol_off, dtype=np.int32))
        self.nao_cart = np.asarray([o.nao for o in opts], dtype=np.int32)
        self.ao_base = np.asarray(ao_off, dtype=np.int32)
        self.dm_off = np.append(0, np.cumsum(self.nao_cart.astype(np.int64) ** 2))[:-1].astype(np.int64)
        self.total = int(np.sum(self.nao_cart.astype(np.int64) ** 2))
        lib = mjk.load()
        bp = ctypes.c_void_p()
        mjk.check(lib.mjk_init_basis_prod(
            ctypes.byref(bp), 0.5, mjk.ptr(self.bas_pair2shls), mjk.ptr(self.bas_pairs_locs), self.ncptype,
            mjk.ptr(self.pair_q), mjk.ptr(self.atm), len(self.atm), mjk.ptr(self.bas), len(self.bas),
            mjk.ptr(self.env)), 'mjk_init_basis_prod (batch)')
        batch = ctypes.c_void_p()
        rc = lib.mjk_batch_create(ctypes.byref(batch), ctypes.byref(bp), n_mol, mjk.ptr(self.mol_pair_off),
                                  mjk.ptr(self.dm_off), mjk.ptr(self.nao_cart), mjk.ptr(self.ao_base))
        if rc != mjk.MJK_OK:
            if bp.value:
                lib.mjk_del_basis_prod(ctypes.byref(bp))
            mjk.check(rc, 'mjk_batch_create')
        self.batch = batch
        self.ncp = int(lib.mjk_batch_num_class_pairs(batch))

    def close(self):
        if getattr(self, 'batch', None) is not None and self.batch.value:
            mjk.load().mjk_batch_destroy(ctypes.byref(self.batch))
        self.batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- densities -----------------------------------------------------------------------
    def to_cart(self, dms):
        """List of spherical densities -> one flat Cartesian device array."""
        torch = _torch()
        flat = torch.empty(self.total, dtype=torch.float64, device='cuda')
        for m, (o, d) in enumerate(zip(self.opts, dms)):
            c, ct = o.coeff_dev()
            dd = torch.as_tensor(d, dtype=torch.float64, device='cuda')
            n = int(self.nao_cart[m])
            flat[self.dm_off[m]:self.dm_off[m] + n * n] = (c @ dd @ ct).reshape(-1)
        return flat

    def from_cart(self, flat):
        out = []
        for m, o in enumerate(self.opts):
            c, ct = o.coeff_dev()
            n = int(self.nao_cart[m])
            out.append(ct @ flat[self.dm_off[m]:self.dm_off[m] + n * n].reshape(n, n) @ c)
        return out

    # ---- builds --------------------------------------------------------------------------
    def screen(self, precision_cutoff=0.0):
        return mjk.Screen(self.cutoff, float(precision_cutoff), 0)

    def build_cart(self, dm_flat, vj=None, vk=None, with_j=True, with_k=True, precision_cutoff=0.0,
                   use_graph=None):
        """J/K (flat Cartesian device arrays) of a flat Cartesian density, on self.stream."""
        torch = _torch()
        use_graph = self.use_graph if use_graph is None else use_graph
        if with_j and vj is None:
            vj = torch.empty(self.total, dtype=torch.float64, device='cuda')
        if with_k and vk is None:
            vk = torch.empty(self.total, dtype=torch.float64, device='cuda')
        sc = self.screen(precision_cutoff)
        cur = torch.cuda.current_stream()
        self.stream.wait_stream(cur)
        mjk.check(mjk.load().mjk_batch_build(self.batch, mjk.ptr(vj), mjk.ptr(vk), mjk.ptr(dm_flat),
                                             ctypes.byref(sc), int(bool(use_graph)),
                                             ctypes.c_void_p(self.stream.cuda_stream)), 'mjk_batch_build')
        cur.wait_stream(self.stream)
        return vj, vk

    def get_jk(self, dms, precision_cutoff=0.0):
        """Lists of (vj, vk) in each molecule's spherical basis (device tensors)."""
        flat = self.to_cart(dms)
        vj, vk = self.build_cart(flat, precision_cutoff=precision_cutoff, use_graph=False)
        return self.from_cart(vj), self.from_cart(vk)

    def stats(self):
        out = np.zeros(4)
        mjk.check(mjk.load().mjk_batch_stats(self.batch, mjk.ptr(out)), 'mjk_batch_stats')
        return {'quartets_fp64': int(out[0]), 'quartets_fp32': int(out[1]), 'slots': int(out[3])}

    # ---- autotuner -----------------------------------------------------------------------
    def tiles(self):
        out = np.zeros((self.ncp, 10), dtype=np.int32)
        mjk.check(mjk.load().mjk_batch_tiles(self.batch, mjk.ptr(out), self.ncp), 'mjk_batch_tiles')
        return out

    def class_pair_keys(self):
        """A label per class pair: angular momenta and primitive counts of both classes."""
        t = self.tiles()
        keys = []
        for row in t:
            keys.append(_label(self._class_groups(int(row[0])), self._class_groups(int(row[1]))))
        assert len(set(keys)) == len(keys)
        return keys

    def _class_groups(self, c):
        p = int((np.sqrt(8 * c + 1) - 1) // 2)
        q = c - p * (p + 1) // 2
        return self.groups[p], self.groups[q]

    def set_kc(self, kc):
        kc = np.ascontiguousarray(kc, dtype=np.int32)
        mjk.check(mjk.load().mjk_batch_set_kc(self.batch, mjk.ptr(kc), self.ncp), 'mjk_batch_set_kc')

    def tune(self, dm_flat, reps=3, precision_cutoff=0.0, log=None):
        """Time every candidate ket chunk of every class pair on this batch's own work and
        keep the fastest.  Returns {class-pair label: (kc, ms)}."""
        torch = _torch()
        lib = mjk.load()
        vj = torch.zeros(self.total, dtype=torch.float64, device='cuda')
        vk = torch.zeros(self.total, dtype=torch.float64, device='cuda')
        sc = self.screen(precision_cutoff)
        t = self.tiles()
        kc = t[:, 6].copy()
        keys = self.class_pair_keys()
        result = {}
        torch.cuda.synchronize()
        for i in range(self.ncp):
            best = None
            for cand in KC_CANDIDATES[0 if t[i, 8] == 0 else 1]:
                trial = kc.copy()
                trial[i] = cand
                self.set_kc(trial)
                ms = ctypes.c_float()
                mjk.check(lib.mjk_batch_time_class_pair(self.batch, i, mjk.ptr(vj), mjk.ptr(vk), mjk.ptr(dm_flat),
                                                        ctypes.byref(sc), reps, ctypes.c_void_p(self.stream.cuda_stream),
                                                        ctypes.byref(ms)), 'mjk_batch_time_class_pair')
                if best is None or ms.value < best[1]:
                    best = (cand, ms.value)
            kc[i] = best[0]
            result[keys[i]] = best
            if log:
                log(f'[tune] {keys[i]}: kc {best[0]}  {best[1]:.3f} ms')
        self.set_kc(kc)
        self.tuned = result
        return result

    def save_tuning(self, path):
        keys = self.class_pair_keys()
        table = {k: int(v) for k, v in zip(keys, self.tiles()[:, 6])}
        with open(path, 'w') as f:
            json.dump(table, f, indent=1, sort_keys=True)

    def load_tuning(self, path):
        if not os.path.exists(path):
            return False
        with open(path) as f:
            table = json.load(f)
        t = self.tiles()
        kc = t[:, 6].copy()
        for i, k in enumerate(self.class_pair_keys()):
            if k in table:
                kc[i] = int(table[k])
        self.set_kc(kc)
        return True


def batch_loop_jk(opts, dm_list_cart, precision_cutoff=0.0):
    """The per-molecule baseline: one mjk_build_jk_all per molecule on the current stream."""
    out = []
    for o, d in zip(opts, dm_list_cart):
        out.append(hf.jk_device(o, d, True, True, precision_cutoff, density_screen=False))
    return out
