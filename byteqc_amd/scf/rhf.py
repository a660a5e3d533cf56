"""Restricted Hartree-Fock driver around the MI355X J/K builder.

The reference's `RHF` (cuobc/scf/hf.py:272-274) is PySCF's RHF with two methods
swapped: `get_jk` (J/K on the GPU) and `_eigh` (GPU generalised eigensolver,
hf.py:253-269).  PySCF is not available here, so this module supplies the loop itself:
core Hamiltonian from the library's one-electron integrals, SAD or user guess, Pulay
DIIS, canonical orthogonalisation and eigh on the GPU, and an optional incremental
Fock build (J/K of the density change, as PySCF's direct SCF does).  Any object with
a `get_jk(mol, dm, hermi=1, ...)` method can be plugged in: that is the drop-in contract.
"""
from __future__ import annotations

import time

import numpy as np

from .. import lib as mjk


def _torch():
    import torch
    return torch


def int1e(vhfopt, nthreads=0):
    """(S, T, V) in the caller's AO basis (spherical unless mol.cart)."""
    lib = mjk.load()
    n = int(vhfopt.coeff.shape[0])
    s = np.zeros((n, n))
    t = np.zeros((n, n))
    v = np.zeros((n, n))
    mjk.check(lib.mjk_int1e(mjk.ptr(vhfopt.atm), len(vhfopt.atm), mjk.ptr(vhfopt.bas),
                            len(vhfopt.bas), mjk.ptr(vhfopt.env), mjk.ptr(s), mjk.ptr(t),
                            mjk.ptr(v), nthreads), 'mjk_int1e')
    c = vhfopt.coeff
    return c.T @ s @ c, c.T @ t @ c, c.T @ v @ c


def orthogonalizer(s, threshold=1e-15):
    """Canonical orthogonaliser t (t^T s t = 1) of a GPU overlap matrix; eigenvectors of s
    with eigenvalue below `threshold` are dropped."""
    torch = _torch()
    seig, t = torch.linalg.eigh(s)
    keep = seig >= threshold
    return t[:, keep] / torch.sqrt(seig[keep])


def eig(h, s, threshold=1e-15, t=None):
    """Generalised eigenproblem h c = s c e by canonical orthogonalisation (hf.py:253-265).

    h, s: torch tensors on the GPU.  t: s's orthogonaliser when the caller keeps it (an SCF
    solves with one overlap matrix every cycle), else computed here."""
    torch = _torch()
    if t is None:
        t = orthogonalizer(s, threshold)
    heff = t.T @ h @ t
    e, c = torch.linalg.eigh(heff)
    return e, t @ c


def purify(f, nocc, max_iter=120, tol=1e-11):
    """Projector onto the nocc lowest eigenvectors of a symmetric matrix f (orthonormal basis)
    by second-order spectral projection (SP2: Niklasson, Phys. Rev. B 66, 155115 (2002)).

    x starts as (e_max - f) / (e_max - e_min) with Gershgorin bounds, so its spectrum lies in
    [0, 1] with the occupied states on top; each step applies x^2 or 2x - x^2, whichever moves
    the trace toward nocc, which pushes every eigenvalue to 0 or 1 while keeping the order.
    One GEMM per step.  Returns None when the idempotency error tr(x - x^2) has not fallen
    below tol (no gap between states nocc and nocc + 1, or too many steps).
    Works on the tensor's device (CPU or GPU)."""
    import torch
    n = f.shape[0]
    d = torch.diagonal(f)
    r = torch.sum(torch.abs(f), dim=1) - torch.abs(d)
    lo = float(torch.min(d - r))
    hi = float(torch.max(d + r))
    if not hi > lo:
        return None
    x = -f / (hi - lo)
    x.diagonal().add_(hi / (hi - lo))
    for _ in range(max_iter):
        x2 = x @ x
        tx, tx2 = torch.stack((torch.trace(x), torch.trace(x2))).tolist()
        idem = tx - tx2
        if abs(idem) < tol and abs(tx - nocc) < 1e-6:
            return x
        if abs(tx2 - nocc) < abs(2.0 * tx - tx2 - nocc):
            x = x2
        else:
            x = 2.0 * x - x2
    return None


class DIIS:
    """Pulay DIIS on the Fock matrix with the orthogonalised commutator as error."""

    def __init__(self, space=8):
        self.space = space
        self.focks, self.errs = [], []
        self.gram = None  # error overlaps <e_i, e_j>, one new row per update

    def update(self, f, err):
        torch = _torch()
        self.focks.append(f)
        self.errs.append(err)
        if len(self.focks) > self.space:
            self.focks.pop(0)
            self.errs.pop(0)
            self.gram = self.gram[1:, 1:]
        n = len(self.focks)
        row = torch.stack([torch.sum(x * err) for x in self.errs])
        g = torch.empty((n, n), dtype=torch.float64, device=f.device)
        if n > 1:
            g[:n - 1, :n - 1] = self.gram
        g[n - 1, :] = row
        g[:, n - 1] = row
        self.gram = g
        if n < 2:
            return f
        b = torch.empty((n + 1, n + 1), dtype=torch.float64, device=f.device)
        b[-1, :] = -1
        b[:, -1] = -1
        b[-1, -1] = 0
        b[:n, :n] = g
        rhs = torch.zeros(n + 1, dtype=torch.float64, device=f.device)
        rhs[-1] = -1
        try:
            c = torch.linalg.solve(b, rhs)
        except RuntimeError:
            c = torch.linalg.lstsq(b, rhs[:, None]).solution[:, 0]
        return sum(c[i] * self.focks[i] for i in range(n))


# ---- superposition of atomic densities ----------------------------------------------
_ATOM_CACHE = {}


def _occupations(e, nelec, tol=1e-3):
    """Aufbau occupations (per spatial orbital, up to 2) with degenerate levels sharing."""
    occ = np.zeros(len(e))
    left = float(nelec)
    i = 0
    while left > 1e-12 and i < len(e):
        j = i + 1
        while j < len(e) and abs(e[j] - e[i]) < tol:
            j += 1
        g = j - i
        take = min(left, 2.0 * g)
        occ[i:j] = take / g
        left -= take
        i = j
    return occ


def atomic_density(symbol, basis, get_jk=None, max_cycle=60):
    """Spherically averaged fractional-occupation atomic HF density (spherical AOs)."""
    key = (symbol, str(basis))
    if key in _ATOM_CACHE:
        return _ATOM_CACHE[key]
    from ..gto import Mole
    from . import hf
    atom = Mole([(symbol, (0., 0., 0.))], basis=basis)
    opt = hf._VHFOpt(atom).build()
    s, t, v = int1e(opt)
    h = t + v
    torch = _torch()
    hd = torch.as_tensor(h, device='cuda')
    sd = torch.as_tensor(s, device='cuda')
    nelec = atom.nelectron
    e, c = eig(hd, sd)
    occ = _occupations(e.cpu().numpy(), nelec)
    dm = (c * torch.as_tensor(occ, device='cuda')) @ c.T
    diis = DIIS(6)
    e_old = 0.0
    for it in range(max_cycle):
        vj, vk = hf.get_jk(atom, dm, vhfopt=opt, out_device=True)
        f = hd + vj - 0.5 * vk
        e_tot = float(0.5 * torch.sum(dm * (hd + f)))
        err = f @ dm @ sd - sd @ dm @ f
        f = diis.update(f, err)
        e, c = eig(f, sd)
        occ = _occupations(e.cpu().numpy(), nelec)
        dm_new = (c * torch.as_tensor(occ, device='cuda')) @ c.T
        if abs(e_tot - e_old) < 1e-9 and float(torch.abs(dm_new - dm).max()) < 1e-7:
            dm = dm_new
            break
        dm, e_old = dm_new, e_tot
    opt.clear()
    out = dm.cpu().numpy()
    _ATOM_CACHE[key] = out
    return out


def init_guess_by_atom(mol):
    """Block-diagonal superposition of atomic densities in mol's AO order."""
    n = mol.nao
    dm = np.zeros((n, n))
    aoslice = mol.aoslice_by_atom()
    for ia in range(mol.natm):
        a0, a1 = aoslice[ia, 2], aoslice[ia, 3]
        dm[a0:a1, a0:a1] = atomic_density(mol.atom_symbol(ia), mol.basis)
    return dm


class RHF:
    """Closed-shell SCF.  kernel() returns the converged total energy."""

    def __init__(self, mol, conv_tol=1e-8, conv_tol_grad=None, max_cycle=100, diis_space=8,
                 direct_scf_tol=1e-13, precision_cutoff=0.0, incremental=True,
                 rebuild_every=None, verbose=0, jk=None, early_tol=None, early_until=1e-3,
                 diagnose_jk=False, purification=True):
        self.mol = mol
        self.conv_tol = conv_tol
        self.conv_tol_grad = conv_tol_grad if conv_tol_grad is not None else np.sqrt(conv_tol) * 0.1
        self.max_cycle = max_cycle
        self.diis_space = diis_space
        self.direct_scf_tol = direct_scf_tol
        self.precision_cutoff = precision_cutoff
        self.incremental = incremental
        # incremental Fock builds on dm - dm_last; rebuild_every=n forces a full build after n
        # incremental ones (None: never, as PySCF's direct SCF, which the reference's RHF
        # inherits: cuobc/scf/hf.py:272-274)
        self.rebuild_every = np.inf if rebuild_every is None else rebuild_every
        self.verbose = verbose
        self.jk = jk
        # looser quartet screening while the orbital gradient max|[F, D]| exceeds early_until;
        # the switch to direct_scf_tol forces a full rebuild, so the converged energy is
        # computed at the tight tolerance
        self.early_tol = early_tol
        self.early_until = early_until
        self._tol_now = None
        # diagnose_jk: also build plain FP64 J/K (full, tight screening) of every cycle's density,
        # outside the timings, and record max|dJ|, max|dK| and the energy it implies
        self.diagnose_jk = diagnose_jk
        # the next density from the DIIS Fock matrix by SP2 purification (GEMMs) instead of
        # an eigendecomposition; the eigensolver remains the fallback and gives the final
        # orbitals
        self.purification = purification
        # path of a checkpoint (PySCF's chkfile keys in .npz) written after kernel()
        self.chkfile = None
        self._opt_gpu = None
        self.converged = False
        self.e_tot = None
        self.cycles = 0
        self.history = []
        self.timings = {}

    # ---- pieces --------------------------------------------------------------------
    def build(self):
        from . import hf
        t0 = time.perf_counter()
        if self._opt_gpu is None:
            self._opt_gpu = hf._VHFOpt(self.mol).build(self.direct_scf_tol)
        t1 = time.perf_counter()
        s, t, v = int1e(self._opt_gpu)
        self._s, self._h = s, t + v
        self.timings.update(vhfopt=t1 - t0, int1e=time.perf_counter() - t1)
        return self

    def get_ovlp(self):
        return self._s

    def get_hcore(self):
        return self._h

    def get_init_guess(self, key='sad'):
        if key in ('sad', 'atom', 'minao'):
            return init_guess_by_atom(self.mol)
        if key == 'fragments':
            from ..guess.fragments import fragment_density
            return fragment_density(self.mol)
        if key == 'learned':
            from ..guess.learned import learned_guess
            return learned_guess(self.mol, self._s)
        if key.startswith('learned:'):
            # the model of another basis (e.g. 'learned:cc-pvdz'), projected into this one
            from ..guess.learned import learned_guess
            return learned_guess(self.mol, self._s, via=key.split(':', 1)[1])
        if key in ('1e', 'hcore'):
            torch = _torch()
            e, c = eig(torch.as_tensor(self._h, device='cuda'), torch.as_tensor(self._s, device='cuda'))
            nocc = self.mol.nelectron // 2
            co = c[:, :nocc]
            return (2.0 * co @ co.T).cpu().numpy()
        raise ValueError(key)

    def get_jk(self, mol, dm, hermi=1, with_j=True, with_k=True, omega=None):
        """J/K of a device density; the hook a replacement builder overrides."""
        if self.jk is not None:
            return self.jk(mol, dm, hermi=hermi, with_j=with_j, with_k=with_k)
        from . import hf
        return hf.get_jk(mol, dm, hermi, self._opt_gpu, with_j, with_k, omega,
                         precision_cutoff=self.precision_cutoff, out_device=True,
                         direct_scf_tol=self._tol_now)

    def make_rdm1(self, mo_coeff=None, mo_occ=None):
        torch = _torch()
        c = self.mo_coeff if mo_coeff is None else mo_coeff
        nocc = self.mol.nelectron // 2
        co = c[:, :nocc]
        return 2.0 * co @ co.T

    def energy_nuc(self):
        return self.mol.energy_nuc()

    # ---- the loop ------------------------------------------------------------------
    def kernel(self, dm0=None, init_guess='sad'):
        torch = _torch()
        if self._opt_gpu is None or not hasattr(self, '_h'):
            self.build()
        mol = self.mol
        t_start = time.perf_counter()
        if dm0 is None:
            dm0 = self.get_init_guess(init_guess)
        t_guess = time.perf_counter()
        h = torch.as_tensor(self._h, device='cuda')
        s = torch.as_tensor(self._s, device='cuda')
        dm = torch.as_tensor(np.asarray(dm0), dtype=torch.float64, device='cuda')
        enuc = self.energy_nuc()
        nocc = mol.nelectron // 2
        diis = DIIS(self.diis_space)
        orth = orthogonalizer(s)  # s is fixed through the SCF: one diagonalisation of it
        vj = vk = None
        dm_last = None
        e_last = None
        self.history = []
        self.converged = False
        since_full = 0
        force_full = False
        t_jk = 0.0
        t_diag = 0.0
        loose = self.early_tol is not None and self.early_tol > self.direct_scf_tol
        self._tol_now = self.early_tol if loose else None
        for cycle in range(self.max_cycle):
            tj = time.perf_counter()
            if loose and self.history and self.history[-1]['grad'] < self.early_until:
                loose = False
                self._tol_now = None
                force_full = True  # full rebuild at the tight tolerance
            if self.incremental and dm_last is not None and since_full < self.rebuild_every and not force_full:
                dj, dk = self.get_jk(mol, dm - dm_last)
                vj = vj + dj
                vk = vk + dk
                since_full += 1
            else:
                vj, vk = self.get_jk(mol, dm)
                since_full = 0
                force_full = False
            torch.cuda.synchronize()
            dt_jk = time.perf_counter() - tj
            t_jk += dt_jk
            f = h + vj - 0.5 * vk
            e_tot = float(0.5 * torch.sum(dm * (h + f))) + enuc
            diag = {}
            if self.diagnose_jk:
                from . import hf
                td = time.perf_counter()
                if self.jk is None:
                    # work of the timed build: computed quartets and the wave slots they occupied
                    st = self._opt_gpu.stats()
                    diag = {'quartets': st['quartets_fp64'] + st['quartets_fp32'],
                            'quartets_fp32': st['quartets_fp32'], 'slots': st['slots']}
                rj, rk = hf.get_jk(mol, dm, 1, self._opt_gpu, out_device=True)
                e_ref = float(0.5 * torch.sum(dm * (2 * h + rj - 0.5 * rk))) + enuc
                diag.update(max_dj=float((vj - rj).abs().max()), max_dk=float((vk - rk).abs().max()),
                            de_vs_fp64_fock=e_tot - e_ref)
                t_diag += time.perf_counter() - td  # kept out of the timings
            err = f @ dm @ s - s @ dm @ f
            gnorm = float(torch.abs(err).max())
            de = None if e_last is None else e_tot - e_last
            self.history.append({'cycle': cycle, 'e_tot': e_tot, 'delta_e': de, 'grad': gnorm,
                                 't': time.perf_counter() - t_start - t_diag, 't_jk': dt_jk,
                                 'tol': self._tol_now or self.direct_scf_tol, **diag})
            if self.verbose:
                print(f'cycle {cycle}: E = {e_tot:.12f}  dE = {de}  |[F,D]| = {gnorm:.2e}', flush=True)
            if de is not None and abs(de) < self.conv_tol and gnorm < self.conv_tol_grad and not loose:
                self.converged = True
                e_last = e_tot
                break
            fd = diis.update(f, err)
            x = purify(orth.T @ fd @ orth, nocc) if self.purification else None
            dm_last = dm
            if x is None:
                mo_e, mo_c = eig(fd, s, t=orth)
                co = mo_c[:, :nocc]
                dm = 2.0 * co @ co.T
            else:
                dm = 2.0 * orth @ x @ orth.T
            e_last = e_tot
        torch.cuda.synchronize()
        t_end = time.perf_counter()
        self.e_tot = e_last
        self.cycles = len(self.history)
        self.mo_energy, self.mo_coeff = eig(f, s, t=orth)
        self.mo_occ = np.zeros(self.mo_energy.shape[0])
        self.mo_occ[:nocc] = 2.0
        self.dm = dm
        self.timings.update(guess=t_guess - t_start, scf=t_end - t_guess - t_diag, jk=t_jk,
                            total=t_end - t_start - t_diag)
        if self.chkfile:
            from . import chkfile
            chkfile.dump_scf(self.chkfile, mol, self.e_tot, self.mo_energy, self.mo_coeff, self.mo_occ)
        return self.e_tot
