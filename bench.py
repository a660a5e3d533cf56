#!/usr/bin/env python
"""Benchmark: direct-SCF J/K build on one MI355X.

Workload (default): 64 seeded water molecules in a 16 Angstrom box, def2-TZVP (2752
spherical / 3072 Cartesian AOs), plain FP64, Schwarz cutoff 1e-13.  The density follows
the reference's J/K harness (cuobc/test/jk.py:7815-7825): numpy.random.seed(2341),
dm = rand(nao, nao) + its transpose, prescreen=False (no density-weighted screening);
--density sad uses the SAD guess with density screening instead.  One step = one full
J/K build (mjk_build_jk_all) with the density already in HBM; `value` is seconds per
build (lower is better).

Also reported:
  roofline      the most expensive class kernel, timed per launch with HIP events on the
                stream it runs on; achieved = its algorithmic FLOPs (byteqc_amd/perf/flops.py
                x quartets it computed) / launch time, against the 78.6 TFLOP/s FP64 peak
  cpu_baseline  the CPU oracle's Rys J/K (C, OpenMP, all host cores) on the same shells,
                screening and density, timed on a systematic sample of bra-pair rows and
                scaled to the full build
  scf           (--scf) RHF time-to-1e-8 Eh: FP64 + SAD guess (the reference's path) against
                FP64 + learned guess, mixed (FP32 quartets below a 1e-7 bound, 1e-9
                screening until max|[F,D]| < 1e-4) + SAD, and mixed + learned
  batch         (--batch N) N seeded water clusters of 1-32 molecules, 6-31G*, random PSD
                densities (D = 2 Q Q^T, trace = electron count): all N J/K builds as a
                per-molecule loop of mjk_build_jk_all, as one batched build, as a replayed
                hipGraph, and as a replayed graph after the ket-chunk autotuner
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def run_batch(n_mol, reps, torch):
    """Seconds for all n_mol J/K builds: per-molecule loop vs batched launches."""
    from byteqc_amd.gto.clusters import water_cluster
    from byteqc_amd.guess.train import box_for
    from byteqc_amd.scf import hf
    from byteqc_amd.scf.batch import BatchJK
    rng = np.random.default_rng(2341)
    sizes = rng.integers(1, 33, n_mol)
    mols = [water_cluster(int(n), box_for(int(n)), 3000 + i, basis='6-31gs') for i, n in enumerate(sizes)]
    dms = []
    for m in mols:
        q, _ = np.linalg.qr(rng.normal(size=(m.nao, m.nelectron // 2)))
        dms.append(2.0 * q @ q.T)
    t0 = time.perf_counter()
    b = BatchJK(mols, use_graph=True)
    setup = time.perf_counter() - t0
    flat = b.to_cart(dms)
    log(f'batch: {n_mol} clusters, {int(sizes.sum())} waters, nao_cart total {int(b.nao_cart.sum())}, '
        f'{b.ncp} class pairs, setup {setup:.1f} s')

    def timed(fn, n):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / n

    out = {'n_mol': n_mol, 'waters': int(sizes.sum()), 'basis': '6-31G*', 'setup_s': setup}
    vj = torch.empty_like(flat)
    vk = torch.empty_like(flat)
    out['batched_s'] = timed(lambda: b.build_cart(flat, vj, vk, use_graph=False), reps)
    out['graph_s'] = timed(lambda: b.build_cart(flat, vj, vk, use_graph=True), reps)
    t0 = time.perf_counter()
    table = b.tune(flat, reps=2)
    out['tune_s'] = time.perf_counter() - t0
    out['graph_tuned_s'] = timed(lambda: b.build_cart(flat, vj, vk, use_graph=True), reps)
    out['tuned_kc'] = {k: v[0] for k, v in table.items()}
    st = b.stats()
    out['quartets'] = st['quartets_fp64'] + st['quartets_fp32']
    ref_j, ref_k = vj.clone(), vk.clone()
    b.close()
    opts = [hf._VHFOpt(m).build(1e-13) for m in mols]
    dmc = [flat[b.dm_off[i]:b.dm_off[i] + int(b.nao_cart[i]) ** 2].reshape(1, int(b.nao_cart[i]), -1)
           for i in range(n_mol)]
    res = []

    def loop():
        res.clear()
        for o, d in zip(opts, dmc):
            res.append(hf.jk_device(o, d, True, True, 0.0, density_screen=False))

    out['loop_s'] = timed(loop, reps)
    err = 0.0
    for i, (j, k) in enumerate(res):
        n = int(b.nao_cart[i])
        sl = slice(int(b.dm_off[i]), int(b.dm_off[i]) + n * n)
        err = max(err, float((j.reshape(-1) - ref_j[sl]).abs().max()), float((k.reshape(-1) - ref_k[sl]).abs().max()))
    out['max_abs_diff_vs_loop'] = err
    out['speedup_graph_tuned_vs_loop'] = out['loop_s'] / out['graph_tuned_s']
    for o in opts:
        o.clear()
    log(f'batch: loop {out["loop_s"]:.3f} s, batched {out["batched_s"]:.3f} s, graph {out["graph_s"]:.3f} s, '
        f'graph+tuned {out["graph_tuned_s"]:.3f} s (tuning {out["tune_s"]:.1f} s), max |diff| {err:.1e}')
    return out


def measured_traffic(workload, symbol):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy) of a kernel
    symbol, from the rocprofv3 --pmc runs recorded for this workload; None if not recorded."""
    path = os.path.join(ROOT, 'byteqc_amd', 'perf', f'traffic_{workload.replace("-", "_")}.json')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)['kernels'].get(symbol)
    return None if rec is None else rec['fetch_bytes_per_launch'] + rec['write_bytes_per_launch']


def per_class_profile(vhfopt, dmc, vj, vk, screen, torch, lib, mjk):
    """Time every class pair separately (HIP events on the launch stream)."""
    from byteqc_amd.perf.flops import class_flops, class_label
    stream = torch.cuda.current_stream()
    locs = vhfopt.bas_pairs_locs
    bas = vhfopt.bas
    p2s = vhfopt.bas_pair2shls
    info = []
    for c in range(vhfopt.ncptype):
        if locs[c + 1] > locs[c]:
            i, j = p2s[0, locs[c]], p2s[1, locs[c]]
            info.append((int(bas[i, 1]), int(bas[j, 1]), int(bas[i, 2] * bas[j, 2])))
        else:
            info.append(None)
    sc2 = mjk.Screen(screen.direct_scf_tol, screen.precision_cutoff, 2 if screen.density_screen else 0)
    # block maxima for density_screen == 2
    mjk.check(lib.mjk_build_jk_all(vhfopt.bp, mjk.ptr(vj), mjk.ptr(vk), mjk.ptr(dmc), vhfopt.nao, 1,
                                   ctypes.byref(screen), ctypes.c_void_p(stream.cuda_stream)), 'warm')
    torch.cuda.synchronize()
    out = []
    prev = np.zeros(4)
    mjk.check(lib.mjk_last_stats(vhfopt.bp, mjk.ptr(prev)), 'stats')
    for a in range(vhfopt.ncptype):
        for b in range(a + 1):
            if info[a] is None or info[b] is None:
                continue
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            mjk.check(lib.mjk_build_jk(vhfopt.bp, mjk.ptr(vj), mjk.ptr(vk), mjk.ptr(dmc), vhfopt.nao, 1,
                                       a, b, ctypes.byref(sc2), ctypes.c_void_p(stream.cuda_stream)),
                      'mjk_build_jk')
            e1.record(stream)
            e1.synchronize()
            cur = np.zeros(4)
            mjk.check(lib.mjk_last_stats(vhfopt.bp, mjk.ptr(cur)), 'stats')
            nq = cur[0] - prev[0] + cur[1] - prev[1]
            nslots = cur[3] - prev[3]
            prev = cur
            la, lb, pa = info[a]
            lc, ld, pb = info[b]
            fl = nq * class_flops(la, lb, lc, ld, pa * pb)
            out.append({'class': class_label(la, lb, lc, ld), 'cp': (a, b), 'nprim': pa * pb,
                        'quartets': int(nq), 'slots': int(nslots), 'ms': e0.elapsed_time(e1),
                        'flops': fl})
    return out


def cpu_baseline(vhfopt, dm_cart_host, target_s):
    import oracle
    sb = oracle.SegBasis(vhfopt.atm, vhfopt.bas, vhfopt.env, vhfopt.coeff)
    nb = sb.nbas
    npair = nb * (nb + 1) // 2
    cores = os.cpu_count() or 1
    stride = max(1, npair // 400)
    times = []
    t_all = 0.0
    offset = 0
    while t_all < target_s and offset < stride:
        t0 = time.perf_counter()
        oracle.rys_jk_cart(sb, dm_cart_host, cutoff=vhfopt.direct_scf_tol, nthreads=cores,
                           q=vhfopt.q_cond, stride=stride, offset=offset)
        dt = time.perf_counter() - t0
        times.append(dt)
        t_all += dt
        offset += max(1, stride // 7)
    est = float(np.mean(times)) * stride
    return {'value': est, 'unit': 's/iter', 'cores': cores, 'kind': 'port',
            'sample': f'{len(times)} systematic samples of bra-pair rows (1 in {stride} of {npair} '
                      f'shell pairs each, {t_all:.1f} s measured), scaled by {stride}'}


def run_scf(mol, precision_cutoff=0.0, guess='sad', early_tol=None, early_until=1e-3):
    """RHF to |dE| < 1e-8 Eh and max|[F,D]| < 1e-5; wall time from the guess to convergence
    (guess included; the SAD atomic densities and the learned guess's isolated-fragment density
    are cached beforehand, as for a warm code)."""
    from byteqc_amd.scf.rhf import RHF, init_guess_by_atom
    import torch
    init_guess_by_atom(mol)  # warm the atomic-density cache
    a = torch.rand(mol.nao, mol.nao, dtype=torch.float64, device='cuda')
    torch.linalg.eigh(a + a.T)  # the eigensolver's first call of a size sets up its workspace
    if guess.startswith('learned'):
        from byteqc_amd.guess.fragments import fragment_density
        fragment_density(mol)  # warm the isolated-fragment density cache, as for SAD's atoms
    mf = RHF(mol, conv_tol=1e-8, precision_cutoff=precision_cutoff, early_tol=early_tol, early_until=early_until)
    mf.build()
    e = mf.kernel(init_guess=guess)
    return {'e_tot': e, 'converged': mf.converged, 'cycles': mf.cycles, 'guess': guess,
            'wall_s': mf.timings['guess'] + mf.timings['scf'],
            'time_s': mf.timings['scf'], 'guess_s': mf.timings['guess'], 'jk_s': mf.timings['jk'],
            'setup_s': mf.timings['vhfopt'] + mf.timings['int1e'],
            'precision_cutoff': precision_cutoff, 'early_tol': early_tol, 'early_until': early_until,
            'cycle_jk_s': [round(h['t_jk'], 4) for h in mf.history],
            'cycle_e': [h['e_tot'] for h in mf.history],
            'cycle_grad': [h['grad'] for h in mf.history],
            'device_mem_used_gb': device_mem_used_gb()}


def device_mem_used_gb():
    """Device memory in use (all allocators: torch's cache and the library's own buffers)."""
    import torch
    free, total = torch.cuda.mem_get_info()
    return (total - free) / 1e9


def run_scf_diagnosed(mol, mixed=True):
    """An SCF again with plain FP64 J/K of every cycle's density built beside it (outside the
    timings): max|dJ|, max|dK|, the energy difference and the quartets each build computed.
    mixed: the mixed + learned configuration, else FP64 + SAD."""
    from byteqc_amd.guess.learned import model_path
    from byteqc_amd.scf.rhf import RHF
    if mixed:
        guess = 'learned' if os.path.exists(model_path(mol.basis)) else 'sad'
        mf = RHF(mol, conv_tol=1e-8, precision_cutoff=1e-7, early_tol=1e-9, early_until=1e-4,
                 diagnose_jk=True)
    else:
        guess = 'sad'
        mf = RHF(mol, conv_tol=1e-8, diagnose_jk=True)
    mf.build()
    mf.kernel(init_guess=guess)
    keys = ('cycle', 'tol', 't_jk', 'quartets', 'quartets_fp32', 'slots', 'max_dj', 'max_dk',
            'de_vs_fp64_fock', 'grad')
    return {'guess': guess, 'cycles': mf.cycles, 'e_tot': mf.e_tot,
            'per_cycle': [{k: h[k] for k in keys if k in h} for h in mf.history]}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--workload', default='h2o64-def2tzvp')
    ap.add_argument('--precision-cutoff', type=float, default=0.0)
    ap.add_argument('--density', choices=['random', 'sad'], default='random')
    ap.add_argument('--cpu-seconds', type=float, default=15.0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-classes', action='store_true')
    ap.add_argument('--scf', action='store_true')
    ap.add_argument('--scf-diagnose', action='store_true',
                    help='also rerun mixed + learned with per-cycle |dJ|, |dK|, dE against FP64 J/K')
    ap.add_argument('--scf-runs', default='fp64_sad,fp64_learned,mixed_sad,mixed_learned',
                    help='comma-separated subset of the SCF comparison (fp64_sad always runs)')
    ap.add_argument('--classes-out', default=None, help='write every class timing to this JSON file')
    ap.add_argument('--batch', type=int, default=0, help='also time N batched 6-31G* cluster builds')
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f'bench.py measures one GPU only (got --gpus {args.gpus}); the J/K build has no '
              f'multi-GPU path in this build', file=sys.stderr)
        return 2
    import torch
    if not torch.cuda.is_available():
        print('bench.py needs a GPU', file=sys.stderr)
        return 2
    from byteqc_amd import lib as mjk
    from byteqc_amd.perf.workloads import make_mol, WORKLOADS
    from byteqc_amd.scf import hf
    from byteqc_amd.scf.rhf import init_guess_by_atom
    lib = mjk.load()

    mol = make_mol(args.workload)
    log(f'workload {args.workload}: natm={mol.natm} nao={mol.nao}')
    t0 = time.perf_counter()
    vhfopt = hf._VHFOpt(mol).build(1e-13)
    log(f'_VHFOpt.build {time.perf_counter() - t0:.2f}s: nao_cart={vhfopt.nao} pairs={vhfopt.npairs} '
        f'classes={vhfopt.ncptype}')
    if args.density == 'random':
        # the reference harness: numpy.random.seed(2341); dm = rand(nao, nao); dm + dm.T,
        # get_jk(..., prescreen=False) (cuobc/test/jk.py:7815-7825)
        rs = np.random.RandomState(2341)
        dm = rs.rand(mol.nao, mol.nao)
        dm = dm + dm.T
    else:
        dm = init_guess_by_atom(mol)
    c, ct = vhfopt.coeff_dev()
    dmc = (c @ torch.as_tensor(dm, device='cuda') @ ct).contiguous()
    nao = vhfopt.nao
    vj = torch.empty((1, nao, nao), dtype=torch.float64, device='cuda')
    vk = torch.empty_like(vj)
    screen = mjk.Screen(1e-13, args.precision_cutoff, 0 if args.density == 'random' else 1)
    stream = torch.cuda.current_stream()

    def step():
        mjk.check(lib.mjk_build_jk_all(vhfopt.bp, mjk.ptr(vj), mjk.ptr(vk), mjk.ptr(dmc), nao, 1,
                                       ctypes.byref(screen), ctypes.c_void_p(stream.cuda_stream)),
                  'mjk_build_jk_all')

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log('warmup done')
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sec = (t1 - t0) / args.steps
    st = vhfopt.stats()
    log(f'J/K build: {sec * 1e3:.1f} ms/step, quartets fp64={st["quartets_fp64"]} fp32={st["quartets_fp32"]} '
        f'slots={st["slots"]}')

    result = {
        'metric': 'J/K build s/iter (plain FP64)',
        'value': sec, 'unit': 's/iter', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': sec * 1e3, 'higher_is_better': False, 'scaling': 'strong',
        'vs_baseline': None, 'dtype': 'f64' if args.precision_cutoff == 0 else 'f64+f32',
        'data': f'synthetic (seeded water cluster, {args.density} density)',
        'config': {'workload': args.workload, 'nao': mol.nao, 'nao_cart': nao,
                   'basis': WORKLOADS[args.workload][3], 'schwarz_cutoff': 1e-13,
                   'density': args.density, 'density_screen': bool(screen.density_screen),
                   'precision_cutoff': args.precision_cutoff,
                   'quartets': st['quartets_fp64'] + st['quartets_fp32']},
    }
    if not args.no_classes:
        prof = per_class_profile(vhfopt, dmc, vj, vk, screen, torch, lib, mjk)
        tot_ms = sum(p['ms'] for p in prof)
        tot_fl = sum(p['flops'] for p in prof)
        # one kernel symbol per angular-momentum class: aggregate its launches (one per
        # primitive-count combination) the way rocprofv3 --stats does
        sym = {}
        for p in prof:
            a = sym.setdefault(p['class'], {'ms': 0.0, 'flops': 0.0, 'launches': 0, 'quartets': 0})
            a['ms'] += p['ms']
            a['flops'] += p['flops']
            a['launches'] += 1
            a['quartets'] += p['quartets']
        dname, dom = max(sym.items(), key=lambda kv: kv[1]['ms'])
        ach = dom['flops'] / (dom['ms'] * 1e-3) / 1e12
        from byteqc_amd.perf.flops import kernel_symbol
        sym_name = kernel_symbol(dname)
        result['roofline'] = {
            'bound': 'mfma', 'achieved': ach, 'peak': FP64_PEAK_TFLOPS, 'unit': 'TFLOP/s',
            'frac': ach / FP64_PEAK_TFLOPS, 'traffic': measured_traffic(args.workload, sym_name),
            'kernel': sym_name,
            'launches': dom['launches'], 'avg_launch_ms': dom['ms'] / dom['launches'],
            'quartets': dom['quartets'], 'flops_per_build': dom['flops'],
            'note': 'FP64 vector-ALU roof (Rys recursion + digestion); MI355X FP64 vector and '
                    'matrix peaks are both 78.6 TFLOP/s.  achieved = algorithmic FLOPs of all '
                    'launches of this kernel symbol in one build / their summed HIP-event time, '
                    'measured in a pass that runs the class pairs one at a time on one stream; '
                    'traffic = HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE '
                    '(calibrated on a 1 GiB copy: x2.0 and x1.0), mostly FP64 atomics to J/K',
        }
        # the whole build (class pairs overlapped on four streams) against its wall time, and the
        # sum of the class pairs run one at a time
        result['build_tflops'] = tot_fl / sec / 1e12
        result['per_class_sum_ms'] = tot_ms
        if args.classes_out:
            with open(args.classes_out, 'w') as f:
                json.dump([{k: v for k, v in p.items() if k != 'cp'} for p in
                           sorted(prof, key=lambda p: -p['ms'])], f, indent=1)
        top = sorted(prof, key=lambda p: -p['ms'])[:8]
        result['top_classes'] = [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in p.items()
                                  if k != 'cp'} for p in top]
        log(f'per-class sum {tot_ms:.1f} ms; build {result["build_tflops"]:.2f} TFLOP/s; '
            f'dominant {dname} {dom["ms"]:.2f} ms over {dom["launches"]} launches, {ach:.2f} TFLOP/s')
    if not args.no_cpu:
        dmh = dmc.reshape(nao, nao).cpu().numpy()
        result['cpu_baseline'] = cpu_baseline(vhfopt, dmh, args.cpu_seconds)
        log(f'cpu baseline {result["cpu_baseline"]}')
    if args.scf:
        vhfopt.clear()
        torch.cuda.empty_cache()
        from byteqc_amd.guess.learned import model_path
        learned = 'learned' if os.path.exists(model_path(mol.basis)) else 'sad'
        runs = {'fp64_sad': dict(),
                'fp64_learned': dict(guess=learned),
                'mixed_sad': dict(precision_cutoff=1e-7, early_tol=1e-9, early_until=1e-4),
                'mixed_learned': dict(precision_cutoff=1e-7, early_tol=1e-9, early_until=1e-4, guess=learned)}
        wanted = set(args.scf_runs.split(',')) | {'fp64_sad'}
        runs = {k: v for k, v in runs.items() if k in wanted}
        r = result['scf'] = {}
        for name, kw in runs.items():
            r[name] = run_scf(mol, **kw)
            log(f'scf {name}: {r[name]["cycles"]} cycles, {r[name]["wall_s"]:.2f} s '
                f'(J/K {r[name]["jk_s"]:.2f} s), E = {r[name]["e_tot"]:.10f}')
        base = r['fp64_sad']
        for name in runs:
            r[name]['delta_e_vs_fp64_sad'] = r[name]['e_tot'] - base['e_tot']
            r[name]['speedup_vs_fp64_sad'] = base['wall_s'] / r[name]['wall_s']
        if 'mixed_learned' in r:
            log(f'scf mixed+learned vs fp64+SAD: {r["mixed_learned"]["speedup_vs_fp64_sad"]:.2f}x, '
                f'dE = {r["mixed_learned"]["delta_e_vs_fp64_sad"]:.1e}')
        if args.scf_diagnose:
            r['mixed_learned_per_cycle'] = run_scf_diagnosed(mol)
            r['fp64_sad_per_cycle'] = run_scf_diagnosed(mol, mixed=False)
            for name in ('fp64_sad_per_cycle', 'mixed_learned_per_cycle'):
                log(f'{name}: ' + ' '.join(f'{c["quartets"] / max(c["slots"], 1):.2f}'
                                           for c in r[name]['per_cycle']) + ' (lanes busy per wave slot)')
            worst = max(r['mixed_learned_per_cycle']['per_cycle'], key=lambda c: c['max_dk'])
            log(f'scf mixed+learned per cycle vs FP64 J/K: worst max|dJ| '
                f'{max(c["max_dj"] for c in r["mixed_learned_per_cycle"]["per_cycle"]):.1e}, '
                f'max|dK| {worst["max_dk"]:.1e}')
    if args.batch:
        vhfopt.clear()
        torch.cuda.empty_cache()
        result['batch'] = run_batch(args.batch, 3, torch)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
