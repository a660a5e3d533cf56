"""bench.py's contract on the host: one GPU only, and the algorithmic work model it reports."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'], cwd=ROOT,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert 'one GPU' in r.stderr and r.stdout == ''


def test_flops_model_closed_forms():
    """class_flops against the per-root term counts written out for two classes."""
    sys.path.insert(0, ROOT)
    from byteqc_amd.perf.flops import F_COEF, F_PRIM, F_ROOT, class_flops, kernel_symbol
    # (ss|ss): one root, no recursion terms, one Cartesian product; digestion 12 (J and K)
    assert class_flops(0, 0, 0, 0) == F_PRIM + (F_ROOT + F_COEF + 3) + 12
    # (ps|ss): one root; VRR 3 * 4 * ((1+1)(0+1) - 1) = 12; three products
    assert class_flops(1, 0, 0, 0) == F_PRIM + (F_ROOT + F_COEF + 12 + 9) + 36
    # (pp|ss): two roots; VRR 3*4*(3*1-1) = 24, HRR_j 3*2*(2*1) = 12, nine products
    assert class_flops(1, 1, 0, 0) == F_PRIM + 2 * (F_ROOT + F_COEF + 24 + 12 + 27) + 12 * 9
    # J only: the two J FMAs of each element (4 FLOPs) remain
    assert class_flops(0, 0, 0, 0, with_k=False) == F_PRIM + (F_ROOT + F_COEF + 3) + 4
    # a label in either orientation names the kernel instance that runs it
    assert kernel_symbol('(ss|ps)') == kernel_symbol('(ps|ss)') == 'mjk::jk_reg_kernel<double, 1, 0, 0, 0>'
    assert kernel_symbol('(ds|pp)').startswith('mjk::jk_grp_kernel<double, 1, 1, 2, 0>')
