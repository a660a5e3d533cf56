"""Algorithmic FP64 work of a Rys J/K build (closed form per angular-momentum class).

For a class (la, lb | lc, ld) with N = (la+lb+lc+ld)/2 + 1 roots, nf the number of
Cartesian products, n = la + lb, m = lc + ld, counting an FMA as 2 FLOPs:

  per primitive quartet        T and prefactor                     F_PRIM
  per primitive quartet, root  node/weight (two Clenshaw series)   F_ROOT
                               recursion coefficients              F_COEF
                               vertical recursion   3 * 4 ((n+1)(m+1) - 1)
                               transfer to j        3 * 2 sum_{j=1..lb} (n-j+1)(m+1)
                               transfer to l        3 * 2 sum_{l=1..ld} (la+1)(lb+1)(m-l+1)
                               assembly             3 nf
  per contracted quartet, density  digestion        12 nf (six FMAs: J_ij J_kl K_ik K_il K_jk K_jl)

These are the FLOPs the algorithm needs, not what a given kernel issues (redundant
recomputation is not counted), so the achieved rate reads against the FP64 peak.
"""
from __future__ import annotations

F_PRIM = 30
F_ROOT = 2 * 11 * 4 + 6   # degree-11 Clenshaw for x and w: 11 steps of (mul + fma + add)
F_COEF = 20


def ncart(l):
    return (l + 1) * (l + 2) // 2


def class_flops(la, lb, lc, ld, nprim_quartet=1, n_dm=1, with_j=True, with_k=True):
    nr = (la + lb + lc + ld) // 2 + 1
    nf = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)
    n, m = la + lb, lc + ld
    vrr = 3 * 4 * ((n + 1) * (m + 1) - 1)
    hrr_j = 3 * 2 * sum((n - j + 1) * (m + 1) for j in range(1, lb + 1))
    hrr_l = 3 * 2 * sum((la + 1) * (lb + 1) * (m - l + 1) for l in range(1, ld + 1))
    per_root = F_ROOT + F_COEF + vrr + hrr_j + hrr_l + 3 * nf
    dig = (4 * int(with_j) + 8 * int(with_k)) * nf
    return nprim_quartet * (F_PRIM + nr * per_root) + n_dm * dig


def class_label(la, lb, lc, ld):
    s = 'spdfg'
    return f'({s[la]}{s[lb]}|{s[lc]}{s[ld]})'


def kernel_symbol(label, fp32=False):
    """Kernel template instance that runs a class label such as '(dp|ps)'."""
    import re
    la, lb, lc, ld = ['spdfg'.index(x) for x in re.findall('[spdfg]', label)]
    nc = ncart
    if nc(la) * nc(lb) < nc(lc) * nc(ld):  # the kernels take the pair with more products as bra
        la, lb, lc, ld = lc, ld, la, lb
    gs = (la + 1) * (lb + 1) * (lc + 1) * (ld + 1)
    regs = 3 * gs + nc(la) * nc(lb) * nc(lc) * nc(ld) + nc(la) * nc(lb) + (nc(la) + nc(lb)) * nc(lc)
    group = (la, lb, lc, ld) in ((1, 1, 2, 0), (2, 0, 2, 0), (2, 2, 0, 0))  # measured faster on the group kernel
    name = 'jk_reg_kernel' if regs <= 150 and not group else 'jk_grp_kernel'
    if (la, lb, lc, ld) in ((3, 3, 3, 2), (3, 3, 3, 3)):
        name = 'jk_kernel'  # the block-LDS kernel keeps the two largest classes
    return f'mjk::{name}<{"float" if fp32 else "double"}, {la}, {lb}, {lc}, {ld}>'
