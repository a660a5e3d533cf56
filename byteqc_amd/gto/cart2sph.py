"""Cartesian -> real solid harmonic transforms in the PySCF/libcint AO conventions.

Cartesian components of angular momentum l are ordered with the x power descending,
then the y power descending (xx, xy, xz, yy, yz, zz for d), as in the reference's
digestion loops (cuobc/lib/gvhf/contract_jk.cu:228-260).

Spherical components are ordered m = -l..l except for p, which PySCF keeps as
(x, y, z).  The coefficients are those of the real solid harmonics r^l Y_lm with
int |Y_lm|^2 dOmega = 1 (Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure
Theory", eq. 6.4.47-6.4.50, rescaled by sqrt((2l+1)/(4 pi))).  For l <= 1 the angular
constants 1/(2 sqrt(pi)) and sqrt(3/(4 pi)) live in the Cartesian functions themselves
(pair_data.c:27-28 `CINTcommon_fac_sp`), so the block is the identity.
"""
from __future__ import annotations

from functools import lru_cache
from math import comb, factorial, sqrt, pi

import numpy as np


def cart_components(l: int):
    """[(lx, ly, lz), ...] in libcint order."""
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


def _solid_harmonic(l: int, m: int) -> dict:
    """Coefficients {(lx,ly,lz): c} of the unit-normalised real solid harmonic."""
    am = abs(m)
    vm = 0.5 if m < 0 else 0.0
    nlm = 1. / (2 ** am * factorial(l)) * sqrt(2 * factorial(l + am) * factorial(l - am)
                                               / (2. if m == 0 else 1.))
    nlm *= sqrt((2 * l + 1) / (4 * pi))
    coef = {}
    for t in range((l - am) // 2 + 1):
        for u in range(t + 1):
            # v runs over vm, vm+1, ..., floor(am/2 - vm) + vm
            nv = int(np.floor(am / 2. - vm))
            for iv in range(nv + 1):
                v = vm + iv
                two_v = int(round(2 * v))
                if two_v > am:
                    continue
                c = ((-1) ** int(round(t + v - vm)) * (0.25 ** t) * comb(l, t)
                     * comb(l - t, am + t) * comb(t, u) * comb(am, two_v))
                lx = 2 * t + am - 2 * u - two_v
                ly = 2 * u + two_v
                lz = l - 2 * t - am
                key = (lx, ly, lz)
                coef[key] = coef.get(key, 0.) + c * nlm
    return coef


@lru_cache(maxsize=None)
def _block(l: int) -> np.ndarray:
    comps = cart_components(l)
    if l <= 1:
        return np.eye(len(comps))
    idx = {c: i for i, c in enumerate(comps)}
    out = np.zeros((len(comps), 2 * l + 1))
    for col, m in enumerate(range(-l, l + 1)):
        for key, c in _solid_harmonic(l, m).items():
            out[idx[key], col] += c
    return out


def cart2sph_block(l: int) -> np.ndarray:
    """(ncart, nsph) transform for one shell."""
    return _block(l).copy()
