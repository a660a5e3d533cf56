"""SCF layer: the J/K drop-in (hf.get_jk, hf._VHFOpt) and the RHF driver."""
from .hf import get_jk, _VHFOpt, basis_seg_contraction  # noqa: F401
from .rhf import RHF, DIIS, eig, int1e, init_guess_by_atom  # noqa: F401
