"""byteqc_amd: an MI355X-native direct-SCF Coulomb/exchange (J/K) engine.

The drop-in surface mirrors ByteQC's ``cuobc.scf.hf`` (`get_jk`, `_VHFOpt`, `RHF`);
the integrals run as hand-written gfx950 HIP kernels in ``libmi355jk.so``.
"""
__version__ = '0.1.0'
