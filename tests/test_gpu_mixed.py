"""Mixed FP32/FP64 quartets and density-weighted screening against the FP64 build."""
import pytest

from byteqc_amd.gto import water_cluster
from byteqc_amd.scf import hf
from byteqc_amd.scf.rhf import RHF, init_guess_by_atom

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def cluster():
    mol = water_cluster(6, 8.0, seed=3, basis='def2-tzvp')
    opt = hf._VHFOpt(mol).build()
    return mol, opt


def test_density_screening_is_harmless(cluster):
    mol, opt = cluster
    dm = init_guess_by_atom(mol)
    j0, k0 = hf.get_jk(mol, dm, vhfopt=opt, prescreen=False)
    j1, k1 = hf.get_jk(mol, dm, vhfopt=opt, prescreen=True)
    print('density screening |dJ| %.2e |dK| %.2e' % (abs(j1 - j0).max(), abs(k1 - k0).max()), flush=True)
    assert abs(j1 - j0).max() < 1e-9 and abs(k1 - k0).max() < 1e-9


@pytest.mark.parametrize('cut', [1e-9, 1e-7, 1e-5])
def test_mixed_precision_error_scales_with_cutoff(cluster, cut):
    mol, opt = cluster
    dm = init_guess_by_atom(mol)
    j0, k0 = hf.get_jk(mol, dm, vhfopt=opt)
    j1, k1 = hf.get_jk(mol, dm, vhfopt=opt, precision_cutoff=cut)
    st = opt.stats()
    dj, dk = abs(j1 - j0).max(), abs(k1 - k0).max()
    print(f'cutoff {cut:g}: fp32 quartets {st["quartets_fp32"]} of '
          f'{st["quartets_fp32"] + st["quartets_fp64"]}, |dJ| {dj:.2e} |dK| {dk:.2e}', flush=True)
    assert st['quartets_fp32'] > 0
    # FP32 carries ~1e-7 relative error on quartets whose bound is below the cutoff
    assert dj < 100 * cut and dk < 100 * cut


def test_mixed_precision_scf_energy():
    mol = water_cluster(4, 8.0, seed=5, basis='def2-tzvp')
    e64 = RHF(mol, conv_tol=1e-9).kernel()
    mf = RHF(mol, conv_tol=1e-9, precision_cutoff=1e-7)
    e32 = mf.kernel()
    print(f'FP64 {e64:.10f}  mixed {e32:.10f}  dE {e32 - e64:.2e}', flush=True)
    assert mf.converged
    assert abs(e32 - e64) < 1e-6
