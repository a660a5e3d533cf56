"""SCF host logic that runs on CPU tensors: the SP2 density purification."""
import pytest
import torch

from byteqc_amd.scf.rhf import purify


def _spectrum_matrix(e, seed=0):
    g = torch.Generator().manual_seed(seed)
    n = e.shape[0]
    q, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, generator=g))
    return (q * e) @ q.T, q


@pytest.mark.parametrize('n,nocc', [(40, 7), (200, 60), (120, 1), (120, 119)])
def test_purify_is_the_occupied_projector(n, nocc):
    """SP2 gives the projector onto the nocc lowest eigenvectors (eigh's, to 1e-12), for a
    spectrum as wide as a def2-TZVP Fock matrix's (-20 to 60 Eh) with a 0.4 Eh gap."""
    e = torch.cat([torch.linspace(-20.0, -0.3, nocc), torch.linspace(0.1, 60.0, n - nocc)]).double()
    f, q = _spectrum_matrix(e)
    x = purify(f, nocc)
    assert x is not None
    p = q[:, :nocc] @ q[:, :nocc].T
    assert float((x - p).abs().max()) < 1e-12
    assert abs(float(torch.trace(x)) - nocc) < 1e-9


def test_purify_order_does_not_depend_on_basis():
    """A permuted basis gives the permuted projector (the algorithm sees only the matrix)."""
    e = torch.cat([torch.linspace(-5.0, -1.0, 5), torch.linspace(1.0, 9.0, 25)]).double()
    f, _ = _spectrum_matrix(e, seed=3)
    perm = torch.randperm(30, generator=torch.Generator().manual_seed(1))
    x = purify(f, 5)
    xp = purify(f[perm][:, perm], 5)
    assert float((xp - x[perm][:, perm]).abs().max()) < 1e-12


def test_purify_without_a_gap_gives_up():
    """Degenerate states nocc and nocc + 1: no idempotent matrix of trace nocc commutes with f
    and separates them, so purify returns None (the SCF then diagonalises)."""
    e = torch.tensor([-3.0, -2.0, -1.0, -1.0, 2.0, 4.0], dtype=torch.float64)
    f, _ = _spectrum_matrix(e, seed=5)
    assert purify(f, 3) is None
