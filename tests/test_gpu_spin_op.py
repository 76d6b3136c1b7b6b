"""GPU parity of the spin ladder operators (fcis_spin_ladder behind
spin_op.contract_sladder / mdown / mup, pyscf/csf_fci/spin_op.py) against
the oracle (oracle/spin_op.py, pinned by the CSF coupling and S^2), and the
reference's own checks (csf_fci/test/test_spin_op.py:25-80): along a ladder
S -> -S the norm, <S^2> and <H> are kept, S+ of M = S and S- of M = -S
vanish, mup(mdown(c)) = c.  Tolerances: 1e-13 against the oracle, 1e-9 for
the reference's quantities (its own 9 decimals)."""

from itertools import product

import numpy
import pytest

import oracle
import oracle.csf as ocsf
import oracle.spin_op as osp
from conftest import has_gpu, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()
so = pkg.spin_op


@pytest.mark.parametrize("n,nelec,op", [(4, (2, 2), -1), (4, (2, 2), 1), (5, (3, 2), -1),
                                        (6, (4, 1), 1), (6, (5, 0), -1), (7, (3, 4), 1),
                                        (7, (4, 3), -1), (3, (3, 0), -1), (3, (1, 2), 1)])
def test_ladder_vs_oracle(n, nelec, op):
    rng = numpy.random.default_rng(n * 10 + nelec[0] + op)
    na, nb = nelec
    shape = (oracle.num_strings(n, na), oracle.num_strings(n, nb))
    c = rng.standard_normal(shape)
    got = so.contract_sladder(c, n, nelec, op)
    ref = osp.contract_sladder(c, n, nelec, op)
    assert got.shape == ref.shape
    numpy.testing.assert_allclose(got, ref, rtol=0, atol=1e-13)


def test_vacuum_annihilation_is_empty():
    c = numpy.ones((4, 4))
    assert so.contract_sup(numpy.ones((1, 4)), 4, (4, 1)).size == 0
    assert so.contract_sdown(numpy.ones((4, 1)), 4, (1, 4)).size == 0
    assert so.contract_sdown(numpy.ones((1, 4)), 4, (0, 1)).size == 0
    assert so.contract_sup(c, 4, (1, 1)).shape == (6, 1)


def _expect(h2e, c, n, ne):
    sig = pkg.contract_2e(h2e, c, n, ne)
    ss, mult = pkg.FCISolver().spin_square(c, n, ne)
    return float(c.ravel() @ c.ravel()), float(c.ravel() @ sig.ravel()), mult


@pytest.mark.parametrize("norb,ntot", list(product((6, 7), repeat=2)))
def test_ladder_keeps_norm_spin_and_energy(norb, ntot):
    """test_spin_op.py:29-64 with a random high-spin vector: ladder down
    to M = -S and back up, checking <c|c>, 2S+1 and <c|H|c> each step."""
    nelec = (min(norb, ntot), ntot - min(norb, ntot))
    smult = nelec[0] - nelec[1] + 1
    rng = numpy.random.default_rng(norb * 100 + ntot)
    c = rng.random((oracle.num_strings(norb, nelec[0]), oracle.num_strings(norb, nelec[1])))
    c /= numpy.linalg.norm(c)
    g = rng.random((norb,) * 4)
    g = g + g.transpose(1, 0, 2, 3)
    g = g + g.transpose(0, 1, 3, 2)
    g = g + g.transpose(2, 3, 0, 1)
    h2e = pkg.absorb_h1e(numpy.zeros((norb, norb)), g, norb, nelec, 0.5)
    cc0, chc0, m0 = _expect(h2e, c, norb, nelec)
    assert abs(m0 - smult) < 1e-9
    assert so.contract_sup(c, norb, nelec).size == 0
    s = (smult - 1) / 2
    ne = nelec
    for _ in numpy.arange(s, -s, -1):
        c = so.contract_sdown(c, norb, ne) / so.norm_sdown(smult, ne)
        ne = (ne[0] - 1, ne[1] + 1)
        cc, chc, m = _expect(h2e, c, norb, ne)
        assert abs(cc - 1.0) < 1e-9 and abs(m - smult) < 1e-9 and abs(chc - chc0) < 1e-9
    assert so.contract_sdown(c, norb, ne).size == 0
    for _ in numpy.arange(-s, s, 1):
        c = so.contract_sup(c, norb, ne) / so.norm_sup(smult, ne)
        ne = (ne[0] + 1, ne[1] - 1)
        cc, chc, m = _expect(h2e, c, norb, ne)
        assert abs(cc - 1.0) < 1e-9 and abs(m - smult) < 1e-9 and abs(chc - chc0) < 1e-9


@pytest.mark.parametrize("norb,ntot", list(product((6, 7), repeat=2)))
def test_mdown_mup_round_trip(norb, ntot):
    """test_spin_op.py:66-80: for array, list and stacked inputs."""
    nelec = (min(norb, ntot), ntot - min(norb, ntot))
    smult = nelec[0] - nelec[1] + 1
    rng = numpy.random.default_rng(norb + ntot)
    c = rng.random((oracle.num_strings(norb, nelec[0]), oracle.num_strings(norb, nelec[1])))
    c /= numpy.linalg.norm(c)
    s = (smult - 1) / 2
    for ci0 in (c, [c], c[None, :, :]):
        ne = nelec
        for _ in numpy.arange(s, -s, -1):
            ci1 = so.mup(so.mdown(ci0, norb, ne, smult), norb, ne, smult)
            numpy.testing.assert_allclose(numpy.asarray(ci1).reshape(-1),
                                          numpy.asarray(ci0).reshape(-1), rtol=0, atol=1e-9)
            ne = (ne[0] - 1, ne[1] + 1)


def test_sdown_keeps_csf_coupling_at_10_orbitals():
    """The CSF transformer of M and M - 1 with the same couplings: S- of
    the determinant expansion of a CSF vector x is norm_sdown times that
    of x, all on the device."""
    import torch
    n, na, nb, sm = 10, 6, 4, 3
    t = pkg.csf_fci.CSFTransformer(n, na, nb, sm)
    t1 = pkg.csf_fci.CSFTransformer(n, na - 1, nb + 1, sm)
    assert t.ncsf == t1.ncsf
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn(t.ncsf, dtype=torch.float64, device="cuda", generator=g)
    d = t.vec_csf2det(x, normalize=False)
    d1 = so.contract_sdown(d, n, (na, nb)) / so.norm_sdown(sm, (na, nb))
    y = t1.vec_det2csf(d1.reshape(-1), normalize=False)
    assert float((y - x).abs().max()) < 1e-12
    t.close()
    t1.close()
