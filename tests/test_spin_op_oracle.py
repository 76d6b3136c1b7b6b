"""Pins the spin-ladder oracle (oracle/spin_op.py, restating
pyscf/csf_fci/spin_op.py:22-71) without pyscf:

* sign convention: S- maps every CSF |S, M> of the CSF oracle (itself
  pinned to the reference's csfstring.c) to +norm_sdown(S, M) |S, M - 1>
  with the same coupling; the opposite single-operator convention fails
  this for an even electron count (checked below by planting it);
* magnitudes: S+ S- + Sz (Sz - 1) equals the brute-force S^2 matrix and
  S+ is the transpose of S-;
* vacuum annihilation returns an empty array, as the reference does."""

import numpy
import pytest

import oracle.brute as brute
import oracle.csf as ocsf
import oracle.spin_op as osp


def _matrix(n, nelec, op):
    na, nb = nelec
    nd = len(brute._strings(n, na)) * len(brute._strings(n, nb))
    cols = [osp.contract_sladder(e, n, nelec, op).reshape(-1) for e in numpy.eye(nd)]
    return numpy.array(cols).T


@pytest.mark.parametrize("n,na,nb,sm", [(4, 3, 1, 3), (5, 3, 2, 2), (5, 4, 1, 4), (6, 4, 2, 3),
                                        (6, 3, 3, 3), (6, 5, 1, 5), (6, 3, 3, 1)])
def test_sdown_preserves_csf_coupling(n, na, nb, sm):
    blocks, ncsf = ocsf.csf_shape(n, na, nb, sm)
    if na - 1 - (nb + 1) < -(sm - 1):
        pytest.skip("M - 1 below -S")
    for k in range(ncsf):
        x = numpy.zeros(ncsf)
        x[k] = 1.0
        d = ocsf.csf2det(x, n, na, nb, sm)
        d1 = osp.contract_sladder(d, n, (na, nb), -1).reshape(-1) / osp.norm_sdown(sm, (na, nb))
        y = ocsf.det2csf(d1, n, na - 1, nb + 1, sm)
        numpy.testing.assert_allclose(y, x, rtol=0, atol=1e-13)


def test_planted_sign_convention_fails():
    """The CSF check catches the other single-operator convention (counting
    occupied orbitals below p), planted here in a copy of the oracle."""
    import oracle
    n, na, nb, sm = 4, 3, 1, 3
    sa, sb = oracle.make_strings(n, na), oracle.make_strings(n, nb)
    da = {int(s): i for i, s in enumerate(oracle.make_strings(n, na - 1))}
    db = {int(s): i for i, s in enumerate(oracle.make_strings(n, nb + 1))}
    ncsf = ocsf.csf_shape(n, na, nb, sm)[1]
    x = numpy.zeros(ncsf)
    x[0] = 1.0
    c = ocsf.csf2det(x, n, na, nb, sm).reshape(len(sa), len(sb))
    out = numpy.zeros((len(da), len(db)))
    for a, A in enumerate(int(v) for v in sa):
        for b, B in enumerate(int(v) for v in sb):
            for p in range(n):
                bit = 1 << p
                if (A & bit) and not (B & bit):
                    below = bin(A & (bit - 1)).count("1") + bin(B & (bit - 1)).count("1")
                    out[da[A ^ bit], db[B ^ bit]] += (-1) ** (na - 1 + below) * c[a, b]
    y = ocsf.det2csf(out.reshape(-1) / osp.norm_sdown(sm, (na, nb)), n, na - 1, nb + 1, sm)
    assert abs(y[0] + 1.0) < 1e-13     # -|S, M-1>: fails the coupling check


@pytest.mark.parametrize("n,nelec", [(4, (2, 2)), (5, (3, 2)), (5, (2, 2)), (6, (4, 2)), (4, (3, 0))])
def test_ladder_products_give_s2(n, nelec):
    na, nb = nelec
    if nb == 0:
        assert osp.contract_sladder(numpy.ones(4), n, nelec, 1).size == 0
        return
    sm = _matrix(n, nelec, -1)                      # (na-1, nb+1) <- (na, nb)
    sp = _matrix(n, (na - 1, nb + 1), 1)            # back up
    numpy.testing.assert_allclose(sp, sm.T, rtol=0, atol=0)
    sz = 0.5 * (na - nb)
    S2 = brute.spin_square_matrix(n, nelec)
    numpy.testing.assert_allclose(sp @ sm + sz * (sz - 1) * numpy.eye(S2.shape[0]), S2,
                                  rtol=0, atol=1e-12)
    assert osp.contract_sladder(numpy.ones(1), 3, (0, 1), -1).size == 0
