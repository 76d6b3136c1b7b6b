"""CPU checks of the autograd backward's algebra: the pair folding of dm2
and the gradient formulas, with oracle RDMs.  E = c^T H c is homogeneous of
degree one in (h1e, eri) (absorb_h1e is linear), so
E = <dE/dh1e, h1e> + <dE/deri, eri> must hold exactly; directional
derivatives must match central finite differences of the oracle energy."""

import numpy
import pytest

import oracle
from conftest import package

pkg = package()
ag = pkg.autograd


def test_pair_fold_matches_loop():
    n = 4
    rng = numpy.random.default_rng(0)
    dm2 = rng.normal(size=(n,) * 4)
    out = ag.pair_fold(dm2, n)
    ref = numpy.zeros_like(out)
    for p in range(n):
        for q in range(n):
            for r in range(n):
                for s in range(n):
                    ref[pkg.direct_spin1.pair_index(p, q), pkg.direct_spin1.pair_index(r, s)] += dm2[p, q, r, s]
    numpy.testing.assert_allclose(out, ref, atol=1e-14)


@pytest.mark.parametrize("n,ne,seed", [(6, (3, 3), 1), (7, (4, 2), 2)])
def test_euler_identity_and_finite_differences(n, ne, seed):
    h1e, eri, c = pkg.synthetic.make_problem(n, ne, seed=seed)
    dm1, dm2 = oracle.make_rdm12(c, n, ne)
    gh, geri = ag.energy_gradients(dm1, dm2, n)
    e = oracle.energy(h1e, eri, c, n, ne)
    assert abs(numpy.sum(gh * h1e) + numpy.sum(geri * eri) - e) < 1e-11
    rng = numpy.random.default_rng(seed + 100)
    eps = 1e-5
    for _ in range(5):
        dh = rng.normal(size=h1e.shape)
        dh = 0.5 * (dh + dh.T)
        de = rng.normal(size=eri.shape)
        de = 0.5 * (de + de.T)
        ep = oracle.energy(h1e + eps * dh, eri + eps * de, c, n, ne)
        em = oracle.energy(h1e - eps * dh, eri - eps * de, c, n, ne)
        fd = (ep - em) / (2 * eps)
        an = numpy.sum(gh * dh) + numpy.sum(geri * de)
        assert abs(fd - an) < 1e-8 * max(1.0, abs(an))
