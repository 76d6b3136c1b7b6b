"""Degenerate determinant spaces through the product path: one orbital,
empty and full shells, one spin without electrons, a single alpha or beta
string, against the oracle (oracle/fci_oracle.c, pinned to the brute-force
Hamiltonian on the same spaces in tests/test_oracle.py).  The reference's
GAS tests run such spaces too: (0, 0) and full-shell (5, 5) electrons in
test_gasci.py:455-502.

Tolerances: FP64 sigma 1e-13 of max|sigma| (the libfci_gas golden-vector
bar), diagonal 1e-12, RDMs 1e-12, energies 1e-10; FP32 sigma 2e-6 and
bf16x3 1e-4 of max|sigma| (the dropped low x low product is 2^-16 of each
term, and these small sigma vectors cancel: 2.1e-5 at 16 orbitals, (1, 0)).
"""

import numpy
import pytest

import oracle
from conftest import has_gpu, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()

CASES = [(1, (1, 0)), (1, (1, 1)), (2, (1, 1)), (4, (0, 0)), (4, (4, 4)), (6, (3, 0)),
         (6, (0, 3)), (5, (5, 2)), (7, (1, 1)), (16, (1, 0)), (16, (15, 14)), (12, (6, 0))]
TOL = {"fp64": 1e-13, "fp32": 2e-6, "bf16x3": 1e-4}


def _ids(c):
    return "%do-%d-%d" % (c[0], c[1][0], c[1][1])


@pytest.mark.parametrize("deterministic", [False, True])
@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_sigma_degenerate_spaces(case, deterministic, monkeypatch):
    n, ne = case
    if deterministic:
        monkeypatch.setenv("FCIS_DETERMINISTIC", "1")
    print("edge sigma", n, ne, "deterministic" if deterministic else "fused", flush=True)
    h1e, eri, c = pkg.synthetic.make_problem(n, ne, seed=3)
    h2e = oracle.absorb_h1e(h1e, eri, n, ne, 0.5)
    ref = oracle.contract_2e(h2e, c, n, ne)
    scale = max(numpy.abs(ref).max(), 1.0)
    numpy.testing.assert_allclose(pkg.absorb_h1e(h1e, eri, n, ne, 0.5), h2e, rtol=0, atol=1e-14)
    for prec, tol in TOL.items():
        if deterministic and prec != "fp64":
            continue
        got = pkg.contract_2e(h2e, c, n, ne, precision=prec)
        assert got.shape == ref.shape
        err = numpy.abs(got - ref).max() / scale
        assert err < tol, (prec, err)


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_diagonal_rdms_energy_degenerate_spaces(case):
    n, ne = case
    print("edge hdiag/rdm/kernel", n, ne, flush=True)
    h1e, eri, c = pkg.synthetic.make_problem(n, ne, seed=4)
    numpy.testing.assert_allclose(pkg.make_hdiag(h1e, eri, n, ne).reshape(-1),
                                  oracle.make_hdiag(h1e, eri, n, ne).reshape(-1), rtol=0, atol=1e-12)
    (a, b), (aa, ab, bb) = pkg.make_rdm12s(c, n, ne)
    (ra, rb), (raa, rab, rbb) = oracle.make_rdm12s(c, n, ne)
    for x, y in ((a, ra), (b, rb), (aa, raa), (ab, rab), (bb, rbb)):
        numpy.testing.assert_allclose(x, y, rtol=0, atol=1e-12)
    nroots = min(2, c.size)
    w = oracle.lowest_eigenvalues(h1e, eri, n, ne, nroots=nroots)[0]
    solver = pkg.FCISolver()
    e, ci = solver.kernel(h1e, eri, n, ne, nroots=nroots)
    e = numpy.atleast_1d(e)
    assert abs(e[0] - w[0]) < 1e-10
    # an excited root converged to the solver's residual bound (sqrt(1e-10))
    # is accurate to |r|^2 / gap: 7.7e-8 at 16 orbitals, (15, 14)
    if nroots > 1:
        assert abs(e[1] - w[1]) < 1e-6
    ci0 = ci[0] if nroots > 1 else ci
    assert abs(pkg.energy(h1e, eri, ci0, n, ne) - w[0]) < 1e-10


# The reference's degenerate GAS spaces (test_gasci.py:469-470): two
# subspaces of 3 and 2 orbitals with no electrons, and with every orbital
# doubly occupied.  One determinant each: sigma and the energy are its
# diagonal element; an orbital rotation within the subspaces multiplies it
# by det(u) per spin (det(u)^2 = 1), so the vector is unchanged.
GAS_EDGE = [((0, 0), ((0, 0, 0, 0),)), ((5, 5), ((3, 2, 3, 2),))]


@pytest.mark.parametrize("native", [False, True])
@pytest.mark.parametrize("nelec,restr", GAS_EDGE, ids=["vacuum", "full"])
def test_gas_degenerate_spaces(nelec, restr, native):
    print("edge gas", nelec, "native" if native else "embedded", flush=True)
    fg = pkg.fci_gas
    n = 5
    solver = fg.FCISolver(gas_orbs=(3, 2), gas_restr=restr, gas_restr_type="spin-supergroup")
    solver.native = native
    h1e, eri, _ = pkg.synthetic.make_problem(n, nelec, seed=5)
    with solver.make_space(n, nelec) as gas:
        assert gas.ndet == 1 and gas.native == native
    c = numpy.ones(1)
    h2e = solver.absorb_h1e(h1e, eri, n, nelec, 0.5)
    ref = oracle.contract_2e(oracle.absorb_h1e(h1e, eri, n, nelec, 0.5), c.reshape(1, 1), n, nelec)
    got = solver.contract_2e(h2e, c, n, nelec)
    assert abs(float(numpy.asarray(got).reshape(-1)[0]) - float(ref.reshape(-1)[0])) < 1e-13 * max(
        abs(float(ref.reshape(-1)[0])), 1.0)
    hd = numpy.asarray(solver.make_hdiag(h1e, eri, n, nelec)).reshape(-1)
    assert abs(hd[0] - oracle.make_hdiag(h1e, eri, n, nelec).reshape(-1)[0]) < 1e-12
    e, ci = solver.kernel(h1e, eri, n, nelec)
    assert abs(e - oracle.lowest_eigenvalues(h1e, eri, n, nelec)[0][0]) < 1e-10
    rng = numpy.random.default_rng(17)
    u = numpy.zeros((n, n))
    u[:3, :3] = numpy.linalg.qr(rng.normal(size=(3, 3)))[0]
    u[3:, 3:] = numpy.linalg.qr(rng.normal(size=(2, 2)))[0]
    perm = numpy.eye(n)[:, [1, 0, 2, 3, 4]]
    for rot in (u, perm):
        out = solver.transform_ci_within_gas(c, n, nelec, rot)
        numpy.testing.assert_allclose(numpy.asarray(out).reshape(-1), c, rtol=0, atol=1e-12)
