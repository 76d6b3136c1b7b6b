"""GPU parity of the Slater-Condon pspace builder (fcis_pspace) and of the
solvers built on it: the CAS / GAS / CSF small-space H blocks and the CSF
solver's pspace preconditioner.

References: a from-scratch Slater-Condon H (oracle/brute.py) for small
spaces, and columns of the oracle's sigma on unit vectors for larger ones.
H elements to 1e-12; energies to 1e-9."""

import numpy
import pytest

import oracle
import oracle.brute as brute
import oracle.csf as ocsf
from oracle import gas as og
from conftest import has_gpu, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()


def _problem(n, seed):
    rng = numpy.random.default_rng(seed)
    h1 = rng.standard_normal((n, n)) * 0.3
    h1 = h1 + h1.T
    g = rng.standard_normal((n, n, n, n)) * 0.1
    g = g + g.transpose(1, 0, 2, 3)
    g = g + g.transpose(0, 1, 3, 2)
    g = g + g.transpose(2, 3, 0, 1)
    return h1, oracle.pair_matrix(g, n)


@pytest.mark.parametrize("n,ne", [(4, (2, 2)), (5, (3, 2)), (6, (3, 3)), (6, (4, 1)), (5, (5, 0)),
                                  (6, (2, 2))])
def test_cas_pspace_vs_slater_condon_brute_force(n, ne):
    h1, eri = _problem(n, 7 * n + ne[0])
    H = brute.hamiltonian(h1, eri, n, ne)
    nd = H.shape[0]
    addr, h0 = pkg.pspace(h1, eri, n, ne, np=nd)
    assert sorted(addr.tolist()) == list(range(nd))
    numpy.testing.assert_allclose(h0, H[numpy.ix_(addr, addr)], rtol=0, atol=1e-12)
    # lowest-diagonal selection and a smaller block
    addr, h0 = pkg.pspace(h1, eri, n, ne, np=min(7, nd))
    hd = numpy.diag(H)
    assert numpy.allclose(numpy.sort(hd)[:len(addr)], numpy.sort(hd[addr]))
    numpy.testing.assert_allclose(h0, H[numpy.ix_(addr, addr)], rtol=0, atol=1e-12)


def test_cas_pspace_vs_oracle_sigma_columns():
    n, ne = 9, (5, 4)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=6, with_ci=False)
    h2e = oracle.absorb_h1e(h1e, eri, n, ne, 0.5)
    addr, h0 = pkg.pspace(h1e, eri, n, ne, np=120)
    nd = oracle.num_strings(n, 5) * oracle.num_strings(n, 4)
    e = numpy.zeros(nd)
    for j, a in enumerate(addr.tolist()):
        e[a] = 1.0
        col = oracle.contract_2e(h2e, e, n, ne).reshape(-1)
        e[a] = 0.0
        numpy.testing.assert_allclose(h0[:, j], col[addr], rtol=0, atol=1e-12)


def test_cas_small_space_exact_path_uses_pspace():
    n, ne = 5, (2, 2)          # 100 determinants, exact path up to 64 -> Davidson
    h1, eri = _problem(n, 3)
    w = numpy.linalg.eigvalsh(brute.hamiltonian(h1, eri, n, ne))
    sol = pkg.FCISolver()
    e, _ = sol.kernel(h1, eri, n, ne, nroots=3)
    numpy.testing.assert_allclose(e, w[:3], rtol=0, atol=1e-9)
    n, ne = 4, (2, 2)          # 36 determinants: exact path
    h1, eri = _problem(n, 4)
    w = numpy.linalg.eigvalsh(brute.hamiltonian(h1, eri, n, ne))
    e, _ = pkg.FCISolver().kernel(h1, eri, n, ne, nroots=4)
    numpy.testing.assert_allclose(e, w[:4], rtol=0, atol=1e-10)


def test_gas_pspace_vs_oracle_projection():
    orbs, ne = (3, 3, 2), (3, 3)
    G = [[4, 2, 0], [3, 3, 0], [3, 2, 1], [2, 3, 1], [2, 4, 0], [4, 1, 1]]
    sol = pkg.fci_gas.FCISolver(gas_orbs=orbs, gas_restr=G, gas_restr_type="supergroup")
    h1e, eri = pkg.synthetic.make_problem(8, ne, seed=2, with_ci=False)
    h2e = oracle.absorb_h1e(h1e, eri, 8, ne, 0.5)
    addr, h0 = sol.pspace(h1e, eri, 8, ne, np=400)
    assert len(addr) == 64                      # the reference's cap (fci_gas.py:38)
    L = og.GasLayout(orbs, ne, sol._space_spec(8, ne)[2])
    e = numpy.zeros(L.ndet)
    for j, a in enumerate(addr.tolist()):
        e[a] = 1.0
        col = og.gas_contract_2e(h2e, e, L)
        e[a] = 0.0
        numpy.testing.assert_allclose(h0[:, j], col[addr], rtol=0, atol=1e-12)


@pytest.mark.parametrize("n,na,nb,sm", [(6, 3, 3, 1), (6, 3, 3, 3), (5, 3, 2, 2), (6, 4, 2, 3)])
def test_csf_pspace_vs_brute_force(n, na, nb, sm):
    h1, eri = _problem(n, 5 * n + sm)
    H = brute.hamiltonian(h1, eri, n, (na, nb))
    sol = pkg.csf_fci.FCISolver(smult=sm)
    addr, h0 = sol.pspace(h1, eri, n, (na, nb), npsp=30)
    t = sol.transformer
    X = numpy.zeros((t.ncsf, len(addr)))
    X[addr, numpy.arange(len(addr))] = 1.0
    D = numpy.stack([ocsf.csf2det(X[:, k], n, na, nb, sm).reshape(-1) for k in range(len(addr))],
                    axis=1)
    numpy.testing.assert_allclose(h0, D.T @ H @ D, rtol=0, atol=1e-11)


@pytest.mark.parametrize("n,na,nb,sm", [(6, 3, 3, 1), (7, 4, 3, 2)])
def test_csf_davidson_with_pspace_preconditioner(n, na, nb, sm):
    h1, eri = _problem(n, 13 * n + sm)
    H = brute.hamiltonian(h1, eri, n, (na, nb))
    S2 = brute.spin_square_matrix(n, (na, nb))
    w, v = numpy.linalg.eigh(0.5 * (S2 + S2.T))
    s = 0.5 * (sm - 1)
    P = v[:, abs(w - s * (s + 1)) < 1e-8]
    ref = numpy.linalg.eigvalsh(P.T @ H @ P)
    sol = pkg.csf_fci.FCISolver(smult=sm)
    e, _ = sol.kernel(h1, eri, n, (na, nb), nroots=2, davidson_only=True, pspace_size=20)
    assert numpy.all(sol.converged)
    numpy.testing.assert_allclose(e, ref[:2], rtol=0, atol=1e-9)


def test_csf_pspace_preconditioner_iterations():
    """(10o,10e) singlet, 19,404 CSFs: the pspace preconditioner (200 CSFs,
    the reference's default) reaches the same energy as the diagonal one."""
    h1, eri = pkg.synthetic.make_problem(10, (5, 5), seed=4, with_ci=False)
    out = {}
    for ps in (0, 200):
        sol = pkg.csf_fci.FCISolver(smult=1)
        e, _ = sol.kernel(h1, eri, 10, (5, 5), max_cycle=300, pspace_size=ps)
        assert sol.converged
        out[ps] = (e, sol.stats["cycles"])
    print("cycles: diagonal %d, pspace %d" % (out[0][1], out[200][1]), flush=True)
    assert abs(out[0][0] - out[200][0]) < 1e-9
