"""GPU parity of the CSF solver with a two-component h1e = (h_alpha, h_beta)
(pyscf/csf_fci/csf.py:28-68, 305-320, 518-529): the alpha / beta one-body
contraction (fcis_contract_1e_ab), the UHF Slater-Condon block
(fcis_pspace_uhf), the determinant diagonal, the tagged absorb_h1e /
contract_2e pair, and the solver's energies, all against a from-scratch
dense Hamiltonian H = sum_pq h_a E^a_pq + h_b E^b_pq + V built from
operator action on bit strings (oracle/brute.py).  The CSF solver works
in the spin-adapted space, so its energies are the eigenvalues of H
projected on the S(S+1) eigenspace of S^2.  Tolerances: 1e-12 for
matrices and vectors, 1e-9 for energies."""

import numpy
import pytest

import oracle.brute as brute
from conftest import has_gpu, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()
csf = pkg.csf_fci
ds = pkg.direct_spin1


def _one_body(n, nelec, ha, hb):
    na, nb = nelec
    Ea, Eb = brute.one_body_ops(n, na), brute.one_body_ops(n, nb)
    Ia, Ib = numpy.eye(Ea.shape[2]), numpy.eye(Eb.shape[2])
    M = numpy.zeros((Ea.shape[2] * Eb.shape[2],) * 2)
    for p in range(n):
        for q in range(n):
            M += ha[p, q] * numpy.kron(Ea[p, q], Ib) + hb[p, q] * numpy.kron(Ia, Eb[p, q])
    return M


def _problem(n, seed, symmetric=True):
    rng = numpy.random.default_rng(seed)
    ha = rng.standard_normal((n, n)) * 0.3
    hb = ha + rng.standard_normal((n, n)) * 0.1
    if symmetric:
        ha, hb = ha + ha.T, hb + hb.T
    g = rng.standard_normal((n, n, n, n)) * 0.1
    g = g + g.transpose(1, 0, 2, 3)
    g = g + g.transpose(0, 1, 3, 2)
    g = g + g.transpose(2, 3, 0, 1)
    return ha, hb, g


def _dense(n, nelec, ha, hb, g):
    return brute.hamiltonian(numpy.zeros((n, n)), g, n, nelec) + _one_body(n, nelec, ha, hb)


@pytest.mark.parametrize("n,nelec", [(4, (2, 2)), (5, (3, 2)), (6, (4, 1)), (5, (0, 3))])
def test_contract_1e_ab_vs_dense(n, nelec):
    import torch
    ha, hb, _ = _problem(n, n, symmetric=False)      # orientation of h[p][q] checked too
    M = _one_body(n, nelec, ha, hb)
    sp = ds.get_space(n, nelec)
    x = numpy.random.default_rng(1).standard_normal(sp.ndet)
    xd = torch.from_numpy(x).cuda()
    out = torch.empty_like(xd)
    ds.contract_1e_ab_dev(sp, ha, hb, xd, out)
    numpy.testing.assert_allclose(out.cpu().numpy(), M @ x, rtol=0, atol=1e-12)
    y0 = numpy.random.default_rng(2).standard_normal(sp.ndet)
    out = torch.from_numpy(y0.copy()).cuda()
    ds.contract_1e_ab_dev(sp, ha, hb, xd, out, accumulate=True)
    numpy.testing.assert_allclose(out.cpu().numpy(), y0 + M @ x, rtol=0, atol=1e-12)


@pytest.mark.parametrize("n,nelec", [(4, (2, 2)), (5, (3, 2)), (6, (3, 3))])
def test_uhf_pspace_block_and_diagonal(n, nelec):
    import torch
    ha, hb, g = _problem(n, 10 + n)
    H = _dense(n, nelec, ha, hb, g)
    sp = ds.get_space(n, nelec)
    addr = torch.arange(sp.ndet, device="cuda")
    h0 = ds._pspace_dev(sp, numpy.stack([ha, hb]), g, addr).cpu().numpy()
    numpy.testing.assert_allclose(h0, H, rtol=0, atol=1e-12)
    hd = csf.make_hdiag_det(None, numpy.stack([ha, hb]), g, n, nelec)
    numpy.testing.assert_allclose(hd, numpy.diag(H), rtol=0, atol=1e-12)
    # one-component h1e: the spin-free block and diagonal unchanged
    h0s = ds._pspace_dev(sp, ha, g, addr).cpu().numpy()
    numpy.testing.assert_allclose(h0s, _dense(n, nelec, ha, ha, g), rtol=0, atol=1e-12)


def test_tagged_absorb_and_contract():
    n, nelec = 6, (3, 2)
    ha, hb, g = _problem(n, 3)
    H = _dense(n, nelec, ha, hb, g)
    sol = csf.FCISolver(smult=2)
    h2 = sol.absorb_h1e(numpy.stack([ha, hb]), g, n, nelec, 0.5)
    assert h2.h1e_s is not None
    c = numpy.random.default_rng(5).standard_normal(H.shape[0])
    hc = sol.contract_2e(h2, c.reshape(ds.get_space(n, nelec).shape), n, nelec)
    numpy.testing.assert_allclose(hc.reshape(-1), H @ c, rtol=0, atol=1e-12)
    h2s = sol.absorb_h1e(ha, g, n, nelec, 0.5)          # spin-free h1e: no tag
    assert getattr(h2s, "h1e_s", None) is None


@pytest.mark.parametrize("n,na,nb,sm", [(5, 3, 2, 2), (6, 3, 3, 1), (6, 3, 3, 3), (6, 4, 2, 3)])
def test_csf_solver_with_spin_dependent_h1e(n, na, nb, sm):
    ha, hb, g = _problem(n, 7 * n + sm)
    H = _dense(n, (na, nb), ha, hb, g)
    S2 = brute.spin_square_matrix(n, (na, nb))
    w, v = numpy.linalg.eigh(0.5 * (S2 + S2.T))
    s = 0.5 * (sm - 1)
    P = v[:, abs(w - s * (s + 1)) < 1e-8]
    ref = numpy.linalg.eigvalsh(P.T @ H @ P)
    h1 = numpy.stack([ha, hb])
    sol = csf.FCISolver(smult=sm)
    e, _ = sol.kernel(h1, g, n, (na, nb), nroots=2)
    numpy.testing.assert_allclose(e, ref[:2], rtol=0, atol=1e-9)
    e2, ci2 = sol.kernel(h1, g, n, (na, nb), nroots=2, davidson_only=True, pspace_size=0)
    numpy.testing.assert_allclose(e2, ref[:2], rtol=0, atol=1e-9)
    e3, _ = sol.kernel(h1, g, n, (na, nb), nroots=2, davidson_only=True, pspace_size=10)
    numpy.testing.assert_allclose(e3, ref[:2], rtol=0, atol=1e-9)
    # the returned vectors are the projected eigenvectors: <c|P H P|c> = e
    for k in range(2):
        cv = ci2[k].reshape(-1)
        pc = P @ (P.T @ cv)
        assert abs(pc @ H @ pc / (pc @ pc) - e2[k]) < 1e-9
