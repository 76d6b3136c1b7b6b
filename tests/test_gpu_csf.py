"""GPU parity of the CSF <-> determinant transform, the CSF diagonal and the
CSF-basis solver against the oracle (oracle/csf.py, itself pinned to the
reference's csfstring.c) and against brute-force spin-adapted spectra.
Tolerances: transforms 1e-13 (sums of at most C(16,8) products of unit-norm
coefficients), hdiag_csf 1e-12, energies 1e-9."""

import numpy
import pytest

import oracle
import oracle.brute as brute
import oracle.csf as ocsf
from conftest import has_gpu, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()
csf = pkg.csf_fci

CASES = [(4, 2, 2, 1), (4, 2, 2, 3), (5, 3, 2, 2), (5, 3, 2, 4), (6, 3, 3, 1), (6, 3, 3, 3),
         (6, 4, 2, 3), (7, 4, 3, 2), (7, 4, 3, 4), (8, 4, 4, 1), (8, 5, 3, 3), (9, 5, 4, 2),
         (6, 6, 0, 7), (6, 1, 1, 1), (6, 5, 5, 1), (3, 3, 3, 1)]


def _problem(n, seed):
    rng = numpy.random.default_rng(seed)
    h1 = rng.standard_normal((n, n)) * 0.3
    h1 = h1 + h1.T
    g = rng.standard_normal((n, n, n, n)) * 0.1
    g = g + g.transpose(1, 0, 2, 3)
    g = g + g.transpose(0, 1, 3, 2)
    g = g + g.transpose(2, 3, 0, 1)
    return h1, oracle.pair_matrix(g, n)


@pytest.mark.parametrize("n,na,nb,sm", CASES)
def test_transform_matches_oracle(n, na, nb, sm):
    import torch
    t = csf.CSFTransformer(n, na, nb, sm)
    blocks, ncsf = ocsf.csf_shape(n, na, nb, sm)
    assert t.ncsf == ncsf
    rng = numpy.random.default_rng(n * 100 + na * 10 + nb + sm)
    c = rng.standard_normal(t.ndet)
    mine = t.vec_det2csf(c, normalize=False)
    ref = ocsf.det2csf(c, n, na, nb, sm)
    numpy.testing.assert_allclose(mine, ref, rtol=0, atol=1e-13 * max(1.0, abs(ref).max()))
    x = rng.standard_normal(ncsf)
    mine = t.vec_csf2det(x, normalize=False)
    ref = ocsf.csf2det(x, n, na, nb, sm)
    numpy.testing.assert_allclose(mine, ref, rtol=0, atol=1e-13 * max(1.0, abs(ref).max()))
    # device tensors stay on the device; det2csf . csf2det = 1
    xd = torch.from_numpy(x).cuda()
    back = t.vec_det2csf(t.vec_csf2det(xd, normalize=False), normalize=False)
    assert back.is_cuda
    numpy.testing.assert_allclose(back.cpu().numpy(), x, rtol=0, atol=1e-12)


def test_batched_transform_and_normalisation():
    n, na, nb, sm = 8, 4, 4, 3
    t = csf.CSFTransformer(n, na, nb, sm)
    rng = numpy.random.default_rng(7)
    c = rng.standard_normal((3, t.ndet))
    batch = t.vec_det2csf(c, normalize=False)
    for k in range(3):
        numpy.testing.assert_allclose(batch[k], ocsf.det2csf(c[k], n, na, nb, sm), atol=1e-13)
    xs, norms = t.vec_det2csf(c, normalize=True, return_norm=True)
    numpy.testing.assert_allclose(numpy.linalg.norm(xs, axis=1), 1.0, atol=1e-14)
    numpy.testing.assert_allclose(norms, numpy.linalg.norm(batch, axis=1), rtol=1e-14)
    # Fortran order: vectors on columns
    colf = t.vec_det2csf(numpy.asfortranarray(c.T), order="F", normalize=False)
    numpy.testing.assert_allclose(colf, batch.T, atol=1e-13)
    dets = t.vec_csf2det(batch, normalize=False)
    assert dets.shape == (3, t.ndet)


def test_mask_is_a_permutation_at_16o16e():
    """Full-size property: csf2det of det2csf is the projector onto the
    singlet sector, idempotent, and keeps the norm of a singlet vector."""
    import torch
    t = csf.CSFTransformer(16, 8, 8, 1)
    assert t.ncsf == ocsf.csf_shape(16, 8, 8, 1)[1]
    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn(t.ncsf, dtype=torch.float64, device="cuda", generator=g)
    d = torch.empty(t.ndet, dtype=torch.float64, device="cuda")
    t.csf2det_dev(x, d)
    assert abs(float(d.norm()) - float(x.norm())) < 1e-10 * float(x.norm())
    y = torch.empty_like(x)
    t.det2csf_dev(d, y)
    assert float((y - x).abs().max()) < 1e-12


@pytest.mark.parametrize("n,na,nb,sm", CASES[:12])
def test_hdiag_csf_matches_oracle(n, na, nb, sm):
    h1, eri = _problem(n, n + na)
    t = csf.CSFTransformer(n, na, nb, sm)
    mine = csf.make_hdiag_csf(h1, eri, n, (na, nb), t).cpu().numpy()
    ref = ocsf.hdiag_csf(h1, eri, n, na, nb, sm)
    numpy.testing.assert_allclose(mine, ref, rtol=0, atol=1e-12 * max(1.0, abs(ref).max()))


def _spin_adapted_levels(h1, eri, n, na, nb, sm):
    """Brute force: eigenvalues of H in the S(S+1) eigenspace of S^2."""
    H = brute.hamiltonian(h1, eri, n, (na, nb))
    S2 = brute.spin_square_matrix(n, (na, nb))
    w, v = numpy.linalg.eigh(0.5 * (S2 + S2.T))
    s = 0.5 * (sm - 1)
    P = v[:, abs(w - s * (s + 1)) < 1e-8]
    return numpy.linalg.eigvalsh(P.T @ H @ P)


@pytest.mark.parametrize("n,na,nb,sm", [(5, 3, 2, 2), (6, 3, 3, 1), (6, 3, 3, 3), (6, 4, 2, 5)])
def test_csf_solver_exact_and_davidson_vs_brute(n, na, nb, sm):
    h1, eri = _problem(n, 11 * n + sm)
    ref = _spin_adapted_levels(h1, eri, n, na, nb, sm)
    sol = csf.FCISolver(smult=sm)
    e, ci = sol.kernel(h1, eri, n, (na, nb), nroots=2)
    numpy.testing.assert_allclose(e, ref[:2], atol=1e-9)
    e2, ci2 = sol.kernel(h1, eri, n, (na, nb), nroots=2, davidson_only=True, pspace_size=0)
    numpy.testing.assert_allclose(e2, ref[:2], atol=1e-9)
    # determinant-basis vectors of the right spin and energy
    for k in range(2):
        c = ci2[k].reshape(-1)
        ss, mult = pkg.FCISolver().spin_square(ci2[k], n, (na, nb))
        assert abs(mult - sm) < 1e-6
        assert abs(pkg.energy(h1, eri, c.reshape(ci2[k].shape), n, (na, nb)) - e2[k]) < 1e-9


def test_csf_solver_davidson_matches_det_solver():
    """(10o,10e) singlet: 19,404 CSFs vs 63,504 determinants; the CSF
    ground state is the lowest singlet among the determinant solver's
    roots."""
    h1, eri = pkg.synthetic.make_problem(10, (5, 5), seed=4, with_ci=False)
    det = pkg.FCISolver()
    e_det, c_det = det.kernel(h1, eri, 10, (5, 5), nroots=4, max_cycle=200)
    mults = [det.spin_square(c, 10, (5, 5))[1] for c in c_det]
    singlets = [e for e, m in zip(e_det, mults) if abs(m - 1.0) < 1e-4]
    assert singlets, mults
    sol = csf.FCISolver(smult=1)
    e, c = sol.kernel(h1, eri, 10, (5, 5), max_cycle=200)
    assert sol.converged
    assert sol.transformer.ncsf == 19404
    _, mult = det.spin_square(c, 10, (5, 5))
    assert abs(mult - 1.0) < 1e-6
    assert abs(e - singlets[0]) < 1e-8
    # mixed precision in the CSF basis ends at the same FP64 answer
    e32, _ = csf.FCISolver(smult=1).kernel(h1, eri, 10, (5, 5), max_cycle=200, precision="fp32")
    assert abs(e32 - e) < 1e-9


def test_impossible_spin_raises():
    with pytest.raises(csf.ImpossibleSpinError):
        csf.CSFTransformer(6, 4, 2, 1)
    with pytest.raises(csf.ImpossibleSpinError):
        csf.CSFTransformer(6, 3, 3, 2)
