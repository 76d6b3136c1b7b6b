"""GPU autograd: E(h1e, eri) = c^T H c with the RDM kernel as backward.
At (10o,10e) the gradients are compared with oracle RDMs; at the
(12o,12e) benchmark size (853,776 dets, c the converged ground state) with
central finite differences on 20 random symmetric perturbations (step 1e-5,
tolerance 1e-7 relative), with the homogeneity identity
E = <dE/dh1e, h1e> + <dE/deri, eri>, and with FP32 integrals."""

import numpy
import pytest

import oracle
from conftest import has_gpu, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()


def _grads(h1e, eri, c, n, ne, dtype=None):
    import torch
    dtype = dtype or torch.float64
    h = torch.tensor(h1e, dtype=dtype, requires_grad=True)
    e = torch.tensor(eri, dtype=dtype, requires_grad=True)
    ct = torch.tensor(numpy.asarray(c).reshape(-1), dtype=torch.float64, device="cuda",
                      requires_grad=True)
    E = pkg.autograd.ci_energy(h, e, ct, n, ne)
    E.backward()
    return float(E), h.grad.double().numpy(), e.grad.double().numpy(), ct.grad


def test_autograd_vs_oracle_rdms():
    n, ne = 10, (5, 5)
    h1e, eri, c = pkg.synthetic.make_problem(n, ne, seed=3)
    E, gh, ge, gc = _grads(h1e, eri, c, n, ne)
    dm1, dm2 = oracle.make_rdm12(c, n, ne)
    rh, re = pkg.autograd.energy_gradients(dm1, dm2, n)
    numpy.testing.assert_allclose(gh, rh, atol=1e-10)
    numpy.testing.assert_allclose(ge, re, atol=1e-10)
    assert abs(E - oracle.energy(h1e, eri, c, n, ne)) < 1e-10
    # dE/dc = 2 H c
    h2e = oracle.absorb_h1e(h1e, eri, n, ne, 0.5)
    hc = oracle.contract_2e(h2e, c, n, ne).reshape(-1)
    numpy.testing.assert_allclose(gc.cpu().numpy(), 2 * hc, atol=1e-11)


def test_autograd_c4_finite_differences():
    import torch
    n, ne = 12, (6, 6)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=0, with_ci=False)
    solver = pkg.FCISolver()
    solver.max_cycle = 200   # the synthetic (12o,12e) problem needs ~60 Davidson cycles
    e0, c = solver.kernel(h1e, eri, n, ne)
    assert solver.converged
    E, gh, ge, _ = _grads(h1e, eri, c, n, ne)
    assert abs(E - e0) < 1e-9
    # homogeneity: E is linear in (h1e, eri) at fixed c
    assert abs(numpy.sum(gh * h1e) + numpy.sum(ge * eri) - E) < 1e-9
    rng = numpy.random.default_rng(2024)
    eps = 1e-5
    ct = torch.tensor(c.reshape(-1), device="cuda")
    for _ in range(20):
        dh = rng.normal(size=h1e.shape)
        dh = 0.5 * (dh + dh.T)
        de = rng.normal(size=eri.shape)
        de = 0.5 * (de + de.T)
        with torch.no_grad():
            ep = float(pkg.autograd.ci_energy(torch.tensor(h1e + eps * dh), torch.tensor(eri + eps * de), ct, n, ne))
            em = float(pkg.autograd.ci_energy(torch.tensor(h1e - eps * dh), torch.tensor(eri - eps * de), ct, n, ne))
        fd = (ep - em) / (2 * eps)
        an = numpy.sum(gh * dh) + numpy.sum(ge * de)
        assert abs(fd - an) < 1e-7 * max(1.0, abs(an)), (fd, an)
    # FP32 integrals: promoted inside; gradient returned in float32
    E32, gh32, ge32, _ = _grads(h1e.astype(numpy.float32), eri.astype(numpy.float32), c, n, ne,
                                dtype=torch.float32)
    assert abs(E32 - E) < 1e-4
    numpy.testing.assert_allclose(gh32, gh, atol=1e-5)
