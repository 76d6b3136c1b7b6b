"""Failure paths forced from the host (no kernel runs in them): device
allocations that cannot succeed return FCIS_ERR_MEMORY (-2, the reference's
GAS_ERR_MEMORY, pyscf/lib/mcscf/fci_gas.h:38-42), free what they had
allocated, and leave the library usable (the reference checks cleanup on
injected failures, test_gasci.py:739-752).  Also the plan cache of the
module-level contract_2e."""

import numpy
import pytest

import oracle
from conftest import has_gpu, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()


def _free():
    """Free device memory plus the plan workspaces the library keeps for
    reuse (a failed allocation releases those, so only their sum is
    invariant across the failure)."""
    import torch
    torch.cuda.synchronize()
    return pkg.direct_spin1.available_bytes()


def _still_works():
    n, ne = 6, (3, 2)
    h1e, eri, c = pkg.synthetic.make_problem(n, ne, seed=4)
    h2e = oracle.absorb_h1e(h1e, eri, n, ne, 0.5)
    ref = oracle.contract_2e(h2e, c, n, ne)
    got = pkg.contract_2e(h2e, c, n, ne)
    assert numpy.abs(got - ref).max() < 1e-13 * numpy.abs(ref).max()


def test_space_too_large_for_the_device_returns_memory_status():
    # 31 orbitals, 15 alpha electrons: 300,540,195 alpha strings x 255 links
    # x 8 B = 613 GB of link table, above the device's memory
    pkg.direct_spin1.release_plans()
    before = _free()
    with pytest.raises(RuntimeError, match="status -2"):
        pkg.direct_spin1.Space(31, (15, 1))
    assert abs(_free() - before) < (64 << 20)
    _still_works()


def test_sigma_workspace_allocation_failure_frees_the_plan():
    n, ne = 18, (9, 9)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=1, with_ci=False)
    h2e = pkg.absorb_h1e(h1e, eri, n, ne, 0.5)
    sp = pkg.direct_spin1.Space(n, ne)
    # no workspace kept from earlier plans: the measurement sees this plan only
    pkg.direct_spin1.release_plans()
    before = _free()
    # a 10 TB budget asks for one alpha tile of all 48,620 strings: t1 of 3.2 TB
    with pytest.raises(RuntimeError, match="status -2"):
        pkg.SigmaPlan(sp, h2e, "fp64", 10 << 40)
    assert abs(_free() - before) < (64 << 20)
    sp.close()
    _still_works()


def test_contract_2e_reuses_plans():
    import torch
    ds = pkg.direct_spin1
    ds.release_plans()
    n, ne = 8, (4, 3)
    h1e, eri, c = pkg.synthetic.make_problem(n, ne, seed=2)
    h2e = oracle.absorb_h1e(h1e, eri, n, ne, 0.5)
    ref = oracle.contract_2e(h2e, c, n, ne)
    a = pkg.contract_2e(h2e, c, n, ne)
    assert len(ds._PLANS) == 1
    plan = next(iter(ds._PLANS.values()))
    b = pkg.contract_2e(h2e, torch.from_numpy(c).cuda(), n, ne)
    assert len(ds._PLANS) == 1 and next(iter(ds._PLANS.values())) is plan
    # same plan; the fused kernel's LDS atomics reproduce sigma to rounding
    assert numpy.abs(a - b.cpu().numpy()).max() < 1e-14 * numpy.abs(a).max()
    assert numpy.abs(a - ref).max() < 1e-13 * numpy.abs(ref).max()
    # new Hamiltonians get their own plans; the least recently used goes
    for k in (2.0, 3.0):
        s = pkg.contract_2e(k * h2e, c, n, ne)
        assert numpy.abs(s - k * ref).max() < 1e-13 * k * numpy.abs(ref).max()
    assert len(ds._PLANS) == ds.PLAN_CACHE_SIZE
    assert all(p is not plan for p in ds._PLANS.values()) and plan.handle is None
    ds.release_plans()
    assert not ds._PLANS
