"""Host logic of the plan caches (no device): the module-level contract_2e
cache and the Newton solver's plan cache, with stand-in plans.  Covers
recency eviction, closing on release, topology changes, and recovery when a
plan cannot be created (the retry after freeing the cached workspaces)."""

from types import SimpleNamespace

import numpy
import pytest

from conftest import package

pkg = package()
ds = pkg.direct_spin1
gs = pkg.gasscf


class FakePlan:
    made = []

    def __init__(self, *args, **kwargs):
        self.closed = False
        self.workspace_bytes = 0
        FakePlan.made.append(self)

    def close(self):
        self.closed = True


def test_contract_plan_cache_recency_and_retry(monkeypatch):
    ds.release_plans()
    FakePlan.made = []
    monkeypatch.setattr(ds, "SigmaPlan", FakePlan)
    sp = SimpleNamespace(norb=2, nelec=(1, 1))
    h = [numpy.eye(3) * k for k in (1.0, 2.0, 3.0)]
    p0 = ds._cached_plan(sp, h[0], "fp64", 1 << 20)
    p1 = ds._cached_plan(sp, h[1], "fp64", 1 << 20)
    assert ds._cached_plan(sp, h[0], "fp64", 1 << 20) is p0      # hit: p0 most recent
    p2 = ds._cached_plan(sp, h[2], "fp64", 1 << 20)              # evicts p1
    assert p1.closed and not p0.closed and not p2.closed
    assert list(ds._PLANS.values()) == [p0, p2]
    # a creation failure frees the cached plans and tries once more
    calls = {"n": 0}

    def flaky(*a, **k):
        calls["n"] += 1
        if calls["n"] == 1:
            raise RuntimeError("fcis_sigma_plan_create failed with status -2")
        return FakePlan()
    monkeypatch.setattr(ds, "SigmaPlan", flaky)
    p3 = ds._cached_plan(sp, h[1], "fp32", 1 << 20)
    assert calls["n"] == 2 and p0.closed and p2.closed
    assert list(ds._PLANS.values()) == [p3]

    def broken(*a, **k):
        raise RuntimeError("fcis_sigma_plan_create failed with status -2")
    monkeypatch.setattr(ds, "SigmaPlan", broken)
    with pytest.raises(RuntimeError, match="status -2"):
        ds._cached_plan(sp, h[2], "fp32", 1 << 20)
    assert not ds._PLANS and p3.closed
    ds.release_plans()


def test_newton_plan_cache():
    cache = gs._NewtonPlanCache(limit=2)
    spaces = []

    def make_space():
        spaces.append(FakePlan())
        return spaces[-1]
    cache.bind("topology A")
    a = cache.contract_plan("ka", make_space, lambda s: FakePlan())
    b = cache.contract_plan("kb", make_space, lambda s: FakePlan())
    assert cache.contract_plan("ka", make_space, lambda s: FakePlan()) is a
    c = cache.contract_plan("kc", make_space, lambda s: FakePlan())
    assert b.closed and not a.closed and list(cache.contract) == ["ka", "kc"]
    assert len(spaces) == 1
    # a failing plan leaves the cache consistent
    with pytest.raises(RuntimeError):
        cache.contract_plan("kd", make_space, lambda s: (_ for _ in ()).throw(RuntimeError("x")))
    assert all(not p.closed for p in cache.contract.values())
    cache.rdm = FakePlan()
    cache.spin = object()
    cache.bind("topology A")                      # same topology: nothing released
    assert not cache.rdm.closed
    rdm = cache.rdm
    cache.bind("topology B")                      # new topology: everything closed
    assert rdm.closed and c.closed and spaces[0].closed
    assert cache.rdm is None and cache.spin is None and not cache.contract
    cache.release()
    cache.release()
