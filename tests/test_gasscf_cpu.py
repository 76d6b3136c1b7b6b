"""Host logic of the GASSCF Newton coupling (pyscf/mcscf/gasscf.py:59-97,
213-230, 541-600) that runs without a device: symmetry rejection, solver
construction and adaptation, and gen_g_hop's pass-through when there is no
spin penalty."""

from types import SimpleNamespace

import numpy
import pytest

from conftest import package

pkg = package()
gs = pkg.gasscf
fg = pkg.fci_gas


def test_symmetry_rejected():
    with pytest.raises(NotImplementedError):
        gs._check_symmetry(SimpleNamespace(fcisolver=SimpleNamespace(orbsym=[0, 1])))
    with pytest.raises(NotImplementedError):
        gs._check_symmetry(SimpleNamespace(), mol=SimpleNamespace(symmetry=True))
    with pytest.raises(NotImplementedError):
        gs._check_symmetry(SimpleNamespace(), mo_coeff=SimpleNamespace(orbsym=[0]))
    gs._check_symmetry(SimpleNamespace(fcisolver=SimpleNamespace(orbsym=None)))


def test_solver_construction_and_adaptation():
    with pytest.raises(ValueError):
        gs.new_gas_solver(None, None)
    with pytest.raises(ValueError):
        gs.new_gas_solver(None, [2, 0])
    s = gs.new_gas_solver(None, [2, 2], cache_plans=False)
    assert isinstance(s, gs._GASFCISolver) and not s.cache_plans
    assert s.gas_orbs == (2, 2)
    assert s.gen_linkstr is None
    base = fg.FCISolver(gas_orbs=(2, 3), gas_restr={"max_holes": 1})
    ad = gs.adapt_solver(base)
    assert type(ad) is gs._GASFCISolver and ad.cache_plans
    assert ad.gas_restr == {"max_holes": 1} and ad is not base
    assert type(base) is fg.FCISolver
    ad2 = gs.adapt_solver(ad, cache_plans=False)
    assert ad2 is not ad and not ad2.cache_plans and ad.cache_plans
    with pytest.raises(NotImplementedError):
        gs.adapt_solver(pkg.FCISolver())


def test_gen_g_hop_without_penalty_is_the_base_driver():
    tot = 10
    g = numpy.arange(tot, dtype=float)
    hd = numpy.ones(tot)

    def hop(x):
        return 2 * x

    def upd(u, ci):
        return numpy.zeros(tot)

    def base(mc, mo, ci0, eris, verbose):
        return g, upd, hop, hd
    sol = gs._GASFCISolver(gas_orbs=(2, 2))
    mc = SimpleNamespace(fcisolver=sol, ncas=4, nelecas=(2, 2))
    out = gs.gen_g_hop(mc, None, numpy.ones(6), None, base=base)
    assert out[0] is g and out[1] is upd and out[2] is hop and out[3] is hd
    fg.fix_spin_(sol, 0.0)
    out = gs.gen_g_hop(mc, None, numpy.ones(6), None, base=base)
    assert out[2] is hop
    fg.fix_spin_(sol, -1.0)
    with pytest.raises(ValueError):
        gs.gen_g_hop(mc, None, numpy.ones(6), None, base=base)
    fg.fix_spin_(sol, 0.2, 0.5)       # 0.5 is no S(S+1)
    with pytest.raises(ValueError):
        gs.gen_g_hop(mc, None, numpy.ones(6), None, base=base)
