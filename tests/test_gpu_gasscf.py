"""GPU parity of GASSCF's CI-side Newton coupling (pyscf/mcscf/gasscf.py:
213-600): the spin-penalty terms gen_g_hop adds to a Newton driver's
gradient / update / Hessian-vector / diagonal, and the Newton-owned solver
shell (cached contraction, RDM and S^2 plans).

The penalty terms are pinned to mathematics, not to their own formula: with
a dense S^2 (oracle/brute.py, projected on the GAS space) the penalty energy
f(c) = w <c|P|c> / <c|c> is differentiated by central differences; the
gradient terms equal grad f, the Hessian-vector terms equal the
finite-difference derivative of grad f on the tangent space of the unit
sphere (v orthogonal to c, result projected), where the reference's
symmetrised form and the exact Hessian agree.  Tolerances: 1e-12 against
the dense operator, 1e-7 against central differences (h = 1e-5)."""

import json
import os
from types import SimpleNamespace

import numpy
import pytest

import oracle.brute as brute
from oracle import gas as og
from conftest import GOLDEN, has_gpu, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()
fg = pkg.fci_gas
gs = pkg.gasscf

NGORB = 7


def load(name):
    with numpy.load(os.path.join(GOLDEN, "gas", name + ".npz")) as f:
        d = {k: f[k] for k in f.files}
    d["gas_orbs"] = tuple(int(x) for x in d["gas_orbs"])
    d["nelec"] = tuple(int(x) for x in d["nelec"])
    d["restr_type"] = str(d["restr_type"])
    d["restr"] = json.loads(str(d["restr"]))
    return d


def _dense_penalty(d, shift, ss):
    n = sum(d["gas_orbs"])
    L = og.GasLayout(d["gas_orbs"], d["nelec"], d["blocks"])
    idx = L.cas_index()
    S = brute.spin_square_matrix(n, d["nelec"])[numpy.ix_(idx, idx)]
    S = 0.5 * (S + S.T)
    na, nb = d["nelec"]
    sz = 0.5 * abs(na - nb)
    minimum = sz * (sz + 1)
    t = minimum if ss is None else ss
    D = S - t * numpy.eye(S.shape[0])
    P = shift * (D @ D if t >= minimum + 0.1 else D)
    dg = numpy.diag(S) - t
    dg = shift * (dg * dg if t >= minimum + 0.1 else dg)
    return P, dg


def _stub_base(sizes, seed):
    """A Newton driver with known orbital/CI rows: random gradient and
    diagonal, a random symmetric Hessian, a zero update."""
    rng = numpy.random.default_rng(seed)
    tot = NGORB + sum(sizes)
    g = rng.standard_normal(tot)
    hd = rng.standard_normal(tot)
    A = rng.standard_normal((tot, tot))
    A = A + A.T

    def base(mc, mo, ci0, eris, verbose):
        return g.copy(), (lambda u, ci: numpy.zeros(tot)), (lambda x: A @ x), hd.copy()
    return base, g, hd, A


def _grad_f(P, w, x):
    """Exact gradient of w <x|P|x>/<x|x> at any x."""
    xx = x.dot(x)
    f = x.dot(P @ x) / xx
    return 2 * w * (P @ x - f * x) / xx


def _solver(d, shift, ss, nroots=1):
    sol = gs._GASFCISolver(gas_orbs=d["gas_orbs"], gas_restr=d["restr"],
                           gas_restr_type=d["restr_type"])
    fg.fix_spin_(sol, shift, ss)
    sol.nroots = nroots
    return sol


@pytest.mark.parametrize("name,shift,ss", [("gas_sg_n4", 0.7, None), ("gas_sg_n4", 0.3, 2.0),
                                           ("gas_cum_n3", 0.5, None), ("gas_sg_n8", 0.2, 2.0)])
def test_penalty_terms_vs_dense_and_finite_differences(name, shift, ss):
    d = load(name)
    n = sum(d["gas_orbs"])
    P, pdiag = _dense_penalty(d, shift, ss)
    nd = P.shape[0]
    rng = numpy.random.default_rng(nd)
    c = rng.standard_normal(nd)
    c /= numpy.linalg.norm(c)
    base, g0, hd0, A = _stub_base([nd], 3)
    sol = _solver(d, shift, ss)
    mc = SimpleNamespace(fcisolver=sol, ncas=n, nelecas=d["nelec"])
    g, update, hop, hd = gs.gen_g_hop(mc, None, c, None, base=base)
    blk = slice(NGORB, NGORB + nd)
    # orbital rows are the base driver's
    numpy.testing.assert_array_equal(g[:NGORB], g0[:NGORB])
    numpy.testing.assert_array_equal(hd[:NGORB], hd0[:NGORB])
    # against the dense operator
    p = c.dot(P @ c)
    r = P @ c - p * c
    scale = max(1.0, abs(P).max())
    numpy.testing.assert_allclose(g[blk] - g0[blk], 2 * r, rtol=0, atol=1e-12 * scale)
    numpy.testing.assert_allclose(hd[blk] - hd0[blk], 2 * (pdiag - p - 2 * r * c), rtol=0,
                                  atol=1e-12 * scale)
    v = numpy.concatenate([rng.standard_normal(NGORB), rng.standard_normal(nd)])
    hv = hop(v)
    vb = v[blk]
    want = 2 * (P @ vb - p * vb - r * c.dot(vb) - c * r.dot(vb))
    numpy.testing.assert_allclose(hv - A @ v, numpy.concatenate([numpy.zeros(NGORB), want]),
                                  rtol=0, atol=1e-11 * scale)
    # pinned by calculus: gradient of f(c) = <c|P|c>/<c|c> by central differences
    h = 1e-5
    f = lambda x: x.dot(P @ x) / x.dot(x)  # noqa: E731
    fd = numpy.array([(f(c + h * e) - f(c - h * e)) / (2 * h) for e in numpy.eye(nd)])
    numpy.testing.assert_allclose(g[blk] - g0[blk], fd, rtol=0, atol=1e-7 * scale)
    # tangent-space Hessian: v orthogonal to c, result projected off c
    t = rng.standard_normal(nd)
    t -= c * c.dot(t)
    t /= numpy.linalg.norm(t)
    full = numpy.concatenate([numpy.zeros(NGORB), t])
    ht = (hop(full) - A @ full)[blk]
    fdh = (_grad_f(P, 1.0, c + h * t) - _grad_f(P, 1.0, c - h * t)) / (2 * h)
    proj = lambda y: y - c * c.dot(y)  # noqa: E731
    numpy.testing.assert_allclose(proj(ht), proj(fdh), rtol=0, atol=1e-7 * scale)
    # update at an unnormalised vector: the gradient at its direction
    u = update(numpy.zeros(NGORB), 1.7 * c)
    numpy.testing.assert_allclose(u[blk], _grad_f(P, 1.0, c), rtol=0, atol=1e-12 * scale)
    numpy.testing.assert_array_equal(u[:NGORB], 0.0)
    sol.close()


def test_state_average_weights_and_zero_shift():
    d = load("gas_sg_n4")
    n = sum(d["gas_orbs"])
    P, pdiag = _dense_penalty(d, 0.4, None)
    nd = P.shape[0]
    rng = numpy.random.default_rng(5)
    cs = [x / numpy.linalg.norm(x) for x in rng.standard_normal((2, nd))]
    base, g0, hd0, A = _stub_base([nd, nd], 9)
    sol = _solver(d, 0.4, None, nroots=2)
    w = (0.75, 0.25)
    mc = SimpleNamespace(fcisolver=sol, ncas=n, nelecas=d["nelec"], weights=w)
    g, update, hop, hd = gs.gen_g_hop(mc, None, cs, None, base=base)
    v = rng.standard_normal(NGORB + 2 * nd)
    hv = hop(v) - A @ v
    for k in range(2):
        blk = slice(NGORB + k * nd, NGORB + (k + 1) * nd)
        c = cs[k]
        p = c.dot(P @ c)
        r = P @ c - p * c
        numpy.testing.assert_allclose(g[blk] - g0[blk], 2 * w[k] * r, rtol=0, atol=1e-12)
        vb = v[blk]
        want = 2 * w[k] * (P @ vb - p * vb - r * c.dot(vb) - c * r.dot(vb))
        numpy.testing.assert_allclose(hv[blk], want, rtol=0, atol=1e-11)
        numpy.testing.assert_allclose(hd[blk] - hd0[blk], 2 * w[k] * (pdiag - p - 2 * r * c),
                                      rtol=0, atol=1e-12)
    # a zero shift returns the base driver untouched
    fg.fix_spin_(sol, 0.0)
    g2, _, hop2, _ = gs.gen_g_hop(mc, None, cs, None, base=base)
    numpy.testing.assert_array_equal(g2, g0)
    numpy.testing.assert_array_equal(hop2(v), A @ v)
    sol.close()


def test_newton_solver_shell_plans():
    d = load("gas_sg_n8")
    n = sum(d["gas_orbs"])
    ne = d["nelec"]
    sol = gs._GASFCISolver(gas_orbs=d["gas_orbs"], gas_restr=d["restr"],
                           gas_restr_type=d["restr_type"])
    ref = fg.FCISolver(gas_orbs=d["gas_orbs"], gas_restr=d["restr"],
                       gas_restr_type=d["restr_type"])
    assert sol.gen_linkstr is None and sol.transform_ci_for_orbital_rotation is None
    ci = d["ci"]
    sig = sol.contract_2e(d["h2e"], ci, n, ne)
    scale = abs(d["sigma"]).max()
    numpy.testing.assert_allclose(sig, d["sigma"], rtol=0, atol=1e-12 * scale)
    numpy.testing.assert_allclose(sig, ref.contract_2e(d["h2e"], ci, n, ne), rtol=0,
                                  atol=1e-13 * scale)
    plan = sol._get_contract_plan(d["h2e"], n, ne)
    assert len(sol._contract_plans) == 1 and sol._get_contract_plan(d["h2e"], n, ne) is plan
    # least-recently-used eviction at three plans
    others = [d["h2e"] * (1.0 + 0.1 * k) for k in range(1, 4)]
    for k, e in enumerate(others):
        s = sol.contract_2e(e, ci, n, ne)
        numpy.testing.assert_allclose(s, sig * (1.0 + 0.1 * (k + 1)), rtol=0,
                                      atol=1e-12 * scale)
    assert len(sol._contract_plans) == 3
    assert all(p is not plan for p in sol._contract_plans.values())
    # densities and S^2 through the cached plans; [ci] unwraps, two roots do not
    dm1, dm2 = sol.make_rdm12([ci], n, ne)
    r1, r2 = ref.make_rdm12(ci, n, ne)
    numpy.testing.assert_allclose(dm1, r1, rtol=0, atol=1e-12)
    numpy.testing.assert_allclose(dm2, r2, rtol=0, atol=1e-12)
    t1, t2 = sol.trans_rdm12([d["ket"]], [ci], n, ne)
    q1, q2 = ref.trans_rdm12(d["ket"], ci, n, ne)
    numpy.testing.assert_allclose(t2, q2, rtol=0, atol=1e-12)
    with pytest.raises(NotImplementedError):
        sol.make_rdm1([ci, ci], n, ne)
    numpy.testing.assert_allclose(sol.contract_ss(ci, n, ne), ref.contract_ss(ci, n, ne),
                                  rtol=0, atol=1e-13)
    assert sol._spin_plan is not None and sol._rdm_plan_cache is not None
    ss, _ = sol.spin_square([ci], n, ne)
    assert abs(ss - ref.spin_square(ci, n, ne)[0]) < 1e-12
    # copies and adapted solvers start with empty caches
    cp = sol.copy()
    assert len(cp._contract_plans) == 0 and cp._spin_plan is None
    assert len(sol._contract_plans) == 3
    ad = gs.adapt_solver(ref)
    assert isinstance(ad, gs._GASFCISolver) and ad.cache_plans and not ad._contract_plans
    numpy.testing.assert_allclose(ad.contract_2e(d["h2e"], ci, n, ne), sig, rtol=0,
                                  atol=1e-13 * scale)
    sol.close()
    sol.close()
    assert not sol._contract_plans and sol._spin_plan is None
    ad.close()

