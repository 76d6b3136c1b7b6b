"""GPU parity of the S^2 contraction (fcis_spin_plan_*, fcis_gas_contract_ss)
and of the GAS solver's spin penalty (fix_spin_), against a from-scratch
S^2 matrix (oracle/brute.py: S_z(S_z+1) + S_- S_+ built from operator
action on bit strings) and dense restricted spectra.

Tolerances: S^2 elements 1e-13 (integer-valued signs), energies 1e-9,
physical energies of penalised solves 1e-8."""

import json
import os

import numpy
import pytest

import oracle
import oracle.brute as brute
from oracle import gas as og
from conftest import GOLDEN, has_gpu, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()
fg = pkg.fci_gas


def load(name):
    with numpy.load(os.path.join(GOLDEN, "gas", name + ".npz")) as f:
        d = {k: f[k] for k in f.files}
    d["gas_orbs"] = tuple(int(x) for x in d["gas_orbs"])
    d["nelec"] = tuple(int(x) for x in d["nelec"])
    d["restr_type"] = str(d["restr_type"])
    d["restr"] = json.loads(str(d["restr"]))
    return d


@pytest.mark.parametrize("n,ne", [(4, (2, 2)), (5, (3, 2)), (6, (3, 3)), (6, (4, 1)),
                                  (5, (5, 0)), (4, (1, 3)), (7, (4, 3))])
def test_cas_contract_ss_vs_brute_force(n, ne):
    import torch
    S = brute.spin_square_matrix(n, ne)
    nd = S.shape[0]
    rng = numpy.random.default_rng(n + 10 * ne[0])
    X = rng.standard_normal((3, nd))
    for x in X:
        numpy.testing.assert_allclose(pkg.contract_ss(x, n, ne), S @ x, rtol=0, atol=1e-13)
    sp = pkg.get_space(n, ne)
    with pkg.SpinPlan(sp) as plan:
        d = plan.diag(torch.empty(nd, dtype=torch.float64, device="cuda"))
    numpy.testing.assert_allclose(d.cpu().numpy(), numpy.diag(S), rtol=0, atol=1e-14)


def test_cas_s2_commutes_with_h_and_matches_rdm_value():
    import torch
    n, ne = 12, (6, 5)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=8, with_ci=False)
    h2e = pkg.absorb_h1e(h1e, eri, n, ne, 0.5)
    sp = pkg.get_space(n, ne)
    g = torch.Generator(device="cuda").manual_seed(2)
    x = torch.randn(sp.ndet, dtype=torch.float64, device="cuda", generator=g)
    hx, sx, shx, hsx = (torch.empty_like(x) for _ in range(4))
    with pkg.SigmaPlan(sp, h2e) as hp, pkg.SpinPlan(sp) as spn:
        hp(x, hx)
        spn.contract(x, sx)
        spn.contract(hx, shx)
        hp(sx, hsx)
    assert float((shx - hsx).abs().max()) < 1e-11 * float(hx.abs().max())
    # <x|S^2|x> / <x|x> equals the RDM expression
    xn = (x / x.norm()).cpu().numpy()
    ss_rdm, _ = pkg.FCISolver().spin_square(xn.reshape(sp.shape), n, ne)
    ss_op = float(torch.dot(x, sx) / torch.dot(x, x))
    assert abs(ss_rdm - ss_op) < 1e-10


@pytest.mark.parametrize("name", ["gas_sg_n4", "gas_ras_n4", "gas_cum_n3", "gas_sg_n8"])
def test_gas_contract_ss_vs_projected_brute_force(name):
    d = load(name)
    sol = fg.FCISolver(gas_orbs=d["gas_orbs"], gas_restr=d["restr"],
                       gas_restr_type=d["restr_type"])
    n = sum(d["gas_orbs"])
    L = og.GasLayout(d["gas_orbs"], d["nelec"], d["blocks"])
    idx = L.cas_index()
    S = brute.spin_square_matrix(n, d["nelec"])[numpy.ix_(idx, idx)]
    got = sol.contract_ss(d["ci"], n, d["nelec"])
    numpy.testing.assert_allclose(got, S @ d["ci"], rtol=0, atol=1e-13)
    plan = sol.make_spin_plan(n, d["nelec"])
    numpy.testing.assert_allclose(plan.diagonal_vector(), numpy.diag(S), rtol=0, atol=1e-14)


def test_gas_contract_ss_requires_spin_complete():
    d = load("gas_ssg_n4")        # D = {(1,1 | 1,0)}: not spin complete
    sol = fg.FCISolver(gas_orbs=d["gas_orbs"], gas_restr=d["restr"],
                       gas_restr_type=d["restr_type"])
    with pytest.raises(ValueError):
        sol.contract_ss(d["ci"], 4, d["nelec"])
    fg.fix_spin_(sol, 0.2)
    with pytest.raises(ValueError):
        sol.kernel(d["h1e"], d["eri"], 4, d["nelec"])


def _spin_levels(H, S, ss):
    w, v = numpy.linalg.eigh(0.5 * (S + S.T))
    P = v[:, abs(w - ss) < 1e-8]
    return numpy.linalg.eigvalsh(P.T @ H @ P)


def test_gas_spin_penalty_exact_small_space():
    d = load("gas_sg_n4")
    n = 4
    L = og.GasLayout(d["gas_orbs"], d["nelec"], d["blocks"])
    idx = L.cas_index()
    H = brute.hamiltonian(d["h1e"], d["eri"], n, d["nelec"])[numpy.ix_(idx, idx)]
    S = brute.spin_square_matrix(n, d["nelec"])[numpy.ix_(idx, idx)]
    for ss, shift in [(None, 0.5), (2.0, 0.5)]:
        sol = fg.fix_spin_(fg.FCISolver(gas_orbs=d["gas_orbs"], gas_restr=d["restr"],
                                        gas_restr_type=d["restr_type"]), shift, ss)
        e, c = sol.kernel(d["h1e"], d["eri"], n, d["nelec"])
        assert sol.spin_penalty_method == "exact-small-space"
        ref = _spin_levels(H, S, 0.0 if ss is None else ss)[0]
        assert abs(sol.e_physical - ref) < 1e-9, (ss, sol.e_physical, ref)
        assert abs(sol.e_spin_penalty) < 1e-9


def test_gas_spin_penalty_davidson_targets_triplet():
    d = load("gas_sg_n8")
    n = 8
    L = og.GasLayout(d["gas_orbs"], d["nelec"], d["blocks"])
    idx = L.cas_index()
    H = og.gas_hamiltonian(d["h2e"], L)
    S = brute.spin_square_matrix(n, d["nelec"])[numpy.ix_(idx, idx)]
    t_ref = _spin_levels(H, S, 2.0)[0]
    s_ref = _spin_levels(H, S, 0.0)[0]
    for ss, ref in [(2.0, t_ref), (None, s_ref)]:
        print("target", ss, flush=True)
        sol = fg.fix_spin_(fg.FCISolver(gas_orbs=d["gas_orbs"], gas_restr=d["restr"],
                                        gas_restr_type=d["restr_type"]), 1.0, ss)
        # the quadratic penalty (S^2 - 2)^2 with its diagonal-only
        # preconditioner is stiff on these random integrals: 24-vector space
        e, c = sol.kernel(d["h1e"], d["eri"], n, d["nelec"], max_cycle=600, max_space=24)
        assert sol.converged
        assert sol.spin_penalty_method == "projected-plus-global-davidson"
        assert abs(sol.e_physical - ref) < 1e-8, (ss, sol.e_physical, ref)
        ssv, _ = sol.spin_square(c, n, d["nelec"])
        assert abs(ssv - (0.0 if ss is None else ss)) < 1e-6
