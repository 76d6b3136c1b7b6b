"""GPU parity of sector-native GAS spaces (no CAS embedding; up to the
reference's 63 orbitals, pyscf/lib/mcscf/fci_gas.h:24-27).

Checkers, all independent of the product path:
  * the reference's libfci_gas fixtures (tests/golden/gas/): layout,
    sigma, diagonal and RDMs of seven spaces;
  * the oracle's Slater-Condon H over an explicit determinant list
    (oracle.det_sigma_rows / det_hamiltonian, pinned in
    tests/test_gas_oracle.py against libfci_gas and the brute-force H):
    sampled sigma rows, dense restricted spectra, diagonals and H blocks at
    24, 40 and 62 orbitals, where no CAS array of 32-bit indices exists;
  * the CAS-embedded path on a 12-orbital space (sigma, diagonal, RDMs);
  * mathematics: hermiticity, linearity, E = h.dm1 + 1/2 eri.dm2,
    tr dm1 = N, <S^2> of a spin eigenstate.

Bars: FP64 sigma 1e-12 of max|sigma| (destination-owned pull, no atomics);
diagonal and RDMs 1e-12; energies 1e-9 (Davidson tolerance 1e-10 on the
energy change); FP32 sigma relative error below 1e-5."""

import glob
import json
import os

import numpy
import pytest

import oracle
from oracle import gas as og
from conftest import GOLDEN, has_gpu, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()
fg = pkg.fci_gas
ag = pkg.addons_gas
syn = pkg.synthetic

CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "gas", "*.npz")))


def load(name):
    with numpy.load(os.path.join(GOLDEN, "gas", name + ".npz")) as f:
        d = {k: f[k] for k in f.files}
    d["gas_orbs"] = tuple(int(x) for x in d["gas_orbs"])
    d["nelec"] = tuple(int(x) for x in d["nelec"])
    d["restr_type"] = str(d["restr_type"])
    d["restr"] = json.loads(str(d["restr"]))
    return d


def _dev(x):
    import torch
    return torch.as_tensor(numpy.ascontiguousarray(x, dtype=numpy.float64)).cuda()


def _sigma(gas, h2e, c, precision="fp64", workspace=0):
    import torch
    with fg.GasContractPlan(gas, h2e, precision, workspace) as p:
        x = _dev(c)
        out = torch.empty_like(x)
        p.execute(x, out)
        return out.cpu().numpy()


def _rdm12s(gas, bra, ket):
    plan = fg.GasRDMPlan(None, None, None, gas=gas)
    return plan.make_rdm12s(bra, ket)


def _space(orbs, nelec, restr, kind, native=True):
    o, D = ag.normalize_gas_spec(orbs, nelec, restr, kind)
    return fg.GasSpace(o, nelec, D, native=native), og.GasLayout(o, nelec, D)


# --- the reference's libfci_gas fixtures ------------------------------------

@pytest.mark.parametrize("name", CASES)
def test_native_matches_libfci_gas(name):
    d = load(name)
    print("case", name, flush=True)
    n = sum(d["gas_orbs"])
    with fg.GasSpace(d["gas_orbs"], d["nelec"], d["blocks"], native=True) as gas:
        assert gas.native and gas.cas is None
        desc = list(gas.block_descriptors())
        for k in ("offset", "sa", "sb", "na", "nb"):
            assert [b[k] for b in desc] == d["block_" + k].tolist(), k
        occ, nstr = gas.sector_occupations()
        numpy.testing.assert_array_equal(occ, d["sector_occ"])
        numpy.testing.assert_array_equal(nstr, d["sector_nstr"])
        scale = max(1.0, numpy.abs(d["sigma"]).max())
        sig = _sigma(gas, d["h2e"], d["ci"])
        assert numpy.abs(sig - d["sigma"]).max() < 1e-12 * scale
        # a workspace of a few closure rows: many tiles, same sigma
        n1, maxrow, _, _ = gas.native_info()
        tiny = 2 * 8 * (n * (n + 1) // 2) * max(maxrow, n1 // 7)
        sig_t = _sigma(gas, d["h2e"], d["ci"], workspace=tiny)
        assert numpy.abs(sig_t - sig).max() < 1e-13 * scale
        hd = fg.FCISolver(gas_orbs=d["gas_orbs"], gas_restr=d["restr"],
                          gas_restr_type=d["restr_type"])._hdiag_dev(d["h1e"], d["eri"], n,
                                                                      d["nelec"], gas)
        numpy.testing.assert_allclose(hd.cpu().numpy(), d["hdiag"], rtol=0, atol=1e-12)
        (a, b), (aa, ab, bb) = _rdm12s(gas, d["ci"], d["ket"])
        for got, key in zip((a, b, aa, ab, bb), ("dm1a", "dm1b", "dm2aa", "dm2ab", "dm2bb")):
            numpy.testing.assert_allclose(got, d[key], rtol=0, atol=1e-12, err_msg=key)
        (a, b), _ = _rdm12s(gas, d["ket"], d["ket"])
        numpy.testing.assert_allclose(numpy.trace(a), d["nelec"][0] * numpy.dot(d["ket"], d["ket"]),
                                      rtol=1e-12)


def test_native_vs_cas_embedded_12_orbitals():
    """RAS(2,8,2) with 2 holes / 2 particles, (6,6): the sector-native
    sigma, diagonal and transition RDMs equal the CAS-embedded ones (fused
    CAS kernel over the restricted rows)."""
    orbs, ne = (2, 8, 2), (6, 6)
    o, D = ag.normalize_gas_spec(orbs, ne, dict(max_holes=2, max_particles=2), "ras")
    h1e, eri = syn.make_problem(12, ne, seed=5, with_ci=False)
    h2e = fg.absorb_h1e(h1e, eri, 12, ne, 0.5)
    with fg.GasSpace(o, ne, D, native=True) as nat, fg.GasSpace(o, ne, D, native=False) as emb:
        assert nat.ndet == emb.ndet
        rng = numpy.random.default_rng(1)
        c = rng.normal(size=nat.ndet)
        c /= numpy.linalg.norm(c)
        k = rng.normal(size=nat.ndet)
        s_nat, s_emb = _sigma(nat, h2e, c), _sigma(emb, h2e, c)
        assert numpy.abs(s_nat - s_emb).max() < 1e-12 * numpy.abs(s_emb).max()
        sol = fg.FCISolver(gas_orbs=orbs)
        numpy.testing.assert_allclose(sol._hdiag_dev(h1e, eri, 12, ne, nat).cpu().numpy(),
                                      sol._hdiag_dev(h1e, eri, 12, ne, emb).cpu().numpy(),
                                      rtol=0, atol=1e-12)
        r_nat, r_emb = _rdm12s(nat, c, k), _rdm12s(emb, c, k)
        for x, y in zip(r_nat[0] + r_nat[1], r_emb[0] + r_emb[1]):
            numpy.testing.assert_allclose(x, y, rtol=0, atol=1e-12)
        # FP32 contraction against FP64
        s32 = _sigma(nat, h2e, c, "fp32")
        err = numpy.linalg.norm(s32 - s_nat) / numpy.linalg.norm(s_nat)
        assert 1e-12 < err < 1e-5


# --- beyond the CAS embedding ----------------------------------------------

def test_native_62_orbitals_dense():
    """62 orbitals in four subspaces (1, 3, 28, 30), (2,2) electrons,
    cumulative occupations: strings use bits up to 61.  Against the
    oracle's dense restricted H: sigma, diagonal, H block, the solver's
    two lowest energies, and the ground state's RDM energy identity."""
    orbs, ne = (1, 3, 28, 30), (2, 2)
    restr = [[1, 2], [3, 4], [3, 4], [4, 4]]
    gas, L = _space(orbs, ne, restr, "cumulative-occ")
    with gas:
        n = 62
        h1e, eri = syn.make_problem(n, ne, seed=7, with_ci=False)
        sa, sb = L.det_strings()
        assert int(max(sa.max(), sb.max())).bit_length() > 32
        H = oracle.det_hamiltonian(h1e, eri, n, sa, sb)
        c = numpy.random.default_rng(2).normal(size=L.ndet)
        h2e = fg.absorb_h1e(h1e, eri, n, ne, 0.5)
        ref = H @ c
        got = _sigma(gas, h2e, c)
        assert numpy.abs(got - ref).max() < 1e-12 * numpy.abs(ref).max()
        sol = fg.FCISolver(gas_orbs=orbs, gas_restr=restr, gas_restr_type="cumulative-occ")
        numpy.testing.assert_allclose(sol.make_hdiag(h1e, eri, n, ne), numpy.diag(H),
                                      rtol=0, atol=1e-12 * numpy.abs(numpy.diag(H)).max())
        pick = numpy.sort(numpy.random.default_rng(3).choice(L.ndet, 60, replace=False))
        Hb = sol._h_block(gas, h1e, eri, pick)
        numpy.testing.assert_allclose(Hb, H[numpy.ix_(pick, pick)], rtol=0,
                                      atol=1e-12 * numpy.abs(H).max())
        w = numpy.linalg.eigvalsh(H)
        sol.conv_tol = 1e-11
        e, ci = sol.kernel(h1e, eri, n, ne, nroots=2, max_cycle=300)
        print("stats", sol.stats, flush=True)
        assert sol.converged.all() and gas.native
        numpy.testing.assert_allclose(e, w[:2], rtol=0, atol=1e-8)
        dm1, dm2 = sol.make_rdm12(ci[0], n, ne)
        e_rdm = numpy.einsum("pq,pq", h1e, dm1) + 0.5 * numpy.einsum("pqrs,pqrs", _eri4(eri, n), dm2)
        assert abs(e_rdm - w[0]) < 1e-8
        assert abs(numpy.trace(dm1) - sum(ne)) < 1e-10


def _eri4(eri, n):
    """(n,n,n,n) integrals from the 4-fold pair matrix."""
    idx = numpy.zeros((n, n), dtype=numpy.int64)
    for p in range(n):
        for q in range(n):
            hi, lo = max(p, q), min(p, q)
            idx[p, q] = hi * (hi + 1) // 2 + lo
    return eri[idx[:, :, None, None], idx[None, None, :, :]]


def test_native_40_orbitals_rows():
    """40 orbitals in (2, 6, 12, 20), (4,4) electrons, 51,345 determinants
    whose closure is 52 times larger: sigma against the oracle's
    Slater-Condon rows at extreme and random determinants, the diagonal at
    sampled determinants, hermiticity and linearity over the whole space,
    FP32 error bound."""
    orbs, ne = (2, 6, 12, 20), (4, 4)
    restr = [[3, 4], [7, 8], [7, 8], [8, 8]]
    gas, L = _space(orbs, ne, restr, "cumulative-occ")
    with gas:
        n = 40
        h1e, eri = syn.make_problem(n, ne, seed=8, with_ci=False)
        h2e = fg.absorb_h1e(h1e, eri, n, ne, 0.5)
        rng = numpy.random.default_rng(4)
        c = rng.normal(size=L.ndet)
        c /= numpy.linalg.norm(c)
        print("ndet", L.ndet, "closure", gas.native_info()[0], flush=True)
        sig = _sigma(gas, h2e, c)
        sa, sb = L.det_strings()
        rows = numpy.unique(numpy.concatenate([[0, L.ndet - 1, int(numpy.argmax(numpy.abs(sig)))],
                                               rng.integers(0, L.ndet, 250)]))
        ref = oracle.det_sigma_rows(h1e, eri, c, n, sa, sb, rows)
        assert numpy.abs(sig[rows] - ref).max() < 1e-12 * numpy.abs(sig).max()
        # hermiticity and linearity
        y = rng.normal(size=L.ndet)
        sy = _sigma(gas, h2e, y)
        assert abs(y @ sig - c @ sy) < 1e-12 * numpy.linalg.norm(sig) * numpy.linalg.norm(y)
        s2 = _sigma(gas, h2e, 2.0 * c - 0.5 * y)
        assert numpy.abs(s2 - (2.0 * sig - 0.5 * sy)).max() < 1e-12 * numpy.abs(s2).max()
        # diagonal at sampled determinants: rows of unit vectors
        sol = fg.FCISolver(gas_orbs=orbs, gas_restr=restr, gas_restr_type="cumulative-occ")
        hd = sol.make_hdiag(h1e, eri, n, ne)
        for dd in rows[:12]:
            e = numpy.zeros(L.ndet)
            e[dd] = 1.0
            ref_d = oracle.det_sigma_rows(h1e, eri, e, n, sa, sb, [dd])[0]
            assert abs(hd[dd] - ref_d) < 1e-12 * abs(ref_d)
        s32 = _sigma(gas, h2e, c, "fp32")
        err = numpy.linalg.norm(s32 - sig) / numpy.linalg.norm(sig)
        assert err < 1e-5


def test_native_solver_24_orbitals_mixed_and_spin(monkeypatch):
    """RAS(4,6,14), 2 holes / 2 particles, (5,5): 315,616 determinants,
    solved by the default solver (sector-native above 16 orbitals, operator
    form).  An FP32-mixed solve through the closure GEMM form (FP32 product,
    FP64 checks) agrees with the FP64 one to 1e-8 Ha; on the operator form,
    which computes in FP64, an FP32 request runs the FP64 solve; sigma rows
    of the ground state against the oracle; the RAS space is spin complete,
    so the ground state's <S^2> from its RDMs is S(S+1)."""
    orbs, ne = (4, 6, 14), (5, 5)
    restr = dict(max_holes=2, max_particles=2)
    n = 24
    h1e, eri = syn.make_problem(n, ne, seed=9, with_ci=False)
    sol = fg.FCISolver(gas_orbs=orbs, gas_restr=restr, gas_restr_type="ras")
    e64, c64 = sol.kernel(h1e, eri, n, ne, max_cycle=300)
    print("fp64", sol.stats, flush=True)
    assert sol.converged and sol._space(n, ne).native
    monkeypatch.setenv("FCIS_GAS_NATIVE_PATH", "t1")
    sol32 = fg.FCISolver(gas_orbs=orbs, gas_restr=restr, gas_restr_type="ras")
    sol32.precision = "fp32"
    e32, c32 = sol32.kernel(h1e, eri, n, ne, max_cycle=300)
    print("fp32", sol32.stats, flush=True)
    assert sol32.converged and abs(e32 - e64) < 1e-8
    assert sol32.stats["nhop_fp64"] >= 1
    monkeypatch.delenv("FCIS_GAS_NATIVE_PATH")
    solx = fg.FCISolver(gas_orbs=orbs, gas_restr=restr, gas_restr_type="ras")
    solx.precision = "fp32"
    ex, cx = solx.kernel(h1e, eri, n, ne, max_cycle=300)
    assert solx.stats.get("precision_requested") == "fp32" and "nhop_fp64" not in solx.stats
    assert solx.converged and abs(ex - e64) < 1e-10
    o, D = ag.normalize_gas_spec(orbs, ne, restr, "ras")
    L = og.GasLayout(o, ne, D)
    sa, sb = L.det_strings()
    rows = numpy.random.default_rng(5).integers(0, L.ndet, 100)
    hc = oracle.det_sigma_rows(h1e, eri, c64, n, sa, sb, rows)
    assert numpy.abs(hc - e64 * c64[rows]).max() < 1e-5
    ss, _ = sol.spin_square(c64, n, ne)
    assert min(abs(ss - S * (S + 1)) for S in range(4)) < 1e-6


def test_native_rejects_cas_only_operations():
    orbs, ne = (2, 8, 2), (6, 6)
    gas, _ = _space(orbs, ne, dict(max_holes=1, max_particles=1), "ras")
    with gas:
        import torch
        x = torch.zeros(gas.ndet, dtype=torch.float64, device="cuda")
        with pytest.raises(RuntimeError):
            fg.GasContractPlan(gas, numpy.zeros((78, 78)), "bf16x3")
        with pytest.raises(RuntimeError):
            gas.embed_dev(x, torch.empty(10, dtype=torch.float64, device="cuda"))


@pytest.mark.skipif(not oracle.ref_available(), reason="reference libfci_gas build absent")
@pytest.mark.parametrize("orbs,ne,restr,kind", [
    ((4, 6, 14), (5, 5), dict(max_holes=2, max_particles=2), "ras"),
    ((2, 6, 12, 20), (4, 4), [[3, 4], [7, 8], [7, 8], [8, 8]], "cumulative-occ"),
    ((1, 3, 28, 30), (2, 2), [[1, 2], [3, 4], [3, 4], [4, 4]], "cumulative-occ"),
])
def test_native_vs_reference_libfci_gas_large(orbs, ne, restr, kind):
    """The reference's own libfci_gas (compiled from its sources into
    oracle/_ref) on the same spaces at 24, 40 and 62 orbitals: layout,
    whole sigma vectors, diagonal and transition RDMs."""
    n = sum(orbs)
    o, D = ag.normalize_gas_spec(orbs, ne, restr, kind)
    R = oracle.RefLib().gas(o, ne, D)
    h1e, eri = syn.make_problem(n, ne, seed=11, with_ci=False)
    h2e = fg.absorb_h1e(h1e, eri, n, ne, 0.5)
    with fg.GasSpace(o, ne, D, native=True) as gas:
        print("case", orbs, ne, "ndet", gas.ndet, flush=True)
        assert gas.ndet == R.ndet and gas.nsector == R.nsector
        desc = list(gas.block_descriptors())
        assert [(b["offset"], b["sa"], b["sb"], b["na"], b["nb"]) for b in desc] == \
            [(b["offset"], b["sa"], b["sb"], b["na"], b["nb"]) for b in R.block_desc]
        rng = numpy.random.default_rng(6)
        c = rng.normal(size=gas.ndet)
        c /= numpy.linalg.norm(c)
        k = rng.normal(size=gas.ndet)
        k /= numpy.linalg.norm(k)
        ref = R.contract_2e(h2e, c)
        got = _sigma(gas, h2e, c)
        assert numpy.abs(got - ref).max() < 1e-12 * numpy.abs(ref).max()
        sol = fg.FCISolver(gas_orbs=orbs)
        hd = sol._hdiag_dev(h1e, eri, n, ne, gas).cpu().numpy()
        hr = R.make_hdiag(h1e, eri)
        assert numpy.abs(hd - hr).max() < 1e-12 * numpy.abs(hr).max()
        (a, b), (aa, ab, bb) = _rdm12s(gas, c, k)
        (ra, rb), (raa, rab, rbb) = R.trans_rdm12s(c, k)
        for x, y in zip((a, b, aa, ab, bb), (ra, rb, raa, rab, rbb)):
            assert numpy.abs(x - y).max() < 1e-12


# --- S^2 and the spin penalty on sector-native spaces ----------------------

@pytest.mark.parametrize("name", ["gas_sg_n4", "gas_ras_n4", "gas_cum_n3", "gas_sg_n8"])
def test_native_contract_ss_vs_ladder_construction(name):
    from oracle import brute
    d = load(name)
    L = og.GasLayout(d["gas_orbs"], d["nelec"], d["blocks"])
    sa, sb = L.det_strings()
    S = brute.det_spin_square_matrix(sa, sb, d["nelec"])
    with fg.GasSpace(d["gas_orbs"], d["nelec"], d["blocks"], native=True) as gas:
        plan = fg.GasSpinPlan(gas)
        numpy.testing.assert_allclose(plan.contract(d["ci"]), S @ d["ci"], rtol=0, atol=1e-13)
        numpy.testing.assert_allclose(plan.diagonal_vector(), numpy.diag(S), rtol=0, atol=1e-14)


def test_native_spin_62_orbitals_penalty():
    """62 orbitals, 1767 determinants: S^2 against the ladder construction
    over the determinant list; the spin-penalised solver (fix_spin_) reaches
    the lowest triplet and singlet of the dense restricted H within each
    spin block."""
    from oracle import brute
    orbs, ne = (1, 3, 28, 30), (2, 2)
    restr = [[1, 2], [3, 4], [3, 4], [4, 4]]
    o, D = ag.normalize_gas_spec(orbs, ne, restr, "cumulative-occ")
    assert ag.is_spin_complete(o, ne, D)
    L = og.GasLayout(o, ne, D)
    sa, sb = L.det_strings()
    S = brute.det_spin_square_matrix(sa, sb, ne)
    n = 62
    h1e, eri = syn.make_problem(n, ne, seed=12, with_ci=False)
    H = oracle.det_hamiltonian(h1e, eri, n, sa, sb)
    with fg.GasSpace(o, ne, D, native=True) as gas:
        x = numpy.random.default_rng(7).normal(size=L.ndet)
        plan = fg.GasSpinPlan(gas)
        numpy.testing.assert_allclose(plan.contract(x), S @ x, rtol=0, atol=1e-12)
    w, v = numpy.linalg.eigh(S)
    levels = {}
    for ss in (0.0, 2.0):
        P = v[:, abs(w - ss) < 1e-8]
        levels[ss] = numpy.linalg.eigvalsh(P.T @ H @ P)[0]
    # these integrals put the lowest triplet 24 Ha below the lowest singlet:
    # the linear singlet penalty needs a shift above half that gap
    for ss, shift in ((2.0, 1.0), (None, 40.0)):
        sol = fg.fix_spin_(fg.FCISolver(gas_orbs=orbs, gas_restr=restr,
                                        gas_restr_type="cumulative-occ"), shift, ss)
        e, c = sol.kernel(h1e, eri, n, ne, max_cycle=600, max_space=24)
        print("target", ss, sol.stats, flush=True)
        assert sol.converged and sol._space(n, ne).native
        ref = levels[0.0 if ss is None else ss]
        assert abs(sol.e_physical - ref) < 1e-8, (ss, sol.e_physical, ref)


def test_native_spin_24_orbitals_commutes_and_matches_rdm():
    """RAS(4,6,14) (5,5): S^2 commutes with H on the (spin-complete) space,
    and <x|S^2|x> equals the RDM formula of spin_square for a random x."""
    import torch
    orbs, ne = (4, 6, 14), (5, 5)
    restr = dict(max_holes=2, max_particles=2)
    n = 24
    o, D = ag.normalize_gas_spec(orbs, ne, restr, "ras")
    h1e, eri = syn.make_problem(n, ne, seed=13, with_ci=False)
    h2e = fg.absorb_h1e(h1e, eri, n, ne, 0.5)
    with fg.GasSpace(o, ne, D, native=True) as gas:
        x = numpy.random.default_rng(8).normal(size=gas.ndet)
        x /= numpy.linalg.norm(x)
        plan = fg.GasSpinPlan(gas)
        sx = plan.contract(x)
        hx = _sigma(gas, h2e, x)
        lhs = plan.contract(hx)
        rhs = _sigma(gas, h2e, sx)
        assert numpy.abs(lhs - rhs).max() < 1e-11 * numpy.abs(lhs).max()
        ss_rdm, _ = fg.FCISolver(gas_orbs=orbs, gas_restr=restr,
                                 gas_restr_type="ras").spin_square(x, n, ne)
        assert abs(float(x @ sx) - ss_rdm) < 1e-10


# --- the two sector-native sigma forms ---------------------------------------

def _sigma_path(monkeypatch, gas, h2e, c, path):
    monkeypatch.setenv("FCIS_GAS_NATIVE_PATH", path)
    import torch
    with fg.GasContractPlan(gas, h2e, "fp64") as p:
        assert p.path == "native-" + path
        x = _dev(c)
        out = torch.empty_like(x)
        p.execute(x, out)
        return out.cpu().numpy()


@pytest.mark.parametrize("path", ["direct", "links"])
@pytest.mark.parametrize("name", CASES)
def test_native_direct_path_matches_libfci_gas(monkeypatch, name, path):
    """The direct forms (operator form, link-pair loops) on the reference's
    fixtures."""
    d = load(name)
    with fg.GasSpace(d["gas_orbs"], d["nelec"], d["blocks"], native=True) as gas:
        sig = _sigma_path(monkeypatch, gas, d["h2e"], d["ci"], path)
        scale = max(1.0, numpy.abs(d["sigma"]).max())
        assert numpy.abs(sig - d["sigma"]).max() < 1e-12 * scale


@pytest.mark.parametrize("orbs,ne,restr,kind", [
    ((2, 8, 2), (6, 6), dict(max_holes=2, max_particles=2), "ras"),
    ((4, 6, 14), (5, 5), dict(max_holes=2, max_particles=2), "ras"),
    ((2, 6, 12, 20), (4, 4), [[3, 4], [7, 8], [7, 8], [8, 8]], "cumulative-occ"),
    ((1, 3, 28, 30), (2, 2), [[1, 2], [3, 4], [3, 4], [4, 4]], "cumulative-occ"),
])
def test_native_direct_and_t1_paths_agree(monkeypatch, orbs, ne, restr, kind):
    """The three sector-native sigma forms give the same restricted sigma
    (12 to 62 orbitals), and it matches the oracle's Slater-Condon rows."""
    n = sum(orbs)
    o, D = ag.normalize_gas_spec(orbs, ne, restr, kind)
    L = og.GasLayout(o, ne, D)
    h1e, eri = syn.make_problem(n, ne, seed=14, with_ci=False)
    h2e = fg.absorb_h1e(h1e, eri, n, ne, 0.5)
    with fg.GasSpace(o, ne, D, native=True) as gas:
        c = numpy.random.default_rng(15).normal(size=gas.ndet)
        s_t1 = _sigma_path(monkeypatch, gas, h2e, c, "t1")
        s_dir = _sigma_path(monkeypatch, gas, h2e, c, "direct")
        s_lnk = _sigma_path(monkeypatch, gas, h2e, c, "links")
        scale = numpy.abs(s_t1).max()
        assert numpy.abs(s_dir - s_t1).max() < 1e-12 * scale
        assert numpy.abs(s_lnk - s_t1).max() < 1e-12 * scale
        sa, sb = L.det_strings()
        rows = numpy.random.default_rng(16).integers(0, gas.ndet, 80)
        ref = oracle.det_sigma_rows(h1e, eri, c, n, sa, sb, rows)
        assert numpy.abs(s_dir[rows] - ref).max() < 1e-12 * scale


def test_native_t1_tiles_within_gemm_span_105m_determinants(monkeypatch):
    """GAS (8,8) (8,8), at least 4 electrons in the first 8 orbitals:
    105,164,282 determinants, closure 141M columns.  A 96 GB workspace would
    give t1 tiles whose GEMM operands span 6.4e9 elements, where rocBLAS
    returned wrong sigma without an error; the plan keeps tiles below 2^32
    elements and the t1 form agrees with the operator form."""
    import torch
    orbs, ne, restr = (8, 8), (8, 8), [[4, 8], [16, 16]]
    n = sum(orbs)
    o, D = ag.normalize_gas_spec(orbs, ne, restr, "cumulative-occ")
    h1e, eri = syn.make_problem(n, ne, seed=0, with_ci=False)
    h2e = fg.absorb_h1e(h1e, eri, n, ne, 0.5)
    with fg.GasSpace(o, ne, D, native=True) as gas:
        assert gas.ndet == 105164282
        g = torch.Generator(device="cuda").manual_seed(1)
        x = torch.randn(gas.ndet, dtype=torch.float64, device="cuda", generator=g)
        monkeypatch.setenv("FCIS_GAS_NATIVE_PATH", "direct")
        with fg.GasContractPlan(gas, h2e, "fp64") as p:
            ref = torch.empty_like(x)
            p.execute(x, ref)
        monkeypatch.setenv("FCIS_GAS_NATIVE_PATH", "t1")
        with fg.GasContractPlan(gas, h2e, "fp64", 96 << 30) as p:
            assert p.path == "native-t1"
            y = torch.empty_like(x)
            p.execute(x, y)
        scale = float(ref.abs().max())
        assert float((y - ref).abs().max()) < 1e-12 * scale


def test_native_operator_form_reproducible_and_chosen():
    """The operator form is destination-owned with a fixed summation order:
    two applications agree bitwise, and a second plan on the same space
    (tables reused, entries recomputed) gives the same bits.  It is the
    plan's own choice at 40 orbitals."""
    import torch
    orbs, ne, restr = (2, 8, 30), (4, 4), dict(max_holes=1, max_particles=2)
    n = sum(orbs)
    o, D = ag.normalize_gas_spec(orbs, ne, restr, "ras")
    h1e, eri = syn.make_problem(n, ne, seed=21, with_ci=False)
    h2e = fg.absorb_h1e(h1e, eri, n, ne, 0.5)
    with fg.GasSpace(o, ne, D, native=True) as gas:
        x = _dev(numpy.random.default_rng(22).normal(size=gas.ndet))
        outs = []
        for _ in range(2):
            with fg.GasContractPlan(gas, h2e, "fp64") as p:
                assert p.path == "native-direct"
                for _ in range(2):
                    y = torch.empty_like(x)
                    p.execute(x, y)
                    outs.append(y)
        for y in outs[1:]:
            assert torch.equal(y, outs[0])


def test_native_transform_ci_within_gas_24_orbitals():
    """transform_ci_within_gas on a sector-native space (RAS(4,6,14) (5,5),
    315,616 determinants): the rotated vector's RDMs (sector-native RDM
    kernels) are the rotated RDMs, and the energy in the rotated integrals
    (sector-native sigma) is unchanged."""
    orbs, ne = (4, 6, 14), (5, 5)
    restr = dict(max_holes=2, max_particles=2)
    n = 24
    h1e, eri = syn.make_problem(n, ne, seed=17, with_ci=False)
    sol = fg.FCISolver(gas_orbs=orbs, gas_restr=restr, gas_restr_type="ras")
    assert sol._space(n, ne).native
    rng = numpy.random.default_rng(18)
    u = numpy.zeros((n, n))
    o = 0
    for k in orbs:
        q, _ = numpy.linalg.qr(rng.standard_normal((k, k)))
        u[o:o + k, o:o + k] = q
        o += k
    c = rng.normal(size=sol._space(n, ne).ndet)
    c /= numpy.linalg.norm(c)
    c1 = sol.transform_ci_within_gas(c, n, ne, u)
    assert abs(numpy.linalg.norm(c1) - 1.0) < 1e-12
    (a0, b0), (aa0, ab0, bb0) = sol.make_rdm12s(c, n, ne)
    (a1, b1), (aa1, ab1, bb1) = sol.make_rdm12s(c1, n, ne)
    rot4 = lambda t: numpy.einsum("ijkl,ip,jq,kr,ls->pqrs", t, u, u, u, u, optimize=True)
    numpy.testing.assert_allclose(a1, u.T @ a0 @ u, rtol=0, atol=1e-12)
    numpy.testing.assert_allclose(b1, u.T @ b0 @ u, rtol=0, atol=1e-12)
    for x1, x0 in ((aa1, aa0), (ab1, ab0), (bb1, bb0)):
        numpy.testing.assert_allclose(x1, rot4(x0), rtol=0, atol=1e-12)
    g = _eri4(eri, n)
    e0 = sol.energy(h1e, g, c, n, ne)
    e1 = sol.energy(u.T @ h1e @ u, rot4(g), c1, n, ne)
    assert abs(e1 - e0) < 1e-10 * max(1.0, abs(e0))
