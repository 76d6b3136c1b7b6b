"""GPU parity of the HIP path against the oracle and the reference's golden
vectors (libfci_gas), through the C ABI.  FP64 tolerances: sigma within
1e-13 of max|sigma| (the benchmark's "1e-13 relative"), hdiag 1e-12,
RDMs 1e-12 (the reference's own GAS-vs-CAS tolerances, test_gasci.py:272-293,
:765-786), energies 1e-10."""

import math

import numpy
import pytest

import oracle
from conftest import golden_cases, has_gpu, load_golden, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()


def _dev(x):
    import torch
    return torch.from_numpy(numpy.ascontiguousarray(x, dtype=numpy.float64).reshape(-1)).cuda()


@pytest.mark.parametrize("n,k", [(1, 0), (1, 1), (4, 2), (6, 1), (8, 4), (9, 5), (12, 6), (16, 8), (16, 1)])
def test_link_tables_bit_exact(n, k):
    sp = pkg.get_space(n, (k, k))
    for tril in (False, True):
        mine = sp.link_index(0, tril).cpu().numpy()
        ref = oracle.gen_linkstr_index(n, k, tril)
        assert mine.shape == ref.shape
        assert numpy.array_equal(mine, ref)


@pytest.mark.parametrize("name", golden_cases())
def test_sigma_hdiag_rdm_vs_golden(name):
    g = load_golden(name)
    n, ne = g["norb"], g["nelec"]
    h2e = pkg.absorb_h1e(g["h1e"], g["eri"], n, ne, 0.5)
    numpy.testing.assert_allclose(h2e, g["h2e"], rtol=0, atol=1e-14)
    sig = pkg.contract_2e(h2e, g["ci"], n, ne)
    scale = numpy.abs(g["sigma"]).max()
    numpy.testing.assert_allclose(sig, g["sigma"], rtol=0, atol=1e-13 * scale)
    hd = pkg.make_hdiag(g["h1e"], g["eri"], n, ne).reshape(g["hdiag"].shape)
    numpy.testing.assert_allclose(hd, g["hdiag"], rtol=0, atol=1e-12)
    (d1a, d1b), (aa, ab, bb) = pkg.direct_spin1._rdm12s(g["ci"], g["ket"], n, ne)
    for mine, key in ((d1a, "dm1a"), (d1b, "dm1b"), (aa, "dm2aa"), (ab, "dm2ab"), (bb, "dm2bb")):
        numpy.testing.assert_allclose(mine, g[key], rtol=0, atol=1e-12)


@pytest.mark.parametrize("n,ne,seed", [(10, (5, 5), 1), (11, (6, 4), 2), (12, (6, 6), 0), (9, (9, 0), 3), (7, (0, 3), 4)])
def test_sigma_vs_oracle_medium(n, ne, seed):
    h1e, eri, c = pkg.synthetic.make_problem(n, ne, seed=seed)
    h2e = oracle.absorb_h1e(h1e, eri, n, ne, 0.5)
    ref = oracle.contract_2e(h2e, c, n, ne)
    sig = pkg.contract_2e(h2e, c, n, ne)
    numpy.testing.assert_allclose(sig, ref, rtol=0, atol=1e-13 * numpy.abs(ref).max())


@pytest.mark.parametrize("deterministic", [False, True])
def test_sigma_tiling_and_determinism(deterministic, monkeypatch):
    """Alpha tiling (1, 84 and nstra tiles) gives the same sigma; the
    FCIS_DETERMINISTIC three-kernel path is bitwise reproducible, the fused
    path (LDS atomics in its beta scatter) reproducible to rounding."""
    import torch
    if deterministic:
        monkeypatch.setenv("FCIS_DETERMINISTIC", "1")
    n, ne = 10, (5, 4)
    h1e, eri, c = pkg.synthetic.make_problem(n, ne, seed=5)
    h2e = pkg.absorb_h1e(h1e, eri, n, ne, 0.5)
    sp = pkg.get_space(n, ne)
    cd = _dev(c)
    outs = []
    npair = n * (n + 1) // 2
    for ws in (1 << 30, 3 * 2 * npair * sp.nstrb * 8, 1):
        with pkg.SigmaPlan(sp, h2e, "fp64", ws) as plan:
            o1 = torch.empty_like(cd)
            o2 = torch.empty_like(cd)
            plan(cd, o1)
            plan(cd, o2)
            if deterministic:
                assert torch.equal(o1, o2), "sigma is not bitwise reproducible"
            else:
                assert (o1 - o2).abs().max().item() < 1e-14 * o1.abs().max().item()
            outs.append((plan.ntiles, o1.cpu().numpy()))
    assert outs[0][0] == 1 and outs[-1][0] == sp.nstra
    for nt, o in outs[1:]:
        numpy.testing.assert_allclose(o, outs[0][1], rtol=0, atol=1e-13 * numpy.abs(outs[0][1]).max())


def test_sigma_full_size_properties_c1():
    """(14o,14e): symmetry <x|Hy> = <y|Hx>, linearity and diagonal
    consistency at the full benchmark size (size-independent checks)."""
    import torch
    n, ne = 14, (7, 7)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=0, with_ci=False)
    h2e = pkg.absorb_h1e(h1e, eri, n, ne, 0.5)
    sp = pkg.get_space(n, ne)
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn(sp.ndet, dtype=torch.float64, device="cuda", generator=g)
    y = torch.randn(sp.ndet, dtype=torch.float64, device="cuda", generator=g)
    hx, hy, hz = torch.empty_like(x), torch.empty_like(x), torch.empty_like(x)
    with pkg.SigmaPlan(sp, h2e, "fp64", 8 << 30) as plan:
        plan(x, hx)
        plan(y, hy)
        plan(2.0 * x - 3.0 * y, hz)
        a, b = float(torch.dot(x, hy)), float(torch.dot(y, hx))
        assert abs(a - b) < 1e-10 * max(abs(a), 1.0)
        lin = (hz - (2.0 * hx - 3.0 * hy)).abs().max().item()
        assert lin < 1e-11 * hx.abs().max().item()
        hd = pkg.direct_spin1._make_hdiag_dev(h1e, eri, n, ne)
        for i in (0, 12345, sp.ndet // 2, sp.ndet - 1):
            e = torch.zeros_like(x)
            e[i] = 1.0
            plan(e, hz)
            assert abs(hz[i].item() - hd[i].item()) < 1e-12


def test_davidson_c0_vs_oracle():
    n, ne = 8, (4, 4)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=0, with_ci=False)
    solver = pkg.FCISolver()
    e, ci = solver.kernel(h1e, eri, n, ne)
    w, v = oracle.lowest_eigenvalues(h1e, eri, n, ne, nroots=3)
    assert solver.converged
    assert abs(e - w[0]) < 1e-10
    assert abs(numpy.linalg.norm(ci) - 1.0) < 1e-12
    assert abs(abs(float(ci.reshape(-1) @ v[0])) - 1.0) < 1e-8
    assert abs(oracle.energy(h1e, eri, ci, n, ne) - e) < 1e-10
    es, cis = solver.kernel(h1e, eri, n, ne, nroots=3)
    numpy.testing.assert_allclose(es, w[:3], atol=1e-9)


def test_davidson_small_space_exact():
    n, ne = 4, (2, 1)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=2, with_ci=False)
    from oracle import brute
    w = numpy.linalg.eigvalsh(brute.hamiltonian(h1e, eri, n, ne))
    solver = pkg.FCISolver()
    e, ci = solver.kernel(h1e, eri, n, ne, nroots=2)
    numpy.testing.assert_allclose(e, w[:2], atol=1e-12)
    e1, _ = solver.kernel(h1e, eri, n, ne, davidson_only=True)
    assert abs(e1 - w[0]) < 1e-10


def test_rdm_identities_c4_size():
    n, ne = 12, (6, 6)
    h1e, eri, c = pkg.synthetic.make_problem(n, ne, seed=0)
    dm1, dm2 = pkg.make_rdm12(c, n, ne)
    N = sum(ne)
    assert abs(numpy.trace(dm1) - N) < 1e-10
    numpy.testing.assert_allclose(numpy.einsum("pqrr->pq", dm2), (N - 1) * dm1, atol=1e-10)
    e_rdm = numpy.einsum("pq,pq", h1e, dm1) + 0.5 * numpy.einsum(
        "pqrs,pqrs", oracle.unpack_pair_matrix(eri, n), dm2)
    assert abs(e_rdm - pkg.energy(h1e, eri, c, n, ne)) < 1e-10


# Low-precision sigma: measured against the FP64 oracle with tolerances
# written here (relative to max|sigma|): FP32 t1 contraction 2e-6,
# bf16 3-term split 2e-4.
@pytest.mark.parametrize("precision,tol", [("fp32", 2e-6), ("bf16x3", 2e-4)])
@pytest.mark.parametrize("n,ne,seed", [(8, (4, 4), 0), (10, (5, 4), 1), (12, (6, 6), 2)])
def test_sigma_low_precision(precision, tol, n, ne, seed):
    h1e, eri, c = pkg.synthetic.make_problem(n, ne, seed=seed)
    h2e = oracle.absorb_h1e(h1e, eri, n, ne, 0.5)
    ref = oracle.contract_2e(h2e, c, n, ne)
    sig = pkg.contract_2e(h2e, c, n, ne, precision=precision)
    err = numpy.abs(sig - ref).max() / numpy.abs(ref).max()
    print("%s n=%d rel err %.3e" % (precision, n, err))
    assert err < tol
    assert err > 1e-12  # it really ran in reduced precision


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_mixed_precision_davidson_matches_fp64(precision):
    n, ne = 10, (5, 5)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=0, with_ci=False)
    s64 = pkg.FCISolver()
    e64, c64 = s64.kernel(h1e, eri, n, ne)
    sm = pkg.FCISolver()
    sm.precision = precision
    em, cm = sm.kernel(h1e, eri, n, ne)
    w, _ = oracle.lowest_eigenvalues(h1e, eri, n, ne)
    assert sm.converged and s64.converged
    assert abs(e64 - w[0]) < 1e-9
    assert abs(em - w[0]) < 1e-8
    assert sm.stats["nhop_low"] > 0 and sm.stats["nhop_fp64"] > 0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fused_and_unfused_paths_agree(precision, monkeypatch):
    """The fused row path (norb <= 16) and the deterministic three-kernel
    path give the same sigma to rounding."""
    import torch
    n, ne = 9, (5, 4)
    h1e, eri, c = pkg.synthetic.make_problem(n, ne, seed=8)
    h2e = pkg.absorb_h1e(h1e, eri, n, ne, 0.5)
    sp = pkg.get_space(n, ne)
    cd = _dev(c)
    with pkg.SigmaPlan(sp, h2e, precision, 1 << 28) as plan:
        a = torch.empty_like(cd)
        plan(cd, a)
    monkeypatch.setenv("FCIS_DETERMINISTIC", "1")
    with pkg.SigmaPlan(sp, h2e, precision, 1 << 28) as plan:
        b = torch.empty_like(cd)
        plan(cd, b)
    tol = 1e-13 if precision == "fp64" else 2e-6
    scale = a.abs().max().item()
    assert (a - b).abs().max().item() < tol * scale


@pytest.mark.parametrize("ne", [(2, 2), (3, 2), (1, 0)])
def test_sigma_16_orbitals_vs_oracle(ne):
    """norb = 16 (npair = 136): the fused kernels run a half-empty 9th
    tile (rows past npair zero); sigma against the oracle."""
    n = 16
    h1e, eri, c = pkg.synthetic.make_problem(n, ne, seed=3)
    h2e = pkg.absorb_h1e(h1e, eri, n, ne, 0.5)
    ref = oracle.contract_2e(oracle.absorb_h1e(h1e, eri, n, ne, 0.5), c, n, ne).reshape(-1)
    for prec, tol in (("fp64", 1e-13), ("fp32", 2e-6)):
        got = pkg.contract_2e(h2e, c, n, ne, precision=prec).reshape(-1)
        assert numpy.abs(got - ref).max() < tol * numpy.abs(ref).max(), prec


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 2e-6)])
def test_sigma_16_orbitals_fused_vs_three_kernel(precision, tol, monkeypatch):
    """The fused kernel (one or many alpha tiles) and the three-kernel path
    agree at norb = 16 (npair = 136, a half-empty 9th tile)."""
    import torch
    n, ne = 16, (5, 4)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=5, with_ci=False)
    h2e = pkg.absorb_h1e(h1e, eri, n, ne, 0.5)
    sp = pkg.get_space(n, ne)
    g = torch.Generator(device="cuda").manual_seed(7)
    c = torch.randn(sp.ndet, dtype=torch.float64, device="cuda", generator=g)
    out = {}
    for name, ws, env in (("fused", 1 << 33, {}), ("fused_tiles", 1 << 28, {}),
                          ("three", 1 << 33, {"FCIS_DETERMINISTIC": "1"})):
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        with pkg.SigmaPlan(sp, h2e, precision, ws) as plan:
            assert plan.fused == (name != "three")
            if name == "fused_tiles":
                assert plan.ntiles > 1
            out[name] = torch.empty_like(c)
            plan(c, out[name])
        for k in env:
            monkeypatch.delenv(k)
    ref = out["three"]
    scale = float(ref.abs().max())
    for name in ("fused", "fused_tiles"):
        assert float((out[name] - ref).abs().max()) < tol * scale, name


def test_sigma_16_orbitals_bf16x3_fused_vs_fp64_three_kernel(monkeypatch):
    """The bf16x3 fused kernel (seven gathering builders, register-held pair
    maps, pre-split B tile) over one and many alpha tiles against the FP64
    three-kernel sigma at norb = 16: |ds|/|s| < 2e-5 (about 6x the error of
    the 3-term split measured at (16o,16e))."""
    import torch
    n, ne = 16, (5, 4)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=6, with_ci=False)
    h2e = pkg.absorb_h1e(h1e, eri, n, ne, 0.5)
    sp = pkg.get_space(n, ne)
    g = torch.Generator(device="cuda").manual_seed(9)
    c = torch.randn(sp.ndet, dtype=torch.float64, device="cuda", generator=g)
    monkeypatch.setenv("FCIS_DETERMINISTIC", "1")
    ref = torch.empty_like(c)
    with pkg.SigmaPlan(sp, h2e, "fp64", 1 << 33) as plan:
        assert not plan.fused
        plan(c, ref)
    monkeypatch.delenv("FCIS_DETERMINISTIC")
    nref = float(torch.linalg.vector_norm(ref))
    for ws in (1 << 33, 1 << 28):
        with pkg.SigmaPlan(sp, h2e, "bf16x3", ws) as plan:
            assert plan.fused
            out = torch.empty_like(c)
            plan(c, out)
            rel = float(torch.linalg.vector_norm(out - ref)) / nref
            print("bf16x3 fused (%d tiles) vs FP64: %.3e" % (plan.ntiles, rel))
            assert 1e-12 < rel < 2e-5


def test_plan_workspace_kept_and_reused():
    """A freed plan's workspace (256 MiB and more) stays allocated, the next
    plan of that size takes it over and gives the same sigma, and
    release_plans frees what is kept."""
    import ctypes
    import torch
    lib = pkg._lib.load()
    dev = torch.cuda.current_device()

    def kept():
        b = ctypes.c_uint64()
        assert lib.fcis_kept_memory(dev, ctypes.byref(b)) == 0
        return b.value

    n, ne = 14, (7, 7)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=3, with_ci=False)
    h2e = pkg.absorb_h1e(h1e, eri, n, ne, 0.5)
    sp = pkg.get_space(n, ne)
    c = torch.randn(sp.ndet, dtype=torch.float64, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(4))
    pkg.direct_spin1.release_plans()
    assert kept() == 0
    ys = []
    for _ in range(2):
        with pkg.direct_spin1.SigmaPlan(sp, h2e, "fp64", 8 << 30) as p:
            assert kept() == 0
            y = torch.empty_like(c)
            p(c, y)
            ys.append(y)
        assert kept() >= 1 << 28
    assert torch.equal(ys[0], ys[1]) or float((ys[0] - ys[1]).abs().max()) < 1e-13 * float(ys[0].abs().max())
    pkg.direct_spin1.release_plans()
    assert kept() == 0
