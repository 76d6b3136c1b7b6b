"""GPU parity of the batched path: one persistent launch computing sigma for
many CAS problems and vector slots (against the single-problem plan and
the oracle), its hipGraph replays, the autotuner, the block vector
kernels and the lockstep batched Davidson (against dense / ARPACK spectra
from the oracle and against the single-problem solver).  FP64 sigma
tolerance 1e-13 of max|sigma| (the batch accumulates with atomics, so
only rounding order differs); energies 1e-9."""

import numpy
import pytest

import oracle
from conftest import has_gpu, package

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]

pkg = package()
batch = pkg.batch

# (norb, (na, nb), seed, nvec): uneven sizes, empty-ish and full slot counts
MIX = [(6, (3, 3), 11, 4), (12, (7, 6), 7, 2), (9, (5, 4), 1, 3), (7, (4, 3), 14, 1),
       (11, (6, 5), 0, 4), (8, (5, 4), 9, 2), (10, (5, 5), 12, 4), (12, (6, 6), 18, 1),
       (4, (2, 1), 3, 2), (5, (1, 1), 4, 1)]


def _batch_inputs(mix, seed=0):
    import torch
    spaces, h2es, refs = [], [], []
    for n, ne, s, nv in mix:
        h1e, eri = pkg.synthetic.make_problem(n, ne, seed=s, with_ci=False)
        spaces.append(pkg.get_space(n, ne))
        h2es.append(pkg.absorb_h1e(h1e, eri, n, ne, 0.5))
    bs = batch.BatchSigma(spaces, h2es, [m[3] for m in mix])
    g = torch.Generator(device="cuda").manual_seed(seed)
    cin = torch.randn(bs.total, dtype=torch.float64, device="cuda", generator=g)
    return bs, spaces, h2es, cin


def _single_sigma(sp, h2e, c, precision="fp64"):
    import torch
    out = torch.empty_like(c)
    with pkg.SigmaPlan(sp, h2e, precision, 1 << 28) as plan:
        plan(c.contiguous(), out)
    return out


def _check_against_single(bs, spaces, h2es, cin, cout, mask=None, tol=1e-13):
    for p, (sp, h2e) in enumerate(zip(spaces, h2es)):
        for v in range(bs.nvec[p]):
            got = bs.slot(cout, p, v)
            if mask is not None and not mask[bs.slot_base[p] + v]:
                assert float(got.abs().max()) == 0.0
                continue
            ref = _single_sigma(sp, h2e, bs.slot(cin, p, v))
            scale = max(float(ref.abs().max()), 1e-300)
            assert float((got - ref).abs().max()) < tol * scale, (p, v)


def test_batch_sigma_matches_single_problem_plans():
    bs, spaces, h2es, cin = _batch_inputs(MIX)
    cout = bs.empty()
    bs.set_graph(False)
    bs(cin, cout)
    _check_against_single(bs, spaces, h2es, cin, cout)


def test_batch_sigma_vs_oracle_small():
    """Direct oracle check of two slots (independent of the single plan)."""
    mix = [(6, (3, 3), 11, 2), (8, (5, 4), 9, 1)]
    bs, spaces, h2es, cin = _batch_inputs(mix, seed=3)
    cout = bs.empty()
    bs(cin, cout)
    for p, (n, ne, s, nv) in enumerate(mix):
        h1e, eri = pkg.synthetic.make_problem(n, ne, seed=s, with_ci=False)
        h2e = oracle.absorb_h1e(h1e, eri, n, ne, 0.5)
        for v in range(nv):
            c = bs.slot(cin, p, v).cpu().numpy()
            ref = oracle.contract_2e(h2e, c, n, ne).reshape(-1)
            got = bs.slot(cout, p, v).cpu().numpy()
            assert numpy.abs(got - ref).max() < 1e-13 * numpy.abs(ref).max()


def test_graph_replays_and_active_mask():
    """Replays of the captured graph, with eager work and synchronisation
    between them, each agree with the eager launch; inactive slots are 0."""
    import torch
    bs, spaces, h2es, cin = _batch_inputs(MIX, seed=1)
    eager = bs.empty()
    bs.set_graph(False)
    bs(cin, eager)
    bs.set_graph(True)
    out = bs.empty()
    for rep in range(4):
        out.fill_(float("nan"))
        bs(cin, out)
        torch.cuda.synchronize()
        scale = float(eager.abs().max())
        assert float((out - eager).abs().max()) < 1e-13 * scale, rep
        _ = (cin * 2.0).sum().item()     # unrelated eager work between replays
    rng = numpy.random.default_rng(5)
    mask = (rng.random(bs.nslots) < 0.5).astype(numpy.uint8)
    mask[0] = 1
    bs.set_active(mask)
    for rep in range(2):
        bs(cin, out)
        torch.cuda.synchronize()
    _check_against_single(bs, spaces, h2es, cin, out, mask=mask)


def test_autotuner_schedule_keeps_results():
    bs, spaces, h2es, cin = _batch_inputs(MIX, seed=2)
    cout = bs.empty()
    ms = bs.tune(cin, cout, reps=2)
    rows, grid = bs.schedule()
    assert ms > 0 and rows in (1, 2, 4, 8, 16) and grid > 0
    bs(cin, cout)
    _check_against_single(bs, spaces, h2es, cin, cout)


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-6), ("bf16x3", 2e-4)])
def test_batch_low_precision(precision, tol):
    mix = MIX[:6]
    spaces, h2es = [], []
    for n, ne, s, nv in mix:
        h1e, eri = pkg.synthetic.make_problem(n, ne, seed=s, with_ci=False)
        spaces.append(pkg.get_space(n, ne))
        h2es.append(pkg.absorb_h1e(h1e, eri, n, ne, 0.5))
    bs = batch.BatchSigma(spaces, h2es, [m[3] for m in mix], precision=precision)
    import torch
    g = torch.Generator(device="cuda").manual_seed(9)
    cin = torch.randn(bs.total, dtype=torch.float64, device="cuda", generator=g)
    cout = bs.empty()
    bs(cin, cout)
    for p, (sp, h2e) in enumerate(zip(spaces, h2es)):
        for v in range(bs.nvec[p]):
            ref = _single_sigma(sp, h2e, bs.slot(cin, p, v))
            got = bs.slot(cout, p, v)
            err = float((got - ref).abs().max()) / float(ref.abs().max())
            assert 1e-12 < err < tol, (p, v, err)


def test_block_ops_match_torch():
    import torch
    ops = batch.BlockOps()
    g = torch.Generator(device="cuda").manual_seed(4)
    lens = [1, 300, 65536, 65537, 200000]
    nx, ny = [3, 9, 1, 12, 5], [2, 4, 5, 1, 3]
    X = [torch.randn(a, n, dtype=torch.float64, device="cuda", generator=g) for a, n in zip(nx, lens)]
    Y = [torch.randn(b, n, dtype=torch.float64, device="cuda", generator=g) for b, n in zip(ny, lens)]

    def rows(t):
        return [t[i].data_ptr() for i in range(t.shape[0])]

    G = ops.gram([(rows(x), rows(y), n) for x, y, n in zip(X, Y, lens)])
    for x, y, gm in zip(X, Y, G):
        ref = (x @ y.T).cpu().numpy()
        assert numpy.abs(gm - ref).max() < 1e-12 * max(1.0, numpy.abs(ref).max())
    # Y = 0.5 Y + X C
    C = [numpy.random.default_rng(k).standard_normal((a, b)) for k, (a, b) in enumerate(zip(nx, ny))]
    ref = [0.5 * y + torch.from_numpy(c).cuda().T @ x for x, y, c in zip(X, Y, C)]
    ops.gemm([(rows(x), rows(y), n, c, 0.5) for x, y, n, c in zip(X, Y, lens, C)])
    for y, r in zip(Y, ref):
        assert float((y - r).abs().max()) < 1e-12 * float(r.abs().max())
    # in-place block transform X[:m] <- X C (Y aliases X), beta = 0 ignores NaN in Y
    x = X[3]
    c = numpy.random.default_rng(9).standard_normal((nx[3], 6))
    r = torch.from_numpy(c).cuda().T @ x
    ops.gemm([(rows(x), rows(x)[:6], lens[3], c, 0.0)])
    assert float((x[:6] - r).abs().max()) < 1e-12 * float(r.abs().max())
    z = torch.full((1, lens[2]), float("nan"), dtype=torch.float64, device="cuda")
    ops.gemm([(rows(X[2]), rows(z), lens[2], numpy.ones((1, 1)), 0.0)])
    assert torch.equal(z, X[2])
    hd = torch.linspace(-1.0, 1.0, lens[3], dtype=torch.float64, device="cuda")
    rv = torch.randn(lens[3], dtype=torch.float64, device="cuda", generator=g)
    out = torch.empty_like(hd)
    ops.precond([(hd.data_ptr(), rv.data_ptr(), out.data_ptr(), lens[3], 0.25)], 1e-3)
    den = hd - 0.25
    den = torch.where(den.abs() < 1e-3, torch.full_like(den, 1e-3), den)
    assert float((out - rv / den).abs().max()) < 1e-12 * float(out.abs().max())


def test_solve_batch_small_vs_dense_spectra():
    mix = [(6, (3, 3), 11), (7, (4, 3), 14), (8, (5, 4), 9), (6, (4, 3), 21), (9, (5, 4), 1)]
    probs = []
    for n, ne, s in mix:
        h1e, eri = pkg.synthetic.make_problem(n, ne, seed=s, with_ci=False)
        probs.append((h1e, eri, n, ne))
    res = batch.solve_batch(probs, nroots=4, max_cycle=200)
    for p, (h1e, eri, n, ne) in enumerate(probs):
        w, _ = oracle.lowest_eigenvalues(h1e, eri, n, ne, nroots=4)
        assert res.converged[p].all(), (p, res.rnorm[p])
        numpy.testing.assert_allclose(res.e[p], w, atol=1e-9)
        # returned vectors are orthonormal eigenvectors
        ci = res.ci[p].cpu().numpy()
        numpy.testing.assert_allclose(ci @ ci.T, numpy.eye(4), atol=1e-8)


def test_solve_batch_matches_single_solver():
    mix = [(10, (5, 5), 12), (11, (6, 5), 0), (12, (7, 6), 7)]
    probs = []
    for n, ne, s in mix:
        h1e, eri = pkg.synthetic.make_problem(n, ne, seed=s, with_ci=False)
        probs.append((h1e, eri, n, ne))
    res = batch.solve_batch(probs, nroots=4, max_cycle=300)
    for p, (h1e, eri, n, ne) in enumerate(probs):
        sol = pkg.FCISolver()
        e, _ = sol.kernel(h1e, eri, n, ne, nroots=4, max_cycle=300)
        assert res.converged[p].all() and numpy.all(sol.converged)
        numpy.testing.assert_allclose(res.e[p], e, atol=1e-8)
