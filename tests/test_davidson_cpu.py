"""The Davidson driver's algorithm (pyscf.lib.davidson1 as the reference's
solvers call it, fci_gas.py:1344-1350) on the CPU: the solver in
``davidson.py`` is written against a small vector-algebra interface, and
here that interface is numpy, so the subspace logic (Gram-Schmidt with the
Daniel-Gragg-Kaufman-Stewart second pass, Rayleigh-Ritz, thick restart,
convergence tests) is checked without a GPU.  Pinned by mathematics: the
eigenvalues of dense symmetric matrices from numpy.linalg.eigh (1e-9),
the Ritz vectors' residual norms, and the orthonormality of the basis."""

import numpy
import pytest

from conftest import package

dav = package().davidson


class NumpyOps:
    """The solver's vector interface on host arrays."""

    def empty(self, shape, like):
        return numpy.empty(shape)

    def clone(self, v):
        return numpy.array(v, dtype=numpy.float64).reshape(-1)

    def copy(self, dst, src):
        dst[...] = src

    def dots(self, m, n, X, v):
        return numpy.atleast_2d(X)[:m].reshape(m, n) @ v.reshape(-1)

    def lincomb(self, m, n, X, coef, beta, y):
        Xm = numpy.atleast_2d(X)[:m].reshape(m, n)
        y[...] = (beta * y.reshape(-1) if beta != 0.0 else 0.0) + numpy.asarray(coef).reshape(m) @ Xm

    def combine(self, n, vecs, coef, y):
        y[...] = sum(c * v.reshape(-1) for c, v in zip(coef, vecs))

    def dots_list(self, n, vecs, v):
        return numpy.array([float(u.reshape(-1) @ v.reshape(-1)) for u in vecs])

    def rotate(self, X, m, n, W):
        Xm = numpy.atleast_2d(X)[:m].reshape(m, n)
        Xm[...] = numpy.asarray(W).T @ Xm.copy()

    def combine_many(self, n, tasks):
        for vecs, outs, coef in tasks:
            V = numpy.stack([v.reshape(-1) for v in vecs])
            for j, o in enumerate(outs):
                o[...] = numpy.asarray(coef)[:, j] @ V


class PlainOps(NumpyOps):
    """The same without the one-pass Ritz/residual combination and the
    one-pass Gram-Schmidt (the driver's fallbacks)."""
    combine_many = None
    dots_list = None
    rotate = None


def _problem(n, seed, spread=10.0):
    rng = numpy.random.default_rng(seed)
    d = numpy.sort(rng.uniform(0.0, spread, n))
    B = rng.standard_normal((n, n)) * 0.05
    H = numpy.diag(d) + 0.5 * (B + B.T)
    return H


def _run(H, nroots, max_space, tol=1e-12, ops=None):
    n = H.shape[0]
    hdiag = numpy.diag(H).copy()

    def hop(x, out):
        out[...] = H @ x

    def precond(r, e, out, x0):
        dd = hdiag - e
        dd[numpy.abs(dd) < 1e-8] = 1e-8
        out[...] = r / dd

    guesses = []
    for k in numpy.argsort(hdiag)[:nroots]:
        g = numpy.zeros(n)
        g[k] = 1.0
        guesses.append(g)
    return dav.davidson(hop, precond, guesses, nroots=nroots, tol=tol, max_cycle=200,
                        max_space=max_space, ops=ops if ops is not None else NumpyOps())


@pytest.mark.parametrize("n,nroots,max_space,seed,ops", [(200, 1, 12, 1, NumpyOps), (300, 3, 12, 2, NumpyOps),
                                                          (150, 4, 6, 3, NumpyOps), (400, 2, 8, 4, NumpyOps),
                                                          (300, 3, 12, 2, PlainOps)])
def test_davidson_eigenvalues_vs_eigh(n, nroots, max_space, seed, ops):
    H = _problem(n, seed)
    res = _run(H, nroots, max_space, ops=ops())
    w = numpy.linalg.eigh(H)[0][:nroots]
    assert res.converged.all()
    numpy.testing.assert_allclose(res.e, w, rtol=0, atol=1e-9)
    X = numpy.asarray(res.x).reshape(nroots, n)
    AX = numpy.asarray(res.ax).reshape(nroots, n)
    # Ritz vectors: orthonormal, images consistent, residuals below sqrt(tol)
    numpy.testing.assert_allclose(X @ X.T, numpy.eye(nroots), rtol=0, atol=1e-10)
    numpy.testing.assert_allclose(AX, X @ H.T, rtol=0, atol=1e-10)
    r = AX - res.e[:, None] * X
    assert numpy.linalg.norm(r, axis=1).max() < 1e-5


def test_davidson_restart_keeps_applications_low():
    """A tiny subspace forces thick restarts; the images of the kept
    vectors come from AX, so H applications stay one per new vector."""
    H = _problem(250, 7)
    res = _run(H, 2, 5)
    assert res.converged.all()
    numpy.testing.assert_allclose(res.e, numpy.linalg.eigh(H)[0][:2], rtol=0, atol=1e-9)
    assert res.nhop <= 2 + res.cycles * 2


def test_davidson_linearly_dependent_guesses():
    """Guesses that span fewer directions than roots raise, as the reference's
    solver refuses a singular initial space."""
    H = _problem(50, 9)
    g = numpy.zeros(50)
    g[0] = 1.0

    def hop(x, out):
        out[...] = H @ x

    with pytest.raises(RuntimeError):
        dav.davidson(hop, lambda r, e, out, x0: out.__setitem__(Ellipsis, r), [g, 2.0 * g],
                     nroots=2, ops=NumpyOps())


def _rounded(x, bits):
    """x with its mantissa rounded to ``bits`` bits (a nonlinear,
    input-dependent error like that of a low-precision sigma)."""
    m, ex = numpy.frexp(x)
    return numpy.ldexp(numpy.round(m * 2.0 ** bits) / 2.0 ** bits, ex)


@pytest.mark.parametrize("n,nroots,bits,seed", [(300, 1, 16, 11), (400, 3, 16, 12),
                                                 (300, 2, 22, 13)])
def test_mixed_refinement_keeps_the_subspace(n, nroots, bits, seed):
    """Mixed precision inside one run: the cheap operator rounds its input
    and its output to ``bits`` mantissa bits (16: about the bf16x3 sigma's
    error, 22: about FP32's); the exact one is checked only at the
    acceptance points.  The result meets the exact criteria (eigenvalues of
    eigh to 1e-9, exact residual below sqrt(tol)) with a few exact
    applications per root (at most four), fewer than a restart from the
    low-precision Ritz vectors needs."""
    H = _problem(n, seed)
    hdiag = numpy.diag(H).copy()
    nexact = [0]

    def hop_lo(x, out):
        out[...] = _rounded(H @ _rounded(x, bits), bits)

    def hop(x, out):
        nexact[0] += 1
        out[...] = H @ x

    def precond(r, e, out, x0):
        dd = hdiag - e
        dd[numpy.abs(dd) < 1e-8] = 1e-8
        out[...] = r / dd

    guesses = []
    for k in numpy.argsort(hdiag)[:nroots]:
        g = numpy.zeros(n)
        g[k] = 1.0
        guesses.append(g)
    tol = 1e-12
    res = dav.davidson(hop_lo, precond, guesses, nroots=nroots, tol=tol, max_cycle=200,
                       max_space=12, ops=NumpyOps(), hop_check=hop, tol_low=1e-6,
                       tol_residual_low=5e-4)
    w = numpy.linalg.eigh(H)[0][:nroots]
    assert res.converged.all()
    numpy.testing.assert_allclose(res.e, w, rtol=0, atol=1e-9)
    X = numpy.asarray(res.x).reshape(nroots, n)
    r = X @ H.T - res.e[:, None] * X
    assert numpy.linalg.norm(r, axis=1).max() < 1e-6
    assert res.nhop_check == nexact[0]
    # the restart alternative: low-precision Ritz vectors into a plain exact solve
    lo = dav.davidson(hop_lo, precond, guesses, nroots=nroots, tol=1e-6, tol_residual=5e-4,
                      max_cycle=200, max_space=12, ops=NumpyOps())
    ref = dav.davidson(lambda x, out: out.__setitem__(Ellipsis, H @ x), precond, list(lo.x),
                       nroots=nroots, tol=tol, max_cycle=200, max_space=12, ops=NumpyOps())
    assert res.nhop_check <= 4 * nroots
    assert res.nhop_check < ref.nhop


@pytest.mark.parametrize("n,nroots,seed", [(300, 1, 21), (300, 3, 22)])
def test_follow_state_restores_after_residual_jump(n, nroots, seed):
    """davidson1's follow_state guard (the reference switches it on,
    fci_gas.py:1350, csf.py:489): when the largest residual norm exceeds 1
    and grew more than threefold, the subspace restarts from the previous
    Ritz vectors.  Here one H application returns a grossly wrong image
    (a planted fault of the operator); with the guard the solve drops that
    direction and still reaches the eigenvalues of eigh, and on a clean
    operator the guard never fires and changes nothing."""
    H = _problem(n, seed)
    hdiag = numpy.diag(H).copy()
    calls = [0]

    def hop_bad(x, out):
        calls[0] += 1
        out[...] = H @ x
        if calls[0] == nroots + 3:
            out[...] += 50.0 * numpy.sign(x)

    def hop(x, out):
        out[...] = H @ x

    def precond(r, e, out, x0):
        dd = hdiag - e
        dd[numpy.abs(dd) < 1e-8] = 1e-8
        out[...] = r / dd

    guesses = []
    for k in numpy.argsort(hdiag)[:nroots]:
        g = numpy.zeros(n)
        g[k] = 1.0
        guesses.append(g)
    w = numpy.linalg.eigh(H)[0][:nroots]
    res = dav.davidson(hop_bad, precond, guesses, nroots=nroots, tol=1e-12, max_cycle=200,
                       max_space=12, ops=NumpyOps(), follow_state=True)
    assert res.nrestore >= 1
    assert res.converged.all()
    numpy.testing.assert_allclose(res.e, w, rtol=0, atol=1e-9)
    X = numpy.asarray(res.x).reshape(nroots, n)
    assert numpy.linalg.norm(X @ H.T - res.e[:, None] * X, axis=1).max() < 1e-5
    # without the guard the wrong image stays in the subspace
    calls[0] = 0
    off = dav.davidson(hop_bad, precond, guesses, nroots=nroots, tol=1e-12, max_cycle=60,
                       max_space=12, ops=NumpyOps())
    assert off.nrestore == 0 and not off.converged.all()

    clean = dav.davidson(hop, precond, guesses, nroots=nroots, tol=1e-12, max_cycle=200,
                         max_space=12, ops=NumpyOps(), follow_state=True)
    plain = dav.davidson(hop, precond, guesses, nroots=nroots, tol=1e-12, max_cycle=200,
                         max_space=12, ops=NumpyOps())
    assert clean.nrestore == 0
    assert clean.cycles == plain.cycles
    numpy.testing.assert_allclose(clean.e, w, rtol=0, atol=1e-9)
