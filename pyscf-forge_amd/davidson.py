"""Block Davidson eigensolver whose vectors never leave the device.

Follows the algorithm of pyscf.lib.davidson1 as the reference's solvers use
it (called at pyscf/mcscf/fci_gas.py:1344-1350 with tol=1e-10,
max_cycle=50, max_space=12, lindep=1e-14; tol_residual defaults to
sqrt(tol)): the trial vectors X and their images AX = H X live in HBM; only
the small subspace matrix (at most max_space + 3*(nroots-1) on a side) is
brought to the host for its FP64 eigendecomposition.  All O(N) work (dot
products, Ritz combinations, residuals, preconditioning, Gram-Schmidt) runs
in the library's HIP kernels; PyTorch only allocates the device buffers.

``hop(x, out)`` writes H x into ``out`` (device float64 vectors of length
``n``); ``precond(r, e, out)`` writes the preconditioned residual.
"""

import math

import numpy

from . import _lib


class DavidsonResult:
    def __init__(self, converged, e, x, ax, cycles, nhop, rnorm):
        self.converged = converged
        self.e = e
        self.x = x
        self.ax = ax
        self.cycles = cycles
        self.nhop = nhop
        self.rnorm = rnorm


def _dots(lib, m, n, X, v, stream):
    out = numpy.empty(max(m, 1))
    if m:
        _lib.check(lib.fcis_dots(m, n, _lib.vptr(X), n, _lib.vptr(v),
                                 _lib.dptr(out), stream), "fcis_dots")
    return out[:m]


def _lincomb(lib, m, n, X, coef, beta, y, stream):
    coef = numpy.ascontiguousarray(coef, dtype=numpy.float64)
    _lib.check(lib.fcis_lincomb(m, n, _lib.vptr(X), n, _lib.dptr(coef),
                                float(beta), _lib.vptr(y), stream), "fcis_lincomb")


def davidson(hop, precond, x0, nroots=1, tol=1e-10, tol_residual=None,
             max_cycle=50, max_space=12, lindep=1e-14, verbose=False,
             callback=None):
    """Lowest ``nroots`` eigenpairs.  x0: list of device vectors (guesses).

    Returns DavidsonResult with e (numpy), x / ax (device tensors of shape
    (nroots, n)), per-root convergence flags, cycle and H-application counts.
    """
    import torch

    lib = _lib.load()
    stream = _lib.stream_ptr()
    n = int(x0[0].numel())
    dev = x0[0].device
    if tol_residual is None:
        tol_residual = math.sqrt(tol)
    cap = int(max_space) + (nroots - 1) * 3
    cap = max(cap, nroots + len(x0))
    cap = min(cap, 64, n)
    X = torch.empty((cap, n), dtype=torch.float64, device=dev)
    AX = torch.empty((cap, n), dtype=torch.float64, device=dev)
    xr = torch.empty((nroots, n), dtype=torch.float64, device=dev)
    axr = torch.empty((nroots, n), dtype=torch.float64, device=dev)
    r = torch.empty(n, dtype=torch.float64, device=dev)
    heff = numpy.zeros((cap, cap))
    space = 0
    nhop = 0

    def add_vector(t):
        """Orthonormalise t against X[:space] (Gram-Schmidt, twice) and
        append it with its image.  Returns False if t is linearly dependent."""
        nonlocal space, nhop
        norm0 = math.sqrt(_dots(lib, 1, n, t, t, stream)[0])
        if norm0 == 0.0:
            return False
        for _ in range(2):
            if space:
                c = _dots(lib, space, n, X, t, stream)
                _lincomb(lib, space, n, X, -c, 1.0, t, stream)
        norm = math.sqrt(max(_dots(lib, 1, n, t, t, stream)[0], 0.0))
        if norm * norm < lindep or norm < 1e-7 * norm0:
            return False
        _lincomb(lib, 1, n, t, numpy.array([1.0 / norm]), 0.0, X[space], stream)
        hop(X[space], AX[space])
        nhop += 1
        col = _dots(lib, space + 1, n, X, AX[space], stream)
        heff[:space + 1, space] = col
        heff[space, :space + 1] = col
        space += 1
        return True

    for g in x0:
        t = g.detach().clone().to(torch.float64)
        if space < cap:
            add_vector(t)
    if space < nroots:
        raise RuntimeError("Davidson: initial guesses are linearly dependent")

    e_old = numpy.full(nroots, numpy.inf)
    conv = numpy.zeros(nroots, dtype=bool)
    rnorm = numpy.full(nroots, numpy.inf)
    cycles = 0
    for icyc in range(max_cycle):
        cycles = icyc + 1
        h = 0.5 * (heff[:space, :space] + heff[:space, :space].T)
        w, v = numpy.linalg.eigh(h)
        e = w[:nroots].copy()
        v = v[:, :nroots]
        for k in range(nroots):
            _lincomb(lib, space, n, X, v[:, k], 0.0, xr[k], stream)
            _lincomb(lib, space, n, AX, v[:, k], 0.0, axr[k], stream)
        de = e - e_old
        new_vectors = []
        for k in range(nroots):
            _lincomb(lib, 1, n, axr[k], numpy.array([1.0]), 0.0, r, stream)
            _lincomb(lib, 1, n, xr[k], numpy.array([-e[k]]), 1.0, r, stream)
            rnorm[k] = math.sqrt(max(_dots(lib, 1, n, r, r, stream)[0], 0.0))
            conv[k] = abs(de[k]) < tol and rnorm[k] < tol_residual
            if not conv[k]:
                t = torch.empty(n, dtype=torch.float64, device=dev)
                precond(r, e[k], t)
                new_vectors.append(t)
        if verbose:
            print("davidson %d  space %d  e %s  de %s  |r| %s" % (
                icyc, space, e, de, rnorm))
        if callback is not None:
            callback(icyc, e, rnorm)
        e_old = e
        if conv.all():
            break
        if space + len(new_vectors) > cap:
            # restart from the current Ritz vectors, keeping their images
            X[:nroots].copy_(xr)
            AX[:nroots].copy_(axr)
            space = nroots
            heff[:] = 0.0
            heff[:nroots, :nroots] = numpy.diag(e)
        added = 0
        for t in new_vectors:
            if space < cap and add_vector(t):
                added += 1
        if added == 0:
            break
    return DavidsonResult(conv.copy(), e_old, xr, axr, cycles, nhop, rnorm.copy())
