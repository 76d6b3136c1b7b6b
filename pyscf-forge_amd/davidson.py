"""Block Davidson eigensolver whose vectors never leave the device.

Follows the algorithm of pyscf.lib.davidson1 as the reference's solvers use
it (called at pyscf/mcscf/fci_gas.py:1344-1350 with tol=1e-10,
max_cycle=50, max_space=12, lindep=1e-14; tol_residual defaults to
sqrt(tol)): the trial vectors X and their images AX = H X live in HBM; only
the small subspace matrix (at most max_space + 3*(nroots-1) on a side) is
brought to the host for its FP64 eigendecomposition.  All O(N) work (dot
products, Ritz combinations, residuals, preconditioning, Gram-Schmidt) runs
in the library's HIP kernels (block gram / gemm, preconditioner); PyTorch
only allocates the device buffers.

``hop(x, out)`` writes H x into ``out`` (device float64 vectors of length
``n``); ``precond(r, e, out, x0)`` writes the preconditioned residual of
the Ritz pair (e, x0).
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


class DeviceOps:
    """Vector algebra of the solver on device float64 vectors, through the
    library's block kernels: X^T v as one gram launch and beta y + X c as
    one gemm launch, each streaming every vector of X once."""

    def __init__(self):
        from .blockops import BlockOps
        self.ops = BlockOps()

    def empty(self, shape, like):
        import torch
        return torch.empty(shape, dtype=torch.float64, device=like.device)

    def clone(self, v):
        return v.detach().clone()

    def copy(self, dst, src):
        dst.copy_(src)

    @staticmethod
    def _rows(X, m, n):
        if not X.is_contiguous():
            raise ValueError("vector blocks must be contiguous")
        return X.data_ptr() + 8 * int(n) * numpy.arange(m, dtype=numpy.int64)

    def dots(self, m, n, X, v):
        if m == 0:
            return numpy.zeros(0)
        g = self.ops.gram([(self._rows(X, m, n), [v.data_ptr()], int(n))])[0]
        return g[:, 0].copy()

    def lincomb(self, m, n, X, coef, beta, y):
        coef = numpy.ascontiguousarray(coef, dtype=numpy.float64).reshape(m, 1)
        self.ops.gemm([(self._rows(X, m, n), [y.data_ptr()], int(n), coef, float(beta))])

    def combine(self, n, vecs, coef, y):
        """y = sum_i coef[i] vecs[i] in one pass (y may not be one of vecs)."""
        rows = numpy.array([v.data_ptr() for v in vecs], dtype=numpy.int64)
        coef = numpy.ascontiguousarray(coef, dtype=numpy.float64).reshape(len(vecs), 1)
        self.ops.gemm([(rows, [y.data_ptr()], int(n), coef, 0.0)])

    def rotate(self, X, m, n, W):
        """X[:m] <- (X[:m] as rows) transformed by W: row j = sum_i W[i, j] X_i,
        in place (one launch; m <= 16)."""
        rows = self._rows(X, m, n)
        self.ops.gemm([(rows, rows, int(n), numpy.ascontiguousarray(W, dtype=numpy.float64), 0.0)])

    def dots_list(self, n, vecs, v):
        """[u . v for u in vecs] in one launch (vecs anywhere in memory)."""
        rows = numpy.array([u.data_ptr() for u in vecs], dtype=numpy.int64)
        return self.ops.gram([(rows, [v.data_ptr()], int(n))])[0][:, 0].copy()

    def combine_many(self, n, tasks):
        """For each (vecs, outs, coef): outs[j] = sum_i coef[i, j] vecs[i]
        (outputs distinct from the inputs), all tasks in one launch that
        streams each task's inputs once."""
        blocks = []
        for vecs, outs, coef in tasks:
            rows = numpy.array([v.data_ptr() for v in vecs], dtype=numpy.int64)
            cols = [o.data_ptr() for o in outs]
            c = numpy.ascontiguousarray(coef, dtype=numpy.float64).reshape(len(vecs), len(outs))
            blocks.append((rows, cols, int(n), c, 0.0))
        self.ops.gemm(blocks)


def davidson(hop, precond, x0, nroots=1, tol=1e-10, tol_residual=None,
             max_cycle=50, max_space=12, lindep=1e-14, verbose=False,
             callback=None, ops=None, hop_check=None, tol_low=None,
             tol_residual_low=None, check_every=6, check_max=20, follow_state=False):
    """Lowest ``nroots`` eigenpairs.  x0: list of device vectors (guesses).

    Returns DavidsonResult with e (numpy), x / ax (device tensors of shape
    (nroots, n)), per-root convergence flags, cycle and H-application counts.

    State matching, as pyscf.lib.davidson1 does it: each root's energy
    change is taken against the previous iteration's Ritz value of the state
    it overlaps most (the subspace eigenvectors may swap order between
    iterations).  ``follow_state`` (the reference passes it at
    fci_gas.py:1350 and csf.py:489) adds davidson1's guard against a state
    flip: when the largest residual norm exceeds 1 and grew more than three
    times over the previous iteration, with more than nroots + 2 vectors in
    the subspace, the subspace restarts from the previous iteration's Ritz
    vectors (formed from X and AX, no new H applications) and the new
    directions are dropped; ``nrestore`` counts these restarts.

    Mixed precision: with ``hop_check`` (the exact operator; ``hop`` is then
    an approximate, cheaper one) every new subspace vector is applied with
    ``hop``.  The first stage runs to (tol_low, tol_residual_low).  Then, and
    whenever the iterations meet (tol, tol_residual), or at least
    ``check_every`` iterations have passed and the largest residual fell
    by less than a quarter over the last three (a low-precision plateau;
    a slow but steady decrease is not one), or
    ``check_max`` iterations have passed, the Ritz vectors are applied with
    ``hop_check``:
    they are accepted if their exact residuals and the change of their exact
    Rayleigh quotients from the iteration's Ritz values meet (tol,
    tol_residual); otherwise the subspace restarts from them with their exact
    images (plus the previous Ritz directions) and goes on.  The subspace is
    kept through the switch: later corrections enter the Ritz vectors with
    coefficients of the size of the remaining error, so the approximate
    images perturb the residual only in proportion to that error, and the
    solve needs about two exact applications per root.  ``nhop`` counts
    ``hop``, ``nhop_check`` ``hop_check``.
    """
    ops = ops if ops is not None else DeviceOps()
    n = int(x0[0].reshape(-1).shape[0])
    if tol_residual is None:
        tol_residual = math.sqrt(tol)
    mixed = hop_check is not None
    if mixed:
        stage_tol = tol if tol_low is None else max(tol, tol_low)
        stage_res = tol_residual if tol_residual_low is None else max(tol_residual, tol_residual_low)
    else:
        stage_tol, stage_res = tol, tol_residual
    refining = False
    cap = int(max_space) + (nroots - 1) * 3
    cap = max(cap, nroots + len(x0))
    cap = min(cap, 64, n)
    X = ops.empty((cap, n), x0[0])
    AX = ops.empty((cap, n), x0[0])
    xr = ops.empty((nroots, n), x0[0])
    axr = ops.empty((nroots, n), x0[0])
    xp = ops.empty((nroots, n), x0[0])      # previous Ritz directions (restart)
    axp = ops.empty((nroots, n), x0[0])
    R = ops.empty((nroots, n), x0[0])      # residuals of the Ritz pairs
    v_prev = None
    heff = numpy.zeros((cap, cap))
    space = 0
    nhop = 0
    nhop_check = 0
    last_check = 0
    # X[:nexact] carry exact images (mixed precision, after the first check):
    # heff[a, b] = X_a . AX_b, and the coupling of an exact vector i with an
    # approximate j is taken from the exact image, heff[j, i]
    nexact = 0

    fused_gs = getattr(ops, "dots_list", None) is not None and \
        getattr(ops, "combine_many", None) is not None

    def project(t, norm, passes):
        """The remaining classical Gram-Schmidt passes on t (two in all, the
        second only where the first cancelled much of t: Daniel-Gragg-
        Kaufman-Stewart, |t'| < |t| / sqrt(2)); returns |t|."""
        for _ in range(passes, 2):
            if not space:
                break
            c = ops.dots(space, n, X, t)
            ops.lincomb(space, n, X, -c, 1.0, t)
            before, norm = norm, math.sqrt(max(ops.dots(1, n, t, t)[0], 0.0))
            if norm >= before * 0.7071067811865476:
                break
        return norm

    def add_vector(t):
        """Orthonormalise t against X[:space] (Gram-Schmidt, twice) and
        append it with its image.  Returns False if t is linearly dependent."""
        nonlocal space, nhop
        written = False
        if fused_gs and space:
            # X^T t and |t|^2 in one pass.  Where one projection suffices
            # (|t - X c|^2 = |t|^2 - |c|^2 >= |t|^2 / 2, the test of the
            # second pass) the normalised projection is written in one more
            # pass; otherwise the first projection is applied and the second
            # pass follows as usual.
            g = ops.dots_list(n, [X[i] for i in range(space)] + [t], t)
            c, tt = g[:space], float(g[space])
            if tt <= 0.0:
                return False
            norm0 = math.sqrt(tt)
            rest = tt - float(c @ c)
            if rest >= 0.5 * tt:
                norm = math.sqrt(rest)
                if norm * norm < lindep or norm < 1e-7 * norm0:
                    return False
                coef = numpy.concatenate([-c, [1.0]]) / norm
                ops.combine_many(n, [([X[i] for i in range(space)] + [t], (X[space],), coef[:, None])])
                written = True
            else:
                ops.lincomb(space, n, X, -c, 1.0, t)
                norm = project(t, math.sqrt(max(ops.dots(1, n, t, t)[0], 0.0)), 1)
        else:
            norm0 = math.sqrt(ops.dots(1, n, t, t)[0])
            if norm0 == 0.0:
                return False
            norm = project(t, norm0, 0)
        if not written:
            if norm * norm < lindep or norm < 1e-7 * norm0:
                return False
            ops.lincomb(1, n, t, numpy.array([1.0 / norm]), 0.0, X[space])
        hop(X[space], AX[space])
        nhop += 1
        col = ops.dots(space + 1, n, X, AX[space])
        heff[:space + 1, space] = col
        heff[space, :space + 1] = col
        if nexact:
            heff[space, :nexact] = ops.dots(nexact, n, AX, X[space])
        space += 1
        return True

    def subspace_matrix():
        h = 0.5 * (heff[:space, :space] + heff[:space, :space].T)
        if nexact:
            h[nexact:, :nexact] = heff[nexact:space, :nexact]
            h[:nexact, nexact:] = heff[nexact:space, :nexact].T
        return h

    def ritz(v, e):
        """xr_k = X v_k, axr_k = AX v_k and R_k = axr_k - e_k xr_k, one pass
        over X and AX for all roots (combine_many) where the ops have it."""
        if getattr(ops, "combine_many", None) is not None:
            vecs = [X[i] for i in range(space)] + [AX[i] for i in range(space)]
            tasks = []
            for k in range(nroots):
                c = numpy.zeros((2 * space, 3))
                c[:space, 0] = v[:, k]
                c[space:, 1] = v[:, k]
                c[:space, 2] = -e[k] * v[:, k]
                c[space:, 2] = v[:, k]
                tasks.append((vecs, (xr[k], axr[k], R[k]), c))
            ops.combine_many(n, tasks)
        else:
            for k in range(nroots):
                ops.lincomb(space, n, X, v[:, k], 0.0, xr[k])
                ops.lincomb(space, n, AX, v[:, k], 0.0, axr[k])
                ops.combine(n, (axr[k], xr[k]), numpy.array([1.0, -e[k]]), R[k])

    def residual_norms(e, recompute):
        """|R_k| (R_k = axr_k - e_k xr_k formed again when ``recompute``)."""
        out = numpy.empty(nroots)
        for k in range(nroots):
            if recompute:
                ops.combine(n, (axr[k], xr[k]), numpy.array([1.0, -e[k]]), R[k])
            out[k] = math.sqrt(max(ops.dots(1, n, R[k], R[k])[0], 0.0))
        return out

    for g in x0:
        t = ops.clone(g)
        if space < cap:
            add_vector(t)
    if space < nroots:
        raise RuntimeError("Davidson: initial guesses are linearly dependent")

    def restore(vp):
        """Restart the subspace from the vectors X vp (orthonormal columns
        vp in subspace coordinates, padded to the current space) with the
        images AX vp: one pass over X and AX, no H applications."""
        nonlocal space, nexact
        k = vp.shape[1]
        for j in range(k):
            ops.lincomb(space, n, X, vp[:, j], 0.0, xp[j])
            ops.lincomb(space, n, AX, vp[:, j], 0.0, axp[j])
        hk = vp.T @ subspace_matrix() @ vp
        for j in range(k):
            ops.copy(X[j], xp[j])
            ops.copy(AX[j], axp[j])
        space = k
        heff[:] = 0.0
        heff[:k, :k] = 0.5 * (hk + hk.T)
        # combinations of exact and approximate images are neither: the
        # next check anchors exact ones again
        nexact = 0

    e_old = numpy.full(nroots, numpy.inf)
    conv = numpy.zeros(nroots, dtype=bool)
    rnorm = numpy.full(nroots, numpy.inf)
    rmax_prev = numpy.inf
    rhist = []       # largest residual of each iteration since the last check
    nrestore = 0
    cycles = 0
    for icyc in range(max_cycle):
        cycles = icyc + 1
        h = subspace_matrix()
        w_all, v_all = numpy.linalg.eigh(h)
        e = w_all[:nroots].copy()
        v = v_all[:, :nroots]
        ritz(v, e)
        e_ref = e_old
        if v_prev is not None and numpy.isfinite(e_old).all():
            # compare each root with the previous Ritz value of the state it
            # overlaps most (v_prev: previous Ritz vectors in the current
            # subspace coordinates, new rows zero)
            head = v_prev.shape[0]
            e_ref = e_old[numpy.argmax(numpy.abs(v[:head].T @ v_prev), axis=1)]
        de = e - e_ref
        rnorm[:] = residual_norms(e, False)
        rmax = float(rnorm.max())
        if (follow_state and v_prev is not None and rmax > 1.0 and rmax > 3.0 * rmax_prev
                and space > nroots + 2
                and not ((numpy.abs(de) < stage_tol) & (rnorm < stage_res)).all()):
            vp = numpy.zeros((space, v_prev.shape[1]))
            vp[:v_prev.shape[0]] = v_prev
            restore(vp)
            nrestore += 1
            v_prev = numpy.eye(space)[:, :nroots]
            if verbose:
                print("davidson %d  large |r| %.3g: restart from the previous Ritz vectors"
                      % (icyc, rmax))
            continue
        rmax_prev = rmax
        conv = (numpy.abs(de) < stage_tol) & (rnorm < stage_res)
        checked = False
        since = icyc - last_check
        stalled = (refining and since >= check_every and len(rhist) >= 3
                   and rmax > 0.8 * rhist[-3])
        if refining:
            rhist.append(rmax)
        if mixed and (conv.all() or stalled or (refining and since >= check_max)
                      or (not refining and icyc + 1 >= max_cycle // 2)):
            # exact images of the Ritz vectors: the acceptance test, and the
            # anchor of the subspace from here on
            for k in range(nroots):
                hop_check(xr[k], axr[k])
            nhop_check += nroots
            e_chk = numpy.array([ops.dots(1, n, xr[k], axr[k])[0] for k in range(nroots)])
            rnorm[:] = residual_norms(e_chk, True)
            conv = (numpy.abs(e_chk - e) < tol) & (rnorm < tol_residual) if refining \
                else numpy.zeros(nroots, dtype=bool)
            de = e_chk - e_old
            e = e_chk
            refining, checked, last_check = True, True, icyc
            stage_tol, stage_res = tol, tol_residual
            rhist = [float(rnorm.max())]
        new_vectors = []
        for k in range(nroots):
            if not conv[k]:
                t = ops.empty((n,), x0[0])
                precond(R[k], e[k], t, xr[k])
                new_vectors.append(t)
        if verbose:
            print("davidson %d  space %d  e %s  de %s  |r| %s%s" % (
                icyc, space, e, de, rnorm, "  (checked)" if checked else ""))
        if callback is not None:
            callback(icyc, e, rnorm)
        e_old = e
        if conv.all() and (checked or not mixed):
            break
        if (checked and space + len(new_vectors) <= cap and space <= 16
                and getattr(ops, "rotate", None) is not None):
            # After a check with room left: rotate the whole subspace to the
            # Ritz basis (nothing is dropped) and give the Ritz vectors their
            # exact images; the low-precision block becomes diag(w)
            ops.rotate(X, space, n, v_all)
            ops.rotate(AX, space, n, v_all)
            ops.copy(X[:nroots], xr)
            ops.copy(AX[:nroots], axr)
            heff[:] = 0.0
            heff[:space, :space] = numpy.diag(w_all)
            for i in range(nroots):
                heff[:space, i] = ops.dots(space, n, X, AX[i])
            nexact = nroots
            v = numpy.eye(space)[:, :nroots]
        elif checked or space + len(new_vectors) > cap:
            # Thick restart: keep the current Ritz vectors and the part of the
            # previous iteration's Ritz vectors orthogonal to them (the
            # locally optimal direction), formed in subspace coordinates so
            # their images come from AX without new H applications.  After a
            # check the Ritz vectors carry their exact images.
            keep = v.copy()
            if v_prev is not None and v_prev.shape[0] <= space:
                prev = numpy.zeros((space, v_prev.shape[1]))
                prev[:v_prev.shape[0]] = v_prev
                for k in range(prev.shape[1]):
                    p = prev[:, k] - keep @ (keep.T @ prev[:, k])
                    p -= keep @ (keep.T @ p)
                    nrm = numpy.linalg.norm(p)
                    if nrm > 1e-3 and keep.shape[1] < min(2 * nroots, cap - len(new_vectors)):
                        keep = numpy.hstack([keep, (p / nrm)[:, None]])
            nk = keep.shape[1]
            for k in range(nk - nroots):
                ops.lincomb(space, n, X, keep[:, nroots + k], 0.0, xp[k])
                ops.lincomb(space, n, AX, keep[:, nroots + k], 0.0, axp[k])
            hk = keep.T @ h @ keep
            ops.copy(X[:nroots], xr)
            ops.copy(AX[:nroots], axr)
            for k in range(nk - nroots):
                ops.copy(X[nroots + k], xp[k])
                ops.copy(AX[nroots + k], axp[k])
            space = nk
            heff[:] = 0.0
            if checked:
                for j in range(nk):
                    heff[:nk, j] = ops.dots(nk, n, X, AX[j])
            else:
                heff[:nk, :nk] = 0.5 * (hk + hk.T)
            # after a check the kept Ritz vectors are the exact ones (later
            # restarts keep Ritz vectors whose approximate part is of the
            # size of the remaining error)
            nexact = nroots if refining else 0
            v = numpy.eye(nk)[:, :nroots]
        v_prev = v
        added = 0
        for t in new_vectors:
            if space < cap and add_vector(t):
                added += 1
        if added == 0:
            break
    res = DavidsonResult(conv.copy(), e_old, xr, axr, cycles, nhop, rnorm.copy())
    res.nhop_check = nhop_check
    res.nrestore = nrestore
    return res


def make_pspace_precond(hdiag, pw, pv, addr, level_shift=1e-3, ops=None):
    """Preconditioner with an exact pspace block and Olsen's correction, the
    one the reference's CSF solver builds from its pspace eigenpairs through
    fci.make_precond(hdiag, pw, pv, addr) (pyscf/csf_fci/csf.py:431-440;
    J. Olsen, P. Jorgensen, J. Simons, Chem. Phys. Lett. 169, 463 (1990)).
    With e' = e - level_shift, D = hdiag - e' (|D| < 1e-8 -> 1e-8) and
    M^-1 = 1/D off the pspace, (H0 - e')^-1 = pv (pw - e')^-1 pv^T on it:
        e1 = <x0|M^-1 r> / <x0|M^-1 x0>,   t = (r - e1 x0) / D.
    hdiag: device vector; pw, pv: eigenpairs of the pspace block (host);
    addr: pspace addresses (device int64)."""
    import torch
    lib = _lib.load()
    ops = ops if ops is not None else DeviceOps()
    n = int(hdiag.numel())
    pw = numpy.asarray(pw, dtype=numpy.float64)
    pv = numpy.asarray(pv, dtype=numpy.float64)
    addr = addr.to(device=hdiag.device, dtype=torch.int64)
    d1 = torch.empty_like(hdiag)
    blocks = getattr(ops, "ops", None)      # DeviceOps: its BlockOps (precond lists)

    def diag_solve(v, e, out):
        _lib.check(lib.fcis_precond(n, _lib.vptr(hdiag), float(e), 1e-8, _lib.vptr(v),
                                    _lib.vptr(out), _lib.stream_ptr()), "fcis_precond")

    def precond(r, e, out, x0):
        es = float(e) - level_shift
        if blocks is not None:
            # both diagonal solves in one launch, both dots in another
            hp = hdiag.data_ptr()
            blocks.precond([(hp, x0.data_ptr(), d1.data_ptr(), n, es),
                            (hp, r.data_ptr(), out.data_ptr(), n, es)], 1e-8)
        else:
            diag_solve(x0, es, d1)
            diag_solve(r, es, out)
        g = torch.stack([x0[addr], r[addr], d1[addr], out[addr]]).cpu().numpy()
        xa, ra, d1a, d2a = g
        inv = (pv / (pw - es)) @ pv.T
        if getattr(ops, "dots_list", None) is not None:
            x0_out, x0_d1 = ops.dots_list(n, [out, d1], x0)
        else:
            x0_out, x0_d1 = ops.dots(1, n, x0, out)[0], ops.dots(1, n, x0, d1)[0]
        x0_h0r = x0_out - xa @ d2a + xa @ (inv @ ra)
        x0_h0x0 = x0_d1 - xa @ d1a + xa @ (inv @ xa)
        e1 = x0_h0r / x0_h0x0
        ops.lincomb(1, n, d1, numpy.array([-e1]), 1.0, out)
    return precond
