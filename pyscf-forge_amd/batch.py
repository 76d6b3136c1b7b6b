"""Many independent CAS problems per launch: batched sigma and Davidson.

``BatchSigma`` wraps fcis_batch_*: sigma = H_p c for every active vector
slot of every problem in one persistent HIP kernel, replayed from a
hipGraph.  ``solve_batch`` runs the Davidson iterations of all problems in
lockstep (each problem keeps its own subspace, Ritz pairs and convergence
flags, as in one fcisolver.kernel call per problem, fci_gas.py:1150-1399),
so that one iteration of the whole batch costs one batched sigma launch
plus a handful of list-form vector kernels (fcis_dots_list,
fcis_lincomb_list, fcis_precond_list) whatever the number of problems.
"""

import ctypes
import math
import time

import numpy

from . import _lib
from . import direct_spin1 as _ds

DOT_DTYPE = numpy.dtype([("x", "<u8"), ("y", "<u8"), ("n", "<i8")])
LIN_DTYPE = numpy.dtype([("y", "<u8"), ("n", "<i8"), ("beta", "<f8"), ("m", "<i4"), ("off", "<i4")])
PRE_DTYPE = numpy.dtype([("hdiag", "<u8"), ("r", "<u8"), ("out", "<u8"), ("n", "<i8"),
                         ("e", "<f8")])


class ListOps:
    """Vector algebra over many device vectors of different lengths, one
    kernel launch per call.  Vectors are given by device address and length."""

    def __init__(self):
        self.lib = _lib.load()

    def dots(self, x, y, n):
        """[<x_i, y_i>] for address arrays x, y and lengths n."""
        m = len(n)
        if m == 0:
            return numpy.zeros(0)
        ent = numpy.empty(m, DOT_DTYPE)
        ent["x"] = x
        ent["y"] = y
        ent["n"] = n
        out = numpy.empty(m)
        _lib.check(self.lib.fcis_dots_list(m, ent.ctypes.data, out.ctypes.data,
                                           _lib.stream_ptr()), "fcis_dots_list")
        return out

    def lincomb(self, entries):
        """entries: list of (y_addr, n, beta, [(x_addr, coef), ...]):
        y = beta*y + sum coef*x."""
        if not entries:
            return
        nterm = sum(len(e[3]) for e in entries)
        ent = numpy.empty(len(entries), LIN_DTYPE)
        xs = numpy.empty(max(nterm, 1), numpy.uint64)
        cf = numpy.empty(max(nterm, 1))
        off = 0
        for i, (y, n, beta, terms) in enumerate(entries):
            ent[i] = (y, n, beta, len(terms), off)
            for x, c in terms:
                xs[off] = x
                cf[off] = c
                off += 1
        _lib.check(self.lib.fcis_lincomb_list(len(entries), ent.ctypes.data, nterm,
                                              xs.ctypes.data, cf.ctypes.data, _lib.stream_ptr()),
                   "fcis_lincomb_list")

    def precond(self, entries, level_shift):
        """entries: list of (hdiag_addr, r_addr, out_addr, n, e)."""
        if not entries:
            return
        ent = numpy.array(entries, dtype=PRE_DTYPE)
        _lib.check(self.lib.fcis_precond_list(len(entries), ent.ctypes.data, float(level_shift),
                                              _lib.stream_ptr()), "fcis_precond_list")


class BatchSigma:
    """sigma for nvec[p] vector slots of each of many CAS problems.

    spaces: list of direct_spin1.Space; h2es: absorbed pair matrices.
    Vectors are packed: slot v of problem p is buf[offsets[p] + v*ndet_p :
    ... + ndet_p] of a flat float64 device tensor of length ``total``."""

    def __init__(self, spaces, h2es, nvec, precision="fp64"):
        self.lib = _lib.load()
        n = len(spaces)
        if n == 0:
            raise ValueError("empty batch")
        self.spaces = list(spaces)
        self._h2e = [numpy.ascontiguousarray(h, dtype=numpy.float64) for h in h2es]
        self.nvec = [int(v) for v in nvec]
        sp_arr = (ctypes.c_void_p * n)(*[s.handle.value for s in self.spaces])
        h_arr = (ctypes.c_void_p * n)(*[h.ctypes.data for h in self._h2e])
        nv = (ctypes.c_int * n)(*self.nvec)
        h = ctypes.c_void_p()
        _lib.check(self.lib.fcis_batch_create(ctypes.byref(h), n, sp_arr, h_arr, nv,
                                              _lib.PRECISIONS[precision], _lib.stream_ptr()),
                   "fcis_batch_create")
        self.handle = h
        self.precision = precision
        total = ctypes.c_int64()
        self.offsets = numpy.zeros(n, dtype=numpy.int64)
        _lib.check(self.lib.fcis_batch_info(h, ctypes.byref(total), self.offsets.ctypes.data,
                                            None, None), "fcis_batch_info")
        self.total = int(total.value)
        self.ndet = numpy.array([s.ndet for s in self.spaces], dtype=numpy.int64)
        self.slot_base = numpy.concatenate([[0], numpy.cumsum(self.nvec)[:-1]]).astype(numpy.int64)
        self.nslots = int(sum(self.nvec))
        self.tuned_ms = None

    def schedule(self):
        rows, grid = ctypes.c_int(), ctypes.c_int()
        _lib.check(self.lib.fcis_batch_info(self.handle, None, None, ctypes.byref(rows),
                                            ctypes.byref(grid)), "fcis_batch_info")
        return rows.value, grid.value

    def empty(self):
        import torch
        return torch.empty(self.total, dtype=torch.float64, device="cuda")

    def slot(self, buf, p, v=0):
        o = int(self.offsets[p] + v * self.ndet[p])
        return buf[o:o + int(self.ndet[p])]

    def set_active(self, mask):
        mask = numpy.ascontiguousarray(mask, dtype=numpy.uint8).reshape(-1)
        if mask.size != self.nslots:
            raise ValueError("active mask needs %d entries" % self.nslots)
        _lib.check(self.lib.fcis_batch_set_active(self.handle, mask.ctypes.data,
                                                  _lib.stream_ptr()), "fcis_batch_set_active")

    def set_graph(self, enable):
        _lib.check(self.lib.fcis_batch_set_graph(self.handle, 1 if enable else 0),
                   "fcis_batch_set_graph")

    def tune(self, cin, cout, reps=3):
        ms = ctypes.c_double()
        _lib.check(self.lib.fcis_batch_tune(self.handle, _lib.vptr(cin), _lib.vptr(cout), int(reps),
                                            ctypes.byref(ms), _lib.stream_ptr()), "fcis_batch_tune")
        self.tuned_ms = ms.value
        return ms.value

    def __call__(self, cin, cout):
        if cin.numel() != self.total or cout.numel() != self.total:
            raise ValueError("batch buffers must have %d elements" % self.total)
        _lib.check(self.lib.fcis_batch_sigma(self.handle, _lib.vptr(cin), _lib.vptr(cout),
                                             _lib.stream_ptr()), "fcis_batch_sigma")

    def close(self):
        if getattr(self, "handle", None) is not None:
            self.lib.fcis_batch_free(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchResult:
    def __init__(self, e, converged, cycles, nhop, rnorm, ci, stats):
        self.e = e                      # list of arrays (nroots_p,)
        self.converged = converged      # list of bool arrays
        self.cycles = cycles            # per problem
        self.nhop = nhop
        self.rnorm = rnorm
        self.ci = ci                    # list of (nroots_p, ndet_p) device tensors (views)
        self.stats = stats


def _prepare(problems):
    """problems: list of (h1e, eri, norb, nelec) -> spaces, h2e, h1e, eri."""
    spaces, h2es, ints = [], [], []
    for h1e, eri, norb, nelec in problems:
        nelec = _ds._unpack_nelec(nelec)
        sp = _ds.get_space(norb, nelec)
        h1 = numpy.ascontiguousarray(h1e, dtype=numpy.float64).reshape(sp.norb, sp.norb)
        er = _ds.pair_matrix(eri, sp.norb)
        spaces.append(sp)
        h2es.append(_ds.absorb_h1e(h1, er, norb, nelec, 0.5))
        ints.append((h1, er))
    return spaces, h2es, ints


def solve_batch(problems, nroots=4, tol=1e-10, tol_residual=None, max_cycle=100, max_space=12,
                lindep=1e-14, level_shift=1e-3, tune=True, use_graph=True, verbose=False):
    """Lowest ``nroots`` states of every problem (h1e, eri, norb, nelec).

    Davidson per problem as direct_spin1.FCISolver.kernel (initial guess of
    fci_gas.py:881-902, preconditioner :1337-1342, convergence |dE| < tol
    and |r| < tol_residual = sqrt(tol), thick restart), all problems advancing
    together.  Returns a BatchResult (energies without ecore)."""
    import torch
    t_setup = time.perf_counter()
    lib = _lib.load()
    ops = ListOps()
    stream = _lib.stream_ptr()
    if tol_residual is None:
        tol_residual = math.sqrt(tol)
    spaces, h2es, ints = _prepare(problems)
    P = len(spaces)
    nd = numpy.array([s.ndet for s in spaces], dtype=numpy.int64)
    kk = numpy.minimum(int(nroots), nd).astype(int)
    cap = numpy.array([min(max(int(max_space) + (k - 1) * 3, 2 * k), 64, n)
                       for k, n in zip(kk, nd)], dtype=int)
    bs = BatchSigma(spaces, h2es, kk)
    tin, tout = bs.empty(), bs.empty()

    # packed device storage
    def packed(counts):
        off = numpy.concatenate([[0], numpy.cumsum(counts * nd)[:-1]]).astype(numpy.int64)
        buf = torch.empty(int((counts * nd).sum()), dtype=torch.float64, device="cuda")
        return buf, off

    X, xoff = packed(cap)
    AX, _ = packed(cap)
    XR, roff = packed(kk)
    AXR, _ = packed(kk)
    XP, _ = packed(kk)
    AXP, _ = packed(kk)
    HD, hoff = packed(numpy.ones(P, dtype=int))
    b8 = 8

    def addr(buf, off, p, i):
        return buf.data_ptr() + b8 * int(off[p] + i * nd[p])

    tin_addr = lambda p, i: tin.data_ptr() + b8 * int(bs.offsets[p] + i * nd[p])   # noqa: E731
    tout_addr = lambda p, i: tout.data_ptr() + b8 * int(bs.offsets[p] + i * nd[p])  # noqa: E731

    # hdiag and initial guesses (lowest diagonal + splitmix64 noise)
    for p, (sp, (h1, er)) in enumerate(zip(spaces, ints)):
        hd = HD[int(hoff[p]):int(hoff[p] + nd[p])]
        _lib.check(lib.fcis_make_hdiag(sp.handle, _lib.dptr(h1), _lib.dptr(er), _lib.vptr(hd),
                                       stream), "fcis_make_hdiag")
        addrs = torch.topk(hd, int(kk[p]), largest=False).indices.cpu().numpy()
        for r, a in enumerate(addrs):
            _lib.check(lib.fcis_init_guess(int(nd[p]), int(a), r, _ds.INIT_NOISE_NORM,
                                           ctypes.c_void_p(tin_addr(p, r)), stream),
                       "fcis_init_guess")
    torch.cuda.synchronize()
    if tune:
        bs.tune(tin, tout)
    bs.set_graph(use_graph)
    t_setup = time.perf_counter() - t_setup

    space = numpy.zeros(P, dtype=int)
    heff = [numpy.zeros((c, c)) for c in cap]
    e_old = [numpy.full(k, numpy.inf) for k in kk]
    conv = [numpy.zeros(k, dtype=bool) for k in kk]
    rnorm = [numpy.full(k, numpy.inf) for k in kk]
    done = numpy.zeros(P, dtype=bool)
    cycles = numpy.zeros(P, dtype=int)
    nhop = numpy.zeros(P, dtype=int)
    v_prev = [None] * P
    nsigma = 0
    t_sigma = 0.0

    def orth_and_apply(newcnt):
        """The first newcnt[p] vectors of tin (problem p) are candidates:
        project out X[p][:space], orthonormalise within the block (dropping
        dependent directions), append to X, apply H, append to AX, update
        heff.  Returns the number of vectors added per problem."""
        nonlocal nsigma, t_sigma
        act = [p for p in range(P) if newcnt[p] > 0]
        norm0 = {}
        if act:
            d = ops.dots([tin_addr(p, j) for p in act for j in range(newcnt[p])],
                         [tin_addr(p, j) for p in act for j in range(newcnt[p])],
                         [nd[p] for p in act for j in range(newcnt[p])])
            i = 0
            for p in act:
                norm0[p] = numpy.sqrt(numpy.maximum(d[i:i + newcnt[p]], 0.0))
                i += newcnt[p]
        for _ in range(2):
            pr = [p for p in act if space[p] > 0]
            if not pr:
                break
            xa, ya, ln, idx = [], [], [], []
            for p in pr:
                for j in range(newcnt[p]):
                    for i in range(space[p]):
                        xa.append(addr(X, xoff, p, i))
                        ya.append(tin_addr(p, j))
                        ln.append(nd[p])
                        idx.append((p, j, i))
            c = ops.dots(xa, ya, ln)
            coef = {}
            for (p, j, i), v in zip(idx, c):
                coef.setdefault((p, j), []).append((addr(X, xoff, p, i), -v))
            ops.lincomb([(tin_addr(p, j), nd[p], 1.0, coef[(p, j)]) for p in pr
                         for j in range(newcnt[p])])
        # Gram matrices of the blocks
        xa, ya, ln, idx = [], [], [], []
        for p in act:
            for j in range(newcnt[p]):
                for l in range(j + 1):
                    xa.append(tin_addr(p, j))
                    ya.append(tin_addr(p, l))
                    ln.append(nd[p])
                    idx.append((p, j, l))
        g = ops.dots(xa, ya, ln)
        G = {p: numpy.zeros((newcnt[p], newcnt[p])) for p in act}
        for (p, j, l), v in zip(idx, g):
            G[p][j, l] = G[p][l, j] = v
        added = numpy.zeros(P, dtype=int)
        ent = []
        for p in act:
            w, u = numpy.linalg.eigh(G[p])
            keep = [q for q in range(len(w)) if w[q] > lindep and
                    math.sqrt(max(w[q], 0.0)) > 1e-7 * float(norm0[p].max())]
            keep = keep[::-1][:cap[p] - space[p]]      # largest first
            for qi, q in enumerate(keep):
                terms = [(tin_addr(p, j), u[j, q] / math.sqrt(w[q])) for j in range(newcnt[p])]
                ent.append((addr(X, xoff, p, space[p] + qi), nd[p], 0.0, terms))
            added[p] = len(keep)
        ops.lincomb(ent)
        # stage the new vectors for sigma
        ops.lincomb([(tin_addr(p, q), nd[p], 0.0, [(addr(X, xoff, p, space[p] + q), 1.0)])
                     for p in act for q in range(added[p])])
        mask = numpy.zeros(bs.nslots, dtype=numpy.uint8)
        for p in act:
            mask[bs.slot_base[p]:bs.slot_base[p] + added[p]] = 1
        if mask.any():
            bs.set_active(mask)
            t0 = time.perf_counter()
            bs(tin, tout)
            torch.cuda.synchronize()
            t_sigma += time.perf_counter() - t0
            nsigma += 1
            ops.lincomb([(addr(AX, xoff, p, space[p] + q), nd[p], 0.0, [(tout_addr(p, q), 1.0)])
                         for p in act for q in range(added[p])])
            nhop[:] += added
            # heff columns
            xa, ya, ln, idx = [], [], [], []
            for p in act:
                for q in range(added[p]):
                    col = space[p] + q
                    for i in range(col + 1):
                        xa.append(addr(X, xoff, p, i))
                        ya.append(addr(AX, xoff, p, col))
                        ln.append(nd[p])
                        idx.append((p, i, col))
            h = ops.dots(xa, ya, ln)
            for (p, i, col), v in zip(idx, h):
                heff[p][i, col] = heff[p][col, i] = v
        space[:] += added
        return added

    t_iter = time.perf_counter()
    added = orth_and_apply(kk.copy())
    if (space < kk).any():
        raise RuntimeError("Davidson: initial guesses are linearly dependent")

    for icyc in range(int(max_cycle)):
        act = [p for p in range(P) if not done[p]]
        if not act:
            break
        V = {}
        ritz = []
        for p in act:
            s = space[p]
            h = 0.5 * (heff[p][:s, :s] + heff[p][:s, :s].T)
            w, v = numpy.linalg.eigh(h)
            V[p] = (w[:kk[p]].copy(), v[:, :kk[p]].copy(), h)
            cycles[p] = icyc + 1
        # Ritz vectors and their images
        ent = []
        for p in act:
            e, v, _ = V[p]
            for k in range(kk[p]):
                ent.append((addr(XR, roff, p, k), nd[p], 0.0,
                            [(addr(X, xoff, p, i), v[i, k]) for i in range(space[p])]))
                ent.append((addr(AXR, roff, p, k), nd[p], 0.0,
                            [(addr(AX, xoff, p, i), v[i, k]) for i in range(space[p])]))
        ops.lincomb(ent)
        # residuals r = A x - e x into tin
        ops.lincomb([(tin_addr(p, k), nd[p], 0.0, [(addr(AXR, roff, p, k), 1.0),
                                                   (addr(XR, roff, p, k), -V[p][0][k])])
                     for p in act for k in range(kk[p])])
        rr = ops.dots([tin_addr(p, k) for p in act for k in range(kk[p])],
                      [tin_addr(p, k) for p in act for k in range(kk[p])],
                      [nd[p] for p in act for k in range(kk[p])])
        i = 0
        newcnt = numpy.zeros(P, dtype=int)
        pre = []
        for p in act:
            e = V[p][0]
            de = e - e_old[p]
            rn = numpy.sqrt(numpy.maximum(rr[i:i + kk[p]], 0.0))
            i += kk[p]
            rnorm[p] = rn
            conv[p] = (numpy.abs(de) < tol) & (rn < tol_residual)
            e_old[p] = e
            if conv[p].all():
                done[p] = True
                continue
            for k in numpy.nonzero(~conv[p])[0]:
                pre.append((HD.data_ptr() + b8 * int(hoff[p]), tin_addr(p, k),
                            tin_addr(p, newcnt[p]), nd[p], float(e[k])))
                newcnt[p] += 1
        if verbose:
            print("batch davidson %d: %d active, max |r| %.2e" % (
                icyc, len(act), max(float(rnorm[p].max()) for p in act)))
        # preconditioned residuals land in tout at their root's slot, then
        # move to the compacted candidate slots of tin (no entry reads a
        # slot another entry writes)
        if pre:
            shift = tout.data_ptr() - tin.data_ptr()
            ops.precond([(hd_a, r_a, r_a + shift, n, e) for (hd_a, r_a, _, n, e) in pre],
                        level_shift)
            ops.lincomb([(o_a, n, 0.0, [(r_a + shift, 1.0)]) for (_, r_a, o_a, n, _) in pre])
        if done.all():
            break
        # thick restart where the new block does not fit
        for p in act:
            if done[p] or space[p] + newcnt[p] <= cap[p]:
                continue
            e, v, h = V[p]
            keep = v.copy()
            if v_prev[p] is not None:
                prev = numpy.zeros((space[p], v_prev[p].shape[1]))
                prev[:v_prev[p].shape[0]] = v_prev[p]
                for k in range(prev.shape[1]):
                    q = prev[:, k] - keep @ (keep.T @ prev[:, k])
                    q -= keep @ (keep.T @ q)
                    nrm = numpy.linalg.norm(q)
                    if nrm > 1e-3 and keep.shape[1] < min(2 * kk[p], cap[p] - newcnt[p]):
                        keep = numpy.hstack([keep, (q / nrm)[:, None]])
            nk = keep.shape[1]
            ent = []
            for k in range(nk - kk[p]):
                ent.append((addr(XP, roff, p, k), nd[p], 0.0,
                            [(addr(X, xoff, p, i), keep[i, kk[p] + k]) for i in range(space[p])]))
                ent.append((addr(AXP, roff, p, k), nd[p], 0.0,
                            [(addr(AX, xoff, p, i), keep[i, kk[p] + k]) for i in range(space[p])]))
            ops.lincomb(ent)
            ent = []
            for k in range(kk[p]):
                ent.append((addr(X, xoff, p, k), nd[p], 0.0, [(addr(XR, roff, p, k), 1.0)]))
                ent.append((addr(AX, xoff, p, k), nd[p], 0.0, [(addr(AXR, roff, p, k), 1.0)]))
            for k in range(nk - kk[p]):
                ent.append((addr(X, xoff, p, kk[p] + k), nd[p], 0.0, [(addr(XP, roff, p, k), 1.0)]))
                ent.append((addr(AX, xoff, p, kk[p] + k), nd[p], 0.0, [(addr(AXP, roff, p, k), 1.0)]))
            ops.lincomb(ent)
            hk = keep.T @ h @ keep
            heff[p][:] = 0.0
            heff[p][:nk, :nk] = 0.5 * (hk + hk.T)
            space[p] = nk
            V[p] = (e, numpy.eye(nk)[:, :kk[p]], None)
        for p in act:
            if not done[p]:
                v_prev[p] = V[p][1]
        added = orth_and_apply(newcnt)
        for p in act:
            if not done[p] and newcnt[p] > 0 and added[p] == 0:
                done[p] = True      # no new direction: stop this problem
    torch.cuda.synchronize()
    t_iter = time.perf_counter() - t_iter
    ci = [XR[int(roff[p]):int(roff[p] + kk[p] * nd[p])].view(int(kk[p]), int(nd[p]))
          for p in range(P)]
    rows, grid = bs.schedule()
    stats = dict(time_setup=t_setup, time_iterations=t_iter, sigma_launches=nsigma,
                 time_sigma=t_sigma, rows_per_item=rows, grid=grid, tuned_ms=bs.tuned_ms,
                 ndet_total=int(nd.sum()))
    res = BatchResult([e.copy() for e in e_old], [c.copy() for c in conv], cycles.copy(),
                      nhop.copy(), [r.copy() for r in rnorm], ci, stats)
    res._keep = (bs, X, AX, XR, AXR, XP, AXP, HD, tin, tout)
    return res
