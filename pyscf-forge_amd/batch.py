"""Many independent CAS problems per launch: batched sigma and Davidson.

``BatchSigma`` wraps fcis_batch_*: sigma = H_p c for every active vector
slot of every problem in one persistent HIP kernel, replayed from a
hipGraph.  ``solve_batch`` runs the Davidson iterations of all problems in
lockstep (each problem keeps its own subspace, Ritz pairs and convergence
flags, as in one fcisolver.kernel call per problem, fci_gas.py:1150-1399),
so that one iteration of the whole batch costs one batched sigma launch
plus a handful of block vector kernels (fcis_gram_list, fcis_gemm_list,
fcis_precond_list) whatever the number of problems.
"""

import ctypes
import math
import time

import numpy

from . import _lib
from . import direct_spin1 as _ds
from .blockops import BlockOps  # noqa: F401  (re-exported)

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

    def tune(self, cin, cout, reps=2):
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
                lindep=1e-14, level_shift=1e-3, tune=True, use_graph=True, verbose=False,
                nguard=None):
    """Lowest ``nroots`` states of every problem (h1e, eri, norb, nelec).

    Davidson per problem as direct_spin1.FCISolver.kernel (initial guess of
    fci_gas.py:881-902, preconditioner :1337-1342, convergence |dE| < tol
    and |r| < tol_residual = sqrt(tol), thick restart, the reference's
    max_space 12 + 3 per extra root), all problems advancing together.
    ``nguard`` extra roots (default: one when nroots > 1) are iterated
    beside the requested ones but not required to converge: their
    correction vectors keep the subspace from settling on the state above a
    root it has not yet found (without the guard, one problem of the
    64-problem synthetic set converges its fourth root to the fifth state).
    Returns a BatchResult (energies without ecore)."""
    if nguard is None:
        nguard = 1 if int(nroots) > 1 else 0
    import torch
    t_setup = time.perf_counter()
    lib = _lib.load()
    ops = BlockOps()
    stream = _lib.stream_ptr()
    if tol_residual is None:
        tol_residual = math.sqrt(tol)
    spaces, h2es, ints = _prepare(problems)
    P = len(spaces)
    nd = numpy.array([s.ndet for s in spaces], dtype=numpy.int64)
    kreq = numpy.minimum(int(nroots), nd).astype(int)
    kk = numpy.minimum(int(nroots) + int(nguard), nd).astype(int)
    if kk.max() > 8:
        raise ValueError("at most 8 roots (with guards) per problem")
    cap = numpy.array([min(max(int(max_space) + (k - 1) * 3, 2 * k), 64, n)
                       for k, n in zip(kk, nd)], dtype=int)
    t_prep = time.perf_counter() - t_setup
    bs = BatchSigma(spaces, h2es, kk)
    t_create = time.perf_counter() - t_setup - t_prep
    tin, tout = bs.empty(), bs.empty()

    def packed(counts):
        off = numpy.concatenate([[0], numpy.cumsum(counts * nd)[:-1]]).astype(numpy.int64)
        buf = torch.empty(int((counts * nd).sum()), dtype=torch.float64, device="cuda")
        return buf, off

    X, xoff = packed(cap)
    AX, _ = packed(cap)
    HD, hoff = packed(numpy.ones(P, dtype=int))
    stride = 8 * nd
    xbase = X.data_ptr() + 8 * xoff
    axbase = AX.data_ptr() + 8 * xoff
    tbase = tin.data_ptr() + 8 * bs.offsets
    obase = tout.data_ptr() + 8 * bs.offsets

    def vecs(base, p, i0, i1):
        return base[p] + stride[p] * numpy.arange(i0, i1, dtype=numpy.int64)

    for p, (sp, (h1, er)) in enumerate(zip(spaces, ints)):
        hd = HD[int(hoff[p]):int(hoff[p] + nd[p])]
        _lib.check(lib.fcis_make_hdiag(sp.handle, _lib.dptr(h1), _lib.dptr(er), _lib.vptr(hd),
                                       stream), "fcis_make_hdiag")
        addrs = _ds.lowest_addresses(hd, int(kk[p]))
        for r, a in enumerate(addrs):
            _lib.check(lib.fcis_init_guess(int(nd[p]), int(a), r, _ds.INIT_NOISE_NORM,
                                           ctypes.c_void_p(int(tbase[p] + stride[p] * r)), stream),
                       "fcis_init_guess")
    torch.cuda.synchronize()
    t_tune = time.perf_counter()
    if tune:
        # one timed launch per candidate schedule: a launch with every slot
        # active takes ~24 ms at the 64-problem workload, long enough that one
        # timing ranks the candidates (a second repetition cost 0.14 s per solve)
        bs.tune(tin, tout, reps=1)
    bs.set_graph(use_graph)
    t_tune = time.perf_counter() - t_tune
    t_setup = time.perf_counter() - t_setup

    space = numpy.zeros(P, dtype=int)
    heff = [numpy.zeros((c, c)) for c in cap]
    e_old = [numpy.full(k, numpy.inf) for k in kk]
    conv = [numpy.zeros(k, dtype=bool) for k in kk]
    rnorm = [numpy.full(k, numpy.inf) for k in kk]
    vfin = [None] * P
    done = numpy.zeros(P, dtype=bool)
    cycles = numpy.zeros(P, dtype=int)
    nhop = numpy.zeros(P, dtype=int)
    v_prev = [None] * P
    tstat = dict(nsigma=0, t_sigma=0.0)

    def project_out(act, cnt):
        """T[p][0..cnt) -= X[p] (X[p]^T T[p]) (one classical Gram-Schmidt pass)."""
        pr = [p for p in act if space[p] > 0 and cnt[p] > 0]
        if not pr:
            return
        G = ops.gram([(vecs(xbase, p, 0, space[p]), vecs(tbase, p, 0, cnt[p]), nd[p]) for p in pr])
        ops.gemm([(vecs(xbase, p, 0, space[p]), vecs(tbase, p, 0, cnt[p]), nd[p], -g, 1.0)
                  for p, g in zip(pr, G)])

    def orthonormalise(act, cnt, thresh):
        """T[p][0..cnt) -> orthonormal T[p][0..a) spanning the directions of
        the block Gram matrix above thresh[p] (in place); returns a."""
        pr = [p for p in act if cnt[p] > 0]
        G = ops.gram([(vecs(tbase, p, 0, cnt[p]), vecs(tbase, p, 0, cnt[p]), nd[p]) for p in pr])
        got = numpy.zeros(P, dtype=int)
        blocks = []
        for p, g in zip(pr, G):
            w, u = numpy.linalg.eigh(0.5 * (g + g.T))
            keep = numpy.nonzero(w > thresh[p])[0][::-1]
            keep = keep[:cap[p] - space[p]]
            if len(keep):
                blocks.append((vecs(tbase, p, 0, cnt[p]), vecs(tbase, p, 0, len(keep)), nd[p],
                               u[:, keep] / numpy.sqrt(w[keep]), 0.0))
            got[p] = len(keep)
        ops.gemm(blocks)
        return got

    def orth_and_apply(cnt):
        """Candidates T[p][0..cnt): CGS2 against X[p], two-pass block
        orthonormalisation, append to X, apply H, append to AX, extend heff."""
        act = [p for p in range(P) if cnt[p] > 0]
        if not act:
            return numpy.zeros(P, dtype=int)
        pr = [p for p in act if space[p] > 0]
        # norms of the raw candidates, for the relative dependence test
        G0 = ops.gram([(vecs(tbase, p, 0, cnt[p]), vecs(tbase, p, 0, cnt[p]), nd[p]) for p in act])
        n0 = {p: float(numpy.diag(g).max()) for p, g in zip(act, G0)}
        project_out(pr, cnt)
        project_out(pr, cnt)
        th1 = {p: max(lindep, 1e-14 * n0[p]) for p in act}
        a1 = orthonormalise(act, cnt, th1)
        project_out(pr, a1)
        a2 = orthonormalise(act, a1, {p: 1e-8 for p in act})
        act = [p for p in act if a2[p] > 0]
        # append to X (copy; T keeps the vectors as the sigma input)
        ops.gemm([(vecs(tbase, p, 0, a2[p]), vecs(xbase, p, space[p], space[p] + a2[p]), nd[p],
                   numpy.eye(a2[p]), 0.0) for p in act])
        if act:
            mask = numpy.zeros(bs.nslots, dtype=numpy.uint8)
            for p in act:
                mask[bs.slot_base[p]:bs.slot_base[p] + a2[p]] = 1
            bs.set_active(mask)
            t0 = time.perf_counter()
            bs(tin, tout)
            torch.cuda.synchronize()
            tstat["t_sigma"] += time.perf_counter() - t0
            tstat["nsigma"] += 1
            ops.gemm([(vecs(obase, p, 0, a2[p]), vecs(axbase, p, space[p], space[p] + a2[p]), nd[p],
                       numpy.eye(a2[p]), 0.0) for p in act])
            Hc = ops.gram([(vecs(xbase, p, 0, space[p] + a2[p]),
                            vecs(axbase, p, space[p], space[p] + a2[p]), nd[p]) for p in act])
            for p, hcol in zip(act, Hc):
                s0, s1 = space[p], space[p] + a2[p]
                heff[p][:s1, s0:s1] = hcol
                heff[p][s0:s1, :s1] = hcol.T
                nhop[p] += a2[p]
                space[p] = s1
        return a2

    t_iter = time.perf_counter()
    orth_and_apply(kk.copy())
    if (space < kk).any():
        raise RuntimeError("Davidson: initial guesses are linearly dependent")

    for icyc in range(int(max_cycle)):
        act = [p for p in range(P) if not done[p]]
        if not act:
            break
        V = {}
        for p in act:
            sp_ = space[p]
            h = 0.5 * (heff[p][:sp_, :sp_] + heff[p][:sp_, :sp_].T)
            w, v = numpy.linalg.eigh(h)
            V[p] = (w[:kk[p]].copy(), v[:, :kk[p]].copy(), h)
            cycles[p] = icyc + 1
        # residuals r_j = sum_i V_ij (AX_i - e_j X_i) into tout
        ops.gemm([(numpy.concatenate([vecs(xbase, p, 0, space[p]), vecs(axbase, p, 0, space[p])]),
                   vecs(obase, p, 0, kk[p]), nd[p],
                   numpy.vstack([-V[p][1] * V[p][0][None, :], V[p][1]]), 0.0) for p in act])
        R = ops.gram([(vecs(obase, p, 0, kk[p]), vecs(obase, p, 0, kk[p]), nd[p]) for p in act])
        newcnt = numpy.zeros(P, dtype=int)
        pre = []
        for p, rr in zip(act, R):
            e = V[p][0]
            de = e - e_old[p]
            rnorm[p] = numpy.sqrt(numpy.maximum(numpy.diag(rr), 0.0))
            conv[p] = (numpy.abs(de) < tol) & (rnorm[p] < tol_residual)
            e_old[p] = e
            if conv[p][:kreq[p]].all():
                done[p] = True
                vfin[p] = V[p][1][:, :kreq[p]].copy()
                continue
            for k in numpy.nonzero(~conv[p])[0]:
                pre.append((int(HD.data_ptr() + 8 * hoff[p]), int(obase[p] + stride[p] * k),
                            int(tbase[p] + stride[p] * newcnt[p]), int(nd[p]), float(e[k])))
                newcnt[p] += 1
        if verbose:
            print("batch davidson %d: %d active, max |r| %.2e" % (
                icyc, len(act), max(float(rnorm[p].max()) for p in act)))
        ops.precond(pre, level_shift)
        if done.all():
            break
        # thick restart where the new block does not fit: X <- X keep in place
        rs = []
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
            rs.append((p, keep))
        if rs:
            ops.gemm([b for p, keep in rs for b in (
                (vecs(xbase, p, 0, space[p]), vecs(xbase, p, 0, keep.shape[1]), nd[p], keep, 0.0),
                (vecs(axbase, p, 0, space[p]), vecs(axbase, p, 0, keep.shape[1]), nd[p], keep, 0.0))])
            for p, keep in rs:
                nk = keep.shape[1]
                hk = keep.T @ V[p][2] @ keep
                heff[p][:] = 0.0
                heff[p][:nk, :nk] = 0.5 * (hk + hk.T)
                space[p] = nk
                V[p] = (V[p][0], numpy.eye(nk)[:, :kk[p]], None)
        for p in act:
            if not done[p]:
                v_prev[p] = V[p][1]
        added = orth_and_apply(newcnt)
        for p in act:
            if not done[p] and newcnt[p] > 0 and added[p] == 0:
                done[p] = True      # no new direction: stop this problem
                vfin[p] = V[p][1][:, :kreq[p]].copy()
    # problems still running at max_cycle: Ritz vectors of their last step
    for p in range(P):
        if vfin[p] is None:
            sp_ = space[p]
            h = 0.5 * (heff[p][:sp_, :sp_] + heff[p][:sp_, :sp_].T)
            w, v = numpy.linalg.eigh(h)
            vfin[p] = v[:, :kreq[p]]
    XR, roff = packed(kreq)
    rbase = XR.data_ptr() + 8 * roff
    ops.gemm([(vecs(xbase, p, 0, vfin[p].shape[0]), vecs(rbase, p, 0, kreq[p]), nd[p], vfin[p], 0.0)
              for p in range(P)])
    torch.cuda.synchronize()
    t_iter = time.perf_counter() - t_iter
    ci = [XR[int(roff[p]):int(roff[p] + kreq[p] * nd[p])].view(int(kreq[p]), int(nd[p]))
          for p in range(P)]
    rows, grid = bs.schedule()
    stats = dict(time_setup=t_setup, time_prepare=t_prep, time_create=t_create, time_tune=t_tune,
                 time_iterations=t_iter, sigma_launches=tstat["nsigma"],
                 time_sigma=tstat["t_sigma"], rows_per_item=rows, grid=grid, tuned_ms=bs.tuned_ms,
                 ndet_total=int(nd.sum()))
    res = BatchResult([e[:k].copy() for e, k in zip(e_old, kreq)],
                      [c[:k].copy() for c, k in zip(conv, kreq)], cycles.copy(), nhop.copy(),
                      [r[:k].copy() for r, k in zip(rnorm, kreq)], ci, stats)
    res._keep = (bs, X, AX, HD, tin, tout)
    return res
