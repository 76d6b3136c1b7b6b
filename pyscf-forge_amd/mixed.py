"""Davidson drivers: plain FP64, and mixed precision with FP64 refinement.

Mixed precision (precision "fp32" or "bf16x3"): the Davidson iterations use
a sigma whose t1 contraction runs in FP32 MFMA or in a 3-term bf16 split
(FP32 accumulation), while residuals, orthogonalisation and the subspace
eigenproblem stay FP64.  When the low-precision residual reaches its noise
floor, the Ritz vectors are applied with the FP64 sigma and the same
subspace goes on from them (iterative refinement inside one Davidson run,
davidson.davidson with hop_check): new corrections are still applied in low
precision, and every acceptance is decided on FP64 images against the FP64
criteria of the reference's Davidson call (tol on the energy change,
sqrt(tol) on the residual, fci_gas.py:1344-1350), so the returned energy and
vectors satisfy the same convergence test as a plain FP64 solve.
"""

import time

import numpy

from . import _lib
from . import davidson as _dv

# low-precision stage: residual floor and energy-change tolerance
LOW_STAGE = {
    "fp32": dict(tol=1e-8, tol_residual=2e-5),
    "bf16x3": dict(tol=1e-6, tol_residual=5e-4),
}


def _davidson_bytes(ndet, nroots, max_space):
    cap = min(int(max_space) + (nroots - 1) * 3, 64)
    cap = max(cap, 2 * nroots)
    return (2 * cap + 2 * nroots + 2 + nroots) * ndet * 8


def _plan(sp, basis, h2e, precision, ws):
    """The sigma plan of a solve: the basis's own (``basis.make_plan``, a
    space without a CAS string space) or a SigmaPlan on ``sp``."""
    plan = basis.make_plan(h2e, precision, ws) if hasattr(basis, "make_plan") else None
    if plan is not None:
        return plan
    from .direct_spin1 import SigmaPlan
    return SigmaPlan(sp, h2e, precision, ws)


def solve(solver, sp, h2e, hdiag, nroots, ci0, precision, tol=1e-10, tol_residual=None,
          max_cycle=50, max_space=12, lindep=1e-14, level_shift=1e-3, verbose=0,
          basis=None, precond=None, follow_state=True):
    """Davidson solve in the determinant basis, or, with ``basis`` (an object
    with ``n`` and ``hop(plan)``, e.g. a CSF transformer), in that basis with
    H x = basis.hop(plan)(x, out); hdiag and ci0 are then in that basis.
    ``precond(r, e, out, x0)`` replaces the diagonal preconditioner
    r / (hdiag - e) (|hdiag - e| < level_shift -> level_shift).
    ``follow_state`` is davidson1's large-residual restart, which the
    reference's solvers switch on (fci_gas.py:1350 without a spin penalty,
    csf.py:489)."""
    import torch

    lib = _lib.load()
    n = sp.ndet if basis is None else int(basis.n)
    stream = _lib.stream_ptr()

    def diag_precond(r, e, out, x0=None):
        _lib.check(lib.fcis_precond(n, _lib.vptr(hdiag), float(e), float(level_shift),
                                    _lib.vptr(r), _lib.vptr(out), stream), "fcis_precond")

    if precond is None:
        precond = diag_precond

    if ci0 is None:
        guesses = solver._init_guess_dev(hdiag, nroots)
    else:
        if isinstance(ci0, (list, tuple)):
            vecs = list(ci0)
        else:
            arr = ci0
            vecs = [arr] if (getattr(arr, "ndim", 1) <= 2 and nroots == 1) else list(arr)
        guesses = []
        for v in vecs:
            t = torch.as_tensor(numpy.asarray(v) if not isinstance(v, torch.Tensor) else v)
            t = t.to(device="cuda", dtype=torch.float64).reshape(-1).contiguous()
            if t.numel() != n:
                raise ValueError("initial vector size %d does not match the basis size %d"
                                 % (t.numel(), n))
            guesses.append(t)
        if len(guesses) < nroots:
            guesses += solver._init_guess_dev(hdiag, nroots)[len(guesses):]

    from .direct_spin1 import available_bytes
    free = available_bytes()
    need = _davidson_bytes(n, nroots, max_space)
    if basis is not None:
        # determinant-space buffers of each plan's basis operator
        need += int(basis.work_bytes(precision)) if hasattr(basis, "work_bytes") else \
            (2 if precision == "fp64" else 4) * sp.ndet * 8
    ws = int(solver.workspace_bytes) or max(int((free - need) * 0.85), 1 << 26)
    stats = {"precision": precision}
    verbose_level = verbose
    verbose = bool(verbose and verbose > 4)

    if precision == "fp64":
        t0 = time.perf_counter()
        with _plan(sp, basis, h2e, "fp64", ws) as plan:
            hop = plan if basis is None else basis.hop(plan)
            res = _dv.davidson(hop, precond, guesses, nroots=nroots, lindep=lindep,
                               max_space=max_space, verbose=verbose, tol=tol,
                               tol_residual=tol_residual, max_cycle=max_cycle,
                               follow_state=follow_state)
            torch.cuda.synchronize()
            stats.update(time=time.perf_counter() - t0, nhop=res.nhop, cycles=res.cycles,
                         ntiles=getattr(plan, "ntiles", None), nrestore=res.nrestore)
        return res.e, list(res.x), res.converged, stats

    if precision not in LOW_STAGE:
        raise ValueError("unknown precision %r" % (precision,))
    if basis is not None and hasattr(basis, "make_plan"):
        # a basis whose low-precision plan computes in FP64 anyway (the
        # sector-native direct forms): the FP64 solve is the same product
        # without the acceptance checks
        probe = basis.make_plan(h2e, precision, max(ws // 3, 1 << 26))
        exact = probe is not None and bool(getattr(probe, "computes_fp64", False))
        if probe is not None:
            probe.close()
        if exact:
            e, x, conv, st = solve(solver, sp, h2e, hdiag, nroots, ci0, "fp64", tol=tol,
                                   tol_residual=tol_residual, max_cycle=max_cycle,
                                   max_space=max_space, lindep=lindep, level_shift=level_shift,
                                   verbose=verbose_level, basis=basis, precond=precond,
                                   follow_state=follow_state)
            st["precision_requested"] = precision
            return e, x, conv, st
    lo = LOW_STAGE[precision]
    t0 = time.perf_counter()
    # both plans live through the solve; the FP64 one (twice the walpha
    # bytes) gets two thirds of the workspace budget
    with _plan(sp, basis, h2e, "fp64", max(ws * 2 // 3, 1 << 26)) as plan64, \
            _plan(sp, basis, h2e, precision, max(ws // 3, 1 << 26)) as plan_lo:
        hop_lo = plan_lo if basis is None else basis.hop(plan_lo)
        hop64 = plan64 if basis is None else basis.hop(plan64)
        res = _dv.davidson(hop_lo, precond, guesses, nroots=nroots, lindep=lindep,
                           max_space=max_space, verbose=verbose, tol=tol,
                           tol_residual=tol_residual, max_cycle=max_cycle, hop_check=hop64,
                           tol_low=lo["tol"], tol_residual_low=lo["tol_residual"],
                           follow_state=follow_state)
        torch.cuda.synchronize()
        stats.update(time=time.perf_counter() - t0, nhop_low=res.nhop,
                     nhop_fp64=res.nhop_check, cycles=res.cycles, nrestore=res.nrestore,
                     ntiles=getattr(plan64, "ntiles", None),
                     ntiles_low=getattr(plan_lo, "ntiles", None))
    return res.e, list(res.x), res.converged, stats
