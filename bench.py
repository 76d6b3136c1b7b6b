#!/usr/bin/env python
"""sigma = H c throughput and Davidson time at (16o,16e) on one MI355X.

Workload: CAS (16o, 16e), nelec = (8, 8): 12870 x 12870 = 165,636,900
determinants, npair = 136.  Integrals from the synthetic recipe
(pyscf-forge_amd/synthetic.py, seed 0); the sigma input vector is a seeded
normal vector generated on the device (the recipe's host draw of 165.6M
normals is slow and irrelevant to the timing).

One "step" is one full FP64 sigma = H c with c and sigma resident in HBM.
`value` = FP64 sigma/s.  The same line reports the FP32 and bf16x3
sigma rates with their error against the FP64 sigma, the FP64 and
mixed-precision Davidson times to the ground state with their energy
difference, the roofline of the dominant kernel (the contraction, timed with
HIP events on its own stream inside this run), a CPU baseline, the
batched workload: 64 uneven CAS problems (norb 6..12, 4 roots each) solved
together with one batched sigma launch per Davidson iteration, the
singlet ground state in the CSF basis (csf2det, sigma, det2csf per step),
and a restricted RAS(4,8,4) space inside the same CAS (restricted sigma in
three precisions, the reference libfci_gas sigma on it, Davidson), and a
sector-native RAS(4,12,4) space of 20 orbitals (no CAS array: sigma in FP64
and FP32 beside the reference libfci_gas on the same space).

    python bench.py [--steps K] [--warmup W] [--no-davidson] [--no-cpu] [--no-batch]
                    [--no-csf] [--no-gas] [--no-gas-native] [--only-gas-native]
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}   # dense MFMA TFLOP/s
HBM_PEAK = 8000.0                                          # GB/s


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def kernel_roofline(plan, sp, prec, calls, sync):
    """Per-kernel averages from the plan's HIP-event timers, with the
    algorithmic work per launch:
      contraction  F = 2 npair^2 dets                          (MFMA bound)
      fused path   reads c (8 B/det), writes walpha (s (nlinka - nalpha)
                   B/det: the off-diagonal alpha links) and sigma (8 B/det);
                   alpha pull reads walpha and sigma and writes sigma:
                   (s (nlinka - nalpha) + 16) B/det              (HBM bound)
      3-kernel     gather  8 + s npair B/det, scatter s npair + 8 B/det
    with s = 8 (FP64) or 4 (FP32 / bf16x3 intermediates)."""
    plan.set_timing(True)
    out = plan._bench_out
    for _ in range(calls):
        plan(plan._bench_in, out)
    sync()
    ms, nl = plan.timing()
    plan.set_timing(False)
    npair = sp.norb * (sp.norb + 1) // 2
    s = 8 if prec == "fp64" else 4
    per_sigma_launch = nl / calls
    avg = ms / nl                      # ms per launch
    dets = sp.ndet / per_sigma_launch[1]
    flop = 2.0 * npair * npair * dets
    out = {"launches_per_sigma": per_sigma_launch.tolist(), "fused": bool(plan.fused)}
    if plan.fused:
        nws = sp.nlinka - sp.nelec[0]
        fused_bytes = (16.0 + s * nws) * dets
        pull_bytes = (s * nws + 16.0) * dets
        out["k_fused_row"] = {"ms": avg[1], "TFLOP/s": flop / avg[1] / 1e9,
                              "frac_mfma": flop / avg[1] / 1e9 / PEAK[prec],
                              "flop_per_launch": flop, "dets_per_launch": dets,
                              "alg_bytes_per_launch": fused_bytes,
                              "GB/s": fused_bytes / avg[1] / 1e6}
        out["k_alpha_pull"] = {"ms": avg[2], "GB/s": pull_bytes / avg[2] / 1e6,
                               "frac_hbm": pull_bytes / avg[2] / 1e6 / HBM_PEAK,
                               "alg_bytes_per_launch": pull_bytes}
        out["contract"] = out["k_fused_row"]
    else:
        gather_bytes = (8.0 + s * npair) * dets
        scatter_bytes = (s * npair + 8.0) * dets
        out["k_gather"] = {"ms": avg[0], "GB/s": gather_bytes / avg[0] / 1e6,
                           "frac_hbm": gather_bytes / avg[0] / 1e6 / HBM_PEAK,
                           "alg_bytes_per_launch": gather_bytes}
        out["k_contract"] = {"ms": avg[1], "TFLOP/s": flop / avg[1] / 1e9,
                             "frac_mfma": flop / avg[1] / 1e9 / PEAK[prec],
                             "flop_per_launch": flop, "dets_per_launch": dets}
        out["k_scatter"] = {"ms": avg[2], "GB/s": scatter_bytes / avg[2] / 1e6,
                            "frac_hbm": scatter_bytes / avg[2] / 1e6 / HBM_PEAK,
                            "alg_bytes_per_launch": scatter_bytes}
        out["contract"] = out["k_contract"]
    return out


class HostOps:
    """The Davidson driver's vector interface on host arrays (the CPU
    reference leg runs the same algorithm over the reference's sigma)."""

    def empty(self, shape, like):
        import numpy
        return numpy.empty(shape)

    def clone(self, v):
        import numpy
        return numpy.array(v, dtype=numpy.float64).reshape(-1)

    def copy(self, dst, src):
        dst[...] = src

    def dots(self, m, n, X, v):
        import numpy
        return numpy.atleast_2d(X)[:m].reshape(m, n) @ v.reshape(-1)

    def lincomb(self, m, n, X, coef, beta, y):
        import numpy
        Xm = numpy.atleast_2d(X)[:m].reshape(m, n)
        y[...] = (beta * y.reshape(-1) if beta != 0.0 else 0.0) + numpy.asarray(coef).reshape(m) @ Xm

    def combine(self, n, vecs, coef, y):
        y[...] = sum(c * v.reshape(-1) for c, v in zip(coef, vecs))

    def dots_list(self, n, vecs, v):
        import numpy
        return numpy.array([float(u.reshape(-1) @ v.reshape(-1)) for u in vecs])

    def combine_many(self, n, tasks):
        import numpy
        for vecs, outs, coef in tasks:
            for j, o in enumerate(outs):
                o[...] = sum(c * v.reshape(-1) for c, v in zip(numpy.asarray(coef)[:, j], vecs))


def rel(a, b):
    import numpy
    a = numpy.asarray(a).reshape(-1)
    b = numpy.asarray(b).reshape(-1)
    return float(numpy.linalg.norm(a - b) / numpy.linalg.norm(b))


def cpu_baseline(pkg, gpu, sync):
    """The reference's own libfci_gas (oracle/_ref, built from the reference
    sources) on the host cores, beside the GPU results of the same inputs:
      * headline: one full FP64 sigma of (16o,16e) on the GPU run's c
        (timed; its difference from the GPU FP64 sigma);
      * C0 (8o,8e): sigma difference, and the ground-state energy of ARPACK
        over the reference sigma against the GPU Davidson;
      * C1 (14o,14e): sigma time and difference, and 12 Davidson iterations
        (the same algorithm, host vectors) over the reference sigma;
      * C4 (12o,12e): the reference RDMs of the GPU ground state against the
        GPU RDM kernel."""
    import numpy
    import oracle
    if not oracle.ref_available():
        return None
    ref = oracle.RefLib()
    cores = int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0))
    out = {"unit": "sigma/s", "cores": cores, "kind": "reference", "configs": {}}

    hd = gpu.get("headline")
    if hd is not None:
        n, ne = hd["norb"], hd["nelec"]
        c = hd["c"]
        y = numpy.empty_like(c)
        plan = ref.plan(hd["h2e"], n, ne)
        t0 = time.perf_counter()
        plan(c, y)
        dt = time.perf_counter() - t0
        runs = [dt]
        if dt < 10.0:
            t0 = time.perf_counter()
            plan(c, y)
            runs.append(time.perf_counter() - t0)
        plan.close()
        dt = min(runs)
        out["value"] = 1.0 / dt
        out["sample"] = ("libfci_gas (reference C, OpenMP on %d threads, -O3 -march=x86-64-v3) FP64 "
                         "sigma of CAS(16o,16e) [165.6M dets] on the GPU run's c: %d run(s), "
                         "best %.2f s" % (cores, len(runs), dt))
        out["sigma_s_16o"] = dt
        out["gpu_fp64_sigma_rel_diff"] = rel(hd["sigma64"], y)
        log("CPU reference sigma (16o,16e): %.2f s on %d threads, GPU FP64 rel diff %.2e" % (
            dt, cores, out["gpu_fp64_sigma_rel_diff"]))
        del y

    c0 = gpu.get("c0")
    if c0 is not None:
        import scipy.sparse.linalg as sla
        n, ne = c0["norb"], c0["nelec"]
        y = ref.contract_2e(c0["h2e"], c0["c"], n, ne)
        nd = c0["c"].size
        op = sla.LinearOperator((nd, nd), dtype=numpy.float64,
                                matvec=lambda x: ref.contract_2e(c0["h2e"], x, n, ne).reshape(-1))
        w = sla.eigsh(op, k=1, which="SA", tol=1e-13)[0]
        out["configs"]["c0"] = {"sigma_rel_diff_gpu_fp64": rel(c0["sigma64"], y),
                                "energy_ref_arpack": float(w[0]),
                                "dE_gpu_fp64_vs_ref": abs(c0["e64"] - float(w[0]))}
        log("C0 vs reference: %s" % out["configs"]["c0"])

    c1 = gpu.get("c1")
    if c1 is not None:
        n, ne = c1["norb"], c1["nelec"]
        c = c1["c"]
        y = numpy.empty_like(c)
        plan = ref.plan(c1["h2e"], n, ne)
        plan(c, y)
        t0 = time.perf_counter()
        plan(c, y)
        dt = time.perf_counter() - t0
        rec = {"sigma_s": dt, "sigma_rel_diff_gpu_fp64": rel(c1["sigma64"], y)}
        hdiag = ref.make_hdiag(c1["h1e"], c1["eri"], n, ne).reshape(-1)
        # the reference solver's guess (fci_gas.py:46-88) and diagonal
        # preconditioner (:1337-1342), as the GPU run uses them
        _, guess = oracle.init_guess(hdiag, 1)

        def hop(x, o):
            plan(numpy.ascontiguousarray(x), o)

        def precond(r, e, o, x0):
            o[...] = oracle.precond(r, hdiag, e, 1e-3)

        t0 = time.perf_counter()
        res = pkg.davidson.davidson(hop, precond, list(guess), nroots=1, tol=1e-10, max_cycle=12,
                                    max_space=12, ops=HostOps())
        rec["davidson12_s"] = time.perf_counter() - t0
        rec["davidson12_energy"] = float(res.e[0])
        rec["davidson12_dE_vs_gpu"] = abs(float(res.e[0]) - c1["e12"])
        plan.close()
        out["configs"]["c1"] = rec
        log("C1 vs reference: %s" % rec)

    c4 = gpu.get("c4")
    if c4 is not None:
        n, ne = c4["norb"], c4["nelec"]
        (a, b), (aa, ab, bb) = ref.trans_rdm12s(c4["ci"], c4["ci"], n, ne)
        dm1 = a + b
        dm2 = aa + bb + ab + ab.transpose(2, 3, 0, 1)
        out["configs"]["c4"] = {"rdm1_max_abs_diff": float(numpy.abs(dm1 - c4["dm1"]).max()),
                                "rdm2_max_abs_diff": float(numpy.abs(dm2 - c4["dm2"]).max())}
        log("C4 vs reference: %s" % out["configs"]["c4"])
    return out


def sigma_precisions(pkg, sp, h2e, c, steps, warmup, sync, with_roofline=False,
                     precisions=("fp64", "fp32", "bf16x3")):
    """FP64 / FP32 / bf16x3 sigma of c: time per sigma and relative error
    against the FP64 sigma (FP64 first when it is among the precisions);
    returns (records, FP64 sigma or None, kernel rooflines)."""
    import torch
    out = torch.empty_like(c)
    npair = sp.norb * (sp.norb + 1) // 2
    F = 2.0 * npair * npair * sp.ndet
    recs, kern, s64 = {}, {}, None
    order = [p for p in ("fp64", "fp32", "bf16x3") if p in precisions]
    for prec in order:
        free = torch.cuda.mem_get_info()[0]
        plan = pkg.SigmaPlan(sp, h2e, prec, int(free * 0.85))
        for _ in range(warmup):
            plan(c, out)
        sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            plan(c, out)
        sync()
        dt = (time.perf_counter() - t0) / steps
        if with_roofline:
            plan._bench_in, plan._bench_out = c, out
            kern[prec] = kernel_roofline(plan, sp, prec, 2, sync)
            kern[prec]["ntiles"] = plan.ntiles
        if prec == "fp64":
            s64 = out.clone()
        recs[prec] = {"ms": dt * 1e3, "sigma_per_s": 1.0 / dt, "TFLOP/s": F / dt / 1e12,
                      "rel_err_vs_fp64": 0.0 if prec == "fp64" else None if s64 is None else
                      float(torch.linalg.vector_norm(out - s64) / torch.linalg.vector_norm(s64))}
        plan.close()
        del plan
    del out
    return recs, s64, kern


def davidson_precisions(pkg, sync, h1e, eri, n, ne, max_cycle, **opts):
    """FP64 and FP32 / bf16x3-mixed Davidson ground states: time, energy,
    difference from the FP64 energy."""
    import torch
    out, e64 = {}, None
    for prec in ("fp64", "fp32", "bf16x3"):
        s = pkg.FCISolver()
        s.precision = prec
        s.max_cycle = max_cycle
        for k, v in opts.items():
            setattr(s, k, v)
        sync()
        t0 = time.perf_counter()
        e, ci = s.kernel(h1e, eri, n, ne)
        sync()
        dt = time.perf_counter() - t0
        e64 = e if prec == "fp64" else e64
        out[prec] = {"seconds": dt, "energy": e, "converged": bool(s.converged),
                     "dE_vs_fp64": abs(e - e64),
                     "stats": {k: v for k, v in getattr(s, "stats", {}).items()
                               if isinstance(v, (int, float, str))}}
        del ci, s
        torch.cuda.empty_cache()
    return out


def config0(pkg, sync):
    """C0: CAS(8o,8e), (4,4), 4900 dets, c from the synthetic recipe (seed 0)."""
    import torch
    n, ne = 8, (4, 4)
    h1e, eri, c = pkg.synthetic.make_problem(n, ne, seed=0)
    h2e = pkg.absorb_h1e(h1e, eri, n, ne, 0.5)
    sp = pkg.get_space(n, ne)
    cd = torch.from_numpy(c.reshape(-1).copy()).cuda()
    recs, s64, _ = sigma_precisions(pkg, sp, h2e, cd, 5, 1, sync)
    dav = davidson_precisions(pkg, sync, h1e, eri, n, ne, 100, davidson_only=True)
    out = {"workload": "CAS(8o,8e) nelec (4,4), 4900 dets, synthetic integrals and c, seed 0",
           "sigma": recs, "davidson": dav}
    log("C0: sigma %s, Davidson dE %s" % (
        {p: (round(r["ms"], 3), r["rel_err_vs_fp64"]) for p, r in recs.items()},
        {p: r["dE_vs_fp64"] for p, r in dav.items()}))
    host = {"norb": n, "nelec": ne, "h2e": h2e, "c": c.reshape(-1).copy(),
            "sigma64": s64.cpu().numpy(), "e64": dav["fp64"]["energy"]}
    return out, host


def config1(pkg, sync):
    """C1: CAS(14o,14e), (7,7), 11.8M dets: FP64 / FP32 / bf16x3 sigma, 12
    FP64 Davidson iterations (nroots 1), and the converged FP64 and mixed
    ground states."""
    import torch
    n, ne = 14, (7, 7)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=0, with_ci=False)
    h2e = pkg.absorb_h1e(h1e, eri, n, ne, 0.5)
    sp = pkg.get_space(n, ne)
    g = torch.Generator(device="cuda").manual_seed(0)
    c = torch.randn(sp.ndet, dtype=torch.float64, device="cuda", generator=g)
    c /= torch.linalg.vector_norm(c)
    recs, s64, _ = sigma_precisions(pkg, sp, h2e, c, 5, 1, sync)
    # 12 iterations from the lowest-diagonal determinant, diagonal
    # preconditioner: the same algorithm the CPU reference leg runs
    s = pkg.FCISolver()
    sync()
    t0 = time.perf_counter()
    e12, _ = s.kernel(h1e, eri, n, ne, max_cycle=12, pspace_size=0)
    sync()
    d12 = time.perf_counter() - t0
    dav = davidson_precisions(pkg, sync, h1e, eri, n, ne, 100)
    out = {"workload": "CAS(14o,14e) nelec (7,7), 11,778,624 dets, synthetic integrals seed 0",
           "sigma": recs, "davidson12_fp64_s": d12, "davidson12_energy": float(e12),
           "davidson": dav}
    log("C1: sigma %s, 12 iterations %.3f s, Davidson %s" % (
        {p: (round(r["ms"], 2), r["rel_err_vs_fp64"]) for p, r in recs.items()}, d12,
        {p: (round(r["seconds"], 2), r["dE_vs_fp64"]) for p, r in dav.items()}))
    host = {"norb": n, "nelec": ne, "h1e": h1e, "eri": eri, "h2e": h2e, "c": c.cpu().numpy(),
            "sigma64": s64.cpu().numpy(), "e12": float(e12)}
    del c, s64
    torch.cuda.empty_cache()
    return out, host


def config4(pkg, sync):
    """C4: CAS(12o,12e), (6,6), 853,776 dets: E(h1e, eri) = c^T H c through
    the autograd Function at the FP64 ground state, backward = RDM kernel;
    gradient against a central finite difference along one random
    symmetric perturbation (step 1e-5), FP32 integrals too."""
    import numpy
    import torch
    n, ne = 12, (6, 6)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=0, with_ci=False)
    s = pkg.FCISolver()
    e, ci = s.kernel(h1e, eri, n, ne)
    c = torch.from_numpy(numpy.ascontiguousarray(ci).reshape(-1)).cuda()
    out = {"workload": "CAS(12o,12e) nelec (6,6), 853,776 dets, synthetic integrals seed 0",
           "energy": float(e)}
    reps = 3
    for dt_name, dtype in (("fp64", torch.float64), ("fp32", torch.float32)):
        # the first call creates the sigma and RDM plans (timed apart); then
        # the mean of `reps` warm forward / backward passes
        times = []
        for it in range(reps + 1):
            h = torch.tensor(h1e, dtype=dtype, requires_grad=True)
            g = torch.tensor(eri, dtype=dtype, requires_grad=True)
            sync()
            t0 = time.perf_counter()
            E = pkg.autograd.ci_energy(h, g, c, n, ne)
            sync()
            t1 = time.perf_counter()
            E.backward()
            sync()
            t2 = time.perf_counter()
            times.append((t1 - t0, t2 - t1))
        first, warm = times[0], numpy.mean(times[1:], axis=0)
        rng = numpy.random.default_rng(5)
        dh = rng.normal(size=h1e.shape)
        dh = 0.5 * (dh + dh.T)
        step = 1e-5
        ep = float(pkg.autograd.ci_energy(torch.tensor(h1e + step * dh), torch.tensor(eri), c, n, ne))
        em = float(pkg.autograd.ci_energy(torch.tensor(h1e - step * dh), torch.tensor(eri), c, n, ne))
        fd = (ep - em) / (2 * step)
        an = float((h.grad.double() * torch.tensor(dh)).sum())
        out[dt_name] = {"forward_s": float(warm[0]), "backward_rdm_s": float(warm[1]),
                        "first_call_forward_s": first[0], "first_call_backward_s": first[1],
                        "E_minus_solver": abs(float(E.detach()) - float(e)),
                        "dE_dh_fd_rel_diff": abs(an - fd) / max(abs(fd), 1e-30)}
    dm1, dm2 = pkg.make_rdm12(ci, n, ne)
    log("C4: %s" % out)
    host = {"norb": n, "nelec": ne, "ci": numpy.ascontiguousarray(ci).reshape(-1),
            "dm1": numpy.asarray(dm1), "dm2": numpy.asarray(dm2)}
    return out, host


def batch_solve(pkg, sync):
    """64 uneven CAS problems (norb 6..12, nroots = 4, integrals from seed i)
    solved together: one batched sigma launch per Davidson iteration.
    problems/s counts the whole solve (string spaces, diagonals, guesses,
    schedule tuning and iterations)."""
    import numpy
    man = pkg.synthetic.batch_manifest()
    probs = []
    F = 0.0
    for n, ne, s in man:
        h1e, eri = pkg.synthetic.make_problem(n, ne, seed=s, with_ci=False)
        probs.append((h1e, eri, n, ne))
        nd = math.comb(n, ne[0]) * math.comb(n, ne[1])
        # four roots and the solver's guard root per problem
        F += 5 * 2.0 * (n * (n + 1) // 2) ** 2 * nd
    sync()
    t0 = time.perf_counter()
    res = pkg.batch.solve_batch(probs, nroots=4, max_cycle=300)
    sync()
    dt = time.perf_counter() - t0
    st = res.stats
    nconv = int(sum(bool(c.all()) for c in res.converged))
    tf = F / (st["tuned_ms"] * 1e-3) / 1e12 if st.get("tuned_ms") else None
    out = {
        "workload": "64 CAS problems, norb 6..12, nelec (n//2+d, n//2), d in {0,1}, "
                    "nroots 4, %d dets in total" % st["ndet_total"],
        "problems_per_s": len(probs) / dt, "seconds": dt,
        "seconds_setup": st["time_setup"], "seconds_batch_create": st["time_create"],
        "seconds_tune": st["time_tune"], "seconds_iterations": st["time_iterations"],
        "seconds_sigma": st["time_sigma"], "sigma_launches": st["sigma_launches"],
        "converged": nconv, "max_cycles": int(numpy.max(res.cycles)),
        "schedule": {"rows_per_item": st["rows_per_item"], "grid": st["grid"]},
        "sigma_all_slots_ms": st["tuned_ms"],
        "sigma_all_slots_TFLOP/s": tf,
        "sigma_frac_of_fp64_peak": None if tf is None else tf / PEAK["fp64"],
    }
    fx = os.path.join(ROOT, "tests", "golden", "batch", "manifest_eigs.npz")
    if os.path.exists(fx):
        # lowest eigenvalues of every problem (dense eigh / Lanczos, committed data)
        with numpy.load(fx) as f:
            exact = f["e"]
        if numpy.isfinite(exact).all():
            d = numpy.array([numpy.abs(numpy.asarray(e) - exact[p, :len(e)]).max()
                             for p, e in enumerate(res.e)])
            out["dE_vs_exact_max"] = float(d.max())
            out["problems_off_exact_1e-6"] = int((d > 1e-6).sum())
    log("batch: %d problems in %.2f s (%.2f problems/s), %d converged, sigma(all slots) %.2f ms = "
        "%.1f TFLOP/s, max |dE| vs exact %s" % (len(probs), dt, len(probs) / dt, nconv,
                                                st["tuned_ms"], tf or 0.0, out.get("dE_vs_exact_max")))
    return out


def csf_solve(pkg, sync, h1e, eri, n, ne, max_cycle):
    """Singlet ground state in the CSF basis (the reference's csf_fci
    solver): Davidson on CSF vectors, each H application = csf2det, sigma,
    det2csf on the device; FP64 and bf16x3-mixed."""
    import torch
    out = {}
    t0 = time.perf_counter()
    t = pkg.csf_fci.CSFTransformer(n, ne[0], ne[1], 1)
    g = torch.Generator(device="cuda").manual_seed(1)
    d = torch.randn(t.ndet, dtype=torch.float64, device="cuda", generator=g)
    x = torch.empty(t.ncsf, dtype=torch.float64, device="cuda")
    sync()
    out["transformer_s"] = time.perf_counter() - t0
    out["ncsf"], out["ndet"] = t.ncsf, t.ndet
    for name, fn in [("det2csf_ms", lambda: t.det2csf_dev(d, x)), ("csf2det_ms", lambda: t.csf2det_dev(x, d))]:
        fn()
        sync()
        t1 = time.perf_counter()
        for _ in range(3):
            fn()
        sync()
        out[name] = (time.perf_counter() - t1) / 3 * 1e3
    del d, x
    e64 = None
    for prec in ["fp64", "bf16x3"]:
        s = pkg.csf_fci.FCISolver(smult=1)
        s.precision = prec
        s.max_cycle = max_cycle
        sync()
        t1 = time.perf_counter()
        e, ci = s.kernel(h1e, eri, n, ne)
        sync()
        dt = time.perf_counter() - t1
        e64 = e if prec == "fp64" else e64
        out[prec] = {"seconds": dt, "energy": e, "converged": bool(s.converged),
                     "dE_vs_fp64": abs(e - e64),
                     "stats": {k: v for k, v in s.stats.items() if isinstance(v, (int, float, str))}}
        log("CSF Davidson %s: %.2f s, E = %.12f, %s" % (prec, dt, e, out[prec]["stats"]))
        del ci, s
        torch.cuda.empty_cache()
    return out


def gas_solve(pkg, sync, h1e, eri, steps, with_cpu, max_cycle, davidson=True):
    """Restricted (GAS) CI inside the (16o,16e) CAS: RAS(4,8,4) with (8,8)
    electrons, at most 2 holes in RAS1 and 2 electrons in RAS3 (4,870,404
    determinants, 2.9% of the CAS).  FP64 / FP32 / bf16x3 restricted sigma
    (fused kernel over the alpha rows and beta chunks within one excitation
    of the space), the same space sector-native (operator form, FP64), the
    reference's libfci_gas sigma on the same space on the host cores, and
    the FP64 and bf16x3-mixed Davidson ground states."""
    import numpy
    import torch
    fg = pkg.fci_gas
    orbs, ne, restr = (4, 8, 4), (8, 8), {"max_holes": 2, "max_particles": 2}
    sol = fg.FCISolver(gas_orbs=orbs, gas_restr=restr, gas_restr_type="ras")
    t0 = time.perf_counter()
    gas = sol._space(16, ne)
    sync()
    out = {"workload": "RAS(4,8,4) (8,8) max_holes 2 max_particles 2 in CAS(16o,16e), "
                       "synthetic integrals seed 0",
           "ndet": gas.ndet, "space_s": time.perf_counter() - t0}
    nrows, nch, ncas = gas.rows_info()
    out["rows"], out["chunks"], out["chunk_frac_of_cas"] = nrows, nch, nch / ncas
    h2e = fg.absorb_h1e(h1e, eri, 16, ne)
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn(gas.ndet, dtype=torch.float64, device="cuda", generator=g)
    ref = None
    for prec in ("fp64", "fp32", "bf16x3"):
        with fg.GasContractPlan(gas, h2e, prec) as plan:
            y = torch.empty_like(x)
            plan.execute(x, y)
            sync()
            t1 = time.perf_counter()
            for _ in range(steps):
                plan.execute(x, y)
            sync()
            dt = (time.perf_counter() - t1) / steps
        rec = {"ms": dt * 1e3, "sigma_per_s": 1.0 / dt}
        if prec == "fp64":
            ref = y.clone()
        else:
            rec["rel_err_vs_fp64"] = float((y - ref).norm() / ref.norm())
        out[prec] = rec
        log("GAS sigma %s: %.2f ms" % (prec, dt * 1e3))
    # the same space sector-native (the operator form, FP64), beside the
    # CAS-embedded kernels above
    with fg.GasSpace(orbs, ne, gas.blocks, native=True) as gn:
        with fg.GasContractPlan(gn, h2e, "fp64") as plan:
            y = torch.empty_like(x)
            plan.execute(x, y)
            sync()
            t1 = time.perf_counter()
            for _ in range(steps):
                plan.execute(x, y)
            sync()
            dt = (time.perf_counter() - t1) / steps
            path = plan.path
            out["fp64_sector_native"] = {"ms": dt * 1e3, "sigma_per_s": 1.0 / dt, "path": path,
                                         "max_rel_diff_vs_cas_embedded":
                                             float((y - ref).abs().max() / ref.abs().max())}
        log("GAS sigma fp64 sector-native (%s): %.2f ms" % (path, dt * 1e3))
    if with_cpu:
        import oracle
        if oracle.ref_available():
            D = gas.blocks
            R = oracle.RefLib().gas(orbs, ne, D)
            P = R.plan(h2e)
            xc = x.cpu().numpy()
            yc = numpy.empty_like(xc)
            P(xc, yc)
            t1 = time.perf_counter()
            reps = 0
            while reps < 2 or time.perf_counter() - t1 < 8.0:
                P(xc, yc)
                reps += 1
            dt = (time.perf_counter() - t1) / reps
            P.close()
            cores = int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0))
            out["cpu_reference"] = {"sigma_s": dt, "sigma_per_s": 1.0 / dt, "cores": cores,
                                    "runs": reps, "kind": "reference libfci_gas",
                                    "max_abs_diff_vs_gpu_fp64":
                                        float(numpy.abs(yc - ref.cpu().numpy()).max())}
            log("GAS sigma libfci_gas on %d cores: %.3f s" % (cores, dt))
    del x, ref
    e64 = None
    for prec in ("fp64", "bf16x3") if davidson else ():
        s = fg.FCISolver(gas_orbs=orbs, gas_restr=restr, gas_restr_type="ras")
        s.precision = prec
        s.max_cycle = max_cycle
        sync()
        t1 = time.perf_counter()
        e, _ = s.kernel(h1e, eri, 16, ne)
        sync()
        dt = time.perf_counter() - t1
        e64 = e if prec == "fp64" else e64
        out["davidson_" + prec] = {"seconds": dt, "energy": e, "converged": bool(s.converged),
                                   "dE_vs_fp64": abs(e - e64),
                                   "stats": {k: v for k, v in s.stats.items()
                                             if isinstance(v, (int, float, str))}}
        log("GAS Davidson %s: %.2f s, E = %.12f" % (prec, dt, e))
        torch.cuda.empty_cache()
    return out


GAS_NATIVE_WORKLOADS = [
    ("ras20", (4, 12, 4), (6, 6), {"max_holes": 2, "max_particles": 2}, "ras",
     "RAS(4,12,4) (6,6) max_holes 2 max_particles 2, 20 orbitals"),
    ("ras40", (2, 8, 30), (4, 4), {"max_holes": 1, "max_particles": 2}, "ras",
     "RAS(2,8,30) (4,4) max_holes 1 max_particles 2, 40 orbitals"),
    ("gas62", (2, 6, 24, 30), (4, 4), [[3, 4], [6, 8], [7, 8], [8, 8]], "cumulative-occ",
     "GAS(2,6,24,30) (4,4) cumulative occupations [3,4] [6,8] [7,8] [8,8], 62 orbitals"),
]


def gas_native_sigma(pkg, sync, steps, with_cpu):
    """Sector-native restricted sigma on three spaces whose CAS is never
    formed: RAS(4,12,4) (6,6) in 20 orbitals (6,989,020 determinants),
    RAS(2,8,30) (4,4) in 40 orbitals (1,543,896) and a 4-subspace GAS in 62
    orbitals (1,352,121), each in the form the plan picks (the operator form
    on all three), plus the GEMM form (t1 over the single-excitation
    closure) on the 20-orbital space.  FP64 and FP32 plans, and the
    reference's libfci_gas sigma on the same space on the host cores."""
    import numpy
    import torch
    fg, ag = pkg.fci_gas, pkg.addons_gas
    res = {}
    for key, orbs, ne, restr, kind, label in GAS_NATIVE_WORKLOADS:
        n = sum(orbs)
        o, D = ag.normalize_gas_spec(orbs, ne, restr, kind)
        h1e, eri = pkg.synthetic.make_problem(n, ne, seed=0, with_ci=False)
        h2e = fg.absorb_h1e(h1e, eri, n, ne, 0.5)
        t0 = time.perf_counter()
        gas = fg.GasSpace(o, ne, D, native=True)
        sync()
        n1, maxrow, nsa, nsb = gas.native_info()
        out = {"workload": label + ", sector-native, synthetic integrals seed 0",
               "ndet": gas.ndet, "closure": n1, "space_s": time.perf_counter() - t0}
        g = torch.Generator(device="cuda").manual_seed(5)
        x = torch.randn(gas.ndet, dtype=torch.float64, device="cuda", generator=g)
        ref = None
        npair = n * (n + 1) // 2
        runs = [("fp64", None), ("fp32", None)]
        if key == "ras20":
            runs.append(("fp64", "t1"))
        for prec, force in runs:
            if force:
                os.environ["FCIS_GAS_NATIVE_PATH"] = force
            with fg.GasContractPlan(gas, h2e, prec) as plan:
                y = torch.empty_like(x)
                plan.execute(x, y)
                sync()
                t1 = time.perf_counter()
                for _ in range(steps):
                    plan.execute(x, y)
                sync()
                dt = (time.perf_counter() - t1) / steps
                path = plan.path
            os.environ.pop("FCIS_GAS_NATIVE_PATH", None)
            rec = {"ms": dt * 1e3, "sigma_per_s": 1.0 / dt, "path": path}
            if path == "native-t1":
                rec["gemm_TFLOP/s"] = 2.0 * npair * npair * n1 / dt / 1e12
            if ref is None:
                ref = y.clone()
            else:
                rec["rel_err_vs_fp64"] = float((y - ref).norm() / ref.norm())
            out[prec if force is None else prec + "_" + force] = rec
            log("native GAS %s sigma %s (%s): %.2f ms" % (key, prec, path, dt * 1e3))
        if with_cpu:
            import oracle
            if oracle.ref_available():
                R = oracle.RefLib().gas(o, ne, D)
                P = R.plan(h2e)
                xc = x.cpu().numpy()
                yc = numpy.empty_like(xc)
                P(xc, yc)
                t1 = time.perf_counter()
                reps = 0
                while reps < 2 or time.perf_counter() - t1 < 6.0:
                    P(xc, yc)
                    reps += 1
                dt = (time.perf_counter() - t1) / reps
                P.close()
                R.close()
                cores = int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0))
                out["cpu_reference"] = {"sigma_s": dt, "cores": cores, "runs": reps,
                                        "kind": "reference libfci_gas",
                                        "max_rel_diff_vs_gpu_fp64":
                                            float(numpy.abs(yc - ref.cpu().numpy()).max() /
                                                  numpy.abs(yc).max())}
                log("native GAS %s sigma libfci_gas on %d cores: %.3f s" % (key, cores, dt))
        del x, ref
        gas.close()
        torch.cuda.empty_cache()
        res[key] = out
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--norb", type=int, default=16)
    ap.add_argument("--no-davidson", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precisions", default="fp64,fp32,bf16x3")
    ap.add_argument("--max-cycle", type=int, default=100)
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 64-problem batched solve (nroots=4)")
    ap.add_argument("--no-csf", action="store_true",
                    help="skip the CSF-basis (singlet) Davidson at the same size")
    ap.add_argument("--no-gas", action="store_true",
                    help="skip the restricted (RAS) space inside the same CAS")
    ap.add_argument("--no-gas-native", action="store_true",
                    help="skip the sector-native RAS(4,12,4) space of 20 orbitals")
    ap.add_argument("--only-gas-native", action="store_true",
                    help="run only the sector-native GAS section (profiling)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the CAS(8o,8e), CAS(14o,14e) and CAS(12o,12e) autograd sections")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 4 GiB device copy first (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: there is no multi-GPU path "
              "(the sigma plan's alpha tiles are the unit a sharded version would "
              "split across GPUs); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import importlib
    import numpy
    import torch
    pkg = importlib.import_module("pyscf-forge_amd")
    torch.cuda.set_device(0)
    sync = torch.cuda.synchronize

    if args.copy_calibration:
        x = torch.ones(1 << 29, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(2):
            y.copy_(x)
        sync()
        log("copy calibration: 2 x 4 GiB device copies")
        del x, y
        torch.cuda.empty_cache()

    if args.only_gas_native:
        print(json.dumps({"gas_native": gas_native_sigma(pkg, sync, args.steps, not args.no_cpu)}),
              flush=True)
        return

    n = args.norb
    ne = (n // 2, n // 2)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=0, with_ci=False)
    h2e = pkg.absorb_h1e(h1e, eri, n, ne, 0.5)
    sp = pkg.get_space(n, ne)
    log("space (%do,%de): %d dets" % (n, sum(ne), sp.ndet))
    g = torch.Generator(device="cuda").manual_seed(0)
    c = torch.randn(sp.ndet, dtype=torch.float64, device="cuda", generator=g)
    c /= torch.linalg.vector_norm(c)
    precs = [p for p in args.precisions.split(",") if p]
    result, sigma64, kern = sigma_precisions(pkg, sp, h2e, c, args.steps, args.warmup, sync,
                                             with_roofline=True, precisions=precs)
    for prec, r in result.items():
        log("%s sigma %.2f ms (%.2f TFLOP/s eff.), contraction %.2f TFLOP/s, err %s" % (
            prec, r["ms"], r["TFLOP/s"], kern[prec]["contract"]["TFLOP/s"], r["rel_err_vs_fp64"]))
    gpu_host = {}
    if not args.no_cpu and sigma64 is not None:
        gpu_host["headline"] = {"norb": n, "nelec": ne, "h2e": h2e, "c": c.cpu().numpy(),
                                "sigma64": sigma64.cpu().numpy()}
    del c, sigma64
    torch.cuda.empty_cache()

    dav = None
    if not args.no_davidson:
        dav = {}
        e64 = None
        for prec in ["fp64"] + [p for p in precs if p != "fp64"]:
            solver = pkg.FCISolver()
            solver.precision = prec
            solver.max_cycle = args.max_cycle
            sync()
            t0 = time.perf_counter()
            e, ci = solver.kernel(h1e, eri, n, ne)
            sync()
            dt = time.perf_counter() - t0
            if prec == "fp64":
                e64 = e
            dav[prec] = {"seconds": dt, "energy": e, "converged": bool(solver.converged),
                         "dE_vs_fp64": None if e64 is None else abs(e - e64),
                         "stats": {k: v for k, v in solver.stats.items()
                                   if isinstance(v, (int, float, str))}}
            log("Davidson %s: %.2f s, E = %.12f, dE = %s, %s" % (
                prec, dt, e, dav[prec]["dE_vs_fp64"], dav[prec]["stats"]))
            del ci, solver
            torch.cuda.empty_cache()

    configs = {}
    if not args.no_configs:
        configs["c0"], gpu_host["c0"] = config0(pkg, sync)
        configs["c1"], gpu_host["c1"] = config1(pkg, sync)
        configs["c4"], gpu_host["c4"] = config4(pkg, sync)

    batch = None
    if not args.no_batch:
        batch = batch_solve(pkg, sync)

    csf = None
    if not args.no_csf and not args.no_davidson:
        csf = csf_solve(pkg, sync, h1e, eri, n, ne, args.max_cycle)

    gas = None
    if not args.no_gas and n == 16:
        gas = gas_solve(pkg, sync, h1e, eri, args.steps, not args.no_cpu,
                        max(args.max_cycle, 400), not args.no_davidson)

    gas_native = None
    if not args.no_gas_native:
        gas_native = gas_native_sigma(pkg, sync, args.steps, not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (reference libfci_gas on host cores)")
        cpu = cpu_baseline(pkg, gpu_host, sync)
    del gpu_host

    main_k = kern["fp64"]["contract"] if "fp64" in kern else None
    main_name = None
    if main_k is not None:
        main_name = "k_fused_row<fp64>" if kern["fp64"]["fused"] else "k_contract_f64"
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_sigma_f64_16o16e.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        if tj.get("kernel") == main_name:
            traffic = tj.get("hbm_bytes_per_launch")
    fp64 = result.get("fp64", {})
    line = {
        "metric": "sigma=H*c per second (FP64) at (16o,16e); Davidson time FP64 vs mixed",
        "value": fp64.get("sigma_per_s"),
        "unit": "sigma/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fp64.get("ms"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "CAS(16o,16e) sigma=H*c, nelec=(8,8), 165636900 dets, "
                               "npair=136, synthetic integrals seed 0",
                   "norb": n, "nelec": list(ne), "ndet": sp.ndet},
        "roofline": None if main_k is None else {
            "bound": "mfma", "kernel": main_name,
            "achieved": main_k["TFLOP/s"], "peak": PEAK["fp64"], "unit": "TFLOP/s",
            "frac": main_k["TFLOP/s"] / PEAK["fp64"], "traffic": traffic},
        "cpu_baseline": cpu,
        "precisions": result,
        "kernels": kern,
        "davidson": dav,
        "batch": batch,
        "csf": csf,
        "gas": gas,
        "gas_native": gas_native,
        "configs": configs,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
