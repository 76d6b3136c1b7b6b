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
three precisions, the reference libfci_gas sigma on it, Davidson).

    python bench.py [--steps K] [--warmup W] [--no-davidson] [--no-cpu] [--no-batch]
                    [--no-csf] [--no-gas]
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
      fused path   reads c (8 B/det), writes walpha (s nlinka B/det) and
                   sigma (8 B/det); alpha pull reads walpha and sigma and
                   writes sigma: (s nlinka + 16) B/det          (HBM bound)
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
        fused_bytes = (16.0 + s * sp.nlinka) * dets
        pull_bytes = (s * sp.nlinka + 16.0) * dets
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


def cpu_baseline(pkg, budget_s=12.0):
    """The reference's own libfci_gas (oracle/_ref, built from the reference
    sources) on the host cores: full FP64 sigma of the (14o,14e) problem,
    repeated for ~budget_s, scaled to (16o,16e) sigma/s by the algorithmic
    work ratio F_sigma(14o)/F_sigma(16o)."""
    import numpy
    import oracle
    if not oracle.ref_available():
        return None
    ref = oracle.RefLib()
    n, ne = 14, (7, 7)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=0, with_ci=False)
    h2e = oracle.absorb_h1e(h1e, eri, n, ne, 0.5)
    nstr = math.comb(n, 7)
    c = numpy.random.default_rng(0).standard_normal(nstr * nstr)
    c /= numpy.linalg.norm(c)
    out = numpy.empty_like(c)
    plan = ref.plan(h2e, n, ne)
    plan(c, out)
    t0 = time.perf_counter()
    reps = 0
    while True:
        plan(c, out)
        reps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = (time.perf_counter() - t0) / reps
    plan.close()
    f1 = 2.0 * (n * (n + 1) // 2) ** 2 * nstr * nstr
    f2 = 2.0 * 136 ** 2 * math.comb(16, 8) ** 2
    cores = int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0))
    return {"value": (1.0 / dt) * f1 / f2, "unit": "sigma/s", "cores": cores,
            "kind": "reference",
            "sample": "libfci_gas (reference C, OpenMP, -O3 -march=x86-64-v3) FP64 sigma of "
                      "(14o,14e) [11.8M dets], %d runs of %.2f s, scaled to (16o,16e) by "
                      "F_sigma ratio %.4f" % (reps, dt, f1 / f2),
            "sigma_s_14o": dt}


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
        F += 4 * 2.0 * (n * (n + 1) // 2) ** 2 * nd
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
    log("batch: %d problems in %.2f s (%.2f problems/s), %d converged, sigma(all slots) %.2f ms = "
        "%.1f TFLOP/s" % (len(probs), dt, len(probs) / dt, nconv, st["tuned_ms"], tf or 0.0))
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
    of the space), the reference's libfci_gas sigma on the same space on the
    host cores, and the FP64 and bf16x3-mixed Davidson ground states."""
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

    n = args.norb
    ne = (n // 2, n // 2)
    h1e, eri = pkg.synthetic.make_problem(n, ne, seed=0, with_ci=False)
    h2e = pkg.absorb_h1e(h1e, eri, n, ne, 0.5)
    sp = pkg.get_space(n, ne)
    log("space (%do,%de): %d dets" % (n, sum(ne), sp.ndet))
    g = torch.Generator(device="cuda").manual_seed(0)
    c = torch.randn(sp.ndet, dtype=torch.float64, device="cuda", generator=g)
    c /= torch.linalg.vector_norm(c)
    out = torch.empty_like(c)
    npair = n * (n + 1) // 2
    F = 2.0 * npair * npair * sp.ndet

    result = {}
    kern = {}
    sigma64 = None
    precs = [p for p in args.precisions.split(",") if p]
    for prec in precs:
        free = torch.cuda.mem_get_info()[0]
        plan = pkg.SigmaPlan(sp, h2e, prec, int(free * 0.85))
        plan._bench_in, plan._bench_out = c, out
        for _ in range(args.warmup):
            plan(c, out)
        sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan(c, out)
        sync()
        dt = (time.perf_counter() - t0) / args.steps
        kern[prec] = kernel_roofline(plan, sp, prec, 2, sync)
        kern[prec]["ntiles"] = plan.ntiles
        if prec == "fp64":
            sigma64 = out.clone()
            err = 0.0
        else:
            err = float(torch.linalg.vector_norm(out - sigma64) / torch.linalg.vector_norm(sigma64)) \
                if sigma64 is not None else None
        result[prec] = {"ms": dt * 1e3, "sigma_per_s": 1.0 / dt,
                        "TFLOP/s": F / dt / 1e12, "rel_err_vs_fp64": err}
        log("%s sigma %.2f ms (%.2f TFLOP/s eff.), contraction %.2f TFLOP/s, err %s" % (
            prec, dt * 1e3, F / dt / 1e12, kern[prec]["contract"]["TFLOP/s"], err))
        plan.close()
        del plan
    del out, c, sigma64
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

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (reference libfci_gas on host cores)")
        cpu = cpu_baseline(pkg)

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
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
