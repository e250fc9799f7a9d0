#!/usr/bin/env python3
"""Benchmark: converged velocity solves per second at 1e-6 relative residual,
plain vs chosen form, on one GPU.

Workload: the marine channel (640 x 80 km, bed -100 -> -700 m, H 1500 -> 200 m
plus a seeded Gaussian random field, mesh refined 16 km -> 2 km at the grounding
line, ~200k vertices / ~394k triangles, DIVA with 10 layers, Glen n = 3,
A = 1e-16, Weertman m = 3, beta_sq = 1e4), synthetic and seeded.

One step = one converged DIVA velocity solve from a zero initial guess (zero
velocity and zero DIVA state, set on the device) until the FP64 block-scaled
nonlinear residual ||D^-1(b(u) - A(u)u)|| / ||D^-1 b(u)|| <= 1e-6, A and b
assembled at the final u.  Inputs are uploaded to HBM before the timed region.

  chosen form: Anderson-mixed Picard (depth 2, mixing 1.0, history restarted
               when the residual exceeds 5x the best so far) after a 5-decade
               regularisation continuation, warm-started FP64 BiCGStab with the
               aggregation-AMG preconditioner, rtol 0.01; on the channel the
               V-cycle smooths with l1 block Jacobi (--smoother)
  plain form:  the reference iteration (relaxation 0.2 with the adaptive
               relaxation rule, zero Krylov initial guess, rtol 1e-7), FP64
               BiCGStab with the same preconditioner

`value` is the chosen form's solves/s; the plain form is timed in the same run
(`plain`), and both are checked against each other (max relative velocity
difference).  `roofline` is the largest single kernel of the chosen form, the
multigrid preconditioner's finest-level prolongation + smoothing sweep (a
pass over the fine matrix, two per BiCGStab iteration), each launch inside the
timed solves bracketed by HIP events on the solver's stream; `traffic` is its
HBM bytes per launch from the committed rocprofv3 counter summary.
`roofline_krylov_product` is the same for the FP64 block products with their
dot-product epilogues (two per iteration), with the plain product timed
back-to-back after the solves beside it.
`cpu_baseline` is the oracle's FP64 restatement of the reference iteration
(GMRES(30) + ILU(0)) on the host cores, timed on a bounded sample of Picard
iterations of the same workload and scaled to solves/s with the plain form's
Picard count; beside it, full measured CPU solves of the dome (the
CPU-runnable configuration, 2,205 vertices) with the same stopping test, and
the GPU's plain and chosen forms on the same dome.
Other keys: `roofline_out_of_cache` (the Krylov block product on the
continental mesh, whose 326 MB per product exceed the 256 MB Infinity Cache),
`oracle_parity_at_scale` (the chosen form against the committed oracle
solution of a 20k-vertex DIVA channel: maximum, median and share of
triangles within 1e-5), `thickness_step` (the semi-implicit
ice-thickness step on the same mesh with the velocity the solve left on the
device, host arrays for the other fields), `bpa` (the BPA 3-D stress
balance on ISMIP-HOM experiment A, the reference's BPA integrated test: a
1 km mesh on the GPU with its block product's roofline, and the reference's
8 km mesh on the CPU oracle and the GPU), and the unscaled residual
||b - A u|| / ||b|| of every solve next to the block-scaled one.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BASE = dict(include_crossterms=0, vel_max=1e5, abstol=0.0, stop_rule=1, nonlinear_rtol=1e-6, visc_it_nit=1000,
            krylov_maxit=20000)
CHOSEN = dict(precision=0, precond=3, warm_start=1, rtol=0.01, accel_depth=2, accel_mixing=1.0, accel_continuation=5,
              accel_restart=5.0)
PLAIN = dict(precision=0, precond=3, warm_start=0, rtol=1e-7)
# On the 200k-vertex channel the chosen form also smooths with l1 block Jacobi
# (40% fewer Krylov iterations there); on the continental mesh that smoother
# slows the Picard iteration down and on the ensemble members it needs more
# Krylov iterations, so those keep plain block Jacobi (autotune searches both).
L1_WORKLOADS = ("channel",)


def chosen_for(workload, smoother="auto"):
    """The chosen form of a workload; smoother overrides the V-cycle smoother
    ("bj" block Jacobi, "l1" l1 block Jacobi)."""
    l1 = workload in L1_WORKLOADS if smoother == "auto" else smoother == "l1"
    return dict(CHOSEN, amg_smoother=1 if l1 else 0)
METRIC = "converged velocity solves/s at 1e-6 rel. residual, plain vs chosen form, 1 GPU"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(name):
    """HBM bytes per launch of a kernel from its committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def solve_once(s, stress="DIVA", info_check=True):
    s.set_velocity(None, None)
    if stress == "DIVA":
        s.reset_diva_state()
    info = s.run(stress)
    if info_check and not info.rel_residual <= BASE["nonlinear_rtol"]:
        raise RuntimeError(f"solve did not converge: residual {info.rel_residual}")
    return info


def cpu_info():
    """CPU model and the cores this process may use (the oracle's OpenMP threads)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", avail))
    return {"model": model, "nproc": avail, "threads": threads}


def dome_full_solves(V, W, reps=3):
    """The CPU-runnable configuration measured end to end: full converged SSA
    solves of the dome (block-scaled residual <= 1e-6, relaxed Picard with the
    reference rule) by the oracle on the host cores (median of `reps`), and by
    the GPU's plain and chosen forms on the same inputs."""
    import numpy as np
    from ufemism2_amd import _lib as L
    from oracle import binding as OB
    w = W.dome(h=40e3)
    inp = w.inputs()
    p = L.default_params()
    for k, v in dict(BASE, rtol=1e-7).items():
        setattr(p, k, v)
    o = OB.Oracle(w.descriptor())
    times, info = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        rc, uo, vo, info = o.solve_ssa(p, inp)
        times.append(time.perf_counter() - t0)
        if rc != 0 or not info.rel_residual <= BASE["nonlinear_rtol"]:
            raise RuntimeError(f"oracle dome solve failed: rc {rc} residual {info.rel_residual}")
    t_cpu = float(np.median(times))
    out = {"workload": f"dome_ssa_40km ({w.mesh.nV} vertices, {w.mesh.nT} triangles)",
           "cpu_oracle": {"solves_per_s": 1.0 / t_cpu, "seconds": t_cpu, "picard": info.n_visc_its,
                          "krylov": info.n_Axb_its, "residual": info.rel_residual,
                          "residual_unscaled": info.rel_residual_unscaled,
                          "method": "oracle: relaxed Picard (reference rule), GMRES(30) + ILU(0), rtol 1e-7"}}
    for label, form, reps_g in (("gpu_plain", PLAIN, 3), ("gpu_chosen", CHOSEN, 10)):
        s = V.VelocitySolver(w.descriptor(), 0, **BASE, **form)
        s.upload(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"], inp["mask_grounded"])
        solve_once(s, "SSA")
        t0 = time.perf_counter()
        for _ in range(reps_g):
            gi = solve_once(s, "SSA")
        t = (time.perf_counter() - t0) / reps_g
        u, v = s.get_velocity()
        s.close()
        sc = max(np.abs(uo).max(), np.abs(vo).max())
        out[label] = {"solves_per_s": 1.0 / t, "seconds": t, "picard": gi.n_visc_its, "krylov": gi.n_Axb_its,
                      "residual": gi.rel_residual, "residual_unscaled": gi.rel_residual_unscaled,
                      "max_rel_vel_diff_vs_cpu_oracle": float(max(np.abs(u - uo).max(), np.abs(v - vo).max()) / sc)}
    out["gpu_chosen_vs_cpu"] = out["gpu_chosen"]["solves_per_s"] / out["cpu_oracle"]["solves_per_s"]
    return out


def oracle_parity_at_scale(V, W):
    """The chosen form on the 20k-vertex DIVA channel (seed 2) against the
    committed oracle solution of the reference iteration (nonlinear residual
    1e-8): relative velocity differences, normalised by max |u_oracle|.  Two
    converged states of these equations differ along a weakly determined mode
    of the floating shelf (the engine's own reference iteration ends 1.8e-2
    from the fixture, each state with residual 1e-8 in both implementations),
    so the median and the share of triangles within 1e-5 are reported beside
    the maximum; the test suite checks the oracle's residual at the engine's
    state."""
    import numpy as np
    path = os.path.join(ROOT, "tests", "golden", "oracle_channel_diva_20k.npz")
    if not os.path.exists(path):
        return None
    f = np.load(path)
    w = W.channel(seed=2, target_nV=20000, nz=10, stress_balance="DIVA")
    if str(f["digest"]) != w.mesh.digest():
        return {"error": "mesh digest differs from the fixture's"}
    inp = w.inputs()
    s = V.VelocitySolver(w.descriptor(), 0, **dict(BASE, nonlinear_rtol=1e-8), **chosen_for("channel"))
    s.upload(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"], inp["mask_grounded"])
    s.set_velocity(None, None)
    s.reset_diva_state()
    info = s.run("DIVA")
    u, v = s.get_velocity()
    s.close()
    fu, fv = f["u"].astype(np.float64), f["v"].astype(np.float64)
    sc = max(np.abs(fu).max(), np.abs(fv).max())
    d = np.hypot(u - fu, v - fv) / sc
    return {"workload": f"channel_diva seed 2 ({w.mesh.nV} vertices, {w.mesh.nT} triangles, nz 10)",
            "max_rel_vel_err_vs_oracle": float(max(np.abs(u - fu).max(), np.abs(v - fv).max()) / sc),
            "median_rel_vel_err_vs_oracle": float(np.median(d)),
            "share_of_triangles_within_1e-5": float((d <= 1e-5).mean()),
            "chosen_residual": info.rel_residual, "chosen_residual_unscaled": info.rel_residual_unscaled,
            "oracle_residual": float(f["rel_residual"]), "stop": "block-scaled nonlinear residual <= 1e-8"}


def out_of_cache_roofline(V, W, steps=1):
    """The Krylov block products with their dot epilogues, timed live (HIP
    events on the solver's stream) inside chosen-form solves of the
    continental mesh, whose matrix stream (326 MB per product) exceeds the
    256 MB Infinity Cache; and the same product launched back to back."""
    w = W.continental()
    inp = w.inputs()
    s = V.VelocitySolver(w.descriptor(), 0, **BASE, **CHOSEN)
    s.upload(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"], inp["mask_grounded"])
    solve_once(s, "DIVA")
    s.kernel_timing(True)
    for _ in range(steps):
        info = solve_once(s, "DIVA")
    s.kernel_timing(False)
    n, ms, by = s.kernel_timing_read(0)
    spmv_ms, spmv_bytes = s.profile_spmv(100)
    s.close()
    rate = by / (ms * 1e-3) / 1e9 if n else 0.0
    return {"kernel": "k_bsr_spmv<double, OP_DOT1|OP_DOT2> (timed live)", "workload":
            f"continental_diva_400k ({w.mesh.nV} vertices, {w.mesh.nT} triangles)",
            "bound": "hbm", "achieved": rate, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": rate / HBM_PEAK_GBS,
            "traffic": read_traffic("bsr_spmv_dot_continental400k_pmc.txt"), "avg_ms": ms, "launches": n,
            "algorithmic_bytes": by, "solve_seconds": info.seconds,
            "plain_product_back_to_back": {"avg_ms": spmv_ms, "algorithmic_bytes": spmv_bytes,
                                           "achieved": spmv_bytes / (spmv_ms * 1e-3) / 1e9}}


def cpu_baseline(w, seconds, n_picard_plain):
    """Oracle (CPU restatement of the reference's linear solve, GMRES(30) with
    left ILU(0), one rank) on a bounded sample of the same workload.  A full
    CPU solve of this size takes hours, so the sample is the first Picard
    iteration's linear system (assembled at zero velocity): two truncated
    GMRES runs give the time per iteration and the residual reduction rate,
    extrapolated to the reference tolerance (rtol 1e-7) and multiplied by
    the plain form's Picard count."""
    import math
    import numpy as np
    from ufemism2_amd import _lib as L
    from oracle import binding as OB
    o = OB.Oracle(w.descriptor())
    inp = w.inputs()
    p = L.default_params()
    for k, v in BASE.items():
        setattr(p, k, v)
    nT = w.mesh.nT
    A, b = o.system(p, inp, np.zeros(nT), np.zeros(nT))
    A.sort_indices()
    t0 = time.perf_counter()
    OB.gmres_ilu0_stats(A, b, rtol=1e-30, maxit=30)
    t_it = (time.perf_counter() - t0) / 30
    m1 = max(30, int(seconds / 3.0 / t_it / 30) * 30)
    t0 = time.perf_counter()
    _, i1, r1 = OB.gmres_ilu0_stats(A, b, rtol=1e-30, maxit=m1)
    t1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, i2, r2 = OB.gmres_ilu0_stats(A, b, rtol=1e-30, maxit=2 * m1)
    t2 = time.perf_counter() - t0
    per_it = t2 / i2
    rate = (r2 / r1) ** (1.0 / max(i2 - i1, 1))
    r0 = r1 / rate ** i1
    its_needed = math.log(1e-7) / math.log(rate) if rate < 1.0 else float("inf")
    t_solve = per_it * its_needed * n_picard_plain
    cores = cpu_info()["threads"]
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"oracle GMRES(30)+ILU(0) on the first Picard system ({2 * nT} unknowns): {i1}+{i2} "
                      f"iterations in {t1 + t2:.1f} s, {per_it * 1e3:.1f} ms/iteration, residual rate "
                      f"{rate:.5f}/iteration -> {its_needed:.0f} iterations to rtol 1e-7 (r0 {r0:.3g}), times "
                      f"{n_picard_plain} Picard iterations of the plain form (estimate, not a full solve)"}


WORKLOADS = ("channel", "dome", "continental", "ensemble")


def build_workloads(W, args):
    """The benchmark workloads: the channel (default, the metric's workload),
    the dome (the CPU-runnable case, SSA), the continental sheet (DIVA, plus
    its heat columns) and the ensemble of uneven channel members."""
    if args.workload == "channel":
        return "channel_diva_200k", [W.channel(target_nV=args.nv, stress_balance="DIVA")]
    if args.workload == "dome":
        return "dome_ssa_40km", [W.dome(h=40e3)]
    if args.workload == "continental":
        return "continental_diva_400k", [W.continental()]
    return f"ensemble_{args.members}", [W.ensemble_member(k) for k in range(args.members)]


def time_form(V, ws, form, steps, warmup, label, concurrency=1, autotune=False):
    """Warm-up and timed converged solves of one form over the workloads;
    returns (seconds for `steps` passes, infos, last velocities per workload).
    autotune: each solver first runs its correctness-gated coordinate search
    (outside the timed region) and keeps the winner."""
    solvers = []
    for w in ws:
        inp = w.inputs()
        s = V.VelocitySolver(w.descriptor(), 0, **BASE, **form)
        s.upload(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"], inp["mask_grounded"])
        if autotune:
            best, table = s.autotune(w.stress_balance)
            log(f"[bench] autotune {w.name}: " + "; ".join(
                f"{cfg} {'%.2fs' % t if t is not None else 'rejected'}" for cfg, t in table) + f" -> {best}")
        solvers.append((w, s))
    for i in range(warmup):
        for w, s in solvers:
            info = solve_once(s, w.stress_balance)
        log(f"[bench] {label} warmup {i}: picard {info.n_visc_its} krylov {info.n_Axb_its} "
            f"residual {info.rel_residual:.2e} {info.seconds:.2f}s")
    infos = []
    for _, s in solvers:
        s.kernel_timing(True)
    t0 = time.perf_counter()
    for i in range(steps):
        if concurrency > 1 and len(solvers) > 1:
            # members on `concurrency` host threads, one context (and HIP stream) each;
            # the C ABI releases the interpreter lock during the solve
            import concurrent.futures as cf
            with cf.ThreadPoolExecutor(max_workers=concurrency) as ex:
                infos.extend(ex.map(lambda ws: solve_once(ws[1], ws[0].stress_balance), solvers))
        else:
            for w, s in solvers:
                infos.append(solve_once(s, w.stress_balance))
        log(f"[bench] {label} step {i}: picard {infos[-1].n_visc_its} krylov {infos[-1].n_Axb_its} "
            f"residual {infos[-1].rel_residual:.2e} {infos[-1].seconds:.2f}s")
    t = time.perf_counter() - t0
    vel = [s.get_velocity() for _, s in solvers]
    for _, s in solvers:
        s.kernel_timing(False)
    return t, infos, vel, solvers


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed chosen-form solves")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-steps", type=int, default=1, help="timed plain-form solves")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU baseline sample length")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the legs outside the metric (out-of-cache roofline, oracle parity at scale, "
                         "dome CPU/GPU solves)")
    ap.add_argument("--spmv-reps", type=int, default=200)
    ap.add_argument("--nv", type=int, default=200_000, help="target vertices of the channel mesh")
    ap.add_argument("--workload", choices=WORKLOADS, default="channel")
    ap.add_argument("--smoother", choices=("auto", "bj", "l1"), default="auto",
                    help="V-cycle smoother of the chosen form (auto: l1 on the channel geometry)")
    ap.add_argument("--members", type=int, default=64, help="ensemble members (--workload ensemble)")
    ap.add_argument("--concurrency", type=int, default=8,
                    help="ensemble members solved at once, one HIP stream each (--workload ensemble)")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the chosen form's Anderson depth and multigrid knobs per workload before timing "
                         "(correctness-gated coordinate search; not part of the default measurement)")
    ap.add_argument("--bpa-only", action="store_true",
                    help="only the BPA leg (ISMIP-HOM A), printed as one JSON object (not the bench line)")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="one chosen-form solve, then only the block-product and copy-kernel launches "
                         "(for hardware counter runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the velocity solve runs on a single MI355X here; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import __graft_entry__ as ge
    ge.load_package()
    from ufemism2_amd import velocity as V, workloads as W

    if args.bpa_only:
        print(json.dumps(bpa_leg(V, W, cpu=not args.no_cpu_baseline)), flush=True)
        return

    t0 = time.perf_counter()
    name, ws = build_workloads(W, args)
    log(f"[bench] workload {name}: {len(ws)} mesh(es), {sum(w.mesh.nV for w in ws)} vertices, "
        f"built in {time.perf_counter() - t0:.1f}s")
    w = ws[0]

    if args.profile_kernels:
        inp = w.inputs()
        s = V.VelocitySolver(w.descriptor(), 0, **BASE, **chosen_for(args.workload, args.smoother))
        s.upload(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"], inp["mask_grounded"])
        solve_once(s, w.stress_balance)
        ms, nbytes = s.profile_spmv(args.spmv_reps)
        copy_bytes = int(nbytes) // 2 // 16 * 16
        cms = s.profile_copy(copy_bytes, args.spmv_reps)
        s.close()
        print(json.dumps({"spmv_avg_ms": ms, "spmv_algorithmic_bytes": nbytes, "copy_avg_ms": cms,
                          "copy_read_bytes": copy_bytes, "copy_write_bytes": copy_bytes,
                          "launches_each": args.spmv_reps + 1}), flush=True)
        return

    conc = args.concurrency if args.workload == "ensemble" else 1
    chosen = chosen_for(args.workload, args.smoother)
    t_chosen, infos, vel_c, solvers = time_form(V, ws, chosen, args.steps, args.warmup, "chosen", conc,
                                                autotune=args.autotune)
    # kernels timed live inside the timed solves (HIP events on the solver stream):
    # 1 = the preconditioner's finest-level prolongation + smoothing (the largest
    # single kernel of the solve), 0 = the Krylov block products
    live = {}
    for kern in (0, 1):
        n_t, ms_t, by_t = 0, 0.0, 0.0
        for _, s in solvers:
            n, ms, by = s.kernel_timing_read(kern)
            ms_t = (ms_t * n_t + ms * n) / max(n_t + n, 1)
            n_t += n
            by_t = by
        live[kern] = (n_t, ms_t, by_t)
    spmv_ms, spmv_bytes = solvers[0][1].profile_spmv(args.spmv_reps)
    heat = time_heat(V, W, w.mesh.nV, 12, args.steps, solvers[0][1]) if args.workload == "continental" else None
    thickness = (time_thickness(V, W, w, solvers[0][1])
                 if not args.no_extra and args.workload in ("channel", "continental") else None)
    for _, s in solvers:
        s.close()
    n_solves = args.steps * len(ws)

    pinfos, t_plain, diff = [], 0.0, None
    if args.plain_steps > 0:
        t_plain, pinfos, vel_p, psolvers = time_form(V, ws, PLAIN, args.plain_steps, 0, "plain", conc)
        for _, s in psolvers:
            s.close()
        diff = 0.0
        for (uc, vc), (up, vp) in zip(vel_c, vel_p):
            scale = max(np.abs(up).max(), np.abs(vp).max())
            diff = max(diff, float(max(np.abs(uc - up).max(), np.abs(vc - vp).max()) / scale))

    value = n_solves / t_chosen
    plain_value = args.plain_steps * len(ws) / t_plain if args.plain_steps else None
    def rate(k):
        n, ms, by = live[k]
        return by / (ms * 1e-3) / 1e9 if n else 0.0
    traffic = read_traffic("amg_prolong_smooth_fine_channel200k_pmc.txt") if args.workload == "channel" else None
    traffic_spmv = read_traffic("bsr_spmv_dot_channel200k_pmc.txt") if args.workload == "channel" else None
    out = {
        "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * t_chosen / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": name, "meshes": len(ws), "nV": sum(x.mesh.nV for x in ws),
                   "nTri": sum(x.mesh.nT for x in ws), "nz": w.nz,
                   "unknowns": 2 * sum(x.mesh.nT for x in ws), "seed": w.seed, "mesh": w.mesh.digest(),
                   "stress_balance": w.stress_balance, "concurrent_streams": conc,
                   "stop": "block-scaled nonlinear residual <= 1e-6",
                   "chosen": "Anderson(2, 1.0, history restart at 5x the best residual) Picard after 5-decade "
                             "eps0 continuation, FP64 AMG-BiCGStab warm, "
                             "rtol 0.01, V-cycle smoother "
                             + ("l1 block Jacobi" if chosen["amg_smoother"] else "block Jacobi"),
                   "plain": "relaxed Picard (reference rule), FP64 AMG-BiCGStab zero guess, rtol 1e-7"},
        "chosen": {"picard": [i.n_visc_its for i in infos], "krylov": [i.n_Axb_its for i in infos],
                   "residual": [i.rel_residual for i in infos],
                   "residual_unscaled": [i.rel_residual_unscaled for i in infos]},
        "plain": {"value": plain_value, "unit": "solves/s",
                  "ms_per_step": 1e3 * t_plain / max(args.plain_steps, 1),
                  "picard": [i.n_visc_its for i in pinfos], "krylov": [i.n_Axb_its for i in pinfos],
                  "residual": [i.rel_residual for i in pinfos],
                  "residual_unscaled": [i.rel_residual_unscaled for i in pinfos]},
        "speedup_vs_plain": value / plain_value if plain_value else None,
        "autotuned": ([{**{k: getattr(s.params, k) for k in ("accel_depth",)}, **s.amg_options}
                       for _, s in solvers] if args.autotune else None),
        "max_rel_vel_diff_chosen_vs_plain": diff,
        "roofline": {"kernel": "k_amg_prolong_smooth<double, double> on the finest level (timed live)",
                     "bound": "hbm", "achieved": rate(1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": rate(1) / HBM_PEAK_GBS, "traffic": traffic, "avg_ms": live[1][1],
                     "launches": live[1][0], "algorithmic_bytes": live[1][2]},
        "roofline_krylov_product": {"kernel": "k_bsr_spmv<double, OP_DOT1|OP_DOT2> (timed live)",
                                    "bound": "hbm", "achieved": rate(0), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                    "frac": rate(0) / HBM_PEAK_GBS, "traffic": traffic_spmv,
                                    "avg_ms": live[0][1], "launches": live[0][0],
                                    "algorithmic_bytes": live[0][2],
                                    "plain_product_back_to_back": {
                                        "avg_ms": spmv_ms, "algorithmic_bytes": spmv_bytes,
                                        "achieved": spmv_bytes / (spmv_ms * 1e-3) / 1e9}},
        "cpu_baseline": None,
    }
    if heat is not None:
        out["heat_columns"] = heat
    if thickness is not None:
        out["thickness_step"] = thickness
    if not args.no_extra and args.workload == "channel":
        log("[bench] BPA: ISMIP-HOM A on the GPU (1 km mesh) and GPU vs CPU oracle (the reference's 8 km mesh)")
        out["bpa"] = bpa_leg(V, W, cpu=not args.no_cpu_baseline)
        log("[bench] out-of-cache roofline (continental mesh)")
        out["roofline_out_of_cache"] = out_of_cache_roofline(V, W)
        log("[bench] oracle parity at scale (20k-vertex channel)")
        out["oracle_parity_at_scale"] = oracle_parity_at_scale(V, W)
    if not args.no_cpu_baseline:
        n_plain = pinfos[0].n_visc_its if pinfos else 231
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds, n_plain)
        out["cpu_baseline"]["cpu"] = cpu_info()
        if not args.no_extra:
            log("[bench] full dome solves, CPU oracle and GPU")
            out["cpu_baseline"]["measured_full_solves"] = dome_full_solves(V, W)
    print(json.dumps(out), flush=True)


def bpa_leg(V, W, h=1e3, cpu=True):
    """The BPA 3-D stress balance (solve_BPA) on the reference's own BPA
    integrated test, ISMIP-HOM experiment A at L = 160 km with its settings:
    one solve from zero velocity on a 1 km mesh (timed after one untimed
    solve), its stage times, and the BPA block product timed back to back
    (HIP events; algorithmic bytes: blocks + column indices + row pointers +
    x + y).  With cpu: the oracle's restatement of the reference iteration
    (GMRES(30) + ILU(0)) and the engine on the reference's 8 km mesh, and the
    largest velocity difference between the two."""
    import numpy as np
    from ufemism2_amd.operators import build_aa, zeta_stag
    res = {}
    w = W.ismip_hom_a(h=h)
    aa, zs = build_aa(w.mesh), zeta_stag(w.zeta)
    inp = w.inputs()
    s = V.VelocitySolver(w.descriptor(), **w.params)
    s.zeta_setup(aa, zs)
    args = (inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"], inp["mask_grounded"])
    s.solve_BPA(*args)
    t0 = time.perf_counter()
    u, v, info = s.solve_BPA(*args)
    t = time.perf_counter() - t0
    stages = s.bpa_timings()
    spmv_ms, spmv_bytes = s.bpa_profile_spmv(100)
    s.close()
    nbk = w.mesh.nT * w.nz
    res["gpu"] = {"workload": w.name, "nV": w.mesh.nV, "nTri": w.mesh.nT, "nz": w.nz, "unknowns": 2 * nbk,
                  "seconds_per_solve": t, "solves_per_s": 1.0 / t, "picard": info.n_visc_its,
                  "krylov": info.n_Axb_its, "rel_residual": info.rel_residual, "stage_ms": stages,
                  "max_surface_speed": float(np.hypot(u, v).reshape(w.nz, -1)[0].max())}
    res["roofline_block_product"] = {"kernel": "k_bsr_spmv<double, OP_PLAIN> on the BPA system (back to back)",
                                     "bound": "hbm", "avg_ms": spmv_ms, "algorithmic_bytes": spmv_bytes,
                                     "achieved": spmv_bytes / (spmv_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                                     "unit": "GB/s", "frac": spmv_bytes / (spmv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    if cpu:
        from oracle import binding as OB
        from ufemism2_amd import _lib as L
        w8 = W.ismip_hom_a(h=8e3)
        aa8, zs8 = build_aa(w8.mesh), zeta_stag(w8.zeta)
        i8 = w8.inputs()
        p = L.default_params()
        for k, val in w8.params.items():
            if k in ("bc_u", "bc_v"):
                for i, x in enumerate(val):
                    getattr(p, k)[i] = int(x)
            else:
                setattr(p, k, val)
        t0 = time.perf_counter()
        rc, uo, vo, io = OB.solve_bpa(w8.descriptor(), aa8, zs8, p, i8)
        tc = time.perf_counter() - t0
        s8 = V.VelocitySolver(w8.descriptor(), **w8.params)
        s8.zeta_setup(aa8, zs8)
        a8 = (i8["Hi"], i8["Hs"], i8["A_flow"], i8["beta_sq"], i8["fgr_b"], i8["mask_grounded"])
        s8.solve_BPA(*a8)
        t0 = time.perf_counter()
        ug, vg, ig = s8.solve_BPA(*a8)
        tg = time.perf_counter() - t0
        s8.close()
        scale = float(np.hypot(uo, vo).max())
        res["reference_mesh_8km"] = {
            "nV": w8.mesh.nV, "nTri": w8.mesh.nT, "unknowns": 2 * w8.mesh.nT * w8.nz,
            "cpu_oracle": {"seconds": tc, "picard": io.n_visc_its, "krylov": io.n_Axb_its, "rc": rc,
                           "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                           "kind": "port"},
            "gpu": {"seconds": tg, "picard": ig.n_visc_its, "krylov": ig.n_Axb_its},
            "speedup": tc / tg,
            "max_rel_vel_diff_gpu_vs_oracle": float(max(np.abs(ug - uo).max(), np.abs(vg - vo).max()) / scale)}
    return res


def time_thickness(V, W, w, solver, reps=10):
    """The semi-implicit ice-thickness step after the velocity solve, on the
    same mesh, with the velocity the solve left on the device (no copy) and
    the other fields as host arrays (their copies in and the outputs' copies
    out are inside the time)."""
    import numpy as np
    u, v = solver.get_velocity()
    inp = W.thickness_step_inputs(w, u, v)
    t0 = time.perf_counter()
    solver.thickness_setup()
    t_setup = time.perf_counter() - t0
    kw = dict(inp, u_vav_b=None, v_vav_b=None)
    solver.calc_dHi_dt_semiimplicit(**kw)
    t0 = time.perf_counter()
    for _ in range(reps):
        rate, Hn, divQ, dt_max, its = solver.calc_dHi_dt_semiimplicit(**kw)
    t = (time.perf_counter() - t0) / reps
    return {"vertices": w.mesh.nV, "seconds_per_step": t, "steps_per_s": 1.0 / t, "krylov": its,
            "krylov_method": "GMRES(30), block-Jacobi on vertex pairs", "rtol": 1e-8, "abstol": 1e-6,
            "dt_years": inp["dt"], "max_abs_dHi_dt": float(np.abs(rate).max()), "setup_seconds": t_setup,
            "note": "host arrays in and out (PCIe-inclusive), velocity taken from the solver's device state"}


def time_heat(V, W, nV, nz, steps, solver):
    """Heat columns of the continental configuration: nV columns x nz layers,
    one call = one implicit time step of every column.  Inputs resident in
    device memory (ufe_heat_columns_device through the solver's context), then
    host arrays in and out (ufe_heat_columns, through the context, whose
    stream and device buffers are reused, and without one)."""
    import numpy as np
    from ufemism2_amd.devmem import DeviceArray
    c = W.heat_columns_inputs(nV, nz, seed=3)
    hin, keep = V.make_heat_inputs(nV, nz, c["zeta"], c["fields"], c["dt"])
    res = {}
    for label, sv in (("with_context", solver), ("without_context", None)):
        V.heat_columns(hin, sv)
        t0 = time.perf_counter()
        for _ in range(steps):
            T, unst, nun = V.heat_columns(hin, sv)
        res[label] = (time.perf_counter() - t0) / steps
    dev = {k: DeviceArray.from_host(v) for k, v in c["fields"].items()}
    dev["zeta"] = DeviceArray.from_host(c["zeta"])
    hd = V.make_heat_inputs_device(nV, nz, {k: d.ptr for k, d in dev.items()}, c["dt"])
    Tout = DeviceArray(nV * nz, np.float64)
    un = DeviceArray(nV, np.int32)
    V.heat_columns_device(hd, Tout.ptr, un.ptr, solver)
    reps = max(steps, 50)
    t0 = time.perf_counter()
    for _ in range(reps):
        nd = V.heat_columns_device(hd, Tout.ptr, un.ptr, solver)
    td = (time.perf_counter() - t0) / reps
    same = bool(np.array_equal(Tout.to_host(), T) and nd == nun)
    for d in list(dev.values()) + [Tout, un]:
        d.free()
    t = res["with_context"]
    return {"columns": nV, "nz": nz, "columns_per_s": nV / td, "seconds_per_call": td, "n_unstable": nd,
            "device_result_equals_host_call": same,
            "host_arrays": {"columns_per_s": nV / t, "seconds_per_call": t,
                            "seconds_per_call_without_context": res["without_context"],
                            "note": "includes the host<->device copies of the inputs and outputs"},
            "note": "inputs resident in device memory; one call = launch, reduction and count read-back"}


if __name__ == "__main__":
    main()
