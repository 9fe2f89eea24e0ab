"""Throughput of the HO-CXX time-step hot path on one MI355X.

Metric: Mdof-steps/s per RK4 step including the four Poisson solves, where
Mdof = N_el * K^2 / 1e6.  Workload: lid-driven cavity on the unit square, 512 x 512 = 2^18 bilinear quads stored as
unstructured element/edge lists, K = 5 (P = 4), Re = 100, zero initial
vorticity, RK4 with dt = half the measured explicit-RK4 diffusion limit, Krylov
relative tolerance 1e-10; by default 5 warm-up steps, then steps 6-25 timed.  Inputs live in
HBM for the whole timed region.

Prints ONE JSON line.  `value` is the plain FP64 path; by default the same
workload is then run in mixed precision (the FP64 Krylov method with its V-cycle
preconditioner in FP32) from the same initial state, and the line carries its throughput,
speed-up and relative L2 field error against the FP64 run after the same steps
(valid at <= 1e-8).  The cpu_baseline leg times the FP64 CPU oracle (the
restatement of the reference algorithm) on a bounded sample of the same workload
(a 64 x 64 cavity at the same K, Re, dt rule and tolerance), on 1 thread and on all
host threads, median of 3 runs, with its Krylov iterations, the CPU model and nproc.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("ho-cxx_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
RK4_REAL_LIMIT = 2.785          # |z| on the negative real axis inside the RK4 stability region
LAP_RHO_PER_N2 = {4: 373.9, 5: 799.5}   # rho(J^-1 A) / n^2 on an n x n unit cavity (measured)


def cavity_dt(n, K, Re, frac=0.5):
    rho = LAP_RHO_PER_N2[K] * n * n
    return frac * RK4_REAL_LIMIT / (rho / Re)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads():
    """Host threads available to this process (at most 16, the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(K, Re, tol, n=64, warm=2, runs=3, steps_1=3, steps_n=20, n_bench=512):
    """The FP64 oracle on a bounded sample of the workload (about 30 s of CPU work): the same
    cavity problem at n x n, on 1 thread (the reference's default serial build) and on all host
    threads (OpenMP linear algebra; the FR residual loops, < 2% of the oracle's step, stay
    serial).  After `warm` warm-up steps, `runs` consecutive timed runs of steps_1 / steps_n RK4
    steps each; the median run is reported, with the oracle's Krylov iterations per solve (its
    solver is the reference's: BiCGSTAB warm-started from the previous solution)."""
    import pyoracle
    from hoc import meshes
    from pyoracle import Oracle
    m = meshes.cavity(n, 2)
    dt = cavity_dt(n, K, Re)
    D = n * n * K * K
    res = {}
    for nt in (1, cpu_threads()):
        steps = steps_1 if nt == 1 else steps_n
        pyoracle.set_threads(nt)
        t0 = time.perf_counter()
        o = Oracle(m, K, problem_type=1, Re=Re, dt=dt, time_integration_type=4, advection_type=1, tol=tol, maxiter=150)
        setup = time.perf_counter() - t0
        o.step(warm)
        vals, secs, its = [], [], []
        for _ in range(runs):
            c0 = o.solve_counts()
            t0 = time.perf_counter()
            o.step(steps)
            el = time.perf_counter() - t0
            c1 = o.solve_counts()
            vals.append(D * steps / el / 1e6)
            secs.append(el)
            its.append((c1[1] - c0[1]) / max(1, c1[0] - c0[0]))
        k = int(np.argsort(vals)[len(vals) // 2])
        res[nt] = {"value": vals[k], "runs": vals, "timed_s": secs[k], "steps_per_run": steps,
                   "krylov_iters_per_solve": its[k], "setup_s": setup}
        log(f"cpu baseline, {nt} threads: runs {['%.4f' % v for v in vals]} Mdof-steps/s, "
            f"{its[k]:.1f} Krylov its/solve")
    pyoracle.set_threads(1)
    nt = cpu_threads()
    r1, rn = res[1], res[nt]
    try:
        nproc = os.cpu_count()
    except Exception:
        nproc = None
    return {"value": rn["value"], "unit": "Mdof-steps/s", "cores": nt, "kind": "port",
            "sample": f"oracle (FP64 CPU restatement of the reference algorithm; OpenMP SpMV/vector/dot in the "
                      f"AMG-preconditioned Krylov solve) on a {n}x{n} cavity ({D} dofs, 1/{(n_bench // n) ** 2} of "
                      f"the {n_bench}x{n_bench} workload), K={K}, Re={Re}, dt={dt:.3g}, tol={tol:g}: {warm} warm-up "
                      f"steps, then the median of {runs} runs of {rn['steps_per_run']} RK4 steps on {nt} threads "
                      f"({rn['krylov_iters_per_solve']:.1f} BiCGSTAB iterations per solve)",
            "cpu_model": cpu_model(), "nproc": nproc, "threads_used": nt,
            "runs": rn["runs"], "krylov_iters_per_solve": rn["krylov_iters_per_solve"], "setup_s": rn["setup_s"],
            "single_thread": {"value": r1["value"], "cores": 1, "runs": r1["runs"], "timed_s": r1["timed_s"],
                              "steps_per_run": r1["steps_per_run"],
                              "krylov_iters_per_solve": r1["krylov_iters_per_solve"]}}


def measured_traffic(kernel, n, K):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE + WRITE_SIZE
    (scaled by a plain copy's readings) on this workload, from the committed summary, or
    None when the summary does not cover it."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_cavity512_K5.md")
    if (n, K) != (512, 5) or not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0] == kernel:
            try:
                return float(cells[4]) * 1e6
            except ValueError:
                return None
    return None


def roofline(prof, suffix="", noop=0):
    """Dominant kernel (largest total time among the fine-operator variants of one
    precision): algorithmic bytes per launch over its average launch duration, from HIP
    events recorded on the library's stream around sampled launches; launches predicated off
    after convergence (they return at once) are left out by the library and counted in noop."""
    cand = {k: v for k, v in prof.items()
            if not k.startswith("_") and v["launches"] > 0 and k.endswith("_f32") == (suffix == "_f32")}
    name, v = max(cand.items(), key=lambda kv: kv[1]["total_ms"])
    avg_s = v["total_ms"] / max(1, v["launches"]) / 1e3
    gbs = v["bytes_per_launch"] / avg_s / 1e9
    return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "traffic": None, "traffic_unit": "bytes per launch", "kernel": name, "avg_launch_us": avg_s * 1e6,
            "bytes_per_launch": v["bytes_per_launch"], "launches": v["launches"],
            "noop_launches_excluded": noop,
            "all_variants": {k: {"avg_us": 1e3 * x["total_ms"] / max(1, x["launches"]),
                                 "GBs": x["bytes_per_launch"] / max(1e-12, x["total_ms"] / max(1, x["launches"]) / 1e3) / 1e9,
                                 "launches": x["launches"]} for k, x in cand.items()}}


# Every PROFILE_EVERY-th fine-operator launch of the timed steps is bracketed by HIP events: an
# event pair around every launch costs ~5% (FP64) to ~7% (mixed) of a step, one in eight costs
# < 1%; 8 is coprime with the 7 (FP64) and 9 (mixed) operator launches of a PCG iteration, so all variants and all
# positions in the V-cycle are sampled.
PROFILE_EVERY = 8


def run(hoc, m, n, K, Re, dt, tol, precision, warmup, steps):
    """Setup, warm-up and `steps` timed steps of one handle, with sampled per-launch HIP events
    on the fine-operator kernel over the timed steps."""
    g = hoc.HO2D(K=K, problem_type=1, Re=Re, dt=dt, time_integration_type=4, advection_type=1,
                 tol=tol, maxiter=500, precision=precision)
    g.load_mesh(m)
    t0 = time.perf_counter()
    g.setup()
    setup_s = time.perf_counter() - t0
    st = g.stats()
    tag = "mixed" if precision == hoc.HOC_MIXED else "fp64"
    log(f"[{tag}] setup {setup_s:.1f} s, dofs {st['n_dofs']}, levels {st['n_levels']}, krylov {st['krylov']}")
    for i in range(warmup):
        g.step(1)
        log(f"[{tag}] warmup step {i}: iters/solve {g.stats()['last_iters']}")
    s0 = g.stats()
    g.profile(True, every=PROFILE_EVERY)
    t0 = time.perf_counter()
    g.step(steps)   # synchronizes the device at the end
    el = time.perf_counter() - t0
    prof = g.profile_read()
    prof["_noop"] = g.profile_noop_launches
    g.profile(False)
    st = g.stats()
    ns = max(1, st["n_solves"] - s0["n_solves"])
    iters = (st["total_iters"] - s0["total_iters"]) / ns
    D = st["n_dofs"]
    value = D * steps / el / 1e6
    log(f"[{tag}] timed {steps} steps in {el:.3f} s: {value:.2f} Mdof-steps/s, {iters:.1f} Krylov its/solve, "
        f"last solve: rel residual {st['last_rel_residual']:.2e}")
    return g, {"value": value, "ms_per_step": 1e3 * el / steps, "krylov_iters_per_solve": iters, "setup_s": setup_s,
               "last_rel_residual": st["last_rel_residual"]}, prof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)   # the workload's protocol: steps 6-25 timed
    ap.add_argument("--warmup", type=int, default=5)   # after 5 warm-up steps
    ap.add_argument("--n", type=int, default=512, help="cells per side of the cavity")
    ap.add_argument("--K", type=int, default=5)
    ap.add_argument("--Re", type=float, default=100.0)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both",
                    help="fp64: plain path only; both (default): plain FP64 line plus the mixed-precision "
                         "run from the same initial state, with its field error against the FP64 run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (known bytes), the reference point "
                         "for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE readings")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 27
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only (DESIGN.md)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import hoc
    from hoc import meshes

    n, K, Re = args.n, args.K, args.Re
    dt = cavity_dt(n, K, Re)
    t0 = time.perf_counter()
    m = meshes.cavity(n, 2)
    log(f"mesh {n}x{n}: {time.perf_counter() - t0:.1f} s")
    primary = hoc.HOC_MIXED if args.precision == "mixed" else hoc.HOC_FP64
    g, res, prof = run(hoc, m, n, K, Re, dt, args.tol, primary, args.warmup, args.steps)
    D = g.stats()["n_dofs"]
    out = {
        "metric": "Mdof-steps/s per RK4 step incl. Poisson",
        "value": res["value"], "unit": "Mdof-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if primary == hoc.HOC_FP64 else "f64+f32",
        "data": "synthetic (lid-driven cavity, zero initial vorticity)",
        "config": {"workload": f"lid-driven cavity {n}x{n} quads (unstructured storage), K={K} (P={K - 1}), "
                               f"Re={Re:g}, RK4 dt={dt:.3g}, Krylov tol={args.tol:g}",
                   "n_elem": n * n, "dofs": D, "precision": "fp64" if primary == hoc.HOC_FP64 else "mixed",
                   "krylov_iters_per_solve": res["krylov_iters_per_solve"], "setup_s": res["setup_s"]},
        "roofline": None,
        "cpu_baseline": None,
    }
    timing = (f"HIP events on the library stream around every {PROFILE_EVERY}th launch of the timed steps; "
              f"launches predicated off after convergence excluded")
    out["roofline"] = roofline(prof, "" if primary == hoc.HOC_FP64 else "_f32", prof["_noop"])
    out["roofline"]["traffic"] = measured_traffic(out["roofline"]["kernel"], n, K)
    out["roofline"]["timing"] = timing
    if args.precision == "both":
        g2, res2, prof2 = run(hoc, m, n, K, Re, dt, args.tol, hoc.HOC_MIXED, args.warmup, args.steps)
        a = g.solution_areas()
        err = {}
        for fld, get in (("omega", "get_vorticity"), ("psi", "get_streamfunction")):
            x64, x32 = getattr(g, get)(), getattr(g2, get)()
            dn = np.linalg.norm(x64)
            qd = np.sqrt(np.sum(a * x64 * x64))
            err[fld] = {"rel_l2": float(np.linalg.norm(x32 - x64) / dn) if dn > 0 else 0.0,
                        "rel_l2_quadrature": float(np.sqrt(np.sum(a * (x32 - x64) ** 2)) / qd) if qd > 0 else 0.0}
        worst = max(v for e in err.values() for v in e.values())
        out["mixed"] = {"value": res2["value"], "unit": "Mdof-steps/s", "ms_per_step": res2["ms_per_step"],
                        "speedup_vs_fp64": res2["value"] / res["value"],
                        "field_error_vs_fp64": err, "valid_at_1e-8": worst <= 1e-8,
                        "krylov_iters_per_solve": res2["krylov_iters_per_solve"],
                        "roofline": dict(roofline(prof2, "_f32", prof2["_noop"])),
                        "dtype": "f64 Krylov (operator, residual, vectors) + f32 V-cycle"}
        out["mixed"]["roofline"]["traffic"] = measured_traffic(out["mixed"]["roofline"]["kernel"], n, K)
        out["mixed"]["roofline"]["timing"] = timing
        log(f"mixed: {res2['value']:.2f} Mdof-steps/s ({res2['value'] / res['value']:.2f}x), field error {err}")
        g2.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(K, Re, args.tol)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
