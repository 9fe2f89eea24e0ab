"""The named workloads at their stated lengths on one GPU, FP64 against mixed precision.

`python -m hoc.configs [bfs cylinder ensemble]` runs, for each named workload, the FP64 path and
the mixed-precision path from the same state, and prints one JSON line per workload: steps
reached, seconds and Mdof-steps/s per precision, Krylov iterations per solve, and the relative
L2 errors (plain and quadrature-weighted) of the mixed fields against the FP64 ones (per
instance for the ensemble).  Each run reports its progress on stderr every `--every` steps and
stops early, saying so, when its time budget runs out.

  bfs        backward-facing step 384 x 64 (22 528 elements, jitter 0.1, seed 2), K=5, Re=400,
             dt=5e-4, 2000 RK4 steps from t = 20/3 (the inflow ramp 2|sin(pi t/40)| at 1)
  cylinder   O-grid of quadratic quads around a D=1 cylinder, 256 radial x 512 azimuthal
             (131 072 elements), K=6, Re=100, dt=1e-6, 4000 RK4 steps from t = 20/3, initial
             vorticity 1e-3 U(-1, 1) (seed 3)
  ensemble   64 lid-driven cavities with seeded sizes 16^2..256^2 (1.5 M elements), K=5, Re
             log-uniform in [100, 3200], per-instance dt (half the RK4 diffusion limit, capped by
             an advective CFL of 0.05), 500 RK4 steps
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import HO2D, HOC_FP64, HOC_MIXED, meshes

LAP_RHO_PER_N2 = {4: 373.9, 5: 799.5}   # rho(J^-1 A) / n^2 on an n x n unit cavity (measured by power iteration)
RK4_REAL_LIMIT = 2.785                  # |z| on the negative real axis inside the RK4 stability region


def cavity_dt(n, K, Re, frac=0.5):
    """Half the explicit RK4 diffusion limit of an n x n unit cavity (rho(J^-1 A) ~ c_K n^2)."""
    return frac * RK4_REAL_LIMIT / (LAP_RHO_PER_N2[K] * n * n / Re)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def errors(a, b, areas):
    d = np.linalg.norm(b)
    q = np.sqrt(np.sum(areas * b * b))
    return {"rel_l2": float(np.linalg.norm(a - b) / d) if d > 0 else 0.0,
            "rel_l2_quadrature": float(np.sqrt(np.sum(areas * (a - b) ** 2)) / q) if q > 0 else 0.0}


def workload(name):
    if name == "bfs":
        m = meshes.backward_facing_step(384, 64, 2, jitter=0.1, seed=2)
        return dict(mesh=m, K=5, kw=dict(problem_type=2, Re=400.0, dt=5e-4), steps=2000, t0=20.0 / 3.0, w0=None)
    if name == "cylinder":
        m = meshes.cylinder_ogrid(256, 512, 3)
        return dict(mesh=m, K=6, kw=dict(problem_type=3, Re=100.0, dt=1e-6), steps=4000, t0=20.0 / 3.0,
                    w0=lambda xy: 1e-3 * np.random.default_rng(3).uniform(-1.0, 1.0, len(xy)))
    if name == "ensemble":
        rng = np.random.default_rng(2024)
        ns = rng.integers(16, 257, size=64)
        Re = np.exp(rng.uniform(np.log(100.0), np.log(3200.0), size=64))
        dt = np.array([min(cavity_dt(int(n), 5, r), 0.05 / n) for n, r in zip(ns, Re)])
        m, off = meshes.concat([meshes.cavity(int(n), 2) for n in ns])
        return dict(mesh=m, K=5, kw=dict(problem_type=1, Re=float(Re[0]), dt=float(dt[0])), steps=500, t0=0.0,
                    w0=None, inst=(off, Re, dt))
    raise ValueError(f"unknown workload {name}")


def run(name, precision, tol, every, budget_s):
    c = workload(name)
    g = HO2D(K=c["K"], time_integration_type=4, advection_type=1, tol=tol, maxiter=800, precision=precision,
             **c["kw"])
    g.load_mesh(c["mesh"])
    if "inst" in c:
        off, Re, dt = c["inst"]
        g.set_instances(off, Re, dt)
    g.setup()
    if c["w0"] is not None:
        g.set_vorticity(c["w0"](g.solution_points()))
    if c["t0"]:
        g.set_time(c["t0"])
    tag = "fp64" if precision == HOC_FP64 else "mixed"
    t0 = time.perf_counter()
    done = 0
    while done < c["steps"]:
        n = min(every, c["steps"] - done)
        g.step(n)
        done += n
        el = time.perf_counter() - t0
        log(f"[{name} {tag}] {done}/{c['steps']} steps, {el:.1f} s, t = {g.time:.6g}")
        if el > budget_s and done < c["steps"]:
            log(f"[{name} {tag}] time budget of {budget_s:.0f} s reached")
            break
    el = time.perf_counter() - t0
    st = g.stats()
    out = dict(steps=done, seconds=el, mdof_steps_per_s=g.D * done / el / 1e6,
               krylov_iters_per_solve=st["total_iters"] / max(1, st["n_solves"]),
               w=g.get_vorticity(), psi=g.get_streamfunction(), areas=g.solution_areas(),
               off=c["inst"][0] if "inst" in c else None, K=c["K"], dofs=g.D, n_elem=c["mesh"].n_elem)
    g.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("names", nargs="*", default=["bfs", "cylinder", "ensemble"])
    ap.add_argument("--tol", type=float, default=1e-12)
    ap.add_argument("--every", type=int, default=100, help="steps between progress lines")
    ap.add_argument("--budget", type=float, default=240.0, help="seconds per run before stopping early")
    args = ap.parse_args(argv)
    for name in args.names:
        r64 = run(name, HOC_FP64, args.tol, args.every, args.budget)
        r32 = run(name, HOC_MIXED, args.tol, args.every, args.budget)
        steps = min(r64["steps"], r32["steps"])
        res = {"workload": name, "n_elem": r64["n_elem"], "K": r64["K"], "dofs": r64["dofs"], "tol": args.tol,
               "fp64": {k: r64[k] for k in ("steps", "seconds", "mdof_steps_per_s", "krylov_iters_per_solve")},
               "mixed": {k: r32[k] for k in ("steps", "seconds", "mdof_steps_per_s", "krylov_iters_per_solve")}}
        if r64["steps"] == r32["steps"]:
            res["mixed_vs_fp64"] = {"omega": errors(r32["w"], r64["w"], r64["areas"]),
                                    "psi": errors(r32["psi"], r64["psi"], r64["areas"])}
            if r64["off"] is not None:
                K2 = r64["K"] ** 2
                off = r64["off"]
                worst = 0.0
                for i in range(len(off) - 1):
                    sl = slice(off[i] * K2, off[i + 1] * K2)
                    for f in ("w", "psi"):
                        e = errors(r32[f][sl], r64[f][sl], r64["areas"][sl])
                        worst = max(worst, e["rel_l2"], e["rel_l2_quadrature"])
                res["mixed_vs_fp64"]["worst_instance"] = worst
            res["finite"] = bool(np.all(np.isfinite(r64["w"])) and np.all(np.isfinite(r32["w"])))
        res["steps_compared"] = steps
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
