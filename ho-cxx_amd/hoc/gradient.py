"""Gradient harness: the final enstrophy of a lid-driven cavity run as a differentiable function
of the initial vorticity and of Fourier modes of the lid velocity, evaluated by the library's
discrete adjoint (hoc_adjoint_forward / hoc_adjoint_backward) behind a torch.autograd.Function.

  J(omega_0, a) = 1/2 sum_i A_i omega_N,i^2       (A_i = w_j w_k J at the solution points)
  lid speed u(x) = 1 + sum_{m=1..M} a_m sin(m pi s),   s in [0, 1] along the lid

The lid enters as the wall velocity BC_parl_vel = -u at the lid's flux points (the sign that
setup_IC_BC_SRC gives the unit lid, src/HO_2D.cpp:717-725).

`python -m hoc.gradient` runs the gradient check: a 128 x 128 cavity, K = 5, Re = 1000, 500
RK4 steps with checkpoints every 50; omega_0 a Gaussian random field (squared-exponential
covariance, correlation length 0.1, amplitude 0.1, seed 5) sampled by 2000 random Fourier
features; 32 lid modes (a_m = 0.05 N(0,1)/m, seed 5); the gradient against central differences
(step 1e-6) on 16 seeded random directions (seeds 50-65) in (omega_0, a).  It prints one JSON
line.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import HO2D, lib, meshes

try:
    import torch
except ImportError:  # the harness needs torch; the rest of the package does not
    torch = None


class CavityEnstrophy:
    """One FP64 handle on an n x n unit cavity with the lid profile basis."""

    def __init__(self, n=128, K=5, Re=1000.0, dt=None, nsteps=500, every=50, tol=1e-13, n_modes=32, maxiter=800):
        self.n, self.K, self.Re, self.nsteps, self.every = n, K, Re, nsteps, every
        if dt is None:  # half the RK4 diffusion limit (rho(J^-1 A) ~ c_K n^2), capped by an advective bound
            rho = {4: 373.9, 5: 799.5}.get(K, 800.0) * n * n
            dt = min(0.5 * 2.785 / (rho / Re), 0.05 / n)
        self.dt = dt
        self.mesh = meshes.cavity(n, 2)
        self.g = HO2D(K=K, problem_type=1, Re=Re, dt=dt, time_integration_type=4, advection_type=1, tol=tol,
                      maxiter=maxiter)
        self.g.load_mesh(self.mesh)
        self.g.setup()
        self.areas = self.g.solution_areas()
        self.pv_base = self.g.wall_velocity()
        lid = self.pv_base != 0.0
        bxy = self.g.boundary_points()
        s = np.zeros(len(bxy))
        if lid.any():
            x = bxy[lid, 0]
            s[lid] = (x - x.min()) / max(x.max() - x.min(), 1e-300)
        # d BC_parl_vel / d a_m at every boundary flux point (zero off the lid)
        self.basis = np.array([np.where(lid, self.pv_base * np.sin((m + 1) * np.pi * s), 0.0)
                               for m in range(n_modes)])

    def wall(self, a):
        return self.pv_base + self.basis.T @ np.asarray(a, dtype=np.float64)

    def _start(self, w0, a):
        self.g.set_wall_velocity(self.wall(a))
        self.g.set_vorticity(w0)
        self.g._chk(lib().hoc_set_time(self.g.h, 0.0))

    def value(self, w0, a):
        """J by a plain run (no checkpoints)."""
        self._start(w0, a)
        self.g.step(self.nsteps)
        wN = self.g.get_vorticity()
        return 0.5 * float(np.sum(self.areas * wN * wN))

    def value_and_record(self, w0, a):
        self._start(w0, a)
        self.g.adjoint_forward(self.nsteps, self.every)
        wN = self.g.get_vorticity()
        return 0.5 * float(np.sum(self.areas * wN * wN)), self.areas * wN

    def gradient(self, wbar):
        gw, gp = self.g.adjoint_backward(wbar)
        return gw, self.basis @ gp


if torch is not None:
    class _Enstrophy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, w0, a, prob):
            J, wbar = prob.value_and_record(w0.detach().cpu().double().numpy(), a.detach().cpu().double().numpy())
            ctx.prob, ctx.wbar = prob, wbar
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            gw, ga = ctx.prob.gradient(ctx.wbar)
            s = float(gJ)
            return s * torch.from_numpy(gw), s * torch.from_numpy(ga), None


def enstrophy(w0, a, prob):
    """Differentiable J(w0, a) for torch tensors (float64, host)."""
    if torch is None:
        raise RuntimeError("the gradient harness needs torch")
    return _Enstrophy.apply(w0, a, prob)


def gaussian_field(xy, corr=0.1, amp=0.1, seed=5, n_feat=2000):
    """Random-Fourier-feature sample of a zero-mean Gaussian field with covariance
    amp^2 exp(-|r|^2 / (2 corr^2))."""
    rng = np.random.default_rng(seed)
    k = rng.standard_normal((n_feat, 2)) / corr
    ph = rng.uniform(0.0, 2 * np.pi, n_feat)
    return amp * np.sqrt(2.0 / n_feat) * np.cos(xy @ k.T + ph).sum(axis=1)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--K", type=int, default=5)
    ap.add_argument("--Re", type=float, default=1000.0)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--modes", type=int, default=32)
    ap.add_argument("--dirs", type=int, default=16)
    ap.add_argument("--eps", type=float, default=1e-6)
    ap.add_argument("--tol", type=float, default=1e-13)
    args = ap.parse_args(argv)
    t0 = time.perf_counter()
    prob = CavityEnstrophy(args.n, args.K, args.Re, None, args.steps, args.every, args.tol, args.modes)
    w0 = gaussian_field(prob.g.solution_points())
    a0 = 0.05 * np.random.default_rng(5).standard_normal(args.modes) / np.arange(1, args.modes + 1)
    log = lambda *m: print(*m, file=sys.stderr, flush=True)
    log(f"setup {time.perf_counter() - t0:.1f} s, D = {prob.g.D}, dt = {prob.dt:.4g}")
    wt = torch.tensor(w0, requires_grad=True)
    at = torch.tensor(a0, requires_grad=True)
    t1 = time.perf_counter()
    J = enstrophy(wt, at, prob)
    t2 = time.perf_counter()
    J.backward()
    t3 = time.perf_counter()
    gw, ga = wt.grad.numpy(), at.grad.numpy()
    log(f"J = {float(J.detach()):.12g}; forward {t2 - t1:.1f} s, adjoint {t3 - t2:.1f} s")
    rows = []
    for i in range(args.dirs):
        r = np.random.default_rng(50 + i)
        dw, da = r.standard_normal(prob.g.D), r.standard_normal(args.modes)
        Jp = prob.value(w0 + args.eps * dw, a0 + args.eps * da)
        Jm = prob.value(w0 - args.eps * dw, a0 - args.eps * da)
        fd = (Jp - Jm) / (2 * args.eps)
        ad = float(gw @ dw + ga @ da)
        rows.append({"seed": 50 + i, "adjoint": ad, "central_difference": fd, "rel_err": abs(ad - fd) / abs(fd)})
        log(f"direction {i}: adjoint {ad:.12g} central difference {fd:.12g} rel {rows[-1]['rel_err']:.2e}")
    out = {"config": {"workload": "lid-driven cavity adjoint", "n": args.n, "K": args.K, "Re": args.Re,
                      "steps": args.steps, "checkpoint_every": args.every, "dt": prob.dt, "lid_modes": args.modes,
                      "tol": args.tol, "eps": args.eps, "dofs": prob.g.D},
           "J": float(J.detach()), "forward_s": t2 - t1, "adjoint_s": t3 - t2,
           "max_rel_err": max(r["rel_err"] for r in rows) if rows else None,
           "median_rel_err": float(np.median([r["rel_err"] for r in rows])) if rows else None,
           "directions": rows}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
