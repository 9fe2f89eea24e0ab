"""The adjoint gradient check at its stated size: a 128 x 128 lid-driven cavity, K=5 (D = 409 600),
Re=1000, dt = half the RK4 diffusion limit, 500 RK4 steps with checkpoints every 50; J = the final
enstrophy; inputs omega_0 (a Gaussian random field, correlation length 0.1, amplitude 0.1,
seed 5) and 32 Fourier modes of the lid speed; the gradient (hoc_adjoint_forward/backward behind
the torch harness) against central differences (step 1e-6) on 16 seeded random directions
(seeds 50-65) in (omega_0, a), each within 1e-6.  About three minutes on one MI355X; each
direction's result is written to the terminal as it is known (past pytest's capture), so a
long run keeps reporting."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_adjoint_gradient_cavity128_K5_500_steps(gpu_available, capfd):
    torch = pytest.importorskip("torch")

    def report(*a):
        with capfd.disabled():
            print("[adjoint 128x128 K=5]", *a, flush=True)

    from hoc.gradient import CavityEnstrophy, enstrophy, gaussian_field
    n_modes, eps = 32, 1e-6
    prob = CavityEnstrophy(n=128, K=5, Re=1000.0, dt=None, nsteps=500, every=50, tol=1e-13, n_modes=n_modes)
    assert prob.g.D == 409600
    w0 = gaussian_field(prob.g.solution_points(), corr=0.1, amp=0.1, seed=5)
    a0 = 0.05 * np.random.default_rng(5).standard_normal(n_modes) / np.arange(1, n_modes + 1)
    wt = torch.tensor(w0, requires_grad=True)
    at = torch.tensor(a0, requires_grad=True)
    J = enstrophy(wt, at, prob)
    J.backward()
    gw, ga = wt.grad.numpy(), at.grad.numpy()
    report("J", float(J.detach()), "dt", prob.dt)
    errs = []
    for i in range(16):
        r = np.random.default_rng(50 + i)
        dw, da = r.standard_normal(prob.g.D), r.standard_normal(n_modes)
        fd = (prob.value(w0 + eps * dw, a0 + eps * da) - prob.value(w0 - eps * dw, a0 - eps * da)) / (2 * eps)
        ad = float(gw @ dw + ga @ da)
        errs.append(abs(ad - fd) / abs(fd))
        report("direction", 50 + i, "adjoint", ad, "central difference", fd, "rel", errs[-1])
    assert max(errs) <= 1e-6, errs
