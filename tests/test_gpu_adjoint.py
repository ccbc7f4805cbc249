"""GPU adjoint coefficient gradient through the C ABI:

  * the kernel against the oracle's gradient on the same u, lam (FP64,
    within 1e-12 of the largest entry) at several polynomial orders;
  * dJ/dh1 for J = 1/2 (u-u*)^T B (u-u*) through a torch.autograd.Function
    (test harness only) whose forward is the GPU Helmholtz solve and whose
    backward is the adjoint solve plus the gradient kernel: against central
    finite differences of the GPU solve (1e-6 relative) and against the
    oracle's own forward/adjoint/gradient pipeline (1e-9 relative L2).

Solves use the Jacobi PCG to a 1e-13 relative residual.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from problems import Problem, box

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def host(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("lx1", [3, 5, 8, 12])
def test_gradient_kernel_matches_oracle(gpu, lx1):
    b = box((3, 2, 2), lx1, deform=0.15)
    P = Problem(b)
    F = gpu.box_field(b)
    rng = np.random.default_rng(lx1)
    u = rng.standard_normal(P.shape)
    lam = rng.standard_normal(P.shape)
    go1, go2 = P.field.adj_grad(u, lam)
    g1, g2 = F.zeros(), F.zeros()
    say(f"[adjoint kernel] lx1={lx1}")
    F.adjoint_coef_grad(g1, F.to_dev(u), F.to_dev(lam), g2)
    for a, o in ((host(g1), go1), (host(g2), go2)):
        assert np.max(np.abs(a - o)) <= 1e-12 * np.max(np.abs(o))
    g1b = F.zeros()
    F.adjoint_coef_grad(g1b, F.to_dev(u), F.to_dev(lam))  # without the h2 part
    assert np.array_equal(host(g1b), host(g1))
    F.destroy()


def _fourier_h1(P, nmodes=16, seed=3):
    """h1 = exp(sum of nmodes 0.5-weighted Fourier modes, wavenumbers 1..4)."""
    rng = np.random.default_rng(seed)
    x, y, z = P.x, P.y, P.zc
    s = np.zeros(P.shape)
    for _ in range(nmodes):
        k = rng.integers(1, 5, size=3)
        ph = rng.uniform(0, 2 * np.pi, size=3)
        s += np.cos(np.pi * k[0] * x + ph[0]) * np.cos(np.pi * k[1] * y + ph[1]) * \
            np.cos(np.pi * k[2] * z + ph[2])
    return np.exp(0.5 * s / np.sqrt(nmodes))


def test_autograd_gradient_vs_finite_differences_and_oracle(gpu):
    import torch

    b = box((4, 4, 4), 7, deform=0.1)
    P = Problem(b)
    F = gpu.box_field(b)
    prm = gpu.default_params(istep=20, param22=0.0, restol=0.0)
    h2 = F.zeros()
    # a smooth forcing keeps u O(1), so J and its derivative are well above
    # the solver tolerance and the finite differences resolve 1e-6
    f = P.bm1 * P.rhs_config1(noise=0.0)
    ustar = 0.5 * P.uniform_field(4) * P.mask
    fd_dev, ustar_dev, bm1_dev = F.to_dev(f), F.to_dev(ustar), F.to_dev(P.bm1)

    def solve(rhs, h1):
        u = F.zeros()
        r = rhs.clone()
        out = F.hmholtz("VELX", u, r, h1, h2, F.mask, F.mult, -1e-13, 3000, params=prm)
        assert out["converged"]
        return u

    class HelmholtzSolve(torch.autograd.Function):
        @staticmethod
        def forward(ctx, h1):
            u = solve(fd_dev, h1.detach().contiguous())
            ctx.save_for_backward(h1, u)  # the forward solution is the checkpoint
            return u

        @staticmethod
        def backward(ctx, grad_u):
            h1, u = ctx.saved_tensors
            lam = solve(grad_u.contiguous(), h1.detach().contiguous())
            g1 = F.zeros()
            F.adjoint_coef_grad(g1, u, lam)
            return g1

    def J_of(h1):
        u = HelmholtzSolve.apply(h1)
        d = u - ustar_dev
        return 0.5 * torch.sum(bm1_dev * d * d)

    h1np = _fourier_h1(P)
    h1 = F.to_dev(h1np).requires_grad_(True)
    say("[adjoint] forward + backward on 4^3 N=7")
    J = J_of(h1)
    J.backward()
    g = h1.grad.detach().clone()
    for seed in range(4):
        # random-sign and one-signed perturbations of the coefficient
        w = P.uniform_field(50 + seed)
        dh = F.to_dev((w if seed % 2 == 0 else 1.0 + 0.5 * w) * h1np)
        eps = 1e-5
        with torch.no_grad():
            jp = float(J_of(h1.detach() + eps * dh))
            jm = float(J_of(h1.detach() - eps * dh))
        fd = (jp - jm) / (2 * eps)
        ad = float(torch.sum(g * dh))
        say(f"[adjoint] direction {seed}: adjoint {ad:.10e} fd {fd:.10e}")
        assert abs(fd - ad) <= 1e-6 * abs(ad)
    # the oracle's pipeline: Jacobi PCG forward and adjoint, then its gradient
    po = O.Params(20, 0.0, 0.0, 0.0)
    uo, _, _, _, _ = P.field.hmholtz(f, h1np, np.zeros(P.shape), -1e-13, 3000, po)
    lo, _, _, _, _ = P.field.hmholtz(P.bm1 * (uo - ustar), h1np, np.zeros(P.shape), -1e-13, 3000,
                                    po)
    go, _ = P.field.adj_grad(uo, lo)
    e = float(np.linalg.norm(host(g) - go) / np.linalg.norm(go))
    say(f"[adjoint] gradient vs oracle rel L2 = {e:.3e}")
    assert e < 1e-9
    F.destroy()
