"""Polynomial-order sweep at the high orders: one operator apply plus 100
fixed Jacobi-PCG iterations (hmholtz -> cggo, hmholtz.f:2-69, :611-846) at
N = 9, 11 and 15 on jittered boxes (vertex jitter U(-0.05h, 0.05h) from a
seed, h1 = h2 = 1, f ~ U(-1,1) per unique node), with the FP64 operator and
with the FP32 one (mixed precision), against the oracle on the same inputs.

Bars:
  * FP64: 100 iterations each; the per-iteration residual norms within
    rtol max(1e-8, 4x the oracle's own spread when its rhs is perturbed at
    the 1e-16 level), the solution within max(1e-10, 4x that spread);
  * FP32 operator: 100 iterations; the reported norm is the true FP64
    residual of the returned solution (recomputed with the oracle's
    axhelm, dssum and mask) to 1e-6, and that residual is within 5x the
    FP64 solve's after the same 100 iterations (or below 1e-5 of the
    initial norm, where single precision stops tracking FP64).
"""
import sys

import numpy as np
import pytest

from problems import Problem, box

pytestmark = pytest.mark.gpu

CASES = [
    dict(nel=(3, 3, 2), lx1=10, jitter=0.05, seed=9),
    dict(nel=(3, 2, 2), lx1=12, jitter=0.05, seed=11),
    dict(nel=(2, 2, 2), lx1=16, jitter=0.05, seed=15),
]
ITERS = 100


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def host(t):
    return t.detach().cpu().numpy()


def l2rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def true_rbn2(P, u, f, h1, h2):
    """cggo's norm of f - mask * dssum(A u) in the oracle's FP64."""
    au = P.field.axhelm(u, h1, h2)
    r = f - P.mask * P.field.dssum(au)
    return float(np.sqrt(np.sum(P.field.binv * P.mult * r * r) / P.field.vol))


@pytest.mark.parametrize("case", CASES)
def test_order_sweep_fixed_pcg(gpu, case):
    b = box(**case)
    say(f"[order sweep] {case}")
    P = Problem(b)
    F = gpu.box_field(b)
    one = np.ones(P.shape)
    rhs = P.bm1 * P.uniform_field(case["lx1"])
    # one apply
    u0 = P.uniform_field(case["lx1"] + 1)
    au = F.zeros()
    F.axhelm(au, F.to_dev(u0), F.to_dev(one), F.to_dev(one))
    ref = P.field.axhelm(u0, one, one)
    assert float(np.max(np.abs(host(au) - ref)) / np.max(np.abs(ref))) < 1e-13

    # 100 fixed iterations: a relative tolerance no solve reaches
    tli = -1e-30
    u_o, r_o, n_o, hist_o, rbn0_o = P.field.hmholtz(rhs, one, one, tli, ITERS)
    rng = np.random.default_rng(0)
    spread_u, spread_h = 0.0, 0.0
    for _ in range(4):
        u_p, _, n_p, hist_p, _ = P.field.hmholtz(
            rhs * (1 + 1e-16 * rng.standard_normal(P.shape)), one, one, tli, ITERS)
        spread_u = max(spread_u, l2rel(u_p, u_o))
        k = min(n_p, n_o)
        spread_h = max(spread_h, float(np.max(np.abs(hist_p[:k, 0] / hist_o[:k, 0] - 1))))
    prm64 = gpu.default_params()  # the oracle's defaults: istep 1, param(22) 1e-7
    u = F.zeros()
    r = F.to_dev(rhs)
    out = F.hmholtz("TEMP", u, r, F.to_dev(one), F.to_dev(one), F.mask, F.mult, tli, ITERS,
                    params=prm64, history=True)
    h = out["history"]
    say(f"[order sweep] lx1={case['lx1']} fp64: oracle {n_o} gpu {out['niter']} it, rbn2 "
        f"{hist_o[n_o - 1, 0]:.3e} / {out['rbn2']:.3e}, diff {l2rel(host(u), u_o):.2e}, "
        f"oracle spread u {spread_u:.2e} history {spread_h:.2e}")
    assert n_o == ITERS and out["niter"] == ITERS
    np.testing.assert_allclose(h[:ITERS, 0], hist_o[:ITERS, 0], rtol=max(1e-8, 4 * spread_h))
    assert l2rel(host(u), u_o) < max(1e-10, 4 * spread_u)
    f_asm = host(r)  # dssum'd and masked rhs, as cggo saw it
    rt64 = true_rbn2(P, host(u), f_asm, one, one)

    # the same 100 iterations with the FP32 operator
    prm32 = gpu.default_params(precision=2)
    u32 = F.zeros()
    o32 = F.hmholtz("TEMP", u32, F.to_dev(rhs), F.zeros() + 1.0, F.zeros() + 1.0, F.mask,
                    F.mult, tli, ITERS, params=prm32)
    rt32 = true_rbn2(P, host(u32), f_asm, one, one)
    say(f"[order sweep] lx1={case['lx1']} fp32 operator: {o32['niter']} it, reported "
        f"{o32['rbn2']:.3e}, true {rt32:.3e} (fp64 true {rt64:.3e}, rbn0 {o32['rbn0']:.3e})")
    assert o32["niter"] == ITERS
    assert abs(rt32 - o32["rbn2"]) <= 1e-6 * rt32
    assert rt32 <= max(5.0 * rt64, 1e-5 * o32["rbn0"])
    F.destroy()
