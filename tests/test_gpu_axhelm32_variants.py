"""The single-precision operator of the mixed solvers with its r/s
contractions on the FP32 matrix cores (nkh_axhelm32_set_variant(lx1, 1)),
against the vector kernel and the oracle, through the C ABI:

  * mixed Jacobi PCG (precision = 2): the returned solution's residual,
    recomputed with the oracle's FP64 operator, meets the tolerance with
    either kernel, the iteration counts stay within 10 % (+3) of each other
    and the solutions within 1e-6;
  * mixed flexible CG + h1mg (PRES, precision = 2): the true residual meets
    the tolerance and the solution is within 1e-7 of the oracle's;
  * the variant is checked (0/1; 1 only for lx1 = 8..16).
"""
import sys

import numpy as np
import pytest

from problems import Problem, box, mg_for

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def host(t):
    return t.detach().cpu().numpy()


def l2rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def true_rbn2(P, u, f, h1, h2):
    au = P.field.axhelm(u, np.full(P.shape, h1), np.full(P.shape, h2))
    r = f - P.mask * P.field.dssum(au)
    return float(np.sqrt(np.sum(P.field.binv * P.mult * r * r) / P.field.vol))


@pytest.mark.parametrize("lx1,nel", [(8, (4, 3, 3)), (12, (3, 3, 2)), (16, (3, 2, 2))])
def test_mixed_jacobi_pcg_matrix_cores(gpu, lx1, nel):
    b = box(nel, lx1, deform=0.1)
    P = Problem(b)
    F = gpu.box_field(b)
    h1v, h2v, tli = 0.7, 20.0, -1e-9
    rhs = P.bm1 * P.uniform_field(17)
    prm = gpu.default_params(istep=20, param22=0.0, restol=0.0, precision=2)
    h1, h2 = F.zeros() + h1v, F.zeros() + h2v
    saved = gpu.axhelm32_variant(lx1)
    outs = []
    try:
        for v in (0, 1):
            gpu.set_axhelm32_variant(lx1, v)
            say(f"[axhelm32 variants] lx1={lx1} variant={v}")
            u = F.zeros()
            r = F.to_dev(rhs)
            out = F.hmholtz("VELX", u, r, h1, h2, F.mask, F.mult, tli, 900, params=prm)
            outs.append((host(u), out, host(r)))
    finally:
        gpu.set_axhelm32_variant(lx1, saved)
    (u0, o0, f), (u1, o1, _) = outs
    say(f"[axhelm32 variants] iterations {o0['niter']} / {o1['niter']}")
    for u, o in ((u0, o0), (u1, o1)):
        assert o["converged"]
        assert true_rbn2(P, u, f, h1v, h2v) <= o["tol"] * (1 + 1e-6)
    assert o1["niter"] <= int(1.1 * o0["niter"]) + 3
    assert o0["niter"] <= int(1.1 * o1["niter"]) + 3
    assert l2rel(u1, u0) < 1e-6
    F.destroy()


def test_mixed_flex_cg_matrix_cores(gpu):
    lx1 = 9
    b = box((3, 3, 3), lx1, deform=0.1)
    P = Problem(b)
    mg = mg_for(P)
    F = gpu.box_field(b)
    x, y, z = b.xyz()
    F.h1mg_setup(b.fast_bc(), b.vertex, x, y, z)
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, 10.0)
    bb = P.bm1 * P.uniform_field(79)
    rhs = P.field.dssum(bb) * P.mask
    tol = 1e-10
    uo, ito, _, div0 = mg.flex_cg(rhs, h1, h2, 300, -tol)
    saved = gpu.axhelm32_variant(lx1)
    try:
        gpu.set_axhelm32_variant(lx1, 1)
        u = F.zeros()
        out = F.hmholtz("PRES", u, F.to_dev(bb), F.to_dev(h1), F.to_dev(h2), F.mask, F.mult,
                        -tol, 300, params=gpu.default_params(param42=2, param21=-tol,
                                                             precision=2))
    finally:
        gpu.set_axhelm32_variant(lx1, saved)
    ug = host(u)
    r = rhs - P.field.dssum(P.field.axhelm(ug, h1, h2)) * P.mask
    rn = float(np.sqrt(np.sum(r * r * P.mult) / P.field.vol))
    say(f"[axhelm32 variants] flex-CG {out['niter']} iterations (oracle {ito}), "
        f"true residual {rn / div0:.2e}")
    assert out["converged"]
    assert out["niter"] <= ito + 8
    assert rn < tol * div0 * 1.0001
    assert l2rel(ug, uo) < 1e-7
    F.destroy()


def test_axhelm32_variant_checked(gpu):
    with pytest.raises(gpu.NekAmdError):
        gpu.set_axhelm32_variant(7, 1)
    with pytest.raises(gpu.NekAmdError):
        gpu.set_axhelm32_variant(8, 2)
    with pytest.raises(gpu.NekAmdError):
        gpu.set_axhelm32_variant(17, 0)
