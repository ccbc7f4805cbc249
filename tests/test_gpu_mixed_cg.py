"""Mixed-precision Jacobi PCG (nkh_params.precision = 2 for the non-pressure
hmholtz/cggo): the iteration in FP32 with FP64 inner products, FP64 solution
and FP64 residual refreshes.  Checked through the C ABI:

  * the returned solution's residual, recomputed with the oracle's FP64
    operator (axhelm, dssum, mask), meets the tolerance in cggo's own norm
    sqrt(sum binv*mult*r^2 / vol), and the reported norm is that residual;
  * its distance to the oracle's tightly converged solution is within 3x
    the FP64 GPU solve's at the same tolerance (the FP32 iteration costs
    accuracy only through the tolerance it stops at);
  * iteration counts stay within 10 % (+3) of the FP64 solve's;
  * maxit reached: the correction is folded into u, niter = maxit;
  * batched solves with precision 2 equal the single solves bit for bit;
  * a non-uniform h1 is refused.
"""
import sys

import numpy as np
import pytest

from problems import Problem, box

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def host(t):
    return t.detach().cpu().numpy()


def true_rbn2(P, u, f, h1, h2):
    """cggo's norm of f - mask * dssum(A u), in the oracle's FP64."""
    au = P.field.axhelm(u, np.full(P.shape, h1), np.full(P.shape, h2))
    r = f - P.mask * P.field.dssum(au)
    return float(np.sqrt(np.sum(P.field.binv * P.mult * r * r) / P.field.vol))


CASES = [
    dict(nel=(4, 3, 3), lx1=8, deform=0.1, h1=1.0, h2=0.0, tli=-1e-8),
    dict(nel=(3, 3, 3), lx1=6, deform=0.1, h1=0.7, h2=20.0, tli=-1e-10),
    dict(nel=(4, 4, 2), lx1=5, jitter=0.05, seed=2, h1=1e-3, h2=1833.3, tli=-1e-8),
    dict(nel=(4, 2, 2), lx1=7, periodic_x=True, h1=1e-2, h2=0.0, tli=-1e-9),
]


@pytest.mark.parametrize("case", CASES)
def test_mixed_jacobi_pcg(gpu, case):
    case = dict(case)
    h1v, h2v, tli = case.pop("h1"), case.pop("h2"), case.pop("tli")
    b = box(**case)
    P = Problem(b)
    F = gpu.box_field(b)
    rhs = P.bm1 * P.uniform_field(7)
    prm64 = gpu.default_params(istep=20, param22=0.0, restol=0.0)
    prm32 = gpu.default_params(istep=20, param22=0.0, restol=0.0, precision=2)
    h1, h2 = F.zeros() + h1v, F.zeros() + h2v
    outs = {}
    for tag, prm in (("fp64", prm64), ("mixed", prm32)):
        say(f"[mixed cg] {case} h1={h1v} h2={h2v} tli={tli} {tag}")
        u = F.zeros()
        r = F.to_dev(rhs)
        out = F.hmholtz("VELX", u, r, h1, h2, F.mask, F.mult, tli, 900, params=prm)
        outs[tag] = (host(u), out, host(r))
    u64, o64, f = outs["fp64"]
    u32, o32, _ = outs["mixed"]
    say(f"[mixed cg] iterations fp64 {o64['niter']} mixed {o32['niter']}, "
        f"rbn2 {o64['rbn2']:.3e} / {o32['rbn2']:.3e}, tol {o32['tol']:.3e}")
    assert o64["converged"] and o32["converged"]
    assert o32["niter"] <= int(1.1 * o64["niter"]) + 3
    assert abs(o32["tol"] - o64["tol"]) <= 1e-6 * o64["tol"]
    rt = true_rbn2(P, u32, f, h1v, h2v)
    assert rt <= o32["tol"] * (1 + 1e-6)
    assert abs(rt - o32["rbn2"]) <= 1e-6 * rt + 1e-3 * o32["tol"]
    # the oracle's solution converged to 1e-13 as the reference point
    uo, _, _, _, _ = P.field.hmholtz(rhs, np.full(P.shape, h1v), np.full(P.shape, h2v), -1e-13,
                                     900)
    e64 = np.linalg.norm(u64 - uo) / np.linalg.norm(uo)
    e32 = np.linalg.norm(u32 - uo) / np.linalg.norm(uo)
    say(f"[mixed cg] error vs tight oracle: fp64 {e64:.3e} mixed {e32:.3e}")
    assert e32 <= 3.0 * e64 + 1e-12
    F.destroy()


def test_mixed_maxit_folds_correction(gpu):
    b = box((3, 3, 3), 7, deform=0.1)
    P = Problem(b)
    F = gpu.box_field(b)
    rhs = P.bm1 * P.uniform_field(3)
    prm = gpu.default_params(istep=20, param22=0.0, restol=0.0, precision=2)
    u = F.zeros()
    r = F.to_dev(rhs)
    out = F.hmholtz("VELX", u, r, F.zeros() + 1.0, F.zeros(), F.mask, F.mult, -1e-12, 5,
                    params=prm)
    say(f"[mixed cg maxit] {out}")
    assert out["niter"] == 5 and not out["converged"]
    un = host(u)
    assert np.linalg.norm(un) > 0
    rt = true_rbn2(P, un, host(r), 1.0, 0.0)
    assert abs(rt - out["rbn2"]) <= 1e-6 * rt
    assert out["rbn2"] < out["rbn0"]
    F.destroy()


def test_mixed_batch_equals_single(gpu):
    bA = box((4, 3, 3), 6, deform=0.1)
    bB = box((3, 3, 2), 8, jitter=0.05, seed=3)
    PA, PB = Problem(bA), Problem(bB)
    FA, FB = gpu.box_field(bA), gpu.box_field(bB)
    prm = gpu.default_params(istep=20, param22=0.0, restol=0.0, precision=2)
    cases = [(FA, PA, 1.0, 0.0, 21, -1e-9), (FA, PA, 1e-3, 1833.3, 22, -1e-8),
             (FB, PB, 0.5, 2.0, 23, -1e-10)]
    solves, ref = [], []
    for F, P, h1v, h2v, seed, tli in cases:
        b = F.to_dev(P.bm1 * P.uniform_field(seed))
        h1, h2 = F.zeros() + h1v, F.zeros() + h2v
        u = F.zeros()
        out = F.hmholtz("VELX", u, b.clone(), h1, h2, F.mask, F.mult, tli, 900, params=prm)
        ref.append((host(u), out))
        solves.append(dict(u=F.zeros(), rhs=b.clone(), h1=h1, h2=h2, mask=F.mask, mult=F.mult,
                           tli=tli))
    B = gpu.Batch([c[0] for c in cases])
    for rep in range(2):
        for q in solves:
            q["u"].zero_()
        outs, launches = B.hmholtz([dict(q, rhs=q["rhs"].clone()) for q in solves], 900,
                                   params=prm)
        say(f"[mixed batch] rep {rep}: launches {launches}, iterations "
            f"{[o['niter'] for o in outs]} vs single {[r[1]['niter'] for r in ref]}")
        for k, (q, o) in enumerate(zip(solves, outs)):
            assert o["converged"] and o["niter"] == ref[k][1]["niter"], k
            assert o["rbn2"] == ref[k][1]["rbn2"]
            assert np.array_equal(host(q["u"]), ref[k][0]), k
    B.destroy()
    FA.destroy()
    FB.destroy()


def test_mixed_needs_uniform_coefficients(gpu):
    b = box(2, 5)
    P = Problem(b)
    F = gpu.box_field(b)
    h1 = F.to_dev(1.0 + 0.1 * P.uniform_field(1))
    prm = gpu.default_params(istep=20, param22=0.0, restol=0.0, precision=2)
    with pytest.raises(gpu.NekAmdError, match="uniform"):
        F.hmholtz("VELX", F.zeros(), F.to_dev(P.bm1 * P.uniform_field(2)), h1, F.zeros(),
                  F.mask, F.mult, -1e-8, 100, params=prm)
    F.destroy()
