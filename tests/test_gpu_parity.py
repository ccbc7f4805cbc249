"""GPU parity: libnekamd (HIP, through the C ABI) against the oracle on the
same seeded inputs.

Tolerances (floating point, FP64): operator apply and geometry within 1e-13
relative (max-norm over the field); gather-scatter bit-exact (both sum the
members of a group in increasing local index); Jacobi-PCG iteration counts
within +-1 and solutions within 1e-10 relative L2 of the oracle's -- or, on
meshes where the oracle's own solution moves by more than that when its rhs
is perturbed at the 1e-16 level, within 4x that measured sensitivity.
"""
import sys

import numpy as np
import pytest

from problems import Problem, box

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup(gpu, b):
    say(f"[setup] nel={b.nel} lx1={b.lx1} deform={b.deform} jitter={b.jitter} "
        f"periodic={b.periodic_x}")
    P = Problem(b)
    F = gpu.box_field(b)
    return P, F


def host(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [
    dict(nel=2, lx1=2),
    dict(nel=2, lx1=4),
    dict(nel=(3, 2, 2), lx1=6, deform=0.1),
    dict(nel=2, lx1=8, deform=0.1),
    dict(nel=(2, 3, 2), lx1=5, jitter=0.05, seed=3),
    dict(nel=(4, 2, 2), lx1=8, periodic_x=True, jitter=0.05, seed=2),
    dict(nel=2, lx1=10, deform=0.1),
    dict(nel=2, lx1=12),
    dict(nel=2, lx1=16, deform=0.05),
]


@pytest.mark.parametrize("case", CASES)
def test_setup_geometry_masks(gpu, case):
    P, F = setup(gpu, box(**case))
    bm1, binv = F.mass()
    assert rel(host(bm1), P.bm1) < 1e-13
    assert rel(host(binv), P.field.binv) < 1e-13
    assert abs(F.info()["vol"] - P.field.vol) <= 1e-13 * P.field.vol
    assert np.array_equal(F.glo_num, P.glo)
    assert np.array_equal(host(F.mask), P.mask)
    assert np.array_equal(host(F.mult), P.mult)


@pytest.mark.parametrize("case", CASES)
def test_axhelm(gpu, case):
    P, F = setup(gpu, box(**case))
    rng = np.random.default_rng(5)
    u = rng.standard_normal(P.shape)
    h1f = 1.0 + rng.random(P.shape)
    h2f = rng.random(P.shape)
    one = np.ones(P.shape)
    zero = np.zeros(P.shape)
    for h1, h2, tag in ((one, zero, "h1=1,h2=0"), (one * 0.3, one * 10.0, "uniform h2"),
                        (h1f, zero, "field h1"), (h1f, h2f, "field h1,h2")):
        say(f"[axhelm] {case} {tag}")
        ref = P.field.axhelm(u, h1, h2)
        au = F.zeros()
        F.axhelm(au, F.to_dev(u), F.to_dev(h1), F.to_dev(h2))
        assert rel(host(au), ref) < 1e-13, tag


@pytest.mark.parametrize("case", CASES)
def test_gather_scatter_bit_exact(gpu, case):
    P, F = setup(gpu, box(**case))
    rng = np.random.default_rng(9)
    u = rng.standard_normal(P.shape)
    for op in ("+", "*", "m", "M"):
        say(f"[gs] {case} op={op}")
        ref = P.field.gs.op(np.ascontiguousarray(u.copy()), op)
        t = F.to_dev(u)
        if op == "+":
            F.dssum(t)
        else:
            F.dsop(t, op)
        assert np.array_equal(host(t), ref), op


@pytest.mark.parametrize("case", CASES[1:6])
def test_setprec(gpu, case):
    P, F = setup(gpu, box(**case))
    rng = np.random.default_rng(2)
    h1 = 0.5 + rng.random(P.shape)
    h2 = np.full(P.shape, 3.0)
    ref = P.field.setprec(h1, h2)
    d = F.zeros()
    F.setprec(d, F.to_dev(h1), F.to_dev(h2))
    assert rel(host(d), ref) < 1e-13


def run_both(P, F, rhs, h1, h2, tli, maxit=900):
    u_o, r_o, n_o, hist_o, rbn0_o = P.field.hmholtz(rhs, h1, h2, tli, maxit)
    u = F.zeros()
    r = F.to_dev(rhs)
    out = F.hmholtz("TEMP", u, r, F.to_dev(h1), F.to_dev(h2), F.mask, F.mult, tli, maxit,
                    history=True)
    return (u_o, r_o, n_o, hist_o, rbn0_o), (host(u), host(r), out)


def l2rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_hmholtz_config1(gpu):
    """Config 1: unit-cube Poisson, 8^3 elements, N=7, Jacobi-PCG to 1e-8,
    compared with the oracle's cggo iterate by iterate."""
    P, F = setup(gpu, box(8, 8))
    one = np.ones(P.shape)
    zero = np.zeros(P.shape)
    rhs = P.bm1 * P.rhs_config1(seed=1, noise=1e-3)
    (u_o, r_o, n_o, hist_o, rbn0_o), (u, r, out) = run_both(P, F, rhs, one, zero, -1e-8)
    say(f"[config1] oracle niter={n_o} gpu niter={out['niter']}")
    assert np.array_equal(r, r_o)  # dssum + mask of the rhs: bit-exact
    assert abs(out["niter"] - n_o) <= 1
    assert l2rel(u, u_o) < 1e-10
    h = out["history"]
    k = min(n_o, out["niter"])
    # per-iteration residual norms agree far below the convergence tolerance
    np.testing.assert_allclose(h[:k, 0], hist_o[:k, 0], rtol=1e-8)
    np.testing.assert_allclose(h[:k, 3], hist_o[:k, 3], rtol=1e-8)
    # a global sum of n_L terms: different (deterministic) orders, ~sqrt(n) eps
    assert abs(out["rbn0"] - rbn0_o) <= 1e-11 * rbn0_o


@pytest.mark.parametrize("case,h2v,h1var", [
    (dict(nel=(4, 3, 3), lx1=8, deform=0.1), 10.0, False),
    (dict(nel=(4, 4, 2), lx1=6, jitter=0.05, seed=7), 1.0, True),
    (dict(nel=(5, 3, 3), lx1=8, periodic_x=True), 0.0, False),
    (dict(nel=3, lx1=4), 0.5, True),
])
def test_hmholtz_variants(gpu, case, h2v, h1var):
    P, F = setup(gpu, box(**case))
    rhs = P.bm1 * P.uniform_field(11)
    h1 = np.ones(P.shape)
    if h1var:
        h1 = 1.0 + 0.5 * (P.uniform_field(12) + 1.0)
    h2 = np.full(P.shape, h2v)
    (u_o, r_o, n_o, hist_o, _), (u, r, out) = run_both(P, F, rhs, h1, h2, -1e-8)
    # the spread of FP64 round-off trajectories: the oracle's solve with the
    # rhs perturbed at the 1e-16 level, four times (CG amplifies last-bit
    # differences; the oracle's own OpenMP sums vary at that level too)
    rng = np.random.default_rng(0)
    sens = 0.0
    for _ in range(4):
        u_p = P.field.hmholtz(rhs * (1 + 1e-16 * rng.standard_normal(P.shape)), h1, h2, -1e-8,
                              900)[0]
        sens = max(sens, l2rel(u_p, u_o))
    say(f"[hmholtz] {case} h2={h2v} oracle={n_o} gpu={out['niter']} "
        f"diff={l2rel(u, u_o):.2e} oracle-sensitivity={sens:.2e}")
    assert abs(out["niter"] - n_o) <= 1
    assert l2rel(u, u_o) < max(1e-10, 4 * sens)


def test_hmholtz_absolute_tol_and_maxit(gpu):
    P, F = setup(gpu, box(3, 5))
    rhs = P.bm1 * P.uniform_field(4)
    one = np.ones(P.shape)
    zero = np.zeros(P.shape)
    (u_o, _, n_o, _, _), (u, _, out) = run_both(P, F, rhs, one, zero, 1e-6)
    assert abs(out["niter"] - n_o) <= 1 and l2rel(u, u_o) < 1e-9
    (u_o, _, n_o, _, _), (u, _, out) = run_both(P, F, rhs, one, zero, -1e-14, maxit=7)
    assert n_o == 7 and out["niter"] == 7 and l2rel(u, u_o) < 1e-12
    (u_o, _, n_o, _, _), (u, _, out) = run_both(P, F, rhs * 0, one, zero, -1e-8)
    assert n_o == 0 and out["niter"] == 0 and np.all(u == 0)


def test_large_mesh_residual(gpu):
    """Size-independent property at a larger mesh (24^3 elements, N=7,
    deformed, h2=10): the true residual of the returned solution, recomputed
    from scratch with the operators, is below the relative tolerance."""
    b = box(24, 8, deform=0.1)
    F = gpu.box_field(b)
    import torch

    say("[large] setup done")
    rng = np.random.default_rng(1)
    rhs0 = torch.as_tensor(rng.uniform(-1, 1, F.shape), device="cuda")
    bm1, binv = F.mass()
    rhs = bm1 * rhs0
    F.dssum(rhs)
    rhs = rhs * F.mult  # continuous, then local weighting again
    rhs = rhs * bm1
    r = rhs.clone()
    u = F.zeros()
    h1 = torch.ones_like(u)
    h2 = torch.full_like(u, 10.0)
    out = F.hmholtz("TEMP", u, r, h1, h2, F.mask, F.mult, -1e-8, 900)
    au = F.zeros()
    F.axhelm(au, u, h1, h2)
    F.dssum(au)
    res = r - au * F.mask
    rn = float(torch.sqrt(torch.sum(binv * F.mult * res * res) / F.info()["vol"]))
    say(f"[large] niter={out['niter']} rn/rbn0={rn / out['rbn0']:.3e}")
    assert out["converged"] == 1
    assert rn <= 1.0001e-8 * out["rbn0"]


def test_cggo_unfused_path(gpu):
    """cggo on an rhs that is neither assembled nor masked: the fused
    iteration's preconditions fail, so the unfused sequence (axhelm, dssum,
    mask, weighted dot) runs; it must still match the oracle."""
    P, F = setup(gpu, box((3, 3, 2), 6, deform=0.1))
    f = P.bm1 * P.uniform_field(21)
    f = f * (1.0 + 0.1 * np.cos(np.arange(f.size)).reshape(f.shape))  # discontinuous
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, 2.0)
    x_o, n_o, hist_o, _ = P.field.cggo(f, h1, h2, -1e-8, 900)
    x = F.zeros()
    out = F.cggo(x, F.to_dev(f), F.to_dev(h1), F.to_dev(h2), F.mask, F.mult, -1e-8, 900)
    say(f"[cggo-unfused] oracle={n_o} gpu={out['niter']}")
    assert abs(out["niter"] - n_o) <= 1
    assert l2rel(host(x), x_o) < 1e-9


def test_undeformed_elements_drop_cross_terms(gpu):
    """param(59) = 0: elements flagged undeformed (genxyz.f setdef) keep the
    g4m1..g6m1 that sfastax (hmholtz.f:353-364) wrote -- g1/w_x, g2/w_y,
    g3/w_z -- and the reference's axhelm and setprec skip their cross terms.
    On a box of rectangular elements the true cross terms are zero, so the
    apply and the Jacobi diagonal equal the oracle's on the true factors;
    half the elements are flagged (the others carry the true factors)."""
    b = box((3, 2, 2), 8)
    P = Problem(b)
    F = gpu.Field(b.lx1, b.nel)
    w = P.w
    flags = np.arange(b.nel) % 2  # 0 = undeformed
    g = [a.copy() for a in P.g]
    und = flags == 0
    g[3][und] = (P.g[0] / w[None, None, None, :])[und]
    g[4][und] = (P.g[1] / w[None, None, :, None])[und]
    g[5][und] = (P.g[2] / w[None, :, None, None])[und]
    F.set_geom(g, P.bm1)
    F.set_deformed(flags)
    F.gs_setup(P.glo)
    rng = np.random.default_rng(4)
    u = rng.standard_normal(P.shape)
    h1 = 0.5 + rng.random(P.shape)
    h2 = np.full(P.shape, 2.0)
    say("[undeformed] axhelm and setprec with sfastax factors on flagged elements")
    au = F.zeros()
    F.axhelm(au, F.to_dev(u), F.to_dev(h1), F.to_dev(h2))
    assert rel(host(au), P.field.axhelm(u, h1, h2)) < 1e-13
    d = F.zeros()
    F.setprec(d, F.to_dev(h1), F.to_dev(h2))
    assert rel(host(d), P.field.setprec(h1, h2)) < 1e-13
    F.destroy()
