"""GPU parity of the H1 multigrid (hsmg.f h1mg_setup / h1mg_solve) and the
flexible CG around it (hmholtz.f:2164 hmh_flex_cg), through the C ABI,
against the oracle on the same seeded inputs.

Tolerances (FP64): the hierarchy's integer facts (levels, points per level,
coarse unknowns) exactly; one V-cycle z = M r within 1e-10 relative L2 (the
FDM eigenvectors come from the same Cholesky + Jacobi algorithm compiled by
a different compiler, and above 600 unknowns the coarse problem is solved by
AMG-preconditioned CG to a 1e-13 relative residual instead of a direct
factor); flexible-CG iteration counts
within +-1, residual-norm histories within 1e-6 relative over the common
iterations (plus 1e-15 of the initial residual: below that the last digits
are round-off), solutions within 1e-9 relative L2.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from problems import Problem, box, mask_from_fbc, mg_for

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def host(t):
    return t.detach().cpu().numpy()


def l2rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


MG_CASES = [
    dict(nel=3, lx1=8, deform=0.1),
    dict(nel=(3, 2, 4), lx1=5, jitter=0.05, seed=4),
    dict(nel=3, lx1=4, deform=0.05),
    dict(nel=(4, 3, 3), lx1=7, periodic_x=True, jitter=0.04, seed=7),
    dict(nel=(2, 3, 3), lx1=10, deform=0.1),
    dict(nel=(3, 3, 2), lx1=6, neumann_top=True),
    # more than 600 coarse unknowns: the coarse problem goes through the
    # AMG-preconditioned CG instead of the dense inverse
    dict(nel=(12, 11, 10), lx1=4, deform=0.05),
]


def build(gpu, case):
    case = dict(case)
    neumann_top = case.pop("neumann_top", False)
    neumann_all = case.pop("neumann_all", False)
    b = box(**case)
    fbc = b.fast_bc()
    if neumann_top:
        top = (np.arange(b.nel) // (b.nelx * b.nely)) == b.nelz - 1
        fbc[top, 5] = 2
    if neumann_all:  # no Dirichlet face: the constant null space
        fbc[fbc == 1] = 2
    custom = neumann_top or neumann_all
    say(f"[hsmg setup] {case} neumann_top={neumann_top} neumann_all={neumann_all}")
    P = Problem(b, fbc=fbc if custom else None)
    mg = mg_for(P)
    F = gpu.box_field(b)
    if custom:
        m = F.to_dev(mask_from_fbc(fbc, b.lx1))
        F.dsop(m, "*")
        F.mask = m
    assert np.array_equal(host(F.mask), P.mask)
    x, y, z = b.xyz()
    F.h1mg_setup(fbc, b.vertex, x, y, z)
    return P, mg, F


@pytest.mark.parametrize("case", MG_CASES)
def test_hierarchy_facts(gpu, case):
    P, mg, F = build(gpu, case)
    info = F.h1mg_info()
    assert info["lmax"] == mg.lmax
    assert info["nh"] == mg.nh
    assert info["ncrs"] == O.lib().nko_mg_ncrs(mg.h)
    F.destroy()


@pytest.mark.parametrize("case", MG_CASES)
def test_h1mg_solve(gpu, case):
    P, mg, F = build(gpu, case)
    rng = np.random.default_rng(11)
    r = P.field.dssum(P.bm1 * rng.standard_normal(P.shape))
    zo = mg.solve(r)
    rd = F.to_dev(r)
    zd = F.zeros()
    say(f"[h1mg_solve] {case}")
    F.h1mg_solve(zd, rd)
    zg = host(zd)
    # the input is masked in place at the fine level (hsmg.f h1mg_mask)
    np.testing.assert_array_equal(host(rd), r * mg.mask(mg.lmax))
    e = l2rel(zg, zo)
    say(f"[h1mg_solve] rel L2 = {e:.3e}")
    assert e < 1e-10
    # z is continuous (dsavg) and zero where the fine mask is
    zz = F.zeros()
    zz.copy_(zd)
    F.dsop(zz, "M")
    assert np.max(np.abs(host(zz) - zg)) <= 1e-14 * np.max(np.abs(zg))
    F.destroy()


@pytest.mark.parametrize("case", MG_CASES)
@pytest.mark.parametrize("h2v", [0.0, 10.0])
def test_flex_cg(gpu, case, h2v):
    P, mg, F = build(gpu, case)
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, h2v)
    b = P.bm1 * P.uniform_field(21)
    rhs = P.field.dssum(b) * P.mask
    uo, ito, histo, div0 = mg.flex_cg(rhs, h1, h2, 200, -1e-10)
    say(f"[flex_cg] {case} h2={h2v}: oracle iterations {ito}")
    # through hmholtz('PRES', ...) with param(42)=2 (hmholtz.f:646-653)
    u = F.zeros()
    rd = F.to_dev(b)
    prm = gpu.default_params(param42=2, param21=-1e-10)
    out = F.hmholtz("PRES", u, rd, F.to_dev(h1), F.to_dev(h2), F.mask, F.mult, -1e-10, 200,
                    params=prm, history=True)
    say(f"[flex_cg] gpu iterations {out['niter']} rnorm {out['rbn2']:.3e}")
    np.testing.assert_allclose(host(rd), rhs, rtol=0, atol=1e-14 * np.max(np.abs(rhs)))
    assert abs(out["niter"] - ito) <= 1
    k = min(out["niter"], ito)
    np.testing.assert_allclose(out["history"][:k], histo[:k], rtol=1e-6,
                               atol=1e-15 * histo[0])
    assert abs(out["rbn0"] - div0) <= 1e-12 * div0
    e = l2rel(host(u), uo)
    say(f"[flex_cg] solution rel L2 = {e:.3e}")
    assert e < 1e-9
    F.destroy()


def test_flex_cg_entry_and_tolerances(gpu):
    """nkd_hmh_flex_cg directly: istep=0 forces tolps=1e-4 (absolute), a
    positive param21 caps the chktcg1 tolerance, maxit bounds the count."""
    P, mg, F = build(gpu, dict(nel=3, lx1=6, deform=0.1))
    h1 = np.ones(P.shape)
    h2 = np.zeros(P.shape)
    rhs = P.field.dssum(P.bm1 * P.uniform_field(3)) * P.mask
    for param21, istep, maxit in ((0.0, 0, 100), (1e-6, 1, 100), (-1e-12, 1, 3)):
        uo, ito, histo, div0 = mg.flex_cg(rhs, h1, h2, maxit, param21, istep)
        res = F.to_dev(rhs)
        prm = gpu.default_params(istep=istep, param42=2, param21=param21)
        out = F.hmh_flex_cg(res, F.to_dev(h1), F.to_dev(h2), F.mult, maxit, params=prm,
                            history=True)
        say(f"[flex_cg entry] param21={param21} istep={istep}: {out['niter']} vs {ito}")
        assert abs(out["niter"] - ito) <= 1
        if maxit == 3:
            assert out["niter"] == 3 and not out["converged"]
        else:
            assert out["converged"]
            assert out["rnorm"] < out["tolpss"]
        assert l2rel(host(res), uo) < 1e-8
    F.destroy()


def test_pres_other_solvers_unsupported(gpu):
    b = box(2, 6)
    F = gpu.box_field(b)
    u = F.zeros()
    r = F.zeros() + 1.0
    one = F.zeros() + 1.0
    with pytest.raises(gpu._lib.NekAmdError):  # param(42) = 1: FDM-Jacobi PCG
        F.hmholtz("PRES", u, r, one, F.zeros(), F.mask, F.mult, -1e-8, 10,
                  params=gpu.default_params(param42=1))
    with pytest.raises(gpu._lib.NekAmdError):  # no hierarchy yet
        F.hmholtz("PRES", u, r, one, F.zeros(), F.mask, F.mult, -1e-8, 10,
                  params=gpu.default_params(param42=2))
    F.destroy()


def test_flex_cg_pmask_differs_from_mg_mask(gpu):
    """pmask (all boundary faces Dirichlet) differs from the multigrid's fine
    mask (top faces Neumann, code 2): r is masked again before every V-cycle
    and the denominator is taken after the assembly (the unfused path)."""
    b = box((3, 3, 2), 6, deform=0.1)
    fbc_mg = b.fast_bc()
    top = (np.arange(b.nel) // (b.nelx * b.nely)) == b.nelz - 1
    fbc_mg[top, 5] = 2
    P = Problem(b)
    mg = O.MG(P.field, b.vertex, np.ascontiguousarray(fbc_mg, np.int32), P.x, P.y, P.zc, P.w)
    h1 = np.ones(P.shape)
    h2 = np.zeros(P.shape)
    rhs = P.field.dssum(P.bm1 * P.uniform_field(8)) * P.mask
    uo, ito, histo, div0 = mg.flex_cg(rhs, h1, h2, 200, -1e-10)
    F = gpu.box_field(b)
    x, y, z = b.xyz()
    F.h1mg_setup(fbc_mg, b.vertex, x, y, z)
    res = F.to_dev(rhs)
    out = F.hmh_flex_cg(res, F.to_dev(h1), F.to_dev(h2), F.mult, 200,
                        params=gpu.default_params(param42=2, param21=-1e-10), history=True)
    say(f"[flex_cg pmask] gpu {out['niter']} oracle {ito}")
    assert abs(out["niter"] - ito) <= 1
    k = min(out["niter"], ito)
    np.testing.assert_allclose(out["history"][:k], histo[:k], rtol=1e-6,
                               atol=1e-15 * histo[0])
    assert l2rel(host(res), uo) < 1e-9
    F.destroy()


def test_flex_cg_amg_coarse_vcycle(gpu):
    """param(40) = 1: the coarse problem gets one AMG V-cycle instead of a
    solve to 1e-13 (the reference's AMG coarse-solver option).  The flexible
    CG still converges to the same solution (oracle, exact coarse solve);
    the iteration count may change slightly."""
    P, mg, F = build(gpu, dict(nel=(12, 11, 10), lx1=4, deform=0.05))
    assert F.h1mg_info()["ncrs"] > 600  # the AMG path, not the dense inverse
    h1 = np.ones(P.shape)
    h2 = np.zeros(P.shape)
    rhs = P.field.dssum(P.bm1 * P.uniform_field(31)) * P.mask
    uo, ito, histo, div0 = mg.flex_cg(rhs, h1, h2, 200, -1e-10)
    res = F.to_dev(rhs)
    out = F.hmh_flex_cg(res, F.to_dev(h1), F.to_dev(h2), F.mult, 200,
                        params=gpu.default_params(param42=2, param21=-1e-10, param40=1))
    say(f"[flex_cg param40=1] gpu {out['niter']} oracle (exact coarse) {ito}")
    assert out["converged"]
    assert out["niter"] <= ito + 5
    assert l2rel(host(res), uo) < 1e-7
    F.destroy()


@pytest.mark.parametrize("case", [MG_CASES[0], MG_CASES[3], MG_CASES[6]])
def test_gmres(gpu, case):
    """hmholtz('PRES') with param(42) = 0 -> hmh_gmres (gmres.f:304-545)."""
    P, mg, F = build(gpu, case)
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, 3.0)
    b = P.bm1 * P.uniform_field(41)
    rhs = P.field.dssum(b) * P.mask
    uo, ito, histo, div0 = mg.gmres(rhs, h1, h2, 200, -1e-11, m=30)
    u = F.zeros()
    out = F.hmholtz("PRES", u, F.to_dev(b), F.to_dev(h1), F.to_dev(h2), F.mask, F.mult,
                    -1e-11, 200, params=gpu.default_params(param42=0, param21=-1e-11),
                    history=True)
    say(f"[gmres] {case}: gpu {out['niter']} oracle {ito}")
    assert abs(out["niter"] - ito) <= 1
    k = min(out["niter"], ito)
    np.testing.assert_allclose(out["history"][:k], histo[:k], rtol=1e-6,
                               atol=1e-15 * histo[0])
    assert abs(out["rbn0"] - div0) <= 1e-12 * div0
    assert l2rel(host(u), uo) < 1e-9
    F.destroy()


def test_gmres_restart(gpu):
    """More than 30 Arnoldi steps: the restart (true residual, new basis)."""
    P, mg, F = build(gpu, dict(nel=4, lx1=8, deform=0.3))
    h1 = np.ones(P.shape)
    h2 = np.zeros(P.shape)
    rhs = P.field.dssum(P.bm1 * P.uniform_field(42)) * P.mask
    uo, ito, histo, div0 = mg.gmres(rhs, h1, h2, 200, -1e-11, m=30)
    res = F.to_dev(rhs)
    out = F.hmh_gmres(res, F.to_dev(h1), F.to_dev(h2), F.mult, 200,
                      params=gpu.default_params(param42=0, param21=-1e-11), history=True)
    say(f"[gmres restart] gpu {out['niter']} oracle {ito}")
    assert ito > 30
    assert abs(out["niter"] - ito) <= 1
    k = min(out["niter"], ito)
    np.testing.assert_allclose(out["history"][:k], histo[:k], rtol=1e-5,
                               atol=1e-15 * histo[0])
    assert l2rel(host(res), uo) < 1e-9
    F.destroy()


@pytest.mark.parametrize("case", [MG_CASES[0], MG_CASES[1], MG_CASES[5]])
def test_flex_cg_fp32_smoother(gpu, case):
    """precision = 1: the fine-level Schwarz solves in FP32 inside the FP64
    flexible CG.  The preconditioner changes at the 1e-7 level, so the
    iteration count may move by a little; the solution meets the same FP64
    tolerance and agrees with the oracle's to it."""
    P, mg, F = build(gpu, case)
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, 1.0)
    b = P.bm1 * P.uniform_field(51)
    rhs = P.field.dssum(b) * P.mask
    uo, ito, histo, div0 = mg.flex_cg(rhs, h1, h2, 200, -1e-10)
    for solver_p42 in (2, 0):
        u = F.zeros()
        out = F.hmholtz("PRES", u, F.to_dev(b), F.to_dev(h1), F.to_dev(h2), F.mask, F.mult,
                        -1e-10, 200,
                        params=gpu.default_params(param42=solver_p42, param21=-1e-10,
                                                  precision=1))
        say(f"[fp32 smoother] {case} param42={solver_p42}: gpu {out['niter']} oracle {ito}")
        assert out["converged"]
        assert out["niter"] <= ito + 3
        assert l2rel(host(u), uo) < 1e-8
    F.destroy()


@pytest.mark.parametrize("case", [MG_CASES[0], MG_CASES[1], MG_CASES[6]])
@pytest.mark.parametrize("param40", [0, 1])
def test_flex_cg_mixed_defect_correction(gpu, case, param40):
    """precision = 2: FP64 defect correction around single-precision inner
    flexible-CG solves (FP32 operator, fine-level V-cycle and vectors; FP64
    inner products and coarse solve, the latter to 0.1 relative inside the
    inner V-cycles).  The true FP64 residual, recomputed by
    the oracle's operator on the host, meets the FP64 tolerance; the solution
    agrees with the oracle's FP64 solution to that tolerance; the inner
    iterations (restarted per correction) stay within a few of the FP64
    count."""
    P, mg, F = build(gpu, case)
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, 10.0)
    b = P.bm1 * P.uniform_field(61)
    rhs = P.field.dssum(b) * P.mask
    tol = 1e-10
    uo, ito, _, div0 = mg.flex_cg(rhs, h1, h2, 300, -tol)
    u = F.zeros()
    out = F.hmholtz("PRES", u, F.to_dev(b), F.to_dev(h1), F.to_dev(h2), F.mask, F.mult, -tol,
                    300, params=gpu.default_params(param42=2, param21=-tol, precision=2,
                                                   param40=param40))
    ug = host(u)
    r = rhs - P.field.dssum(P.field.axhelm(ug, h1, h2)) * P.mask
    rn = float(np.sqrt(np.sum(r * r * P.mult) / P.field.vol))
    say(f"[mixed] {case} param40={param40}: inner iterations {out['niter']} (FP64 oracle {ito}), "
        f"true residual {rn / div0:.2e} of the initial")
    assert out["converged"]
    assert abs(out["rbn0"] - div0) <= 1e-12 * div0
    assert rn < tol * div0 * 1.0001
    assert abs(out["rbn2"] - rn) <= 1e-6 * rn + 1e-15 * div0
    if param40 == 0:
        assert out["niter"] <= ito + 8
    assert l2rel(ug, uo) < 1e-7
    F.destroy()


def test_flex_cg_mixed_needs_uniform_coefficients(gpu):
    P, mg, F = build(gpu, MG_CASES[0])
    h1 = 1.0 + 0.1 * P.uniform_field(3)
    u = F.zeros()
    with pytest.raises(gpu._lib.NekAmdError, match="uniform"):
        F.hmholtz("PRES", u, F.to_dev(P.bm1), F.to_dev(h1), F.zeros(), F.mask, F.mult, -1e-8,
                  50, params=gpu.default_params(param42=2, param21=-1e-8, precision=2))
    with pytest.raises(gpu._lib.NekAmdError, match="precision"):
        F.hmholtz("PRES", u, F.to_dev(P.bm1), F.zeros() + 1.0, F.zeros(), F.mask, F.mult, -1e-8,
                  50, params=gpu.default_params(param42=0, param21=-1e-8, precision=2))
    F.destroy()


@pytest.mark.parametrize("case", [MG_CASES[0], MG_CASES[1]])
def test_fp32_smoother_matrix_core_variants(gpu, case):
    """The single-precision local solves with their contractions on the
    matrix cores (FP32 operands, or bf16 operands with FP32 accumulation):
    the FP32-smoother and mixed solves still meet the FP64 tolerance in the
    true residual, within a few iterations of the oracle's FP64 count."""
    P, mg, F = build(gpu, case)
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, 10.0)
    b = P.bm1 * P.uniform_field(71)
    rhs = P.field.dssum(b) * P.mask
    tol = 1e-10
    uo, ito, _, div0 = mg.flex_cg(rhs, h1, h2, 300, -tol)
    assert gpu.fdm_variant() == 0
    try:
        for variant in (1, 2):
            gpu.set_fdm_variant(variant)
            for precision in (1, 2):
                u = F.zeros()
                out = F.hmholtz("PRES", u, F.to_dev(b), F.to_dev(h1), F.to_dev(h2), F.mask,
                                F.mult, -tol, 300,
                                params=gpu.default_params(param42=2, param21=-tol,
                                                          precision=precision))
                ug = host(u)
                r = rhs - P.field.dssum(P.field.axhelm(ug, h1, h2)) * P.mask
                rn = float(np.sqrt(np.sum(r * r * P.mult) / P.field.vol))
                say(f"[fdm variant {variant}] {case} precision={precision}: {out['niter']} "
                    f"iterations (oracle {ito}), true residual {rn / div0:.2e}")
                assert out["converged"]
                assert out["niter"] <= ito + 8
                assert rn < tol * div0 * 1.0001
                assert l2rel(ug, uo) < 1e-7
    finally:
        gpu.set_fdm_variant(0)
    with pytest.raises(gpu.NekAmdError):
        gpu.set_fdm_variant(4)
    F.destroy()


@pytest.mark.parametrize("case", [MG_CASES[0], MG_CASES[1], MG_CASES[4]])
def test_fp64_smoother_matrix_core_variant(gpu, case):
    """The double-precision local solves with their six contractions as FP64
    matrix-core products (variant 3): one V-cycle within 1e-12 of the
    oracle's h1mg_solve, the FP64 flexible CG with the oracle's iteration
    count (+-1) and solution (1e-9); the accumulation order inside a
    product differs from the vector kernel's, so agreement is to rounding."""
    P, mg, F = build(gpu, case)
    rng = np.random.default_rng(13)
    r = P.field.dssum(P.bm1 * rng.standard_normal(P.shape))
    zo = mg.solve(r)
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, 10.0)
    b = P.bm1 * P.uniform_field(73)
    rhs = P.field.dssum(b) * P.mask
    uo, ito, _, _ = mg.flex_cg(rhs, h1, h2, 300, -1e-10)
    try:
        gpu.set_fdm_variant(3)
        assert gpu.fdm_variant() == 3
        say(f"[fdm variant 3] {case} V-cycle")
        zd = F.zeros()
        F.h1mg_solve(zd, F.to_dev(r))
        e = l2rel(host(zd), zo)
        say(f"[fdm variant 3] V-cycle rel L2 = {e:.3e}")
        assert e < 1e-12
        u = F.zeros()
        out = F.hmholtz("PRES", u, F.to_dev(b), F.to_dev(h1), F.to_dev(h2), F.mask, F.mult,
                        -1e-10, 300, params=gpu.default_params(param42=2, param21=-1e-10))
        e = l2rel(host(u), uo)
        say(f"[fdm variant 3] flex-CG {out['niter']} iterations (oracle {ito}), rel L2 {e:.3e}")
        assert out["converged"] and abs(out["niter"] - ito) <= 1
        assert e < 1e-9
    finally:
        gpu.set_fdm_variant(0)
    F.destroy()


@pytest.mark.parametrize("case", [dict(nel=3, lx1=8, deform=0.1, neumann_all=True),
                                  dict(nel=(12, 11, 10), lx1=4, deform=0.05, neumann_all=True)])
def test_pressure_null_space(gpu, case):
    """No Dirichlet face, h2 = 0 (the reference's ifvcor): the coarse problem
    holds its last unknown at zero and removes its solution's mean
    (crs_xxt.c:883-955), flexible CG and GMRES end with ortho
    (hmholtz.f:2272, gmres.f:427,534).  A consistent right-hand side; the
    GPU against the oracle (iterations +-1, histories 1e-6, solutions 1e-9),
    the solution's local-node mean zero.  The second mesh has more than 600
    coarse unknowns: AMG-preconditioned CG on the coarse problem."""
    P, mg, F = build(gpu, case)
    assert O.lib().nko_mg_null_space(mg.h) == 1
    h1 = np.ones(P.shape)
    h2 = np.zeros(P.shape)
    b = P.bm1 * P.uniform_field(51)
    rhs = P.field.dssum(b) * P.mask
    # consistent: the assembled right-hand side sums to zero over unique nodes
    rhs = rhs - np.sum(rhs * P.mult) / np.sum(P.mult)
    b = rhs * P.mult  # hmholtz's dssum of this local split gives rhs back
    for name, p42, ref in (("flex_cg", 2, mg.flex_cg), ("gmres", 0, mg.gmres)):
        uo, ito, histo, div0 = ref(np.ascontiguousarray(rhs), h1, h2, 200, -1e-10)
        u = F.zeros()
        out = F.hmholtz("PRES", u, F.to_dev(b), F.to_dev(h1), F.to_dev(h2), F.mask, F.mult,
                        -1e-10, 200, params=gpu.default_params(param42=p42, param21=-1e-10),
                        history=True)
        ug = host(u)
        say(f"[null space] {case} {name}: gpu {out['niter']} oracle {ito}, "
            f"diff {l2rel(ug, uo):.2e}, mean {np.sum(ug) / ug.size:.2e}")
        assert abs(out["niter"] - ito) <= 1, name
        k = min(out["niter"], ito)
        np.testing.assert_allclose(out["history"][:k], histo[:k], rtol=1e-6,
                                   atol=1e-15 * histo[0], err_msg=name)
        assert l2rel(ug, uo) < 1e-9, name
        assert abs(np.sum(ug)) <= 1e-12 * np.sum(np.abs(ug)), name
    F.destroy()
