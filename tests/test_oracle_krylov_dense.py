"""Pins the oracle's Krylov solvers iterate by iterate against their dense
global restatements (tests/dense_krylov.py): cggo's Jacobi PCG
(hmholtz.f:611-846) with its B^-1-weighted norm, stop rule and null-space
correction; chktcg1's tolerance floors (hmholtz.f:527-609) including the
Neumann branch; hmh_flex_cg (hmholtz.f:2164-2281) and hmh_gmres
(gmres.f:304-545), both preconditioned by the dense multigrid model of
tests/dense_mg.py, so the whole pressure solve is pinned without any of the
oracle's solver code.  Planted mistakes in the oracle make each comparison
fail.
"""
import numpy as np
import pytest

import dense_krylov as K
from dense_mg import DenseMG
from meshes import RotatedBox
from nek5000_amd.mesh import BoxMesh
from oracle import oracle as O
from problems import Problem

SYM = {"dirichlet": (1, 1, 1, 1, 1, 1), "mixed": (1, 1, 2, 2, 1, 2),
       "neumann": (2, 2, 2, 2, 2, 2)}


def build(nel, lx1, codes="dirichlet", seed=0, mg=True, **kw):
    box = BoxMesh(*nel, lx1, **kw)
    rb = RotatedBox(box, seed=seed)
    prep = box.dirichlet_faces()
    c = np.array(SYM[codes])[None, :] * prep[:, list(rb.SYM_TO_PREP)]
    fbc = np.ascontiguousarray(rb.face_codes(c.astype(np.int32)), np.int32)
    P = Problem(rb, fbc=fbc)
    dm = DenseMG(rb, fbc, rb.xyz(), P.field.axhelm) if mg else None
    omg = O.MG(P.field, rb.vertex, fbc, P.x, P.y, P.zc, P.w) if mg else None
    return rb, P, dm, omg


def lev_fine(rb, fbc):
    from dense_mg import Level
    return Level(rb, rb.lx1, fbc)


def fields(P, seed):
    rng = np.random.default_rng(seed)
    h1 = 0.5 + rng.random(P.shape)       # local coefficient fields
    h2 = rng.random(P.shape)
    return h1, h2


CG_CASES = [
    dict(nel=(3, 2, 2), lx1=6, deform=0.1, seed=1, h="fields"),
    dict(nel=(2, 3, 2), lx1=5, jitter=0.1, seed=2, codes="mixed", h="uniform"),
]


@pytest.mark.parametrize("case", CG_CASES)
def test_cggo_history_equals_dense(case):
    case = dict(case)
    hk = case.pop("h")
    rb, P, _, _ = build(mg=False, **case)
    lev = lev_fine(rb, P.fbc)
    if hk == "fields":
        h1, h2 = fields(P, 3)
    else:
        h1, h2 = np.ones(P.shape), np.full(P.shape, 2.0)
    S = K.System(lev, P.field.axhelm, P.bm1, h1, h2, lev.mask)
    S.setprec(P.g)
    u = np.random.default_rng(5).standard_normal(lev.N)
    f_g = lev.mask * lev.assemble(P.bm1 * lev.gather(u))
    x_o, n_o, h_o, rb0_o = P.field.cggo(np.ascontiguousarray(lev.gather(f_g)), h1, h2, -1e-10,
                                        500)
    x_d, n_d, h_d, rb0_d = K.cggo(S, f_g, -1e-10, 500)
    assert n_o == n_d, (n_o, n_d)
    assert abs(rb0_o - rb0_d) <= 1e-13 * rb0_d
    # every scalar (rbn2, rtz1, rho, alpha, beta) of the first 40 iterations
    # to 1e-10 (measured 1e-14): beyond that, finite-precision CG amplifies
    # the round-off of the two orders of summation exponentially (measured
    # 1e-10 by iteration 50, 0.1 in alpha by iteration 80 on the first case)
    # while the count and the solution stay the same
    k = min(40, len(h_d) - 1)
    np.testing.assert_allclose(h_o[:k], h_d[:k], rtol=1e-10, atol=1e-300)
    assert np.linalg.norm(x_o - lev.gather(x_d)) <= 1e-9 * np.linalg.norm(x_d)


def test_cggo_null_space_branch():
    """No Dirichlet face and h2 = 0: cggo removes the mean of r and z
    (hmholtz.f:705-720, 747-750); the histories still agree, and chktcg1
    raised the tolerance by its Neumann branch."""
    rb, P, _, _ = build((2, 2, 2), 5, codes="neumann", seed=3, mg=False, deform=0.1)
    lev = lev_fine(rb, P.fbc)
    assert lev.mask.min() == 1.0
    h1, h2 = np.ones(P.shape), np.zeros(P.shape)
    S = K.System(lev, P.field.axhelm, P.bm1, h1, h2, lev.mask)
    S.setprec(P.g)
    u = np.random.default_rng(6).standard_normal(lev.N)
    f_g = lev.assemble(P.bm1 * lev.gather(u))
    x_o, n_o, h_o, _ = P.field.cggo(np.ascontiguousarray(lev.gather(f_g)), h1, h2, -1e-9, 300)
    x_d, n_d, h_d, _ = K.cggo(S, f_g, -1e-9, 300)
    assert n_o == n_d
    np.testing.assert_allclose(h_o[:8, :2], h_d[:8, :2], rtol=1e-10)
    # chktcg1: pure Neumann and h2 = 0 -> tol >= 10 eps rinit
    tol_o = P.field.chktcg1(1e-20, np.ascontiguousarray(lev.gather(f_g)), h1, h2)
    tol_d = K.chktcg1(S, 1e-20, f_g)
    rinit = np.sqrt(np.sum(f_g * f_g / S.B) / S.vol)
    assert abs(tol_d - 1e-12 * rinit) <= 1e-15 * tol_d
    assert abs(tol_o - tol_d) <= 1e-13 * tol_d


@pytest.mark.parametrize("codes,h2v", [("dirichlet", 0.0), ("neumann", 1.0), ("mixed", 0.0)])
def test_chktcg1_floors(codes, h2v):
    """Outside the Neumann branch chktcg1 raises tol only to eps * rinit."""
    rb, P, _, _ = build((2, 2, 2), 5, codes=codes, seed=4, mg=False, jitter=0.1)
    lev = lev_fine(rb, P.fbc)
    h1, h2 = np.ones(P.shape), np.full(P.shape, h2v)
    S = K.System(lev, P.field.axhelm, P.bm1, h1, h2, lev.mask)
    f_g = lev.mask * lev.assemble(P.bm1 * lev.gather(
        np.random.default_rng(8).standard_normal(lev.N)))
    fl = np.ascontiguousarray(lev.gather(f_g))
    rinit = np.sqrt(np.sum(f_g * f_g / S.B) / S.vol)
    for tol in (1e-20, 1e-6):
        t_o = P.field.chktcg1(tol, fl, h1, h2)
        t_d = K.chktcg1(S, tol, f_g)
        assert abs(t_o - t_d) <= 1e-13 * t_d
        assert abs(t_d - max(tol, 1e-13 * rinit)) <= 1e-15 * t_d


PRES_CASES = [
    dict(nel=(3, 2, 2), lx1=8, deform=0.1, seed=11),
    dict(nel=(2, 2, 3), lx1=5, jitter=0.1, seed=12, codes="mixed"),
]


def pres_setup(case, seed):
    rb, P, dm, omg = build(**case)
    lev = dm.lev[-1]
    h1, h2 = np.ones(P.shape), np.full(P.shape, 3.0)
    S = K.System(lev, P.field.axhelm, P.bm1, h1, h2, lev.mask)
    rhs = lev.mask * lev.assemble(P.bm1 * lev.gather(
        np.random.default_rng(seed).standard_normal(lev.N)))
    return rb, P, dm, omg, lev, h1, h2, S, rhs


@pytest.mark.parametrize("case", PRES_CASES)
def test_flex_cg_history_equals_dense(case):
    rb, P, dm, omg, lev, h1, h2, S, rhs = pres_setup(case, 21)
    x_o, n_o, h_o, d_o = omg.flex_cg(np.ascontiguousarray(lev.gather(rhs)), h1, h2, 200,
                                     -1e-10)
    x_d, n_d, h_d, d_d = K.flex_cg(S, dm.solve, rhs, 200, -1e-10)
    assert n_o == n_d, (n_o, n_d)
    assert abs(d_o - d_d) <= 1e-13 * d_d
    np.testing.assert_allclose(h_o, h_d, rtol=1e-8)
    np.testing.assert_allclose(h_o[:5], h_d[:5], rtol=1e-11)
    assert np.linalg.norm(x_o - lev.gather(x_d)) <= 1e-10 * np.linalg.norm(x_d)


@pytest.mark.parametrize("case", PRES_CASES)
@pytest.mark.parametrize("m", [30, 4])
def test_gmres_history_equals_dense(case, m):
    rb, P, dm, omg, lev, h1, h2, S, rhs = pres_setup(case, 22)
    x_o, n_o, h_o, d_o = omg.gmres(np.ascontiguousarray(lev.gather(rhs)), h1, h2, 200, -1e-10,
                                   m=m)
    x_d, n_d, h_d, d_d = K.gmres(S, dm.solve, rhs, 200, -1e-10, m=m)
    assert n_o == n_d, (n_o, n_d)
    assert abs(d_o - d_d) <= 1e-13 * d_d
    np.testing.assert_allclose(h_o, h_d, rtol=1e-4)
    np.testing.assert_allclose(h_o[:3], h_d[:3], rtol=1e-11)
    assert np.linalg.norm(x_o - lev.gather(x_d)) <= 1e-9 * np.linalg.norm(x_d)


def test_planted_krylov_mistakes_fail():
    """cggo's norm without B^-1, the flexible CG's rho2 sign, GMRES without
    the earlier Givens rotations: each breaks its history comparison."""
    case = dict(CG_CASES[0])
    case.pop("h")
    rb, P, _, _ = build(mg=False, **case)
    lev = lev_fine(rb, P.fbc)
    h1, h2 = np.ones(P.shape), np.full(P.shape, 2.0)
    S = K.System(lev, P.field.axhelm, P.bm1, h1, h2, lev.mask)
    S.setprec(P.g)
    f_g = lev.mask * lev.assemble(P.bm1 * lev.gather(
        np.random.default_rng(5).standard_normal(lev.N)))
    _, _, h_d, _ = K.cggo(S, f_g, -1e-10, 500)
    with O.mutation(128):
        _, _, h_o, _ = P.field.cggo(np.ascontiguousarray(lev.gather(f_g)), h1, h2, -1e-10, 500)
    assert np.max(np.abs(h_o[:5, 0] / h_d[:5, 0] - 1)) > 1e-3
    rb, P, dm, omg, lev, h1, h2, S, rhs = pres_setup(PRES_CASES[0], 21)
    rl = np.ascontiguousarray(lev.gather(rhs))
    _, _, hf_d, _ = K.flex_cg(S, dm.solve, rhs, 200, -1e-10)
    _, _, hg_d, _ = K.gmres(S, dm.solve, rhs, 200, -1e-10, m=30)
    with O.mutation(256):
        _, _, hf_o, _ = omg.flex_cg(rl.copy(), h1, h2, 200, -1e-10)
    k = min(len(hf_o), len(hf_d))
    assert len(hf_o) != len(hf_d) or np.max(np.abs(hf_o[:k] / hf_d[:k] - 1)) > 1e-3
    with O.mutation(512):
        _, _, hg_o, _ = omg.gmres(rl.copy(), h1, h2, 200, -1e-10, m=30)
    k = min(len(hg_o), len(hg_d))
    assert len(hg_o) != len(hg_d) or np.max(np.abs(hg_o[:k] / hg_d[:k] - 1)) > 1e-3


def test_pressure_null_space_equals_dense():
    """No Dirichlet face and h2 = 0 (the reference's ifvcor): the coarse
    problem holds one unknown at zero and removes the mean of its solution
    (crs_xxt.c:883-955), and the flexible CG and GMRES end with ortho
    (hmholtz.f:2272; gmres.f:427, 534).  Oracle and dense restatements agree
    iterate by iterate on a consistent right-hand side; the solution's
    local-node mean is zero."""
    rb, P, dm, omg = build(nel=(3, 2, 2), lx1=6, deform=0.1, seed=13, codes="neumann")
    lev = dm.lev[-1]
    assert dm.null_space and O.lib().nko_mg_null_space(omg.h) == 1
    h1, h2 = np.ones(P.shape), np.zeros(P.shape)
    S = K.System(lev, P.field.axhelm, P.bm1, h1, h2, lev.mask)
    rhs = lev.assemble(P.bm1 * lev.gather(np.random.default_rng(31).standard_normal(lev.N)))
    rhs = rhs - np.mean(rhs)  # consistent: orthogonal to the constant
    count = lev.assemble(np.ones(P.shape))
    for solver in ("flex_cg", "gmres"):
        if solver == "flex_cg":
            x_o, n_o, h_o, d_o = omg.flex_cg(np.ascontiguousarray(lev.gather(rhs)), h1, h2,
                                             200, -1e-10)
            x_d, n_d, h_d, d_d = K.flex_cg(S, dm.solve, rhs, 200, -1e-10, null_count=count)
        else:
            x_o, n_o, h_o, d_o = omg.gmres(np.ascontiguousarray(lev.gather(rhs)), h1, h2, 200,
                                           -1e-10, m=30)
            x_d, n_d, h_d, d_d = K.gmres(S, dm.solve, rhs, 200, -1e-10, m=30,
                                         null_count=count)
        assert n_o == n_d, (solver, n_o, n_d)
        np.testing.assert_allclose(h_o, h_d, rtol=1e-6, err_msg=solver)
        assert np.linalg.norm(x_o - lev.gather(x_d)) <= 1e-9 * np.linalg.norm(x_d), solver
        assert abs(np.sum(x_o)) <= 1e-12 * np.sum(np.abs(x_o)), solver
        # the residual of the returned solution meets the tolerance
        assert np.sqrt(np.sum((rhs - S.Am(x_d)) ** 2) / S.vol) <= 1e-10 * d_d * 1.01, solver
