"""Pins the oracle's residual projection (navier4.f:562-1325: hsolve,
project1/2, proj_ortho, proj_ortho_full_cgs2, iproj_chk) by what the
mathematics fixes, independently of the code under test:

  * after every project2 the stored pairs satisfy B = A X and X^T W B = I
    (W = vmult: the weighted inner product of assembled fields);
  * a right-hand side A x with x in span(X) projects to zero and xbar = x;
  * after the Givens sweep the first column is a multiple of the newest
    solution;
  * a solution already in the space is rejected (rank deficiency), the space
    never exceeds mmx = (mxprev-4)/2 vectors;
  * a change of h1 rebuilds B with the new operator (and the first project1
    after the first solve always does, since h1old starts at zero);
  * hsolve with projection returns the solution of the unprojected solve.
"""
import numpy as np
import pytest

from oracle import oracle as O
from problems import Problem, box


def _problem():
    return Problem(box((3, 3, 2), 5, deform=0.2))


def _A(P, x, h1, h2):
    b = P.field.axhelm(x, h1, h2)
    b = P.field.dssum(b)
    return b * P.mask


def _wdot(P, a, b):
    return float(np.sum(a * P.mult * b))


def _check_space(P, pr, h1, h2, tol=1e-11):
    m = pr.m
    X = [pr.x(k) for k in range(m)]
    B = [pr.b(k) for k in range(m)]
    G = np.array([[_wdot(P, X[i], B[j]) for j in range(m)] for i in range(m)])
    np.testing.assert_allclose(G, np.eye(m), atol=tol)
    for k in range(m):
        ref = _A(P, X[k], h1, h2)
        assert np.max(np.abs(B[k] - ref)) <= tol * max(1.0, np.max(np.abs(ref)))


def _step(P, pr, xt, h1, h2):
    """One hsolve with an exact inner solve: project1(A xt), correction
    u = xt - xbar, project2(u).  Returns (ireset, residual ratio)."""
    b = np.ascontiguousarray(_A(P, xt, h1, h2))
    m0 = pr.m
    ir, bb4, baf = pr.project1(b, h1, h2, P.mask, P.mult)
    u = np.ascontiguousarray(xt - pr.xbar()) if m0 > 0 else np.ascontiguousarray(xt.copy())
    pr.project2(u, h1, h2, P.mask, P.mult)
    np.testing.assert_allclose(u, xt, atol=1e-12 * np.max(np.abs(xt)))
    return ir, (baf / bb4 if m0 > 0 else 1.0)


def test_space_orthonormal_window_and_first_column():
    P = _problem()
    h1 = np.ones(P.shape)
    h2 = np.zeros(P.shape)
    pr = O.Proj(P.field, mxprev=12)  # mmx = 4
    assert pr.mmx == 4
    resets = []
    for k in range(7):
        xt = P.uniform_field(100 + k) * P.mask
        ir, _ = _step(P, pr, xt, h1, h2)
        resets.append(ir)
        assert pr.m == min(k + 1, 4)
        _check_space(P, pr, h1, h2)
        x0 = pr.x(0)
        c = _wdot(P, x0, xt) / _wdot(P, xt, xt)
        assert np.max(np.abs(x0 - c * xt)) < 1e-10 * np.max(np.abs(x0))
    # h1old starts at zero: the second call rebuilds once, then never again
    assert resets == [0, 1, 0, 0, 0, 0, 0]


def test_span_projects_to_zero():
    P = _problem()
    h1 = 1.0 + 0.5 * P.uniform_field(7) ** 2
    h2 = np.full(P.shape, 3.0)
    pr = O.Proj(P.field, mxprev=20)
    xs = [P.uniform_field(200 + k) * P.mask for k in range(3)]
    for xt in xs:
        _step(P, pr, xt, h1, h2)
    assert pr.m == 3
    xt = 0.3 * xs[0] - 1.2 * xs[2] + 0.05 * xs[1]
    b = np.ascontiguousarray(_A(P, xt, h1, h2))
    nb = np.sqrt(_wdot(P, b, b))
    ir, bb4, baf = pr.project1(b, h1, h2, P.mask, P.mult)
    assert ir == 0
    assert abs(bb4 - nb) <= 1e-13 * nb
    assert baf < 1e-11 * bb4
    np.testing.assert_allclose(pr.xbar(), xt, atol=1e-11 * np.max(np.abs(xt)))


def test_repeated_solution_is_rejected():
    P = _problem()
    h1 = np.ones(P.shape)
    h2 = np.zeros(P.shape)
    pr = O.Proj(P.field, mxprev=20)
    x1 = P.uniform_field(300) * P.mask
    _step(P, pr, x1, h1, h2)
    _step(P, pr, P.uniform_field(301) * P.mask, h1, h2)
    assert pr.m == 2
    ir, ratio = _step(P, pr, x1, h1, h2)
    assert ratio < 1e-11
    assert pr.m == 2
    _check_space(P, pr, h1, h2)


def test_h1_change_rebuilds_space():
    P = _problem()
    h1 = np.ones(P.shape)
    h2 = np.zeros(P.shape)
    pr = O.Proj(P.field, mxprev=20)
    for k in range(4):
        _step(P, pr, P.uniform_field(400 + k) * P.mask, h1, h2)
    h1n = 2.0 + P.uniform_field(9)
    h2n = np.full(P.shape, 0.5)
    b = np.ascontiguousarray(_A(P, P.uniform_field(410) * P.mask, h1n, h2n))
    ir, _, _ = pr.project1(b, h1n, h2n, P.mask, P.mult)
    assert ir == 1
    assert pr.m == 4
    _check_space(P, pr, h1n, h2n)
    # moving-mesh flag forces the rebuild with unchanged h1/h2
    b = np.ascontiguousarray(_A(P, P.uniform_field(411) * P.mask, h1n, h2n))
    ir, _, _ = pr.project1(b, h1n, h2n, P.mask, P.mult, ifmvbd=1)
    assert ir == 1
    _check_space(P, pr, h1n, h2n)


def test_hsolve_flags():
    f = O.hsolve_uses_projection
    assert not f("VELX", 10, 20, 5, 0, ifprojfld=False)
    assert f("VELX", 10, 20, 5, 0, ifprojfld=True)
    assert not f("VELX", 4, 20, 5, 0, ifprojfld=True)
    assert not f("PRES", 10, 20, 5, 0, ifprojfld=False)     # param95 = 0: off
    assert f("PRES", 10, 20, 5, 5, ifprojfld=False)
    assert not f("PRES", 10, 0, 5, 5, ifprojfld=False)      # param93 = 0
    assert not f("TEMP", 10, 20, 5, 5, ifprojfld=True, ifield=3, ldimt_proj=1)


def test_hsolve_matches_unprojected_solve():
    P = _problem()
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, 2.0)
    f0 = P.rhs_config1(noise=0.0) * P.bm1
    f1 = P.uniform_field(500) * P.bm1
    pr = O.Proj(P.field, mxprev=20)
    prm = O.Params(20, 0.0, 0.0, 1e-10)
    iters = []
    for k in range(5):
        r = np.ascontiguousarray(f0 + 0.05 * k * f1)
        u, it, (bb4, baf) = O.hsolve(P.field, "VELX", r, h1, h2, P.mask, P.mult, 1e-10, 500, pr,
                                     prm)
        uref, _, _ = O.hsolve(P.field, "VELX", r, h1, h2, P.mask, P.mult, 1e-10, 500, None, prm,
                              use_projection=False)
        assert np.max(np.abs(u - uref)) < 1e-7 * np.max(np.abs(uref))
        iters.append(it)
        if k >= 2:
            assert baf < 0.5 * bb4
    assert iters[-1] < iters[0]
