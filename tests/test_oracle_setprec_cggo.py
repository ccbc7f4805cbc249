"""Pins the oracle's Jacobi diagonal (hmholtz.f:380 setprec) and its CG
(hmholtz.f:611 cggo, hmholtz.f:2 hmholtz).

  * setprec: 1/d equals the diagonal of the brute-force assembled matrix
    Q^T A_L Q (A applied to every global unit vector) on affine meshes at all
    nodes; on deformed meshes at every node that is not on an element edge
    (the reference adds the g4..g6 cross terms only at element corners,
    hmholtz.f:436-468, which is exact there and at face/interior nodes);
  * cggo on the unit-cube Poisson problem (8^3 elements, N=7): converges to
    the analytic solution sin(pi x) sin(pi y) sin(pi z); the returned u has a
    true residual below the requested relative tolerance; the CG scalars
    satisfy rho = p.Ap > 0 and beta = rtz1/rtz2 identities.
"""
import numpy as np
import pytest

from oracle import oracle as O
from problems import Problem, box


def assembled_diag(P, h1, h2):
    glo = P.glo.reshape(P.shape)
    uid = P.uid
    ids, inv = np.unique(uid.reshape(-1), return_inverse=True)
    diag = np.zeros(len(ids))
    for g in range(len(ids)):
        e = (inv == g).astype(np.float64).reshape(P.shape)
        Ae = P.field.dssum(P.field.axhelm(e, h1, h2))
        diag[g] = Ae.reshape(-1)[inv == g][0]
    del glo
    return diag[inv].reshape(P.shape)


@pytest.mark.parametrize("case,exact_everywhere", [
    (dict(nel=2, lx1=4), True),
    (dict(nel=(2, 1, 2), lx1=5, deform=0.1), False),
    (dict(nel=2, lx1=4, jitter=0.08, seed=5), False),
])
def test_setprec_vs_assembled_diagonal(case, exact_everywhere):
    P = Problem(box(**case))
    rng = np.random.default_rng(1)
    # setprec scales by h1 at the node, not at the quadrature point, so the
    # diagonal is exact for uniform h1; h2 enters through the diagonal mass
    h1 = np.full(P.shape, 1.3)
    h2 = 0.7 + 0.2 * rng.random(P.shape)
    h2 = P.field.dssum(h2) * P.mult   # continuous coefficient
    d = P.field.setprec(h1, h2)
    diag = assembled_diag(P, h1, h2)
    n = P.shape[-1]
    # The exact diagonal also holds 2*(g4 Drr Dss + g5 Drr Dtt + g6 Dss Dtt)*h1
    # at every element-edge node; the reference adds it at the 8 corners only.
    # Predict what it leaves out at the other edge nodes and add it back.
    Dd = np.diag(P.D)
    drr = Dd[None, None, None, :]
    dss = Dd[None, None, :, None]
    dtt = Dd[None, :, None, None]
    g4, g5, g6 = P.g[3], P.g[4], P.g[5]
    cross = 2.0 * (g4 * drr * dss + g5 * drr * dtt + g6 * dss * dtt) * h1
    corner = np.zeros((n, n, n), bool)
    for c in np.ndindex(2, 2, 2):
        corner[tuple(np.array(c) * (n - 1))] = True
    cross[:, corner] = 0.0
    missing = P.field.dssum(cross)
    pred = 1.0 / d + missing
    assert np.max(np.abs(pred - diag) / np.abs(diag)) < 1e-13
    if exact_everywhere:
        assert np.max(np.abs(1.0 / d - diag) / np.abs(diag)) < 1e-13
    else:
        # the left-out terms are visible on a piecewise-trilinear (jittered) mesh
        assert np.max(np.abs(missing)) > 1e-8 * np.max(np.abs(diag)) or case.get("deform")


def test_cggo_unit_cube_poisson():
    P = Problem(box(8, 8))
    f = P.rhs_config1(noise=0.0)
    one = np.ones(P.shape)
    zero = np.zeros(P.shape)
    rhs = P.bm1 * f
    u, r_after, niter, hist, rbn0 = P.field.hmholtz(rhs, one, zero, -1e-10, 900)
    exact = np.sin(np.pi * P.x) * np.sin(np.pi * P.y) * np.sin(np.pi * P.zc)
    assert 0 < niter < 900
    assert np.max(np.abs(u - exact)) < 1e-6
    # true residual of the returned iterate (independent recomputation)
    res = r_after - P.mask * P.field.dssum(P.field.axhelm(u, one, zero))
    rn = np.sqrt(np.sum(P.field.binv * P.mult * res * res) / P.field.vol)
    assert rn <= 1e-10 * rbn0 * 1.0001
    # CG scalars: rho > 0 (SPD), alpha = rtz1/rho, beta = rtz1(k)/rtz1(k-1)
    assert np.all(hist[:niter, 2] > 0)
    np.testing.assert_allclose(hist[:niter, 3], hist[:niter, 1] / hist[:niter, 2], rtol=1e-15)
    np.testing.assert_allclose(hist[1:niter, 4], hist[1:niter, 1] / hist[:niter - 1, 1],
                               rtol=1e-15)
    assert hist[niter, 0] <= 1e-10 * rbn0


def test_cggo_noise_rhs_config1():
    """Config-1 right-hand side (with the seeded noise): residual reaches 1e-8."""
    P = Problem(box(8, 8))
    one = np.ones(P.shape)
    zero = np.zeros(P.shape)
    rhs = P.bm1 * P.rhs_config1(seed=1, noise=1e-3)
    u, r_after, niter, hist, rbn0 = P.field.hmholtz(rhs, one, zero, -1e-8, 900)
    res = r_after - P.mask * P.field.dssum(P.field.axhelm(u, one, zero))
    rn = np.sqrt(np.sum(P.field.binv * P.mult * res * res) / P.field.vol)
    assert rn <= 1e-8 * rbn0 * 1.0001
    assert hist[niter - 1, 0] > 1e-8 * rbn0  # first iterate above tolerance
    assert 20 < niter < 400


def test_cggo_zero_rhs_and_maxit():
    P = Problem(box(2, 4))
    one = np.ones(P.shape)
    zero = np.zeros(P.shape)
    u, _, niter, _, _ = P.field.hmholtz(np.zeros(P.shape), one, zero, -1e-8, 900)
    assert niter == 0 and np.all(u == 0)
    rhs = P.bm1 * P.uniform_field(3)
    u, _, niter, hist, _ = P.field.hmholtz(rhs, one, zero, -1e-14, 3)
    assert niter == 3
