"""Pins the oracle's adjoint coefficient gradient (dJ/dh1 = -(D lam)^T G (D u)
per local node, dJ/dh2 = -lam bm1 u) independently of its code:

  * the Helmholtz form is linear in h1: sum_l g1_l h1_l = -lam . A_L(h1, 0) u
    for any u, lam and h1 (A_L = the oracle's axhelm, a separate
    implementation of the same contractions), and likewise for h2;
  * central finite differences of J(h1) = 1/2 (u-u*)^T B (u-u*), with
    A(h1) u = f solved by the oracle's Jacobi PCG to a 1e-13 relative
    residual, agree with the adjoint gradient along random directions.
"""
import numpy as np

from oracle import oracle as O
from problems import Problem, box


def test_gradient_is_the_derivative_of_the_form():
    P = Problem(box((3, 2, 2), 6, deform=0.15))
    rng = np.random.default_rng(5)
    u = rng.standard_normal(P.shape)
    lam = rng.standard_normal(P.shape)
    g1, g2 = P.field.adj_grad(u, lam)
    for seed in (1, 2):
        h1 = 1.0 + 0.5 * P.uniform_field(seed)
        h2 = 2.0 + P.uniform_field(seed + 10)
        au = P.field.axhelm(u, h1, np.zeros(P.shape))
        lhs = float(np.sum(g1 * h1))
        rhs = -float(np.sum(lam * au))
        assert abs(lhs - rhs) <= 1e-12 * np.sum(np.abs(g1 * h1))
        full = P.field.axhelm(u, h1, h2)
        assert abs(float(np.sum(g1 * h1 + g2 * h2)) + float(np.sum(lam * full))) <= \
            1e-12 * (np.sum(np.abs(g1 * h1)) + np.sum(np.abs(g2 * h2)))
    # swapping u and lam leaves the symmetric form unchanged
    g1s, _ = P.field.adj_grad(lam, u)
    np.testing.assert_allclose(g1s, g1, rtol=1e-12, atol=1e-12 * np.max(np.abs(g1)))


def _solve(P, rhs, h1, h2):
    prm = O.Params(20, 0.0, 0.0, 0.0)
    u, _, it, _, _ = P.field.hmholtz(rhs, h1, h2, -1e-13, 2000, prm)
    return u


def test_gradient_matches_central_differences():
    P = Problem(box((2, 2, 2), 6, deform=0.1))
    x, y, z = P.x, P.y, P.zc
    h1 = np.exp(0.3 * np.sin(2 * np.pi * x) * np.cos(np.pi * y) + 0.2 * z)
    h2 = np.zeros(P.shape)
    f = P.bm1 * P.uniform_field(3)
    ustar = P.uniform_field(4) * P.mask

    def J(h):
        u = _solve(P, f, h, h2)
        d = u - ustar
        return 0.5 * float(np.sum(P.bm1 * d * d)), u

    J0, u = J(h1)
    lam = _solve(P, P.bm1 * (u - ustar), h1, h2)
    g1, _ = P.field.adj_grad(u, lam)
    for seed in range(3):
        # a continuous perturbation: the same value at every copy of a node
        dh = P.uniform_field(50 + seed) * h1
        eps = 1e-5
        fd = (J(h1 + eps * dh)[0] - J(h1 - eps * dh)[0]) / (2 * eps)
        ad = float(np.sum(g1 * dh))
        assert abs(fd - ad) <= 1e-6 * abs(ad), (fd, ad)
