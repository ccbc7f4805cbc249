"""Pins the oracle's H1 multigrid pieces (hsmg.f / fast3d.f) and its
flexible CG (hmholtz.f:2164).

  * fd_weights_full / semhat: exact derivatives and interpolation of
    polynomials; ah = D^T W D with D from dgll (a different construction),
    bh = GLL weights;
  * generalev: A S = B S diag(lam), S^T B S = I, eigenvalues equal
    scipy.linalg.eigh(A, B);
  * FDM local solve: on a uniform box, an element with neighbours on all
    sides sees the extended operator A_ext = A(x)B(x)B + B(x)A(x)B + B(x)B(x)A
    built from the assembled three-element 1-D stiffness/mass (numpy,
    independent of the oracle's extension code); S D S^T inverts it exactly;
  * coarse solve: for a random vertex vector x, the Galerkin restriction of
    A P x (P = trilinear interpolation, A the fine operator) solves back to x;
  * flexible CG with the V-cycle reaches the tolerance on the true residual
    and agrees with the Jacobi-PCG solution of the same system.
"""
import numpy as np
import pytest
import scipy.linalg

from oracle import oracle as O
from problems import Problem, box, mg_for


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7])
def test_fd_weights_semhat(n):
    z, _ = O.zwgll(n + 1)
    xx = 0.37
    c = O.fd_weights_full(xx, z, m=1)
    for k in range(n + 1):
        assert abs(c[0] @ z ** k - xx ** k) < 1e-13
        assert abs(c[1] @ z ** k - (k * xx ** (k - 1) if k else 0.0)) < 1e-11
    a, b, d, zz = O.semhat(n)
    assert np.array_equal(zz, z)
    _, w = O.zwgll(n + 1)
    assert np.allclose(b, w, atol=1e-15)
    D = O.dgll(z)
    np.testing.assert_allclose(a, D.T @ np.diag(w) @ D, atol=1e-12 * np.abs(a).max())
    np.testing.assert_allclose(a @ np.ones(n + 1), 0, atol=1e-12)


@pytest.mark.parametrize("nl", [4, 6, 10])
def test_generalev(nl):
    rng = np.random.default_rng(nl)
    M = rng.standard_normal((nl, nl))
    A = M + M.T
    Bm = rng.standard_normal((nl, nl))
    B = Bm @ Bm.T + nl * np.eye(nl)
    S, lam = O.generalev(A, B)
    np.testing.assert_allclose(A @ S, B @ S @ np.diag(lam), atol=1e-11)
    np.testing.assert_allclose(S.T @ B @ S, np.eye(nl), atol=1e-12)
    np.testing.assert_allclose(lam, scipy.linalg.eigh(A, B, eigvals_only=True), rtol=1e-12,
                               atol=1e-12)


def assembled_1d(ah, bh, h, n):
    """Stiffness and mass of three elements of length h, restricted to the
    extended node set (left neighbour's node n-1, element nodes 0..n, right
    neighbour's node 1)."""
    N = 3 * n + 1
    A = np.zeros((N, N))
    B = np.zeros(N)
    for e in range(3):
        s = e * n
        A[s:s + n + 1, s:s + n + 1] += (2.0 / h) * ah
        B[s:s + n + 1] += (h / 2.0) * bh
    idx = [n - 1] + list(range(n, 2 * n + 1)) + [2 * n + 1]
    return A[np.ix_(idx, idx)], np.diag(B[idx])


@pytest.mark.parametrize("lx1", [5, 8])
def test_fdm_inverts_extended_operator(lx1):
    P = Problem(box(5, lx1))
    mg = mg_for(P)
    L = mg.lmax
    nh = mg.nh[L - 1]
    nl = nh + 2
    n = nh - 1
    a, b, _, _ = O.semhat(n)
    h = P.box.h[0]
    A1, B1 = assembled_1d(a, b, h, n)
    Aext = (np.kron(np.kron(B1, B1), A1) + np.kron(np.kron(B1, A1), B1)
            + np.kron(np.kron(A1, B1), B1))
    rng = np.random.default_rng(1)
    nel = P.shape[0]
    u = rng.standard_normal((nel, nl ** 3))
    r = np.ascontiguousarray(np.einsum("ij,ej->ei", Aext, u))
    e = np.zeros_like(r)
    O.lib().nko_mg_apply_fdm(mg.h, L, e.ctypes.data, r.ctypes.data)
    centre = 2 + 5 * (2 + 5 * 2)  # element (2,2,2): neighbours on every side
    np.testing.assert_allclose(e[centre], u[centre], rtol=0, atol=1e-9 * np.abs(u).max())


@pytest.mark.parametrize("case", [dict(nel=4, lx1=8, deform=0.1), dict(nel=(4, 3, 3), lx1=6)])
def test_coarse_solve_inverts_galerkin(case):
    P = Problem(box(**case))
    mg = mg_for(P)
    nel = P.shape[0]
    z = P.z
    phi = np.stack([(1 - z) / 2, (1 + z) / 2])  # phi[c, i]
    V = P.box.vertex.reshape(nel, 8)
    vmask = mg.mask(1).reshape(nel, 8)
    rng = np.random.default_rng(3)
    xv = rng.standard_normal(int(V.max()) + 1)
    xl = xv[V] * vmask  # continuous, zero on Dirichlet vertices
    u = np.einsum("ekji,ir,js,kt->etsr", xl.reshape(nel, 2, 2, 2), phi, phi, phi)
    one = np.ones(P.shape)
    w = P.field.axhelm(np.ascontiguousarray(u), one, np.zeros(P.shape))
    bl = np.einsum("etsr,ir,js,kt->ekji", w, phi, phi, phi).reshape(nel, 8)
    x = np.zeros_like(bl)
    O.lib().nko_mg_coarse_solve(mg.h, x.ctypes.data, np.ascontiguousarray(bl).ctypes.data)
    np.testing.assert_allclose(x, xl, atol=1e-10 * np.abs(xl).max())


def test_flex_cg_vs_jacobi():
    P = Problem(box((4, 4, 3), 8, deform=0.1))
    mg = mg_for(P)
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, 10.0)
    b = P.bm1 * P.uniform_field(5)
    rhs = P.field.dssum(b) * P.mask
    u, it, hist, div0 = mg.flex_cg(rhs, h1, h2, 300, -1e-12)
    assert it < 60
    res = rhs - P.mask * P.field.dssum(P.field.axhelm(u, h1, h2))
    rn = np.sqrt(np.sum(res * res * P.mult) / P.field.vol)
    assert rn < 1.0001e-12 * div0
    uj, _, nj, _, _ = P.field.hmholtz(b, h1, h2, -1e-13, 900)
    assert np.linalg.norm(u - uj) / np.linalg.norm(uj) < 1e-9
    assert it * 4 < nj


@pytest.mark.parametrize("m", [30, 4])
def test_gmres_oracle(m):
    """hmh_gmres (right-preconditioned, restarted): the residual norms it
    reports are the true ones (recomputed from its solution), they do not
    increase within a restart cycle, and the solution agrees with the
    flexible CG's."""
    P = Problem(box((3, 4, 3), 7, deform=0.1))
    mg = mg_for(P)
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, 2.0)
    rhs = P.field.dssum(P.bm1 * P.uniform_field(9)) * P.mask
    u, it, hist, div0 = mg.gmres(rhs, h1, h2, 300, -1e-11, m=m)
    assert it < 200
    res = rhs - P.mask * P.field.dssum(P.field.axhelm(u, h1, h2))
    rn = np.sqrt(np.sum(res * res * P.mult) / P.field.vol)
    assert rn < 1.0001e-11 * div0 * 1.001 + 1e-14 * div0
    np.testing.assert_allclose(rn, hist[-1], rtol=1e-4)
    for c0 in range(0, it, m):
        seg = hist[c0:c0 + m]
        assert np.all(np.diff(seg) <= 1e-12 * seg[0])
    uf, _, _, _ = mg.flex_cg(rhs, h1, h2, 300, -1e-11)
    assert np.linalg.norm(u - uf) / np.linalg.norm(uf) < 1e-8
