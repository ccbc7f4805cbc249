"""Pins the oracle's geometric factors (coef.f) and axhelm (hmholtz.f:72).

Independent routes:
  * affine boxes: the element matrix is the Kronecker sum
        (hy hz / 2hx) K(x)M(x)M + (hx hz / 2hy) M(x)K(x)M + (hx hy / 2hz) M(x)M(x)K
    with K = D^T W D, M = W (1-D GLL stiffness and mass);
  * any mesh: v^T A u equals sum_q w_q det(J_q) (J_q^-T grad_r u).(J_q^-T grad_r v)
    computed with numpy's 3x3 inverse and determinant of the Jacobian matrix
    (not the cofactor formulas the oracle uses);
  * symmetry of A, A*1 = 0 for h2 = 0, mass term = sum w det u v.
"""
import numpy as np
import pytest

from problems import Problem, box


def grads(u, D):
    """(u_r, u_s, u_t) for u[e, t, s, r]."""
    ur = np.einsum("im,etsm->etsi", D, u)
    us = np.einsum("jm,etmr->etjr", D, u)
    ut = np.einsum("km,emsr->eksr", D, u)
    return ur, us, ut


def jac_matrix(P):
    D = P.D
    rows = []
    for c in (P.x, P.y, P.zc):
        rows.append(np.stack(grads(c, D), axis=-1))  # [..., (r,s,t)]
    J = np.stack(rows, axis=-2)  # [..., (x,y,z), (r,s,t)]
    return J


def bilinear(P, u, v, h1=None):
    J = jac_matrix(P)
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)  # [..., (r,s,t), (x,y,z)]
    gu = np.stack(grads(u, P.D), axis=-1)
    gv = np.stack(grads(v, P.D), axis=-1)
    # physical gradient: grad_x u = Jinv^T grad_r u
    pu = np.einsum("...ri,...r->...i", Jinv, gu)
    pv = np.einsum("...ri,...r->...i", Jinv, gv)
    w3 = np.einsum("k,j,i->kji", P.w, P.w, P.w)
    integrand = np.sum(pu * pv, axis=-1) * det * w3[None]
    if h1 is not None:
        integrand = integrand * h1
    return float(np.sum(integrand)), det


CASES = [
    dict(nel=2, lx1=4),
    dict(nel=(3, 2, 2), lx1=6, deform=0.1),
    dict(nel=2, lx1=8, deform=0.1),
    dict(nel=(2, 3, 2), lx1=5, jitter=0.05, seed=3),
    dict(nel=2, lx1=8, jitter=0.05, deform=0.1, seed=4),
]


@pytest.mark.parametrize("case", CASES)
def test_geometry_vs_inverse_jacobian(case):
    P = Problem(box(**case))
    J = jac_matrix(P)
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)
    w3 = np.einsum("k,j,i->kji", P.w, P.w, P.w)[None]
    G = np.einsum("...ai,...bi->...ab", Jinv, Jinv) * (det * w3)[..., None, None]
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    scale = max(np.max(np.abs(G[..., a, a])) for a in range(3))
    for gi, (a, b) in zip(P.g, pairs):
        assert np.max(np.abs(gi - G[..., a, b])) <= 1e-13 * scale
    np.testing.assert_allclose(P.bm1, det * w3, rtol=1e-13)
    assert np.all(det > 0)


@pytest.mark.parametrize("case", CASES)
def test_axhelm_bilinear_form(case):
    P = Problem(box(**case))
    rng = np.random.default_rng(7)
    u = rng.standard_normal(P.shape)
    v = rng.standard_normal(P.shape)
    one = np.ones(P.shape)
    zero = np.zeros(P.shape)
    Au = P.field.axhelm(u, one, zero)
    ref, det = bilinear(P, u, v)
    got = float(np.sum(v * Au))
    assert abs(got - ref) <= 1e-12 * abs(ref) + 1e-12
    # symmetry
    Av = P.field.axhelm(v, one, zero)
    assert abs(np.sum(u * Av) - got) <= 1e-12 * abs(got)
    # constants are in the kernel (h2 = 0)
    A1 = P.field.axhelm(one, one, zero)
    assert np.max(np.abs(A1)) <= 1e-11 * np.max(np.abs(Au))
    # variable h1 and the h2 mass term
    h1 = 1.0 + 0.5 * rng.random(P.shape)
    h2 = 0.3 + rng.random(P.shape)
    Ah = P.field.axhelm(u, h1, h2)
    ref_h, _ = bilinear(P, u, v, h1=h1)
    w3 = np.einsum("k,j,i->kji", P.w, P.w, P.w)[None]
    ref_h += float(np.sum(h2 * det * w3 * u * v))
    assert abs(np.sum(v * Ah) - ref_h) <= 1e-12 * abs(ref_h)


@pytest.mark.parametrize("lx1", [3, 5, 8])
def test_axhelm_affine_kronecker(lx1):
    b = box((2, 2, 2), lx1, lengths=(1.0, 2.0, 0.5))
    P = Problem(b)
    hx, hy, hz = b.h
    K = P.D.T @ np.diag(P.w) @ P.D
    rng = np.random.default_rng(11)
    u = rng.standard_normal(P.shape)
    one = np.ones(P.shape)
    Au = P.field.axhelm(u, one, np.zeros(P.shape))
    # u[e, t, s, r]: K along one direction, the diagonal mass W along the others
    out = []
    for e in range(P.shape[0]):
        ue = u[e]
        a = (hy * hz / (2 * hx)) * np.einsum("ri,tsi->tsr", K, ue) * (P.w[None, :, None] * P.w[:, None, None])
        b_ = (hx * hz / (2 * hy)) * np.einsum("sj,tjr->tsr", K, ue) * (P.w[None, None, :] * P.w[:, None, None])
        c = (hx * hy / (2 * hz)) * np.einsum("tk,ksr->tsr", K, ue) * (P.w[None, None, :] * P.w[None, :, None])
        out.append(a + b_ + c)
    ref = np.stack(out)
    assert np.max(np.abs(Au - ref)) <= 1e-12 * np.max(np.abs(ref))
