"""Pins the oracle's GLL basis (speclib.f restatement).

1. Against the reference's own speclib.f, compiled by flang from the Nek5000
   sources into oracle/_ref (skipped when that build is absent).
2. Against closed forms that fix the basis independently: interior GLL nodes
   are the roots of P_N', weights are 2/(N(N+1) P_N(z_i)^2), D differentiates
   every polynomial of degree <= N exactly, I12 interpolates them exactly.
"""
import ctypes
import os

import numpy as np
import pytest
from numpy.polynomial import legendre as L

NPS = [2, 3, 4, 5, 6, 8, 10, 12, 16]


@pytest.mark.parametrize("np_", NPS)
def test_against_reference_speclib(oracle, np_):
    ref = oracle.ref_speclib()
    if ref is None:
        pytest.skip("oracle/_ref/libspeclib_ref.so not built (needs the Nek5000 sources)")
    z, w = oracle.zwgll(np_)
    D = oracle.dgll(z)
    zr, wr = np.zeros(np_), np.zeros(np_)
    n = ctypes.c_int(np_)
    ref.zwgll_(ctypes.c_void_p(zr.ctypes.data), ctypes.c_void_p(wr.ctypes.data), ctypes.byref(n))
    dr, dtr = np.zeros(np_ * np_), np.zeros(np_ * np_)
    ref.dgll_(ctypes.c_void_p(dr.ctypes.data), ctypes.c_void_p(dtr.ctypes.data),
              ctypes.c_void_p(zr.ctypes.data), ctypes.byref(n), ctypes.byref(n))
    Dr = dr.reshape(np_, np_).T
    assert np.max(np.abs(z - zr)) <= 1e-15
    assert np.max(np.abs(w - wr)) <= 1e-15
    assert np.max(np.abs(D - Dr)) <= 1e-13 * np.max(np.abs(Dr))
    assert np.array_equal(dtr.reshape(np_, np_).T, Dr.T)


@pytest.mark.parametrize("np_", NPS)
def test_gll_closed_form(oracle, np_):
    N = np_ - 1
    z, w = oracle.zwgll(np_)
    cN = np.zeros(N + 1)
    cN[N] = 1.0
    inner = np.sort(L.legroots(L.legder(cN))) if N >= 2 else np.array([])
    assert z[0] == -1.0 and z[-1] == 1.0
    np.testing.assert_allclose(z[1:-1], inner, atol=1e-14)
    PN = L.legval(z, cN)
    np.testing.assert_allclose(w, 2.0 / (N * (N + 1) * PN ** 2), rtol=1e-13)
    assert abs(w.sum() - 2.0) < 1e-14


@pytest.mark.parametrize("np_", NPS)
def test_derivative_exact_on_polynomials(oracle, np_):
    N = np_ - 1
    z, _ = oracle.zwgll(np_)
    D = oracle.dgll(z)
    for k in range(0, N + 1):
        du = D @ z ** k
        exact = k * z ** (k - 1) if k > 0 else np.zeros_like(z)
        assert np.max(np.abs(du - exact)) < 1e-11 * max(1, k * k)
    # a transposed D fails this for any N >= 2
    if N >= 2:
        assert np.max(np.abs(D.T @ z ** 2 - 2 * z)) > 1e-3


@pytest.mark.parametrize("n1,n2", [(8, 4), (4, 2), (8, 10), (6, 3)])
def test_interpolation_exact(oracle, n1, n2):
    z1, _ = oracle.zwgll(n1)
    z2, _ = oracle.zwgll(n2)
    I = oracle.igllm(z1, z2)
    for k in range(n1):
        np.testing.assert_allclose(I @ z1 ** k, z2 ** k, atol=1e-12)
