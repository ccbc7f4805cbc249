"""The Fortran side of the drop-in: nek5000_amd/fortran (the ISO_C_BINDING
module and the shims with Nek's signatures) compiled with flang and Nek's
flags against stand-in commons (tests/fortran), linked with libnekamd.so.

CPU: the bind(c) types have the C sizes; nkh_zwgll and nkh_setvert3d called
from Fortran equal the oracle; dssum before any set-up goes to Nek's own
gather-scatter handle of the field (gsh_fld(ifield), as dssum.f:45-77 picks
it); a failing set-up ends in Nek's EXIT line through exitti.
GPU: a case through the shims (hmholtz -> Jacobi PCG, hmholtz('PRES') ->
flexible CG with h1mg, then axhelm and dssum on the resident element set),
against the oracle; and undeformed elements (IFDFRM false, g4..g6
overwritten the way sfastax leaves them) solved as the reference does.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "fortran"))

import fortran_build as fbuild  # noqa: E402

pytestmark = pytest.mark.skipif(fbuild.flang() is None, reason="no flang")


def _stale():
    """the driver is missing or older than a source it is built from"""
    exe = fbuild.EXE
    if not os.path.exists(exe):
        return True
    t = os.path.getmtime(exe)
    srcs = [os.path.join(fbuild.SRC, f) for f in os.listdir(fbuild.SRC)]
    srcs += [os.path.join(fbuild.HERE, f) for f in ("drive.f", "nek_stubs.f")]
    return any(os.path.getmtime(f) > t for f in srcs if os.path.isfile(f))


def drive(*args, check=True):
    exe = fbuild.EXE
    if _stale():
        fbuild.build()
    r = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=300)
    if check:
        assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    return r


def test_bind_c_type_sizes():
    from nek5000_amd import _lib

    out = drive("sizes").stdout.split()
    assert out[0] == "sizes"
    assert int(out[1]) == ctypes.sizeof(_lib.NkhParams) == 72
    assert int(out[2]) == ctypes.sizeof(_lib.NkhResult) == 32


@pytest.mark.parametrize("np_", [2, 5, 8, 16])
def test_zwgll_through_fortran(oracle, np_):
    rows = [ln.split() for ln in drive("zwgll", np_).stdout.splitlines() if ln.startswith("gll")]
    z = np.array([float(r[2]) for r in rows])
    w = np.array([float(r[3]) for r in rows])
    zo, wo = oracle.zwgll(np_)
    assert np.array_equal(z, zo) and np.array_equal(w, wo)


@pytest.mark.parametrize("case", [dict(nel=(3, 2, 2), lx1=8),
                                  dict(nel=(4, 2, 2), lx1=8, periodic_x=True)])
def test_setvert3d_through_fortran(oracle, tmp_path, case):
    from problems import box

    b = box(**case)
    fin, fout = tmp_path / "vtx.bin", tmp_path / "glo.bin"
    with open(fin, "wb") as fh:
        np.array([b.nel], np.int32).tofile(fh)
        np.ascontiguousarray(b.vertex, np.int64).tofile(fh)
    drive("setvert", fin, fout)
    raw = np.fromfile(fout, np.int64)
    glo_o, ngv_o = oracle.setvert3d(b.vertex, b.lx1)
    assert raw[0] == ngv_o
    assert np.array_equal(raw[1:], np.ascontiguousarray(glo_o).reshape(-1))


def test_dssum_before_setup_uses_nek_handle():
    lines = [ln.split() for ln in drive("dssum0").stdout.splitlines() if ln.startswith("gs_op")]
    # ifield 1 -> gsh_fld(1) = 11, 2 -> gsh_fld(2) = 12, 3 -> gsh_fld(3) = 11;
    # op 1 (+) on real data, not transposed
    assert [int(r[1]) for r in lines] == [11, 12, 11]
    assert all(r[2:] == ["1", "1", "0"] for r in lines)


def test_setup_failure_exits_through_exitti():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present: set-up does not fail")
    r = drive("nogpu", check=False)
    assert r.returncode == 3, (r.stdout, r.stderr)
    assert "EXIT: nekamd nkh_init" in r.stdout


# ---- GPU: a case through the shims ----------------------------------------
def write_case(path, P, name, rhs, h1, h2, tli, maxit, isplit=0, p42=0, idef=None, g=None,
               fbc=None):
    b = P.box
    n = P.bm1.size
    g = P.g if g is None else g
    idef = np.ones(b.nel, np.int32) if idef is None else np.asarray(idef, np.int32)
    fbc = np.zeros((b.nel, 6), np.int32) if fbc is None else np.asarray(fbc, np.int32)
    ds = P.uniform_field(77)
    with open(path, "wb") as fh:
        np.array([b.nel, isplit, p42, maxit], np.int32).tofile(fh)
        np.array([tli], np.float64).tofile(fh)
        fh.write(name.encode()[:4].ljust(4))
        np.ascontiguousarray(b.vertex, np.int64).tofile(fh)
        idef.tofile(fh)
        np.ascontiguousarray(fbc, np.int32).tofile(fh)
        for a in (P.x, P.y, P.zc) + tuple(g) + (P.bm1, rhs, h1, h2, P.mask, P.mult, ds):
            np.ascontiguousarray(a, np.float64).reshape(-1).tofile(fh)
    return ds


def read_out(path, n):
    raw = open(path, "rb").read()
    niter = int(np.frombuffer(raw[:4], np.int32)[0])
    a = np.frombuffer(raw[4:], np.float64)
    return niter, a[:n], a[n:2 * n], a[2 * n:3 * n]


def l2rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.gpu
def test_hmholtz_through_shims(gpu, tmp_path):
    from problems import Problem, box

    P = Problem(box((4, 3, 3), 8, deform=0.1))
    rhs = P.bm1 * P.uniform_field(5)
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, 10.0)
    ds = write_case(tmp_path / "in.bin", P, "TEMP", rhs, h1, h2, -1e-8, 900)
    r = drive("solve", tmp_path / "in.bin", tmp_path / "out.bin")
    print(r.stdout, flush=True)
    niter, u, au, dsg = read_out(tmp_path / "out.bin", P.bm1.size)
    u_o, _, n_o, _, _ = P.field.hmholtz(rhs, h1, h2, -1e-8, 900)
    assert abs(niter - n_o) <= 1
    assert l2rel(u, u_o.reshape(-1)) < 1e-10
    au_o = P.field.axhelm(u.reshape(P.shape), h1, h2).reshape(-1)
    assert np.max(np.abs(au - au_o)) <= 1e-13 * np.max(np.abs(au_o))
    assert np.array_equal(dsg, P.field.dssum(ds).reshape(-1))


@pytest.mark.gpu
def test_pressure_through_shims(gpu, tmp_path):
    from oracle import oracle as O
    from problems import Problem, box, mg_for

    P = Problem(box((3, 3, 2), 8, deform=0.1))
    rhs = P.bm1 * P.uniform_field(6)
    h1 = np.ones(P.shape)
    h2 = np.zeros(P.shape)
    write_case(tmp_path / "in.bin", P, "PRES", rhs, h1, h2, -1e-8, 100, isplit=1, p42=2,
               fbc=O.fbc_from_box(P.box, 1))
    r = drive("solve", tmp_path / "in.bin", tmp_path / "out.bin")
    print(r.stdout, flush=True)
    niter, u, _, _ = read_out(tmp_path / "out.bin", P.bm1.size)
    mg = mg_for(P)
    u_o, n_o, _, _ = mg.flex_cg(P.field.dssum(rhs) * P.mask, h1, h2, 100, -1e-8)
    assert abs(niter - n_o) <= 1
    assert l2rel(u, u_o.reshape(-1)) < 1e-10


@pytest.mark.gpu
def test_undeformed_elements_through_shims(gpu, tmp_path):
    """param(59) = 0 on a box of rectangular elements: IFDFRM false
    everywhere and g4m1..g6m1 overwritten by sfastax (hmholtz.f:353-364,
    g4 = g1/w_x, g5 = g2/w_y, g6 = g3/w_z); the reference's axhelm and setprec
    then drop the cross terms, whose true values are zero here, so the solve
    equals the oracle's on the true factors."""
    from problems import Problem, box

    P = Problem(box((3, 2, 2), 8))
    w = P.w
    g = list(P.g)
    g[3] = g[0] / w[None, None, None, :]
    g[4] = g[1] / w[None, None, :, None]
    g[5] = g[2] / w[None, :, None, None]
    rhs = P.bm1 * P.uniform_field(8)
    h1 = np.ones(P.shape)
    h2 = np.full(P.shape, 2.0)
    write_case(tmp_path / "in.bin", P, "TEMP", rhs, h1, h2, -1e-9, 900,
               idef=np.zeros(P.box.nel), g=g)
    drive("solve", tmp_path / "in.bin", tmp_path / "out.bin")
    niter, u, au, _ = read_out(tmp_path / "out.bin", P.bm1.size)
    u_o, _, n_o, _, _ = P.field.hmholtz(rhs, h1, h2, -1e-9, 900)
    assert abs(niter - n_o) <= 1
    assert l2rel(u, u_o.reshape(-1)) < 1e-10
    au_o = P.field.axhelm(u.reshape(P.shape), h1, h2).reshape(-1)
    assert np.max(np.abs(au - au_o)) <= 1e-13 * np.max(np.abs(au_o))


def test_unbuilt_cases_route_to_reference():
    """Axisymmetric meshes (ifaxis) and the pressure with param(42) = 1 go to
    the reference's own routines under their renamed symbols, before any
    library call (this runs without a GPU: a set-up would end in EXIT)."""
    r = drive("route")
    lines = [ln.split() for ln in r.stdout.splitlines()]
    assert ["nek_hmholtz", "VELX", "1", "1"] in lines
    assert ["nek_axhelm", "1", "1"] in lines
    assert ["nek_hsolve", "TEMP", "2", "1"] in lines
    assert ["nek_hmholtz", "PRES", "1", "1"] in lines
    assert lines[-1] == ["routed"] and "EXIT" not in r.stdout
