"""The library's mesh ingestion (nkh_re2_header/read, nkh_mesh_vertex_ids,
nkh_mesh_xyz_from_corners) on the reference's own .re2 files and on
synthetic ones, against the oracle's numpy restatement and what the meshes
fix:

  * every array the reader returns equals the oracle's reading, bit for bit
    (#v001 4-byte and #v002/#v003 8-byte words, both byte orders);
  * stated facts of the files (element counts, curve and boundary records);
  * vertex ids equal the oracle's brute-force coincidence classes; every
    face without a boundary condition is shared with exactly one other
    element face, every face with one is not;
  * GLL coordinates equal the oracle's linquad; on 'm' edges the GLL node
    at the edge midpoint (odd lx1) is the stored midside node; a box written
    as .re2 reproduces BoxMesh's coordinates and, with periodic faces, its
    vertex classes.
"""
import os
import struct

import numpy as np
import pytest

from nek5000_amd import _lib
from nek5000_amd.mesh import BoxMesh, Re2Mesh
from oracle import mesh_oracle as MO

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "re2")
FILES = {
    "ethier.re2": dict(version="#v003", nel=32, ndim=3, ncurve=96, nfields=1,
                       bcs=[{b"EXO": 24}]),
    "io_test.re2": dict(version="#v002", nel=36, ndim=3, ncurve=0, nfields=4,
                        bcs=[{b"W  ": 66}] + [{b"I  ": 66}] * 3),
    "lpm_one_uniform.re2": dict(version="#v002", nel=125, ndim=3, ncurve=0, nfields=1,
                                bcs=[{b"W  ": 150}]),
    "lowMach_test.re2": dict(version="#v001", nel=64, ndim=2, ncurve=0, nfields=2,
                             bcs=[{b"P  ": 16, b"v  ": 16}, {b"P  ": 16, b"t  ": 16}]),
    "double_shear.re2": dict(version="#v002", nel=64, ndim=2, ncurve=0, nfields=1,
                             bcs=[{b"P  ": 32}]),
}
THREE_D = [k for k, v in FILES.items() if v["ndim"] == 3]


def path(name):
    return os.path.join(GOLD, name)


@pytest.mark.parametrize("name", sorted(FILES))
def test_reader_matches_oracle_and_facts(name):
    m = Re2Mesh(path(name), 4)
    o = MO.read_re2(path(name))
    facts = FILES[name]
    assert o["version"] == facts["version"]
    assert m.nel == o["nelgt"] == facts["nel"]
    assert m.ndim == facts["ndim"] and m.nfields == len(o["cbc"]) == facts["nfields"]
    assert o["ncurve"] == facts["ncurve"]
    assert np.array_equal(m.group, o["group"])
    assert np.array_equal(m.xc, o["xc"]) and np.array_equal(m.yc, o["yc"])
    if m.ndim == 3:
        assert np.array_equal(m.zc, o["zc"])
    assert np.array_equal(m.curve, o["curve"]) and np.array_equal(m.ccurve, o["ccurve"])
    for f in range(m.nfields):
        assert np.array_equal(m.cbc[f], o["cbc"][f])
        assert np.array_equal(m.bc[f], o["bc"][f])
        vals, cnt = np.unique(o["cbc"][f][o["cbc"][f] != b"E  "], return_counts=True)
        assert dict(zip(vals.tolist(), cnt.tolist())) == facts["bcs"][f]


@pytest.mark.parametrize("name", THREE_D)
def test_vertex_ids_and_face_sharing(name):
    m = Re2Mesh(path(name), 4)
    o = MO.read_re2(path(name))
    vo, nvo = MO.vertex_ids(o["xc"], o["yc"], o["zc"], o["cbc"][0], o["bc"][0])
    assert np.array_equal(m.vertex.reshape(m.nel, 8), vo) and m.nvert == nvo
    faces = {}
    for e in range(m.nel):
        for f in range(6):
            key = tuple(sorted(m.vertex.reshape(m.nel, 8)[e, MO.FACE_CORNERS[f]]))
            faces.setdefault(key, []).append((e, f))
    for e in range(m.nel):
        for f in range(6):
            key = tuple(sorted(m.vertex.reshape(m.nel, 8)[e, MO.FACE_CORNERS[f]]))
            shared = len(faces[key])
            if m.cbc[0][e, f] == b"E  ":
                assert shared == 2, (e, f)
            else:
                assert shared == 1, (e, f, m.cbc[0][e, f])


@pytest.mark.parametrize("name", THREE_D)
@pytest.mark.parametrize("lx1", [2, 5, 8])
def test_gll_coordinates(name, lx1):
    m = Re2Mesh(path(name), lx1)
    o = MO.read_re2(path(name))
    xo = MO.xyz_from_corners(lx1, o["xc"], o["yc"], o["zc"], o["ccurve"], o["curve"])
    for a, b in zip(m.xyz(), xo):
        assert np.max(np.abs(a - b)) <= 1e-14 * max(1.0, np.max(np.abs(b)))
    # corners of the GLL grid are the element corners
    x = m.xyz()[0]
    lex = MO.lex_corners(o["xc"], o["yc"], o["zc"])[:, :, 0].reshape(m.nel, 2, 2, 2)
    n = lx1 - 1
    assert np.allclose(x[:, ::n, ::n, ::n], lex, rtol=0, atol=1e-14)


def test_midside_nodes_are_interpolated():
    lx1 = 7  # odd: the GLL points include the edge midpoints
    m = Re2Mesh(path("ethier.re2"), lx1)
    xyz = m.xyz()
    mid = lx1 // 2
    # preprocessor edge -> (k, j, i) of its midpoint in the GLL grid
    pos = {0: (0, 0, mid), 1: (0, mid, -1), 2: (0, -1, mid), 3: (0, mid, 0),
           4: (-1, 0, mid), 5: (-1, mid, -1), 6: (-1, -1, mid), 7: (-1, mid, 0),
           8: (mid, 0, 0), 9: (mid, 0, -1), 10: (mid, -1, -1), 11: (mid, -1, 0)}
    nm = 0
    for e in range(m.nel):
        for k in range(12):
            if m.ccurve[e, k] == b"m":
                kk, jj, ii = pos[k]
                for d in range(3):
                    assert abs(xyz[d][e, kk, jj, ii] - m.curve[e, k, d]) <= 1e-14
                nm += 1
    assert nm == 96


def write_box_re2(fname, box: BoxMesh, big_endian=False, periodic=False):
    """A .re2 (#v002) of a BoxMesh: corners in preprocessor order, 'W  '
    on domain faces, 'P  ' pairs on x faces when periodic."""
    e_ = ">" if big_endian else "<"
    X, Y, Z = box._vertex_coords()
    nel = box.nel
    hdr = f"#v002{nel:9d}{3:3d}{nel:9d} hdr".ljust(80).encode()
    out = bytearray(hdr)
    out += struct.pack(e_ + "f", 6.54321)
    for e in range(nel):
        lex = [X[e].reshape(-1), Y[e].reshape(-1), Z[e].reshape(-1)]
        rec = [0.0]
        for d in range(3):
            prep = np.zeros(8)
            prep[MO.PREP_OF_LEX] = lex[d]
            rec += prep.tolist()
        out += struct.pack(e_ + "25d", *rec)
    out += struct.pack(e_ + "d", 0.0)  # no curves
    recs = []
    f = box.dirichlet_faces()
    ex = np.arange(nel) % box.nelx
    for e in range(nel):
        for fc in range(6):
            if periodic and fc in (1, 3) and ((fc == 1 and ex[e] == box.nelx - 1) or
                                              (fc == 3 and ex[e] == 0)):
                e2 = e - (box.nelx - 1) if fc == 1 else e + (box.nelx - 1)
                f2 = 3 if fc == 1 else 1
                recs.append((e + 1, fc + 1, [e2 + 1, f2 + 1, 0, 0, 0], b"P  "))
            elif f[e, fc]:
                recs.append((e + 1, fc + 1, [0.0] * 5, b"W  "))
    out += struct.pack(e_ + "d", float(len(recs)))
    for eg, fc, bl, cb in recs:
        out += struct.pack(e_ + "7d", float(eg), float(fc), *map(float, bl)) + cb + b"     "
    open(fname, "wb").write(bytes(out))


@pytest.mark.parametrize("big_endian", [False, True])
@pytest.mark.parametrize("periodic", [False, True])
def test_box_written_as_re2(tmp_path, big_endian, periodic):
    # (jitter keys the two images of a periodic vertex apart, so the
    # periodic box stays unjittered)
    box = BoxMesh(4, 3, 2, 6, jitter=0.0 if periodic else 0.1, seed=3, periodic_x=periodic)
    fn = str(tmp_path / "box.re2")
    write_box_re2(fn, box, big_endian, periodic)
    m = Re2Mesh(fn, 6)
    o = MO.read_re2(fn)
    assert np.array_equal(m.xc, o["xc"]) and np.array_equal(m.cbc[0], o["cbc"][0])
    for a, b in zip(m.xyz(), box.xyz()):
        assert np.max(np.abs(a - b)) <= 1e-14
    # same vertex classes as the box's own numbering
    vb = box.vertex.reshape(box.nel, 8)
    vm = m.vertex.reshape(box.nel, 8)
    pairs = {}
    for a, b in zip(vb.reshape(-1), vm.reshape(-1)):
        assert pairs.setdefault(a, b) == b
    assert len(set(pairs.values())) == len(pairs)
    # Dirichlet faces: the 'W  ' ones (periodic x faces are not)
    assert np.array_equal(m.dirichlet_faces(), box.dirichlet_faces())


def test_reader_errors(tmp_path):
    lib = _lib.load()
    bad = tmp_path / "bad.re2"
    bad.write_bytes(b"#x002" + b" " * 100)
    n = __import__("ctypes").c_int()
    assert lib.nkh_re2_header(str(bad).encode(), n, n, n, n) == 1
    data = open(path("lpm_one_uniform.re2"), "rb").read()
    trunc = tmp_path / "trunc.re2"
    trunc.write_bytes(data[:-10])
    assert lib.nkh_re2_header(str(trunc).encode(), n, n, n, n) == 1
    assert b"truncated" in lib.nkh_last_error() or b"trailing" in lib.nkh_last_error()
    m = Re2Mesh(path("ethier.re2"), 4)
    m.ccurve[0, 0] = b"e"  # genxyz.f gensrf: not built
    with pytest.raises(_lib.NekAmdError, match="curve type"):
        m.xyz()
    m = Re2Mesh(path("ethier.re2"), 4)
    m.ccurve[0, 9] = b"C"  # arcs only on edges 1..8
    with pytest.raises(_lib.NekAmdError, match="curve type"):
        m.xyz()


def test_volumes():
    """The ethier mesh is the unit ball (volume 4 pi / 3; its 'm' edges give
    a quadratic approximation of the sphere); lpm_one's is the unit cube."""
    from problems import Problem

    P = Problem(Re2Mesh(path("ethier.re2"), 8))
    assert abs(P.field.vol - 4.0 * np.pi / 3.0) < 0.02 * 4.0 * np.pi / 3.0
    assert np.min(P.jac) > 0
    P = Problem(Re2Mesh(path("lpm_one_uniform.re2"), 6))
    assert abs(P.field.vol - 1.0) < 1e-13
