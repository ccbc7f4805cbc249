"""Writers of Nek case files for tests: .re2 (#v002, 8-byte words) with
curve and boundary records, .ma2 and .map in genmap's layout
(genmap.f:1641-1712 dmp_mapfile), and two curved meshes whose geometry is
known in closed form:

  * annulus_sector: nr x nt x nz elements between radii r0..r1, angles
    t0..t1 and heights z0..z1; 'C' arcs on the edges along the angle
    (positive radius where the centre is on the left of the edge's corner
    order, negative where it is on the right, as Nek's preprocessor writes
    them), so every GLL node of an arc face lies on its cylinder;
  * sphere_shell: a patch of the spherical shell r0..r1 over the cube-face
    directions (a, b, 1)/|.| for |a|, |b| <= w; 's' on the inner and outer
    faces (centre 0, radius r), so every node of those faces lies on its
    sphere.

Corners are in Nek's preprocessor order (1..4 counterclockwise at t = -1,
5..8 above them); element axes r, s, t run along radius/angle/height
(annulus) and a/b/radius (shell).
"""
import struct

import numpy as np

# lexicographic corner -> preprocessor corner slot
PREP_OF_LEX = [0, 1, 3, 2, 4, 5, 7, 6]


def write_re2(fname, xc, yc, zc, curves=(), bcs=(), big_endian=False):
    """xc, yc, zc: (nel, 8) preprocessor corners; curves: (e, edge, vals5,
    char) 0-based e/edge; bcs: (e, face, vals5, cbc3) 0-based e/face."""
    e_ = ">" if big_endian else "<"
    nel = xc.shape[0]
    out = bytearray(f"#v002{nel:9d}{3:3d}{nel:9d} hdr".ljust(80).encode())
    out += struct.pack(e_ + "f", 6.54321)
    for e in range(nel):
        out += struct.pack(e_ + "25d", 0.0, *xc[e], *yc[e], *zc[e])
    out += struct.pack(e_ + "d", float(len(curves)))
    for e, k, v, c in curves:
        out += struct.pack(e_ + "7d", float(e + 1), float(k + 1), *map(float, v))
        out += c.encode() if isinstance(c, str) else c
        out += b" " * 7
    out += struct.pack(e_ + "d", float(len(bcs)))
    for e, f, v, c in bcs:
        out += struct.pack(e_ + "7d", float(e + 1), float(f + 1), *map(float, v))
        out += (c.encode() if isinstance(c, str) else c) + b"     "
    open(fname, "wb").write(bytes(out))


def write_ma2(fname, part, vertex, big_endian=False, nactive=None):
    """vertex (nel, 8) in hypercube corner order; part 0-based."""
    e_ = ">" if big_endian else "<"
    nel = len(part)
    nact = nel if nactive is None else nactive
    hdr = f"#v001{nel:12d}{nact:12d}{0:12d}{0:12d}{8 * nel:12d}{0:12d}{0:12d}".ljust(132)
    out = bytearray(hdr.encode())
    out += struct.pack(e_ + "f", 6.54321)
    for e in range(nel):
        out += struct.pack(e_ + "9i", int(part[e]), *map(int, vertex[e]))
    open(fname, "wb").write(bytes(out))


def write_map(fname, part, vertex):
    nel = len(part)
    with open(fname, "w") as fh:
        fh.write("".join(f"{v:12d}" for v in (nel, nel, 0, 0, 8 * nel, 0, 0)) + "\n")
        for e in range(nel):
            fh.write("".join(f"{int(v):12d}" for v in [part[e], *vertex[e]]) + "\n")


def _prep(lex):
    """(nel, 8) lexicographic -> preprocessor corner order"""
    out = np.zeros_like(lex)
    out[:, PREP_OF_LEX] = lex
    return out


def _wall_bcs(nel, faces):
    return [(e, f, [0.0] * 5, "W  ") for e in range(nel) for f in range(6) if faces[e, f]]


def annulus_sector(nr, nt, nz, r0=0.5, r1=1.0, t0=0.2, t1=1.3, z0=0.0, z1=0.4):
    rr = np.linspace(r0, r1, nr + 1)
    tt = np.linspace(t0, t1, nt + 1)
    zz = np.linspace(z0, z1, nz + 1)
    X, Y, Z, curves, faces = [], [], [], [], []
    e = 0
    for kz in range(nz):
        for jt in range(nt):
            for ir in range(nr):
                lx, ly, lz = [], [], []
                for c in range(8):
                    i, j, k = c & 1, (c >> 1) & 1, c >> 2
                    r, t = rr[ir + i], tt[jt + j]
                    lx.append(r * np.cos(t))
                    ly.append(r * np.sin(t))
                    lz.append(zz[kz + k])
                X.append(lx)
                Y.append(ly)
                Z.append(lz)
                # edges 2, 6 (r = +1, corner 2 -> 3, angle increasing): centre
                # on the left; edges 4, 8 (r = -1, corner 4 -> 1): on the right
                for edge, rad in ((1, rr[ir + 1]), (5, rr[ir + 1]), (3, -rr[ir]), (7, -rr[ir])):
                    curves.append((e, edge, [rad, 0, 0, 0, 0], "C"))
                f = np.zeros(6, dtype=bool)
                f[0], f[2] = jt == 0, jt == nt - 1       # s = -1, +1
                f[3], f[1] = ir == 0, ir == nr - 1       # r = -1, +1
                f[4], f[5] = kz == 0, kz == nz - 1       # t = -1, +1
                faces.append(f)
                e += 1
    nel = e
    X, Y, Z = (_prep(np.array(a)) for a in (X, Y, Z))
    return dict(xc=X, yc=Y, zc=Z, curves=curves, bcs=_wall_bcs(nel, np.array(faces)),
                nel=nel, radii=rr, angles=tt, heights=zz)


def sphere_shell(na, nr, r0=0.6, r1=1.0, w=0.5):
    aa = np.linspace(-w, w, na + 1)
    rr = np.linspace(r0, r1, nr + 1)
    X, Y, Z, curves, faces = [], [], [], [], []
    e = 0
    for kr in range(nr):
        for jb in range(na):
            for ia in range(na):
                lx, ly, lz = [], [], []
                for c in range(8):
                    i, j, k = c & 1, (c >> 1) & 1, c >> 2
                    d = np.array([aa[ia + i], aa[jb + j], 1.0])
                    p = rr[kr + k] * d / np.linalg.norm(d)
                    lx.append(p[0])
                    ly.append(p[1])
                    lz.append(p[2])
                X.append(lx)
                Y.append(ly)
                Z.append(lz)
                curves.append((e, 4, [0, 0, 0, rr[kr], 0], "s"))
                curves.append((e, 5, [0, 0, 0, rr[kr + 1], 0], "s"))
                f = np.zeros(6, dtype=bool)
                f[0], f[2] = jb == 0, jb == na - 1
                f[3], f[1] = ia == 0, ia == na - 1
                f[4], f[5] = kr == 0, kr == nr - 1
                faces.append(f)
                e += 1
    nel = e
    X, Y, Z = (_prep(np.array(a)) for a in (X, Y, Z))
    return dict(xc=X, yc=Y, zc=Z, curves=curves, bcs=_wall_bcs(nel, np.array(faces)),
                nel=nel, radii=rr)


def write_case(fname, m, big_endian=False):
    write_re2(fname, m["xc"], m["yc"], m["zc"], m["curves"], m["bcs"], big_endian)
