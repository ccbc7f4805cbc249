"""numpy restatement of Nek5000's mesh ingestion, the checker for the
library's host-side reader (nek5000_amd/csrc/mesh_io.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/ -- never by the nek5000_amd
package.

  * read_re2: reader_re2.f:2-620 (read_re2_hdr, readp_re2_mesh/curve/bc,
    buf_to_xyz/curve/bc) for #v001..#v004 files in either byte order;
  * read_map: map2.f:712-942 read_map (.map text, .ma2 binary in either
    byte order) on one rank;
  * vertex_ids: corner coincidence by brute force over all corner pairs (an
    independent algorithm from the library's sorted sweep), periodic faces
    joined after the translation between them;
  * xyz_from_corners: genxyz.f:1235-1400 linquad (xyzlin, xyzquad) with
    navier5.f:2422-2560 gh_face_extend_3d, written as dense tensor products;
    then genxyz.f:562-605's 's' faces (sphsrf :677-845) and 'C' edges
    (arcsrf :2-103), restated geometrically: face points where the great
    circles through opposite edge points meet (the intersection on the
    corners' side of the sphere), edge arcs by rotation about the centre
    from the first corner, deflections blended linearly to the opposite face
    (sphere) or with the two faces' linear weights (arc).
"""
import numpy as np

# genxyz.f:1289-1291 indx: lexicographic corner -> preprocessor corner
PREP_OF_LEX = [0, 1, 3, 2, 4, 5, 7, 6]
# genxyz.f:1349-1353 eindx: preprocessor edge -> node of the 3x3x3 array
EDGE_NODE = [1, 5, 7, 3, 19, 23, 25, 21, 9, 11, 17, 15]
# preprocessor faces (s-, r+, s+, r-, t-, t+) -> lexicographic corners
FACE_CORNERS = [[0, 1, 4, 5], [1, 3, 5, 7], [2, 3, 6, 7], [0, 2, 4, 6], [0, 1, 2, 3],
                [4, 5, 6, 7]]


def read_re2(path):
    data = open(path, "rb").read()
    hdr = data[:80].decode("ascii", errors="replace")
    ver = hdr[:5]
    if ver == "#v004":
        parts = hdr.split()
        nelgt, ndim, nelgv = int(parts[1]), int(parts[2]), int(parts[3])
    else:
        nelgt, ndim, nelgv = int(hdr[5:14]), int(hdr[14:17]), int(hdr[17:26])
    ws = 8 if ver in ("#v002", "#v003", "#v004") else 4
    little = abs(np.frombuffer(data[80:84], "<f4")[0] - 6.54321) < 1e-5
    e = "<" if little else ">"
    real = np.dtype(e + ("f8" if ws == 8 else "f4"))
    intw = np.dtype(e + ("f8" if ws == 8 else "i4"))
    nv = 2 ** ndim
    rec = 1 + ndim * nv
    off = 84
    m = np.frombuffer(data[off:off + nelgt * rec * ws], real).reshape(nelgt, rec)
    mi = np.frombuffer(data[off:off + nelgt * rec * ws], intw).reshape(nelgt, rec)
    off += nelgt * rec * ws
    out = dict(version=ver, nelgt=nelgt, ndim=ndim, nelgv=nelgv,
               group=mi[:, 0].astype(np.int64),
               xc=m[:, 1:1 + nv].astype(np.float64), yc=m[:, 1 + nv:1 + 2 * nv].astype(np.float64))
    if ndim == 3:
        out["zc"] = m[:, 1 + 2 * nv:1 + 3 * nv].astype(np.float64)

    def count(o):
        return int(np.frombuffer(data[o:o + ws], intw)[0])

    ncur = count(off)
    off += ws
    curve = np.zeros((nelgt, 12, 5))
    ccurve = np.full((nelgt, 12), b" ", dtype="S1")
    for q in range(ncur):
        r = data[off + q * 8 * ws: off + (q + 1) * 8 * ws]
        ri = np.frombuffer(r, intw)
        rr = np.frombuffer(r, real)
        eg, ed = int(ri[0]), int(ri[1])
        curve[eg - 1, ed - 1] = rr[2:7]
        ccurve[eg - 1, ed - 1] = r[7 * ws:7 * ws + 1]
    off += ncur * 8 * ws
    cbcs, bcs = [], []
    while off < len(data):
        nbc = count(off)
        off += ws
        cbc = np.full((nelgt, 6), b"E  ", dtype="S3")
        bc = np.zeros((nelgt, 6, 5))
        for q in range(nbc):
            r = data[off + q * 8 * ws: off + (q + 1) * 8 * ws]
            ri = np.frombuffer(r, intw)
            rr = np.frombuffer(r, real)
            eg, f = int(ri[0]), int(ri[1])
            bc[eg - 1, f - 1] = rr[2:7]
            cbc[eg - 1, f - 1] = r[7 * ws:7 * ws + 3]
        off += nbc * 8 * ws
        cbcs.append(cbc)
        bcs.append(bc)
    assert off == len(data)
    out.update(curve=curve, ccurve=ccurve, cbc=cbcs, bc=bcs, ncurve=ncur)
    return out


def read_map(path, ndim=3):
    """map2.f:712-942 read_map for one rank: (nel, nactive, part (nel,),
    vertex (nel, 2^ndim)) from a .map (text) or .ma2 (binary) file."""
    nv = 2 ** ndim
    if path.endswith(".ma2"):
        data = open(path, "rb").read()
        hdr = data[:132].decode("ascii")
        nel, nact = int(hdr[5:17]), int(hdr[17:29])
        little = abs(np.frombuffer(data[132:136], "<f4")[0] - 6.54321) <= 2e-4
        w = np.frombuffer(data[136:], ("<" if little else ">") + "i4")
        assert w.size == nel * (1 + nv)
        w = w.reshape(nel, 1 + nv).astype(np.int64)
    else:
        lines = open(path).read().split("\n")
        nel, nact = (int(t) for t in lines[0].split()[:2])
        w = np.array([int(t) for ln in lines[1:] for t in ln.split()], np.int64)
        assert w.size == nel * (1 + nv)
        w = w.reshape(nel, 1 + nv)
    return nel, nact, w[:, 0].astype(np.int32), w[:, 1:]


def lex_corners(xc, yc, zc):
    """(nel, 8, 3) corner coordinates in lexicographic order."""
    c = np.stack([xc, yc, zc], axis=-1)
    return c[:, PREP_OF_LEX, :]


def vertex_ids(xc, yc, zc, cbc=None, bc=None, tol_rel=1e-2):
    """Brute-force corner coincidence; returns (vertex (nel,8) int64, nvert)."""
    P = lex_corners(xc, yc, zc).reshape(-1, 3)
    nel = xc.shape[0]
    edges = []
    for a in range(8):
        for d in range(3):
            b = a ^ (1 << d)
            if b > a:
                edges.append((a, b))
    L = lex_corners(xc, yc, zc)
    hmin = min(np.min(np.linalg.norm(L[:, a] - L[:, b], axis=1)) for a, b in edges)
    tol = tol_rel * hmin
    parent = np.arange(P.shape[0])

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    def unite(a, b):
        a, b = find(a), find(b)
        if a != b:
            parent[max(a, b)] = min(a, b)

    d = np.max(np.abs(P[:, None, :] - P[None, :, :]), axis=2)
    ii, jj = np.nonzero(d <= tol)
    for a, b in zip(ii, jj):
        unite(a, b)
    if cbc is not None:
        for e in range(nel):
            for f in range(6):
                if cbc[e, f] != b"P  ":
                    continue
                e2, f2 = int(bc[e, f, 0]) - 1, int(bc[e, f, 1]) - 1
                A = L[e, FACE_CORNERS[f]]
                B = L[e2, FACE_CORNERS[f2]]
                A0 = A - A.mean(axis=0)
                B0 = B - B.mean(axis=0)
                for k in range(4):
                    k2 = int(np.argmin(np.linalg.norm(B0 - A0[k], axis=1)))
                    unite(e * 8 + FACE_CORNERS[f][k], e2 * 8 + FACE_CORNERS[f2][k2])
    ids = np.zeros(P.shape[0], dtype=np.int64)
    nxt = 0
    for q in range(P.shape[0]):
        r = find(q)
        if ids[r] == 0:
            nxt += 1
            ids[r] = nxt
        ids[q] = ids[r]
    return ids.reshape(nel, 8), nxt


def _lagrange(x, nodes):
    """W[i, j] = l_j(x_i)."""
    x = np.asarray(x, dtype=np.float64)
    W = np.ones((x.size, len(nodes)))
    for j, zj in enumerate(nodes):
        for k, zk in enumerate(nodes):
            if k != j:
                W[:, j] *= (x - zk) / (zj - zk)
    return W


def _gh_edges(q):
    """gh_face_extend_3d, gh_type 2, on q[k, j, i] (3x3x3)."""
    zg = np.array([-1.0, 0.0, 1.0])
    h = np.stack([(1 - zg) / 2, (1 + zg) / 2])  # h[end, i]
    ends = [0, 2]
    v = np.zeros((3, 3, 3))
    for a, kk in enumerate(ends):
        for b, jj in enumerate(ends):
            for c, ii in enumerate(ends):
                v += q[kk, jj, ii] * np.einsum("k,j,i->kji", h[a], h[b], h[c])
    e = np.zeros((3, 3, 3))
    for a, kk in enumerate(ends):      # x-edges
        for b, jj in enumerate(ends):
            e += np.einsum("k,j,i->kji", h[a], h[b], q[kk, jj, :] - v[kk, jj, :])
    for a, kk in enumerate(ends):      # y-edges
        for c, ii in enumerate(ends):
            e += np.einsum("k,j,i->kji", h[a], q[kk, :, ii] - v[kk, :, ii], h[c])
    for b, jj in enumerate(ends):      # z-edges
        for c, ii in enumerate(ends):
            e += np.einsum("k,j,i->kji", q[:, jj, ii] - v[:, jj, ii], h[b], h[c])
    return v + e


# preprocessor face -> symmetric face (connect1.f:200-205), 1-based, and
# symmetric face -> (axis r/s/t = 0/1/2, side -1/+1 = 0/1)
EFACE1 = [3, 2, 4, 1, 5, 6]
SYM_FACE = {1: (0, 0), 2: (0, 1), 3: (1, 0), 4: (1, 1), 5: (2, 0), 6: (2, 1)}


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _slerp(u1, u2, t):
    th = np.arccos(np.clip(np.dot(u1, u2), -1.0, 1.0))
    return (np.sin((1 - t)[:, None] * th) * u1 + np.sin(t[:, None] * th) * u2) / np.sin(th)


def _node(ax, m, ta, p, q):
    idx = [0, 0, 0]
    idx[ax], idx[ta[0]], idx[ta[1]] = m, p, q
    return idx[2], idx[1], idx[0]  # [k, j, i]


def sphere_face(X, z, lexc, crv, ifce):
    """Face ifce (preprocessor, 0-based) of one element onto the sphere
    crv = (xc, yc, zc, R); X = [x, y, z] arrays [k, j, i], changed in place;
    lexc (8, 3) the corners in lexicographic order."""
    n = len(z)
    ax, side = SYM_FACE[EFACE1[ifce]]
    ta = [a for a in range(3) if a != ax]
    c, R = np.asarray(crv[:3], float), float(crv[3])
    assert R > 0

    def corner(p, q):
        idx = [0, 0, 0]
        idx[ax], idx[ta[0]], idx[ta[1]] = side, p, q
        return lexc[idx[0] + 2 * idx[1] + 4 * idx[2]] - c

    u = {(p, q): _unit(corner(p, q)) for p in (0, 1) for q in (0, 1)}
    for p in (0, 1):
        for q in (0, 1):
            assert abs(np.linalg.norm(corner(p, q)) - R) <= 1e-4 * R
    t = (z + 1) / 2
    P = np.zeros((n, n, 3))  # [q, p]
    P[0, :] = _slerp(u[0, 0], u[1, 0], t)
    P[n - 1, :] = _slerp(u[0, 1], u[1, 1], t)
    P[:, 0] = _slerp(u[0, 0], u[0, 1], t)
    P[:, n - 1] = _slerp(u[1, 0], u[1, 1], t)
    mean = sum(u.values())
    for q in range(1, n - 1):
        n1 = np.cross(P[q, 0], P[q, n - 1])
        for p in range(1, n - 1):
            d = np.cross(n1, np.cross(P[0, p], P[n - 1, p]))
            P[q, p] = d if np.dot(d, mean) > 0 else -d
    P = c + R * _unit(P)
    w = (1 + z) / 2 if side == 1 else (1 - z) / 2
    f = side * (n - 1)
    for q in range(n):
        for p in range(n):
            fk = _node(ax, f, ta, p, q)
            dx = [P[q, p, d] - X[d][fk] for d in range(3)]
            for m in range(n):
                nk = _node(ax, m, ta, p, q)
                for d in range(3):
                    X[d][nk] += dx[d] * w[m]


def arc_edge(X, z, xp, yp, radius, isid):
    """Edge isid (1..8, preprocessor) of one element bent into the circle
    of radius |radius| through its corners (centre to the left of corner
    isid -> next for radius > 0); xp, yp the preprocessor corners."""
    n = len(z)
    i1 = isid - 1
    i2 = 0 if isid == 4 else 4 if isid == 8 else isid
    p1, p2 = np.array([xp[i1], yp[i1]]), np.array([xp[i2], yp[i2]])
    g = np.linalg.norm(p2 - p1)
    assert 2 * abs(radius) > g * 1.00001
    tang = (p2 - p1) / g
    left = np.array([-tang[1], tang[0]])
    h = np.sqrt(radius ** 2 - g ** 2 / 4)
    ctr = (p1 + p2) / 2 + np.sign(radius) * h * left
    sweep = 2 * np.arcsin(g / (2 * abs(radius))) * np.sign(radius)
    t = (z + 1) / 2
    a = (p1 - ctr)
    ang = sweep * t
    arc = ctr + np.stack([np.cos(ang) * a[0] - np.sin(ang) * a[1],
                          np.sin(ang) * a[0] + np.cos(ang) * a[1]], axis=1)
    pert = arc - ((1 - t)[:, None] * p1 + t[:, None] * p2)  # along corner isid -> next
    s1 = (isid - 1) % 4  # 0: r up at s=-1, 1: s up at r=+1, 2: r down at s=+1, 3: s down at r=-1
    if s1 >= 2:
        pert = pert[::-1]
    wt = (1 - z) / 2 if isid <= 4 else (1 + z) / 2
    wlo, whi = (1 - z) / 2, (1 + z) / 2
    for d in range(2):
        if s1 == 0:
            X[d] += np.einsum("k,j,i->kji", wt, wlo, pert[:, d])
        elif s1 == 2:
            X[d] += np.einsum("k,j,i->kji", wt, whi, pert[:, d])
        elif s1 == 1:
            X[d] += np.einsum("k,j,i->kji", wt, pert[:, d], whi)
        else:
            X[d] += np.einsum("k,j,i->kji", wt, pert[:, d], wlo)


def xyz_from_corners(lx1, xc, yc, zc, ccurve=None, curve=None, z=None):
    """GLL coordinates (nel, lx1, lx1, lx1) [k, j, i] for x, y, z."""
    from oracle import oracle as O

    if z is None:
        z, _ = O.zwgll(lx1)
    nel = xc.shape[0]
    Jl = _lagrange(z, [-1.0, 1.0])
    Jq = _lagrange(z, [-1.0, 0.0, 1.0])
    Jlq = _lagrange([-1.0, 0.0, 1.0], [-1.0, 1.0])
    out = [np.zeros((nel, lx1, lx1, lx1)) for _ in range(3)]
    L = lex_corners(xc, yc, zc)  # (nel, 8, 3)
    for e in range(nel):
        quad = ccurve is not None and lx1 > 2 and any(c == b"m" for c in ccurve[e])
        for d in range(3):
            cb = L[e, :, d].reshape(2, 2, 2)  # [k, j, i]
            if quad:
                q = np.einsum("ai,bj,ck,kji->cba", Jlq, Jlq, Jlq, cb)
                qf = q.reshape(-1)
                for k in range(12):
                    if ccurve[e, k] == b"m":
                        qf[EDGE_NODE[k]] = curve[e, k, d]
                q = _gh_edges(qf.reshape(3, 3, 3))
                J = Jq
            else:
                q = cb
                J = Jl
            out[d][e] = np.einsum("ia,jb,kc,cba->kji", J, J, J, q)
        if ccurve is None or lx1 < 3:
            continue
        X = [out[d][e] for d in range(3)]
        for f in range(6):
            if ccurve[e, f] == b"s":
                sphere_face(X, z, L[e], curve[e, f, :4], f)
        for k in range(8):
            if ccurve[e, k] == b"C":
                arc_edge(X, z, xc[e], yc[e], curve[e, k, 0], k + 1)
    return out
