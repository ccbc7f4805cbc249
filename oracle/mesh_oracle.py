"""numpy restatement of Nek5000's mesh ingestion, the checker for the
library's host-side reader (nek5000_amd/csrc/mesh_io.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/ -- never by the nek5000_amd
package.

  * read_re2: reader_re2.f:2-620 (read_re2_hdr, readp_re2_mesh/curve/bc,
    buf_to_xyz/curve/bc) for #v001..#v004 files in either byte order;
  * vertex_ids: corner coincidence by brute force over all corner pairs (an
    independent algorithm from the library's sorted sweep), periodic faces
    joined after the translation between them;
  * xyz_from_corners: genxyz.f:1235-1400 linquad (xyzlin, xyzquad) with
    navier5.f:2422-2560 gh_face_extend_3d, written as dense tensor products.
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
    return out
