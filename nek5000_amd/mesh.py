"""Synthetic hexahedral box meshes with Nek's data layout.

Nek reads its element->vertex map and GLL coordinates from case files
(map2.f:118-137 get_vert, genxyz.f); this module produces the same arrays for
box meshes (genbox semantics, tools/genbox/input.genbox): per-element global
vertex ids in hypercube order (vertex(0:1,0:1,0:1,nel)), GLL coordinates
xm1/ym1/zm1 of shape (nel, lz1, ly1, lx1) in C order (= Nek's column-major
(lx1,ly1,lz1,nel)), and per-face boundary conditions.

Counter-based random numbers keyed by a global node id make every local copy
of a shared node carry the same value on CPU and GPU alike.
"""
import numpy as np

from ._lib import load, check, c_int64
import ctypes
import os

# ---------------------------------------------------------------------------
# counter-based RNG (splitmix64 finaliser) keyed by (seed, id)
# ---------------------------------------------------------------------------
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_G = np.uint64(0x9E3779B97F4A7C15)


def _mix(x):
    x = x.copy()
    with np.errstate(over="ignore"):
        x ^= x >> np.uint64(30)
        x *= _M1
        x ^= x >> np.uint64(27)
        x *= _M2
        x ^= x >> np.uint64(31)
    return x


def hash_uniform(seed, ids, stream=0):
    """U[0,1) from (seed, id, stream), 53-bit resolution."""
    ids = np.asarray(ids, dtype=np.uint64)
    with np.errstate(over="ignore"):
        k = ids * _G + np.uint64((seed * 1000003 + stream * 7919) & 0xFFFFFFFFFFFFFFFF)
    h = _mix(k)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def hash_normal(seed, ids, stream=0):
    """N(0,1) by Box-Muller from two keyed uniforms."""
    u1 = hash_uniform(seed, ids, 2 * stream + 101)
    u2 = hash_uniform(seed, ids, 2 * stream + 102)
    u1 = np.maximum(u1, 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


# ---------------------------------------------------------------------------
# GLL basis (host); the library restates speclib.f, this just calls it
# ---------------------------------------------------------------------------
def gll(np_):
    lib = load()
    z = np.zeros(np_)
    w = np.zeros(np_)
    check(lib.nkh_zwgll(z.ctypes.data, w.ctypes.data, np_), "nkh_zwgll")
    d = np.zeros(np_ * np_)
    dt = np.zeros(np_ * np_)
    check(lib.nkh_dgll(d.ctypes.data, dt.ctypes.data, z.ctypes.data, np_), "nkh_dgll")
    # column-major D(i,j) -> numpy D[i, j]
    return z, w, d.reshape(np_, np_).T.copy()


def setvert3d(vertex, lx1):
    """navier8.f:2004-2360 via the library: vertex (nel,2,2,2)[k,j,i] -> glo_num."""
    lib = load()
    vertex = np.ascontiguousarray(vertex, dtype=np.int64)
    nel = vertex.shape[0]
    glo = np.zeros(nel * lx1 ** 3, dtype=np.int64)
    ngv = c_int64(0)
    check(lib.nkh_setvert3d(glo.ctypes.data, ctypes.byref(ngv), lx1, nel, vertex.ctypes.data),
          "nkh_setvert3d")
    return glo, int(ngv.value)


class _HexMesh:
    """What the solve set-up needs from any hexahedral mesh: vertex ids
    (nel,2,2,2) [e,k,j,i] (`vertex`), GLL coordinates (`xyz()`), and
    per-face Dirichlet flags in Nek's preprocessor face order
    (`dirichlet_faces()`)."""

    def glo_num(self):
        return setvert3d(self.vertex, self.lx1)

    # symmetric face order (r-, r+, s-, s+, t-, t+) -> preprocessor face
    # (connect1.f:1009 facind numbering: s-, r+, s+, r-, t-, t+)
    SYM_TO_PREP = (3, 1, 0, 2, 4, 5)

    def fast_bc(self, code=1):
        """get_fast_bc codes (fast3d.f:802-877) per element face in symmetric
        order: `code` on domain-boundary faces (1 = 'O', Dirichlet for the
        pressure), 0 elsewhere.  Shape (nel, 6) int32."""
        prep = self.dirichlet_faces()
        return np.ascontiguousarray(prep[:, list(self.SYM_TO_PREP)].astype(np.int32) * code)

    def face_mask(self):
        """bdry.f:317- BCMASK before its dsop: 1 everywhere, 0 on the nodes of
        Dirichlet faces (facev, connect1.f:1058).  Shape (nel, lz1, ly1, lx1)."""
        n = self.lx1
        m = np.ones((self.nel, n, n, n))
        f = self.dirichlet_faces()
        m[f[:, 0], :, 0, :] = 0.0
        m[f[:, 1], :, :, n - 1] = 0.0
        m[f[:, 2], :, n - 1, :] = 0.0
        m[f[:, 3], :, :, 0] = 0.0
        m[f[:, 4], 0, :, :] = 0.0
        m[f[:, 5], n - 1, :, :] = 0.0
        return m

    def unique_ids(self, glo, ngv):
        """A global id for every local node: glo_num where nonzero, else a
        per-element interior id after ngv (setvert3d ifcenter=.true. order)."""
        n = self.lx1
        uid = glo.reshape(self.nel, n, n, n).copy()
        ni = (n - 2) ** 3
        if ni > 0:
            base = ngv + np.arange(self.nel, dtype=np.int64)[:, None] * ni
            inner = base + 1 + np.arange(ni, dtype=np.int64)[None, :]
            uid[:, 1:-1, 1:-1, 1:-1] = inner.reshape(self.nel, n - 2, n - 2, n - 2)
        return uid


class BoxMesh(_HexMesh):
    """nelx x nely x nelz hexahedra on [0,lx]x[0,ly]x[0,lz].

    deform: amplitude a of the map x_i <- x_i + a*h*sin(2 pi x)sin(2 pi y)sin(2 pi z)
            applied to every GLL node (h = element width in x).
    jitter: amplitude b of U(-b*h, b*h) perturbations of interior vertices.
    periodic_x: x-faces are periodic ('P'); other faces carry `bc` ('v'
            Dirichlet or 'O' outflow -- both zero the mask of the solve).
    """

    def __init__(self, nelx, nely, nelz, lx1, lengths=(1.0, 1.0, 1.0), deform=0.0,
                 jitter=0.0, periodic_x=False, seed=0, bc="v"):
        self.nelx, self.nely, self.nelz = nelx, nely, nelz
        self.lx1 = lx1
        self.nel = nelx * nely * nelz
        self.lengths = lengths
        self.deform = deform
        self.jitter = jitter
        self.periodic_x = periodic_x
        self.seed = seed
        self.bc = bc
        self.h = (lengths[0] / nelx, lengths[1] / nely, lengths[2] / nelz)
        self._vertex = None
        self._xyz = None

    @property
    def n(self):
        return self.nel * self.lx1 ** 3

    # element-corner lattice indices, shape (nel, 2,2,2) [k,j,i] for each dir
    def _corner_lattice(self):
        ex = np.arange(self.nelx)
        ey = np.arange(self.nely)
        ez = np.arange(self.nelz)
        EZ, EY, EX = np.meshgrid(ez, ey, ex, indexing="ij")
        EX = EX.reshape(-1)
        EY = EY.reshape(-1)
        EZ = EZ.reshape(-1)
        o = np.array([0, 1])
        OK, OJ, OI = np.meshgrid(o, o, o, indexing="ij")
        VX = EX[:, None, None, None] + OI[None]
        VY = EY[:, None, None, None] + OJ[None]
        VZ = EZ[:, None, None, None] + OK[None]
        return VX, VY, VZ

    @property
    def vertex(self):
        """Global vertex ids (1-based), shape (nel,2,2,2) indexed [e,k,j,i]."""
        if self._vertex is None:
            VX, VY, VZ = self._corner_lattice()
            nvx = self.nelx if self.periodic_x else self.nelx + 1
            nvy, nvz = self.nely + 1, self.nelz + 1
            vx = VX % nvx if self.periodic_x else VX
            self._vertex = (1 + vx + nvx * (VY + nvy * VZ)).astype(np.int64)
            del nvz
        return self._vertex

    def _vertex_coords(self):
        VX, VY, VZ = self._corner_lattice()
        hx, hy, hz = self.h
        X = VX * hx
        Y = VY * hy
        Z = VZ * hz
        if self.jitter:
            b = self.jitter
            # key by lattice point (unwrapped x index keeps periodic images apart
            # only on the boundary, which is not jittered)
            key = (VX + (self.nelx + 1) * (VY + (self.nely + 1) * VZ)).astype(np.uint64)
            ix = (VX > 0) & (VX < self.nelx)
            iy = (VY > 0) & (VY < self.nely)
            iz = (VZ > 0) & (VZ < self.nelz)
            X = X + ix * (2 * hash_uniform(self.seed, key, 11) - 1) * b * hx
            Y = Y + iy * (2 * hash_uniform(self.seed, key, 12) - 1) * b * hy
            Z = Z + iz * (2 * hash_uniform(self.seed, key, 13) - 1) * b * hz
        return X, Y, Z

    def xyz(self):
        """GLL coordinates, each (nel, lz1, ly1, lx1) float64."""
        if self._xyz is not None:
            return self._xyz
        z, _, _ = gll(self.lx1)
        X, Y, Z = self._vertex_coords()
        # trilinear blend: phi0 = (1-r)/2, phi1 = (1+r)/2
        ph = np.stack([(1 - z) / 2, (1 + z) / 2])  # (2, n)
        out = []
        for C in (X, Y, Z):
            # C[e,k,j,i] corners -> field[e,t,s,r]
            a = np.einsum("ekji,ir->ekjr", C, ph)
            a = np.einsum("ekjr,js->eksr", a, ph)
            a = np.einsum("eksr,kt->etsr", a, ph)
            out.append(np.ascontiguousarray(a))
        x, y, zz = out
        if self.deform:
            a = self.deform * self.h[0]
            s = np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * zz)
            s *= a
            x = x + s
            y = y + s
            zz = zz + s
        self._xyz = (x, y, zz)
        return self._xyz

    def dirichlet_faces(self):
        """(nel, 6) bool in Nek's preprocessor face order (connect1.f:1009
        facind: 1 y=-1, 2 x=+1, 3 y=+1, 4 x=-1, 5 z=-1, 6 z=+1)."""
        e = np.arange(self.nel)
        ex = e % self.nelx
        ey = (e // self.nelx) % self.nely
        ez = e // (self.nelx * self.nely)
        f = np.zeros((self.nel, 6), dtype=bool)
        f[:, 0] = ey == 0
        f[:, 2] = ey == self.nely - 1
        if not self.periodic_x:
            f[:, 1] = ex == self.nelx - 1
            f[:, 3] = ex == 0
        f[:, 4] = ez == 0
        f[:, 5] = ez == self.nelz - 1
        return f


class Re2Mesh(_HexMesh):
    """A Nek case mesh from its .re2 file (reader_re2.f) through the library:
    corners, curve data and boundary conditions as Nek's /INPUT/ holds them,
    vertex ids from the case's .map/.ma2 (map2.f read_map, map_path) or by
    corner coincidence, GLL coordinates by genxyz ('m' midside edges, 'C'
    arcs, 's' sphere faces).

    field: which boundary-condition section drives dirichlet_faces (0 = the
            first, Nek's fluid field when the case has one).
    dirichlet: BC codes whose faces are Dirichlet for the solve; None means
            every face with a boundary condition other than 'E  ' and 'P  '.
    """

    def __init__(self, path, lx1, field=0, dirichlet=None, tol_rel=1e-2, map_path=None):
        lib = load()
        self.path = path
        self.lx1 = lx1
        p = path.encode()
        nelgt, ndim, nelgv, nfld = (ctypes.c_int() for _ in range(4))
        check(lib.nkh_re2_header(p, ctypes.byref(nelgt), ctypes.byref(ndim), ctypes.byref(nelgv),
                                 ctypes.byref(nfld)), "nkh_re2_header")
        self.nel, self.ndim, self.nelgv, self.nfields = (nelgt.value, ndim.value, nelgv.value,
                                                         nfld.value)
        nel, nv = self.nel, 2 ** self.ndim
        self.group = np.zeros(nel, dtype=np.int32)
        self.xc = np.zeros((nel, nv))
        self.yc = np.zeros((nel, nv))
        self.zc = np.zeros((nel, nv))
        self.curve = np.zeros((nel, 12, 5))
        self.ccurve = np.zeros((nel, 12), dtype="S1")
        self.cbc = np.zeros((max(self.nfields, 1), nel, 6), dtype="S3")
        self.bc = np.zeros((max(self.nfields, 1), nel, 6, 5))
        check(lib.nkh_re2_read(p, nel, self.ndim, self.nfields, self.group.ctypes.data,
                               self.xc.ctypes.data, self.yc.ctypes.data, self.zc.ctypes.data,
                               self.curve.ctypes.data, self.ccurve.ctypes.data,
                               self.cbc.ctypes.data, self.bc.ctypes.data), "nkh_re2_read")
        self.field = field
        self.dirichlet = None if dirichlet is None else {c.encode() if isinstance(c, str) else c
                                                         for c in dirichlet}
        self._vertex = None
        self._xyz = None
        self.tol_rel = tol_rel
        # map_path: vertex ids (and partitions) from the case's .map/.ma2
        # (map2.f read_map; "auto": <case>.map, else <case>.ma2, as Nek looks
        # for them) instead of corner coincidence
        if map_path == "auto":
            base = path[:-4] if path.endswith(".re2") else path
            map_path = next((p for p in (base + ".map", base + ".ma2") if os.path.exists(p)),
                            None)
        self.map_path = map_path
        self.part = None

    @property
    def n(self):
        return self.nel * self.lx1 ** 3

    @property
    def vertex(self):
        """Global vertex ids, shape (nel,2,2,2) indexed [e,k,j,i]: the map
        file's when there is one, else 1.. by corner coincidence."""
        if self._vertex is None and self.map_path is not None:
            lib = load()
            p = self.map_path.encode()
            ne, na = ctypes.c_int(), ctypes.c_int()
            check(lib.nkh_map_header(p, ctypes.byref(ne), ctypes.byref(na)), "nkh_map_header")
            v = np.zeros(self.nel * 2 ** self.ndim, dtype=np.int64)
            self.part = np.zeros(self.nel, dtype=np.int32)
            check(lib.nkh_map_read(p, self.nel, self.ndim, self.part.ctypes.data, v.ctypes.data),
                  "nkh_map_read")
            self.nvert = int(len(np.unique(v)))
            self._vertex = v.reshape(self.nel, 2, 2, 2)
        if self._vertex is None:
            lib = load()
            v = np.zeros(self.nel * 8, dtype=np.int64)
            nv = c_int64(0)
            f = self.field if self.nfields else None
            check(lib.nkh_mesh_vertex_ids(
                self.nel, self.ndim, self.xc.ctypes.data, self.yc.ctypes.data,
                self.zc.ctypes.data, self.cbc[f].ctypes.data if f is not None else None,
                self.bc[f].ctypes.data if f is not None else None, float(self.tol_rel),
                v.ctypes.data, ctypes.byref(nv)), "nkh_mesh_vertex_ids")
            self.nvert = int(nv.value)
            self._vertex = v.reshape(self.nel, 2, 2, 2)
        return self._vertex

    def xyz(self):
        """GLL coordinates, each (nel, lz1, ly1, lx1) float64."""
        if self._xyz is None:
            lib = load()
            n = self.lx1
            out = [np.zeros((self.nel, n, n, n)) for _ in range(3)]
            check(lib.nkh_mesh_xyz_from_corners(
                n, self.nel, self.xc.ctypes.data, self.yc.ctypes.data, self.zc.ctypes.data,
                self.ccurve.ctypes.data, self.curve.ctypes.data, *[a.ctypes.data for a in out]),
                "nkh_mesh_xyz_from_corners")
            self._xyz = tuple(out)
        return self._xyz

    def dirichlet_faces(self):
        """(nel, 6) bool in Nek's preprocessor face order."""
        if not self.nfields:
            return np.zeros((self.nel, 6), dtype=bool)
        cb = self.cbc[self.field]
        if self.dirichlet is None:
            return (cb != b"E  ") & (cb != b"P  ")
        return np.isin(cb, list(self.dirichlet))
