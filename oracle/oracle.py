"""ctypes front end of the CPU oracle (oracle/nek_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg -- never by the nek5000_amd package.

Arrays use Nek's layout: a field is a float64 numpy array of shape
(nel, lz1, ly1, lx1) in C order (= (lx1,ly1,lz1,nel) column-major).
Small matrices are numpy arrays D[i, j] = D(i,j).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libnekoracle.so")
REF_SPECLIB = os.path.join(HERE, "_ref", "libspeclib_ref.so")

c_int = ctypes.c_int
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
vp = ctypes.c_void_p


def build(ref=False):
    """Compile the oracle (and, when the Nek5000 sources are present, the
    reference speclib.f) with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", HERE, "libnekoracle.so"], check=True)
    if ref and os.path.isdir(os.environ.get("NEKREF", "/root/reference")):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


class _Field(ctypes.Structure):
    _fields_ = [
        ("n", c_int), ("nel", c_int64),
        ("dxm1", vp), ("dxtm1", vp),
        ("g1", vp), ("g2", vp), ("g3", vp), ("g4", vp), ("g5", vp), ("g6", vp), ("bm1", vp),
        ("binv", vp), ("mask", vp), ("mult", vp),
        ("vol", c_double),
        ("gs", vp),
    ]


class CgRec(ctypes.Structure):
    _fields_ = [("rbn2", c_double), ("rtz1", c_double), ("rho", c_double),
                ("alpha", c_double), ("beta", c_double)]


class Params(ctypes.Structure):
    _fields_ = [("istep", c_int), ("param18", c_double), ("param22", c_double),
                ("restol", c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.nko_zwgll.argtypes = [vp, vp, c_int]
        L.nko_dgll.argtypes = [vp, vp, vp, c_int]
        L.nko_pnleg.argtypes = [c_double, c_int]
        L.nko_pnleg.restype = c_double
        L.nko_igllm.argtypes = [vp, vp, vp, vp, c_int, c_int]
        L.nko_geom.argtypes = [c_int, c_int, vp, vp, vp, vp, vp] + [vp] * 8
        L.nko_setvert3d.argtypes = [vp, c_int, c_int, vp]
        L.nko_setvert3d.restype = c_int64
        L.nko_gs_setup.argtypes = [vp, c_int64]
        L.nko_gs_setup.restype = vp
        L.nko_gs_free.argtypes = [vp]
        L.nko_gs_op.argtypes = [vp, vp, c_int]
        L.nko_axhelm.argtypes = [ctypes.POINTER(_Field), vp, vp, vp, vp]
        L.nko_setprec.argtypes = [ctypes.POINTER(_Field), vp, vp, vp]
        L.nko_chktcg1.argtypes = [ctypes.POINTER(_Field), c_double, vp, vp, vp]
        L.nko_chktcg1.restype = c_double
        L.nko_cggo.argtypes = [ctypes.POINTER(_Field), vp, vp, vp, vp, c_double, c_int,
                               ctypes.POINTER(Params), vp, ctypes.POINTER(c_double)]
        L.nko_cggo.restype = c_int
        L.nko_hmholtz.argtypes = [ctypes.POINTER(_Field), vp, vp, vp, vp, c_double, c_int,
                                  ctypes.POINTER(Params), vp, ctypes.POINTER(c_double)]
        L.nko_hmholtz.restype = c_int
        L.nko_fd_weights_full.argtypes = [c_double, vp, c_int, c_int, vp]
        L.nko_semhat.argtypes = [c_int, vp, vp, vp, vp]
        L.nko_generalev.argtypes = [vp, vp, vp, c_int]
        L.nko_mg_setup.argtypes = [ctypes.POINTER(_Field), vp, vp, vp, vp, vp, vp]
        L.nko_mg_setup.restype = vp
        L.nko_mg_levels.argtypes = [vp, vp]
        L.nko_mg_levels.restype = c_int
        L.nko_mg_solve.argtypes = [vp, vp, vp]
        L.nko_flex_cg.argtypes = [ctypes.POINTER(_Field), vp, vp, vp, vp, c_int, c_double,
                                  c_int, vp, ctypes.POINTER(c_double)]
        L.nko_flex_cg.restype = c_int
        L.nko_gmres.argtypes = [ctypes.POINTER(_Field), vp, vp, vp, vp, c_int, c_int, c_double,
                                c_int, vp, ctypes.POINTER(c_double)]
        L.nko_gmres.restype = c_int
        for nm in ("nko_mg_fdm_s", "nko_mg_fdm_d", "nko_mg_mask", "nko_mg_swt", "nko_mg_rwt",
                   "nko_mg_jh"):
            getattr(L, nm).argtypes = [vp, c_int]
            getattr(L, nm).restype = ctypes.POINTER(c_double)
        L.nko_mg_lengths.argtypes = [vp, c_int]
        L.nko_mg_lengths.restype = ctypes.POINTER(c_double)
        L.nko_mg_ncrs.argtypes = [vp]
        L.nko_mg_ncrs.restype = c_int
        L.nko_mg_null_space.argtypes = [vp]
        L.nko_mg_null_space.restype = c_int
        L.nko_mg_coarse_solve.argtypes = [vp, vp, vp]
        L.nko_mg_schwarz.argtypes = [vp, c_int, vp, vp]
        L.nko_mg_apply_fdm.argtypes = [vp, c_int, vp, vp]
        L.nko_adj_grad.argtypes = [ctypes.POINTER(_Field), vp, vp, vp, vp]
        L.nko_proj_create.argtypes = [c_int64, c_int]
        L.nko_proj_create.restype = vp
        L.nko_proj_free.argtypes = [vp]
        L.nko_proj_m.argtypes = [vp]
        L.nko_proj_m.restype = c_int
        for nm in ("nko_proj_x", "nko_proj_b"):
            getattr(L, nm).argtypes = [vp, c_int]
            getattr(L, nm).restype = ctypes.POINTER(c_double)
        L.nko_proj_xbar.argtypes = [vp]
        L.nko_proj_xbar.restype = ctypes.POINTER(c_double)
        L.nko_project1.argtypes = [ctypes.POINTER(_Field), vp, vp, vp, vp, vp, vp, c_int, vp]
        L.nko_project1.restype = c_int
        L.nko_project2.argtypes = [ctypes.POINTER(_Field), vp, vp, vp, vp, vp, vp]
        L.nko_num_threads.restype = c_int
        L.nko_set_num_threads.argtypes = [c_int]
        L.nko_set_mutation.argtypes = [c_int]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


# ---- speclib ----
def zwgll(np_):
    z = np.zeros(np_)
    w = np.zeros(np_)
    lib().nko_zwgll(_p(z), _p(w), np_)
    return z, w


def dgll(z):
    n = len(z)
    d = np.zeros(n * n)
    dt = np.zeros(n * n)
    lib().nko_dgll(_p(d), _p(dt), _p(_c(z)), n)
    return d.reshape(n, n).T.copy()  # D[i, j]


def pnleg(z, n):
    return lib().nko_pnleg(float(z), int(n))


def igllm(z1, z2):
    n1, n2 = len(z1), len(z2)
    i12 = np.zeros(n1 * n2)
    it12 = np.zeros(n1 * n2)
    lib().nko_igllm(_p(i12), _p(it12), _p(_c(z1)), _p(_c(z2)), n1, n2)
    return i12.reshape(n1, n2).T.copy()  # I[i, j] = h_j(z2_i)


def ref_speclib():
    """The reference's own speclib.f built by flang (oracle/_ref); None if absent."""
    if not os.path.exists(REF_SPECLIB):
        return None
    L = ctypes.CDLL(REF_SPECLIB, mode=os.RTLD_LAZY)
    return L


# ---- geometry ----
def geom(x, y, z, D, w):
    """(g[6], bm1, jac) from GLL coordinates (coef.f xyzrst/glmapm1/geodat1)."""
    x, y, z = _c(x), _c(y), _c(z)
    nel = x.shape[0]
    n = x.shape[-1]
    Dc = _c(np.asarray(D).T)  # column-major
    out = [np.zeros_like(x) for _ in range(8)]
    lib().nko_geom(n, nel, _p(x), _p(y), _p(z), _p(Dc), _p(_c(w)), *[_p(a) for a in out])
    return out[:6], out[6], out[7]


# ---- numbering and gather-scatter ----
def setvert3d(vertex, nx):
    vertex = _c(vertex, np.int64)
    nel = vertex.shape[0]
    glo = np.zeros(nel * nx ** 3, dtype=np.int64)
    ngv = lib().nko_setvert3d(_p(glo), nx, nel, _p(vertex))
    return glo, int(ngv)


class GS:
    def __init__(self, glo_num):
        self.glo = _c(glo_num, np.int64).reshape(-1)
        self.h = lib().nko_gs_setup(_p(self.glo), self.glo.size)

    def op(self, u, op="+"):
        """In place on a contiguous float64 array."""
        k = {"+": 0, "*": 1, "m": 2, "M": 3}[op]
        assert u.dtype == np.float64 and u.flags.c_contiguous and u.size == self.glo.size
        lib().nko_gs_op(self.h, _p(u), k)
        return u

    def __del__(self):
        try:
            lib().nko_gs_free(self.h)
        except Exception:
            pass


class Field:
    """Everything Nek's commons hold for one field, on the host."""

    def __init__(self, D, g, bm1, glo_num, mask, mult=None):
        self.D = _c(D)
        self.Dc = _c(self.D.T)      # dxm1 column-major
        self.Dtc = _c(self.D)       # dxtm1 column-major
        self.g = [_c(a) for a in g]
        self.bm1 = _c(bm1)
        self.gs = GS(glo_num)
        self.mask = _c(mask)
        self.nel = self.bm1.shape[0]
        self.n = self.bm1.shape[-1]
        if mult is None:
            m = np.ones_like(self.bm1)
            self.gs.op(m, "+")
            mult = 1.0 / m
        self.mult = _c(mult)
        b = self.bm1.copy()
        self.gs.op(b, "+")
        self.binv = 1.0 / b
        self.vol = float(np.sum(self.bm1))
        self._s = _Field(self.n, self.nel, _p(self.Dc), _p(self.Dtc),
                         *[_p(a) for a in self.g], _p(self.bm1), _p(self.binv),
                         _p(self.mask), _p(self.mult), self.vol, self.gs.h)

    def axhelm(self, u, h1, h2):
        au = np.zeros_like(self.bm1)
        lib().nko_axhelm(ctypes.byref(self._s), _p(au), _p(_c(u)), _p(_c(h1)), _p(_c(h2)))
        return au

    def dssum(self, u):
        u = _c(u).copy()
        return self.gs.op(u, "+")

    def adj_grad(self, u, lam):
        """Local nodal gradients (g1, g2) of J w.r.t. h1 and h2 given the
        forward solution u and the adjoint lam: g1 = -(D lam)^T G (D u),
        g2 = -lam bm1 u."""
        g1 = np.zeros_like(self.bm1)
        g2 = np.zeros_like(self.bm1)
        lib().nko_adj_grad(ctypes.byref(self._s), _p(g1), _p(g2), _p(_c(u)), _p(_c(lam)))
        return g1, g2

    def setprec(self, h1, h2):
        d = np.zeros_like(self.bm1)
        lib().nko_setprec(ctypes.byref(self._s), _p(d), _p(_c(h1)), _p(_c(h2)))
        return d

    def chktcg1(self, tol, res, h1, h2):
        return lib().nko_chktcg1(ctypes.byref(self._s), float(tol), _p(_c(res)), _p(_c(h1)),
                                 _p(_c(h2)))

    def cggo(self, f, h1, h2, tin, maxit, params=None):
        prm = params or Params(1, 0.0, 1e-7, 1e-7)
        x = np.zeros_like(self.bm1)
        rec = (CgRec * 901)()
        rbn0 = c_double(0)
        niter = lib().nko_cggo(ctypes.byref(self._s), _p(x), _p(_c(f)), _p(_c(h1)),
                               _p(_c(h2)), float(tin), int(maxit), ctypes.byref(prm),
                               ctypes.addressof(rec), ctypes.byref(rbn0))
        hist = np.array([[r.rbn2, r.rtz1, r.rho, r.alpha, r.beta] for r in rec[: niter + 1]])
        return x, niter, hist, rbn0.value

    def hmholtz(self, rhs, h1, h2, tli, maxit, params=None):
        """hmholtz.f:2 (non-PRES): returns (u, rhs_after, niter, history, rbn0)."""
        prm = params or Params(1, 0.0, 1e-7, 1e-7)
        u = np.zeros_like(self.bm1)
        r = _c(rhs).copy()
        rec = (CgRec * 901)()
        rbn0 = c_double(0)
        niter = lib().nko_hmholtz(ctypes.byref(self._s), _p(u), _p(r), _p(_c(h1)), _p(_c(h2)),
                                  float(tli), int(maxit), ctypes.byref(prm),
                                  ctypes.addressof(rec), ctypes.byref(rbn0))
        hist = np.array([[q.rbn2, q.rtz1, q.rho, q.alpha, q.beta] for q in rec[: niter + 1]])
        return u, r, niter, hist, rbn0.value


# ---- fast3d.f / hsmg.f ----
def fd_weights_full(xx, x, m=1):
    x = _c(x)
    n = len(x) - 1
    c = np.zeros((n + 1) * (m + 1))
    lib().nko_fd_weights_full(float(xx), _p(x), n, m, _p(c))
    return c.reshape(m + 1, n + 1)  # c[k, j]: k-th derivative weight of x_j


def semhat(n):
    a = np.zeros((n + 1) ** 2)
    b = np.zeros(n + 1)
    d = np.zeros((n + 1) ** 2)
    z = np.zeros(n + 1)
    lib().nko_semhat(n, _p(a), _p(b), _p(d), _p(z))
    return a.reshape(n + 1, n + 1).T.copy(), b, d.reshape(n + 1, n + 1).T.copy(), z


def generalev(A, B):
    n = A.shape[0]
    a = _c(np.asarray(A).T).copy()
    lam = np.zeros(n)
    lib().nko_generalev(_p(a), _p(_c(np.asarray(B).T)), _p(lam), n)
    return a.reshape(n, n).T.copy(), lam


# face order: symmetric notation (r-, r+, s-, s+, t-, t+); preprocessor
# notation (connect1.f facind) is (s-, r+, s+, r-, t-, t+)
SYM_TO_PREP = [3, 1, 0, 2, 4, 5]


def fbc_from_box(box, code=1):
    """BC codes (get_fast_bc) per element face in symmetric notation: `code`
    (1 = outflow 'O', Dirichlet for the solve) on domain-boundary faces."""
    prep = box.dirichlet_faces()
    return np.ascontiguousarray(prep[:, SYM_TO_PREP].astype(np.int32) * code)


class MG:
    """hsmg.f h1mg_setup / h1mg_solve and hmh_flex_cg on a Field."""

    def __init__(self, field, vertex, fbc, x, y, z, w):
        self.field = field
        self._keep = [_c(vertex, np.int64), _c(fbc, np.int32), _c(x), _c(y), _c(z), _c(w)]
        v, f, xx, yy, zz, ww = self._keep
        self.h = lib().nko_mg_setup(ctypes.byref(field._s), _p(v), _p(f), _p(xx), _p(yy),
                                    _p(zz), _p(ww))
        if not self.h:
            raise RuntimeError("oracle mg setup failed")
        nh = np.zeros(4, dtype=np.int32)
        self.lmax = lib().nko_mg_levels(self.h, _p(nh))
        self.nh = [int(v) for v in nh[: self.lmax]]  # nh[l-1] for level l

    def level_shape(self, l):
        n = self.nh[l - 1]
        return (self.field.nel, n, n, n)

    def _arr(self, fn, l, count):
        ptr = getattr(lib(), fn)(self.h, l)
        return np.ctypeslib.as_array(ptr, shape=(count,)).copy()

    def mask(self, l):
        return self._arr("nko_mg_mask", l, int(np.prod(self.level_shape(l)))).reshape(self.level_shape(l))

    def solve(self, rhs):
        r = _c(rhs).copy()
        z = np.zeros_like(r)
        lib().nko_mg_solve(self.h, _p(z), _p(r))
        return z

    def flex_cg(self, rhs, h1, h2, maxit, param21, istep=1):
        res = _c(rhs).copy()
        hist = np.zeros(maxit)
        div0 = c_double(0)
        it = lib().nko_flex_cg(ctypes.byref(self.field._s), self.h, _p(res), _p(_c(h1)),
                               _p(_c(h2)), int(maxit), float(param21), int(istep),
                               _p(hist), ctypes.byref(div0))
        return res, it, hist[:it], div0.value

    def gmres(self, rhs, h1, h2, maxit, param21, m=30, istep=1):
        """gmres.f:304 hmh_gmres with this V-cycle, restart m."""
        res = _c(rhs).copy()
        hist = np.zeros(maxit)
        div0 = c_double(0)
        it = lib().nko_gmres(ctypes.byref(self.field._s), self.h, _p(res), _p(_c(h1)),
                             _p(_c(h2)), int(maxit), int(m), float(param21), int(istep),
                             _p(hist), ctypes.byref(div0))
        return res, it, hist[:it], div0.value


class Proj:
    """navier4.f:562-1325 residual projection space of one field (hsolve's
    approx/napprox), mmx = (mxprev-4)/2 vectors."""

    def __init__(self, field, mxprev=20):
        self.field = field
        self.h = lib().nko_proj_create(field.bm1.size, int(mxprev))
        if not self.h:
            raise ValueError("mxprev out of range")
        self.mmx = (int(mxprev) - 4) // 2

    @property
    def m(self):
        return lib().nko_proj_m(self.h)

    def _vec(self, ptr):
        return np.ctypeslib.as_array(ptr, shape=(self.field.bm1.size,)).copy().reshape(
            self.field.bm1.shape)

    def x(self, k):
        return self._vec(lib().nko_proj_x(self.h, k))

    def b(self, k):
        return self._vec(lib().nko_proj_b(self.h, k))

    def xbar(self):
        return self._vec(lib().nko_proj_xbar(self.h))

    def project1(self, b, h1, h2, msk, w, ifmvbd=0):
        """In place on b (contiguous float64); returns (ireset, bb4, baf)."""
        assert b.dtype == np.float64 and b.flags.c_contiguous
        norms = np.zeros(2)
        ir = lib().nko_project1(ctypes.byref(self.field._s), self.h, _p(b), _p(_c(h1)),
                                _p(_c(h2)), _p(_c(msk)), _p(_c(w)), int(ifmvbd), _p(norms))
        return ir, norms[0], norms[1]

    def project2(self, x, h1, h2, msk, w):
        assert x.dtype == np.float64 and x.flags.c_contiguous
        lib().nko_project2(ctypes.byref(self.field._s), self.h, _p(x), _p(_c(h1)), _p(_c(h2)),
                           _p(_c(msk)), _p(_c(w)))

    def __del__(self):
        try:
            lib().nko_proj_free(self.h)
        except Exception:
            pass


def hsolve_uses_projection(name, istep, param93, param94, param95, ifprojfld, ifield=1,
                           ldimt_proj=1):
    """navier4.f:589-606: whether hsolve projects (not ifstdh)."""
    ifstdh = not ifprojfld
    p945 = param94
    if name.upper()[:4] == "PRES":
        ifstdh = False
        p945 = param95
    if ifield > ldimt_proj + 1 or param93 == 0 or p945 == 0 or istep < p945:
        ifstdh = True
    return not ifstdh


def hsolve(field, name, r, h1, h2, vmk, vml, tol, maxit, proj, params=None, mg=None,
           param42=2, param21=0.0, use_projection=True, ifmvbd=0):
    """navier4.f:562-634 hsolve for one solve: (u, niter, (bb4, baf)).

    With use_projection the projection branch runs: r*vmk, dssum, project1,
    hmhzpf (chktcg1, then cggo or, for PRES, hmh_gmres/hmh_flex_cg with mg),
    project2.  Otherwise hmholtz.  field.mask must equal vmk and field.mult
    vml (cggo reads them from the field)."""
    prm = params or Params(1, 0.0, 1e-7, 1e-7)
    pres = name.upper()[:4] == "PRES"
    if not use_projection:
        if pres:
            rr = _c(r).copy()
            field.gs.op(rr, "+")
            rr *= vmk
            return _pres_solve(field, mg, rr, h1, h2, maxit, param42, param21, prm.istep) + (
                (0.0, 0.0),)
        u, _, niter, _, _ = field.hmholtz(r, h1, h2, tol, maxit, prm)
        return u, niter, (0.0, 0.0)
    rr = _c(r).copy() * vmk
    field.gs.op(rr, "+")
    _, bb4, baf = proj.project1(rr, h1, h2, vmk, vml, ifmvbd)
    # hmhzpf (navier4.f:513-560)
    t = tol
    if prm.param22 != 0:
        t = abs(prm.param22)
    t = field.chktcg1(t, rr, h1, h2)
    if pres:
        u, niter = _pres_solve(field, mg, rr, h1, h2, maxit, param42, param21, prm.istep)
    else:
        u, niter, _, _ = field.cggo(rr, h1, h2, t, maxit, prm)
    u = _c(u).copy()
    proj.project2(u, h1, h2, vmk, vml)
    return u, niter, (bb4, baf)


def _pres_solve(field, mg, f, h1, h2, maxit, param42, param21, istep):
    """cggo's PRES branches (hmholtz.f:639-653): x = f, then hmh_gmres or
    hmh_flex_cg."""
    if param42 == 0:
        x, it, _, _ = mg.gmres(f, h1, h2, maxit, param21, istep=istep)
    else:
        x, it, _, _ = mg.flex_cg(f, h1, h2, maxit, param21, istep=istep)
    return x, it


def num_threads():
    return lib().nko_num_threads()


def set_num_threads(t):
    lib().nko_set_num_threads(int(t))


class mutation:
    """Context manager planting a mistake in the oracle's multigrid (the bits
    of nko_set_mutation), for tests that show a pin can fail."""

    def __init__(self, bits):
        self.bits = int(bits)

    def __enter__(self):
        lib().nko_set_mutation(self.bits)
        return self

    def __exit__(self, *a):
        lib().nko_set_mutation(0)
