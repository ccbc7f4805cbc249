"""Host-side mirror of Nek5000's Helmholtz solve interface over libnekamd.

Names and argument meanings follow the reference routines:

  * ``Field.axhelm(au, u, h1, h2, imesh=1, isd=1)``  -- hmholtz.f:72
  * ``Field.dssum(u)`` / ``Field.dsop(u, op)``        -- dssum.f:33 / :100
  * ``Field.setprec(d, h1, h2)``                      -- hmholtz.f:380
  * ``Field.hmholtz(name, u, rhs, h1, h2, mask, mult, tli, maxit, isd)``
                                                      -- hmholtz.f:2
  * ``Field.cggo(x, f, h1, h2, mask, mult, tin, maxit)`` -- hmholtz.f:611
  * ``Field.h1mg_setup(fbc, vertex, x, y, z)``        -- hsmg.f:2234
  * ``Field.h1mg_solve(z, rhs)``                      -- hsmg.f:1855
  * ``Field.hmh_flex_cg(res, h1, h2, wt, maxit)``     -- hmholtz.f:2164
  * ``Field.hmh_gmres(res, h1, h2, wt, maxit)``       -- gmres.f:304
  * ``Field.hsolve(name, u, r, h1, h2, vmk, vml, tol, maxit, approx)``
                                                      -- navier4.f:562
  * ``Projection(field, mxprev)`` with ``project1`` / ``project2``
                                                      -- navier4.f:636 / :1123

Arrays are torch tensors on the GPU (float64, Nek layout (nel,lz1,ly1,lx1)
C-order = (lx1,ly1,lz1,nel) column-major).  torch is only the allocator and
the stream: every operation runs in libnekamd's HIP kernels.  Errors raise
NekAmdError (Nek would call exitti).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import NekAmdError, NkhParams, NkhResult, check, load  # noqa: F401
from .mesh import BoxMesh

_initialised = {"device": None}


def _torch():
    import torch

    return torch


def init(device=0):
    lib = load()
    if _initialised["device"] is None:
        torch = _torch()
        if not torch.cuda.is_available():
            raise _lib.NekAmdError("no GPU visible: libnekamd runs on the GPU only")
        torch.cuda.set_device(device)
        check(lib.nkh_init(device), "nkh_init")
        _initialised["device"] = device
    return lib


def _sync_stream(lib):
    torch = _torch()
    check(lib.nkh_set_stream(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
          "nkh_set_stream")


def _ptr(t):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("expected a CUDA tensor")
    if t.dtype != torch.float64 or not t.is_contiguous():
        raise TypeError("expected a contiguous float64 tensor")
    return ctypes.c_void_p(t.data_ptr())


def axhelm_tuning(lx1):
    """(variant, grid, tuned) of the axhelm kernels for order lx1: variant 0
    stages the element in LDS, 1 one t-plane at a time, 2 and 3 the same with
    more thread slabs per block (identical plain applies), 4 and 5 the r/s
    contractions on the FP64 matrix cores with the element or groups of
    t-planes staged (equal to rounding); grid = blocks of the plain apply
    (0: one per element group)."""
    lib = init()
    v, g, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(lib.nkh_axhelm_tuning(int(lx1), ctypes.byref(v), ctypes.byref(g), ctypes.byref(t)),
          "nkh_axhelm_tuning")
    return v.value, g.value, bool(t.value)


def axhelm_tuning_fused(lx1):
    """(CG-fused variant, flexible-CG-fused variant) for order lx1."""
    a, b = ctypes.c_int(), ctypes.c_int()
    check(init().nkh_axhelm_tuning_fused(int(lx1), ctypes.byref(a), ctypes.byref(b)),
          "nkh_axhelm_tuning_fused")
    return a.value, b.value


def set_axhelm_tuning(lx1, variant, grid=0):
    """Fix the axhelm kernel choice for order lx1 -- the plain, CG-fused and
    flexible-CG-fused applies alike -- (no autotuning afterwards)."""
    check(init().nkh_axhelm_set_tuning(int(lx1), int(variant), int(grid)),
          "nkh_axhelm_set_tuning")


def fdm_variant():
    """Kernel of the local FDM solves: 0 vector FMA, 1 FP32 matrix cores and
    2 bf16 matrix cores (FP32 accumulation) for the single-precision
    smoother, 3 FP64 matrix cores for the double-precision one."""
    v = ctypes.c_int()
    check(init().nkh_fdm_variant(ctypes.byref(v)), "nkh_fdm_variant")
    return v.value


def axhelm32_variant(lx1):
    """Kernel of the FP32 operator of the mixed solvers for order lx1: 0
    vector FMA, 1 the r/s contractions on the FP32 matrix cores."""
    v = ctypes.c_int()
    check(init().nkh_axhelm32_variant(int(lx1), ctypes.byref(v)), "nkh_axhelm32_variant")
    return v.value


def set_axhelm32_variant(lx1, variant):
    check(init().nkh_axhelm32_set_variant(int(lx1), int(variant)), "nkh_axhelm32_set_variant")


def set_fdm_variant(variant):
    check(init().nkh_fdm_set_variant(int(variant)), "nkh_fdm_set_variant")


def default_params(istep=1, param18=0, param22=1e-7, restol=1e-7, precision=0, param42=0,
                   param21=0.0, param40=0, param93=20, param94=5, param95=0, ifprojfld=0,
                   ifield=1, ldimt_proj=1, ifmvbd=0):
    """Nek defaults (reader_par.f:26-27,76-79,86-89,159-162): param(22)=1e-7,
    restol=param(22); param(42)=0, param(21)=0, param(93)=mxprev=20,
    param(94)=5, param(95)=0 and ifprojfld=.false. unless the case file sets
    them."""
    return NkhParams(istep, param18, param22, restol, precision, param42, param21, param40,
                     param93, param94, param95, ifprojfld, ifield, ldimt_proj, ifmvbd)


class Field:
    """One element set with geometry, numbering, gs map, mask and multiplicity
    resident on the GPU (Nek's imsh mesh plus its v?mask/vmult)."""

    def __init__(self, lx1, nel, device=0):
        self.lib = init(device)
        self.lx1 = lx1
        self.nel = nel
        self.n = nel * lx1 ** 3
        h = ctypes.c_int(-1)
        check(self.lib.nkh_mesh_create(lx1, nel, ctypes.byref(h)), "nkh_mesh_create")
        self.handle = h.value
        self.mask = None
        self.mult = None
        self.glo_num = None
        self.ngv = None

    @property
    def shape(self):
        return (self.nel, self.lx1, self.lx1, self.lx1)

    def zeros(self):
        torch = _torch()
        return torch.zeros(self.shape, dtype=torch.float64, device="cuda")

    def to_dev(self, a):
        torch = _torch()
        return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape),
                               device="cuda")

    # ---- setup ----------------------------------------------------------
    def geom_from_xyz(self, x, y, z):
        """coef.f glmapm1 + geodat1 on the GPU from GLL coordinates."""
        xd, yd, zd = (self.to_dev(a) if not hasattr(a, "is_cuda") else a for a in (x, y, z))
        _sync_stream(self.lib)
        check(self.lib.nkd_mesh_geom_from_xyz(self.handle, _ptr(xd), _ptr(yd), _ptr(zd)),
              "nkd_mesh_geom_from_xyz")

    def set_geom(self, g, bm1):
        """Upload precomputed g1m1..g6m1 (list of 6 host arrays) and bm1."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in list(g) + [bm1]]
        check(self.lib.nkh_mesh_set_geom(self.handle, *[a.ctypes.data for a in arrs]),
              "nkh_mesh_set_geom")

    def set_deformed(self, ifdfrm):
        """genxyz.f setdef's IFDFRM (nel flags, 0 = undeformed: axhelm and
        setprec drop that element's cross terms), or None for all deformed."""
        if ifdfrm is None:
            check(self.lib.nkh_mesh_set_deformed(self.handle, None), "nkh_mesh_set_deformed")
            return
        f = np.ascontiguousarray(ifdfrm, dtype=np.int32).reshape(-1)
        assert f.size == self.nel
        check(self.lib.nkh_mesh_set_deformed(self.handle, f.ctypes.data),
              "nkh_mesh_set_deformed")

    def gs_setup(self, glo_num):
        glo = np.ascontiguousarray(glo_num, dtype=np.int64).reshape(-1)
        assert glo.size == self.n
        check(self.lib.nkh_mesh_gs_setup(self.handle, glo.ctypes.data), "nkh_mesh_gs_setup")
        self.glo_num = glo

    def info(self):
        n = ctypes.c_int64()
        vol = ctypes.c_double()
        nseg = ctypes.c_int64()
        nsingle = ctypes.c_int64()
        mm = ctypes.c_int()
        check(self.lib.nkh_mesh_info(self.handle, ctypes.byref(n), ctypes.byref(vol),
                                     ctypes.byref(nseg), ctypes.byref(nsingle),
                                     ctypes.byref(mm)), "nkh_mesh_info")
        return dict(n=n.value, vol=vol.value, nseg=nseg.value, nsingle=nsingle.value,
                    max_mult=mm.value)

    def mass(self):
        """(bm1, binvm1) as tensors (copies of the library's device arrays)."""
        pb = ctypes.c_void_p()
        pi = ctypes.c_void_p()
        check(self.lib.nkh_mesh_ptrs(self.handle, ctypes.byref(pb), ctypes.byref(pi)),
              "nkh_mesh_ptrs")
        bm1 = self.zeros()
        binv = self.zeros()
        _copy_raw_to_tensor(pb.value, bm1)
        _copy_raw_to_tensor(pi.value, binv)
        return bm1, binv

    # ---- operators ------------------------------------------------------
    def axhelm(self, au, u, h1, h2, imesh=1, isd=1):
        _sync_stream(self.lib)
        check(self.lib.nkd_axhelm(self.handle, _ptr(au), _ptr(u), _ptr(h1), _ptr(h2), isd),
              "nkd_axhelm")

    def dssum(self, u):
        _sync_stream(self.lib)
        check(self.lib.nkd_dssum(self.handle, _ptr(u)), "nkd_dssum")

    def dsop(self, u, op):
        _sync_stream(self.lib)
        check(self.lib.nkd_dsop(self.handle, _ptr(u), op.encode()[:1]), "nkd_dsop")

    def setprec(self, d, h1, h2):
        _sync_stream(self.lib)
        check(self.lib.nkd_setprec(self.handle, _ptr(d), _ptr(h1), _ptr(h2)), "nkd_setprec")

    def hmholtz(self, name, u, rhs, h1, h2, mask, mult, tli, maxit, isd=1, params=None,
                history=False):
        """hmholtz.f:2: rhs is dssum'd and masked in place; u gets the solution.
        Returns a dict with niter, rbn0, rbn2, tol (and history if asked: per
        Jacobi-PCG iteration (rbn2, rtz1, rho, alpha, beta), or for 'PRES'
        with param42=2 the flexible-CG residual norm per iteration)."""
        prm = params if params is not None else default_params()
        res = NkhResult()
        pres = name[:4] == "PRES"
        hist = np.zeros(max(int(maxit), 1) if pres else 5 * 900) if history else None
        _sync_stream(self.lib)
        check(self.lib.nkd_hmholtz(self.handle, name.encode(), _ptr(u), _ptr(rhs), _ptr(h1),
                                   _ptr(h2), _ptr(mask), _ptr(mult), float(tli), int(maxit),
                                   isd, ctypes.byref(prm), ctypes.byref(res),
                                   ctypes.c_void_p(hist.ctypes.data) if history else None),
              "nkd_hmholtz")
        out = dict(niter=res.niter, converged=res.converged, rbn0=res.rbn0, rbn2=res.rbn2,
                   tol=res.tol)
        if history:
            out["history"] = (hist[: res.niter].copy() if pres else
                              hist.reshape(900, 5)[: max(res.niter + 1, 1)].copy())
        return out

    def cggo(self, x, f, h1, h2, mask, mult, tin, maxit, params=None, history=False,
             noalloc=False):
        """hmholtz.f:611 cggo; noalloc selects nkd_cggo_noalloc (the same solve
        with its buffers in the library's static device array)."""
        prm = params if params is not None else default_params()
        res = NkhResult()
        hist = np.zeros(5 * 900) if history else None
        _sync_stream(self.lib)
        fn, name = ((self.lib.nkd_cggo_noalloc, "nkd_cggo_noalloc") if noalloc else
                    (self.lib.nkd_cggo, "nkd_cggo"))
        check(fn(self.handle, _ptr(x), _ptr(f), _ptr(h1), _ptr(h2), _ptr(mask), _ptr(mult),
                 float(tin), int(maxit), ctypes.byref(prm), ctypes.byref(res),
                 ctypes.c_void_p(hist.ctypes.data) if history else None), name)
        out = dict(niter=res.niter, converged=res.converged, rbn0=res.rbn0, rbn2=res.rbn2,
                   tol=res.tol)
        if history:
            out["history"] = hist.reshape(900, 5)[: max(res.niter + 1, 1)].copy()
        return out

    # ---- H1 multigrid (PnPn pressure) ------------------------------------
    def h1mg_setup(self, fbc, vertex, x, y, z):
        """hsmg.f:2234 h1mg_setup.  fbc: (nel, 6) int codes in symmetric face
        order (0 interior/periodic, 1 Dirichlet, 2 Neumann); vertex: (nel, 8)
        int64; x, y, z: GLL coordinates (host arrays)."""
        arrs = [np.ascontiguousarray(fbc, dtype=np.int32).reshape(-1),
                np.ascontiguousarray(vertex, dtype=np.int64).reshape(-1)]
        arrs += [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (x, y, z)]
        assert arrs[0].size == 6 * self.nel and arrs[1].size == 8 * self.nel
        assert all(a.size == self.n for a in arrs[2:])
        _sync_stream(self.lib)
        check(self.lib.nkh_mg_setup(self.handle, *[a.ctypes.data for a in arrs]), "nkh_mg_setup")

    def h1mg_info(self):
        lmax = ctypes.c_int()
        nh = np.zeros(4, dtype=np.int32)
        ncrs = ctypes.c_int()
        avg = ctypes.c_double()
        check(self.lib.nkh_mg_info(self.handle, ctypes.byref(lmax), nh.ctypes.data,
                                   ctypes.byref(ncrs), ctypes.byref(avg)), "nkh_mg_info")
        return dict(lmax=lmax.value, nh=[int(v) for v in nh[: lmax.value]], ncrs=ncrs.value,
                    crs_avg_iters=avg.value)

    def h1mg_solve(self, z, rhs):
        """hsmg.f:1855 h1mg_solve: z = M rhs (rhs masked in place)."""
        _sync_stream(self.lib)
        check(self.lib.nkd_mg_solve(self.handle, _ptr(z), _ptr(rhs)), "nkd_mg_solve")

    def hmh_flex_cg(self, res, h1, h2, wt, maxit, pmask=None, params=None, history=False):
        """hmholtz.f:2164 hmh_flex_cg: res = rhs in, solution out."""
        prm = params if params is not None else default_params()
        out = NkhResult()
        hist = np.zeros(max(int(maxit), 1)) if history else None
        mask = self.mask if pmask is None else pmask
        _sync_stream(self.lib)
        check(self.lib.nkd_hmh_flex_cg(self.handle, _ptr(res), _ptr(h1), _ptr(h2), _ptr(wt),
                                       _ptr(mask), int(maxit), ctypes.byref(prm),
                                       ctypes.byref(out),
                                       ctypes.c_void_p(hist.ctypes.data) if history else None),
              "nkd_hmh_flex_cg")
        d = dict(niter=out.niter, converged=out.converged, div0=out.rbn0, rnorm=out.rbn2,
                 tolpss=out.tol)
        if history:
            d["history"] = hist[: out.niter].copy()
        return d

    def hmh_gmres(self, res, h1, h2, wt, maxit, pmask=None, params=None, history=False):
        """gmres.f:304 hmh_gmres: res = rhs in, solution out."""
        prm = params if params is not None else default_params()
        out = NkhResult()
        hist = np.zeros(max(int(maxit), 1)) if history else None
        mask = self.mask if pmask is None else pmask
        _sync_stream(self.lib)
        check(self.lib.nkd_hmh_gmres(self.handle, _ptr(res), _ptr(h1), _ptr(h2), _ptr(wt),
                                     _ptr(mask), int(maxit), ctypes.byref(prm),
                                     ctypes.byref(out),
                                     ctypes.c_void_p(hist.ctypes.data) if history else None),
              "nkd_hmh_gmres")
        d = dict(niter=out.niter, converged=out.converged, div0=out.rbn0, rnorm=out.rbn2,
                 tolpss=out.tol)
        if history:
            d["history"] = hist[: out.niter].copy()
        return d

    # ---- residual projection (navier4.f:562) ------------------------------
    def adjoint_coef_grad(self, g1, u, lam, g2=None):
        """dJ/dh1 (and dJ/dh2 into g2) per local node from the forward
        solution u and the adjoint lam."""
        _sync_stream(self.lib)
        check(self.lib.nkd_adjoint_coef_grad(self.handle, _ptr(g1),
                                             _ptr(g2) if g2 is not None else None, _ptr(u),
                                             _ptr(lam)), "nkd_adjoint_coef_grad")

    def hsolve(self, name, u, r, h1, h2, vmk, vml, tol, maxit, approx=None, isd=1, params=None,
               history=False):
        """navier4.f:562 hsolve: hmholtz, or projection + hmhzpf when the
        switches in params select it (approx: a Projection of this field).
        r is modified in place, u gets the solution."""
        prm = params if params is not None else default_params()
        res = NkhResult()
        pres = name[:4] == "PRES"
        hist = np.zeros(max(int(maxit), 1) if pres else 5 * 900) if history else None
        _sync_stream(self.lib)
        check(self.lib.nkd_hsolve(self.handle, name.encode(), _ptr(u), _ptr(r), _ptr(h1),
                                  _ptr(h2), _ptr(vmk), _ptr(vml), float(tol), int(maxit), isd,
                                  approx.handle if approx is not None else -1,
                                  ctypes.byref(prm), ctypes.byref(res),
                                  ctypes.c_void_p(hist.ctypes.data) if history else None),
              "nkd_hsolve")
        out = dict(niter=res.niter, converged=res.converged, rbn0=res.rbn0, rbn2=res.rbn2,
                   tol=res.tol)
        if history:
            out["history"] = (hist[: res.niter].copy() if pres else
                              hist.reshape(900, 5)[: max(res.niter + 1, 1)].copy())
        return out

    def destroy(self):
        if self.handle is not None and self.handle >= 0:
            check(self.lib.nkh_mesh_destroy(self.handle), "nkh_mesh_destroy")
            self.handle = None


class Batch:
    """Several hmholtz solves at once (nkh_batch_create / nkd_hmholtz_batch):
    one entry per solve, each on a Field (a Field may appear more than once)."""

    def __init__(self, fields):
        self.fields = list(fields)
        self.lib = init()
        ids = (ctypes.c_int * max(len(self.fields), 1))(*[f.handle for f in self.fields])
        h = ctypes.c_int(-1)
        check(self.lib.nkh_batch_create(len(self.fields), ids, ctypes.byref(h)),
              "nkh_batch_create")
        self.handle = h.value

    def hmholtz(self, solves, maxit, params=None):
        """solves: list of dicts u, rhs, h1, h2, mask, mult, tli (tensors on
        the GPU).  Returns (list of result dicts, graph launches)."""
        nb = len(self.fields)
        assert len(solves) == nb
        arr = (_lib.NkhBatchSolve * nb)()
        for k, q in enumerate(solves):
            arr[k] = _lib.NkhBatchSolve(_ptr(q["u"]), _ptr(q["rhs"]), _ptr(q["h1"]),
                                        _ptr(q["h2"]), _ptr(q["mask"]), _ptr(q["mult"]),
                                        float(q["tli"]))
        res = (NkhResult * nb)()
        launches = ctypes.c_int(0)
        prm = params if params is not None else default_params()
        _sync_stream(self.lib)
        check(self.lib.nkd_hmholtz_batch(self.handle, arr, int(maxit), ctypes.byref(prm), res,
                                         ctypes.byref(launches)), "nkd_hmholtz_batch")
        out = [dict(niter=r.niter, converged=r.converged, rbn0=r.rbn0, rbn2=r.rbn2, tol=r.tol)
               for r in res]
        return out, launches.value

    def destroy(self):
        if self.handle is not None and self.handle >= 0:
            check(self.lib.nkh_batch_destroy(self.handle), "nkh_batch_destroy")
            self.handle = None


class Projection:
    """hsolve's approx/napprox (navier4.f:1179 proj_get_ivar): up to
    mmx = (mxprev-4)/2 previous solutions of one field and their images, on
    the GPU."""

    def __init__(self, field, mxprev=20):
        self.field = field
        self.lib = field.lib
        h = ctypes.c_int(-1)
        check(self.lib.nkh_proj_create(field.handle, int(mxprev), ctypes.byref(h)),
              "nkh_proj_create")
        self.handle = h.value

    def info(self):
        m, mmx = ctypes.c_int(), ctypes.c_int()
        check(self.lib.nkh_proj_info(self.handle, ctypes.byref(m), ctypes.byref(mmx)),
              "nkh_proj_info")
        return m.value, mmx.value

    @property
    def m(self):
        return self.info()[0]

    def column(self, k, which=0):
        """Copy of column k of X (which 0) or B = A X (which 1), or xbar (2)."""
        ptr = ctypes.c_void_p()
        check(self.lib.nkh_proj_column(self.handle, int(k), int(which), ctypes.byref(ptr)),
              "nkh_proj_column")
        t = self.field.zeros()
        _copy_raw_to_tensor(ptr.value, t)
        return t

    def project1(self, b, h1, h2, msk, w, ifmvbd=0):
        """navier4.f:636 project1 in place on b; returns (ireset, bb4, baf)."""
        norms = np.zeros(2)
        ir = ctypes.c_int()
        _sync_stream(self.lib)
        check(self.lib.nkd_project1(self.handle, _ptr(b), _ptr(h1), _ptr(h2), _ptr(msk), _ptr(w),
                                    int(ifmvbd), norms.ctypes.data, ctypes.byref(ir)),
              "nkd_project1")
        return ir.value, float(norms[0]), float(norms[1])

    def project2(self, x, h1, h2, msk, w):
        """navier4.f:1123 project2 in place on x."""
        _sync_stream(self.lib)
        check(self.lib.nkd_project2(self.handle, _ptr(x), _ptr(h1), _ptr(h2), _ptr(msk), _ptr(w)),
              "nkd_project2")

    def destroy(self):
        if self.handle is not None and self.handle >= 0:
            check(self.lib.nkh_proj_destroy(self.handle), "nkh_proj_destroy")
            self.handle = None


def _copy_raw_to_tensor(ptr, t):
    """Copy n doubles from a library-owned device pointer into tensor t."""
    import torch

    nbytes = t.numel() * 8
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    torch.cuda.synchronize()
    rc = hip.hipMemcpy(ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(ptr), nbytes, 3)
    if rc != 0:
        raise _lib.NekAmdError(f"hipMemcpy failed: {rc}")


def box_field(box: BoxMesh, device=0, multigrid=False):
    """Set up a Field for a BoxMesh the way nek_init does (drive1.f:2-179):
    numbering (setvert3d), geometry (geodat1, on the GPU), gs map, binvm1,
    Dirichlet mask (bcmask: facev then dsop '*'), vmult = 1/dssum(1); with
    multigrid=True also the h1mg hierarchy (Dirichlet faces as 'O')."""
    f = Field(box.lx1, box.nel, device)
    glo, ngv = box.glo_num()
    f.ngv = ngv
    x, y, z = box.xyz()
    f.geom_from_xyz(x, y, z)
    f.gs_setup(glo)
    mask = f.to_dev(box.face_mask())
    f.dsop(mask, "*")
    mult = f.zeros() + 1.0
    f.dssum(mult)
    mult = 1.0 / mult
    f.mask = mask
    f.mult = mult
    if multigrid:
        f.h1mg_setup(box.fast_bc(), box.vertex, x, y, z)
    return f
