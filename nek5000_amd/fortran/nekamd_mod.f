c     ISO_C_BINDING interface to libnekamd (include/nekamd.h) for Nek5000's
c     Fortran.  Compiled with Nek's own flags (fixed form, REAL = 8 bytes);
c     the shims in nekamd_shim.f use it.
      module nekamd
      use iso_c_binding
      implicit none

c     nkh_params / nkh_result of include/nekamd.h, field for field
      type, bind(c) :: nkh_params
        integer(c_int)  :: istep, param18
        real(c_double)  :: param22, restol
        integer(c_int)  :: precision, param42
        real(c_double)  :: param21
        integer(c_int)  :: param40
        integer(c_int)  :: param93, param94, param95, ifprojfld
        integer(c_int)  :: ifield, ldimt_proj, ifmvbd
      end type
      type, bind(c) :: nkh_result
        integer(c_int)  :: niter, converged
        real(c_double)  :: rbn0, rbn2, tol
      end type

      interface
        integer(c_int) function nkh_init(dev) bind(c)
          import
          integer(c_int), value :: dev
        end function
        type(c_ptr) function nkh_last_error() bind(c)
          import
        end function
        integer(c_int) function nkh_zwgll(z, w, np) bind(c)
          import
          real(c_double) :: z(*), w(*)
          integer(c_int), value :: np
        end function
        integer(c_int) function nkh_setvert3d(glo, ngv, nx, nel, vtx)
     $                                        bind(c)
          import
          integer(c_int64_t) :: glo(*), ngv, vtx(*)
          integer(c_int), value :: nx, nel
        end function
        integer(c_int) function nkh_mesh_create(lx1, nel, h) bind(c)
          import
          integer(c_int), value :: lx1, nel
          integer(c_int) :: h
        end function
        integer(c_int) function nkh_mesh_set_geom(h, g1, g2, g3, g4,
     $                                            g5, g6, bm1) bind(c)
          import
          integer(c_int), value :: h
          real(c_double) :: g1(*), g2(*), g3(*), g4(*), g5(*), g6(*)
          real(c_double) :: bm1(*)
        end function
        integer(c_int) function nkh_mesh_set_deformed(h, ifdef) bind(c)
          import
          integer(c_int), value :: h
          integer(c_int) :: ifdef(*)
        end function
        integer(c_int) function nkh_mesh_gs_setup(h, glo) bind(c)
          import
          integer(c_int), value :: h
          integer(c_int64_t) :: glo(*)
        end function
        integer(c_int) function nkh_mg_setup(h, fbc, vtx, x, y, z)
     $                                       bind(c)
          import
          integer(c_int), value :: h
          integer(c_int) :: fbc(*)
          integer(c_int64_t) :: vtx(*)
          real(c_double) :: x(*), y(*), z(*)
        end function
        integer(c_int) function nkh_axhelm(h, au, u, h1, h2, isd)
     $                                     bind(c)
          import
          integer(c_int), value :: h, isd
          real(c_double) :: au(*), u(*), h1(*), h2(*)
        end function
        integer(c_int) function nkh_dssum(h, u) bind(c)
          import
          integer(c_int), value :: h
          real(c_double) :: u(*)
        end function
        integer(c_int) function nkh_hmholtz(h, name, u, rhs, h1, h2,
     $              mask, mult, tli, maxit, isd, prm, res, hist) bind(c)
          import
          integer(c_int), value :: h, maxit, isd
          character(kind=c_char) :: name(*)
          real(c_double) :: u(*), rhs(*), h1(*), h2(*), mask(*)
          real(c_double) :: mult(*)
          real(c_double), value :: tli
          type(nkh_params) :: prm
          type(nkh_result) :: res
          type(c_ptr), value :: hist
        end function
        integer(c_int) function nkh_proj_create(h, mxprev, p) bind(c)
          import
          integer(c_int), value :: h, mxprev
          integer(c_int) :: p
        end function
        integer(c_int) function nkh_hsolve(h, name, u, r, h1, h2, vmk,
     $     vml, tol, maxit, isd, proj, prm, res, hist) bind(c)
          import
          integer(c_int), value :: h, maxit, isd, proj
          character(kind=c_char) :: name(*)
          real(c_double) :: u(*), r(*), h1(*), h2(*), vmk(*), vml(*)
          real(c_double), value :: tol
          type(nkh_params) :: prm
          type(nkh_result) :: res
          type(c_ptr), value :: hist
        end function
      end interface

c     library mesh handles of Nek's two element sets: imsh = 1 (nelv,
c     velocity and fields that share its gather-scatter) and 2 (nelt)
      integer :: nkh_mesh(2) = -1
c     projection spaces, keyed by the address of the caller's approx array
      integer, parameter :: nkh_maxproj = 16
      integer(c_intptr_t) :: nkh_projkey(nkh_maxproj) = 0
      integer :: nkh_projid(nkh_maxproj) = -1

      contains

c     Nek's abort convention for a nonzero return code: exitti with the
c     library's message
c     (exitti reads 132 characters up to a '$')
      subroutine nekamd_check(ierr, what)
      integer ierr
      character(*) what
      character(len=132) buf
      character(len=100) msg
      character(kind=c_char), pointer :: cmsg(:)
      integer k
      if (ierr.eq.0) return
      msg = ' '
      call c_f_pointer(nkh_last_error(), cmsg, [100])
      do k=1,100
         if (cmsg(k).eq.c_null_char) exit
         msg(k:k) = cmsg(k)
      enddo
      buf = 'nekamd '//what//': '//msg
      k = min(len_trim(buf) + 2, 132)
      buf(k:k) = '$'
      call exitti(buf, ierr)
      end subroutine

      end module
