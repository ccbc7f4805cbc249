c     Stand-ins for the Nek5000 routines the nekamd shims call (test
c     driver only): exitti prints Nek's EXIT line and stops with status
c     3; fgslib_gs_op reports the handle it was given; get_fast_bc
c     returns the face codes the driver stored per element; nek_hmholtz,
c     nek_hsolve, nek_axhelm and nek_dssum (the reference's routines
c     under the names the shims route to) report that they were called.
      subroutine exitti(stringi, idata)
      character*1 stringi(132)
      integer idata, k, n
      n = 132
      do k=1,132
         if (stringi(k).eq.'$') then
            n = k - 1
            exit
         endif
      enddo
      write(6,'(a,132a1)',advance='no') 'EXIT: ', (stringi(k),k=1,n)
      write(6,'(i10)') idata
      call flush(6)
      stop 3
      end

      subroutine fgslib_gs_op(h, u, dom, op, tr)
      integer h, dom, op, tr
      real u(1)
      write(6,'(a,4i6)') 'gs_op', h, dom, op, tr
      return
      end

      subroutine get_fast_bc(lbr,rbr,lbs,rbs,lbt,rbt,e,bsym,ierr)
      include 'SIZE'
      integer lbr,rbr,lbs,rbs,lbt,rbt,e,bsym,ierr
      integer ifbc
      common /tfbc/ ifbc(6,lelt)
      lbr = ifbc(1,e)
      rbr = ifbc(2,e)
      lbs = ifbc(3,e)
      rbs = ifbc(4,e)
      lbt = ifbc(5,e)
      rbt = ifbc(6,e)
      ierr = 0
      return
      end

      subroutine nek_hmholtz(name,u,rhs,h1,h2,mask,mult,imsh,tli,maxit,
     $                       isd)
      character*4 name
      real u(1),rhs(1),h1(1),h2(1),mask(1),mult(1),tli
      integer imsh, maxit, isd
      write(6,'(a,1x,a4,2i4)') 'nek_hmholtz', name, imsh, isd
      return
      end

      subroutine nek_hsolve(name,u,r,h1,h2,vmk,vml,imsh,tol,maxit,isd,
     $                      approx,napprox,bi)
      character*4 name
      real u(1),r(1),h1(1),h2(1),vmk(1),vml(1),tol,approx(1),bi(1)
      integer imsh, maxit, isd, napprox(1)
      write(6,'(a,1x,a4,2i4)') 'nek_hsolve', name, imsh, isd
      return
      end

      subroutine nek_axhelm(au,u,helm1,helm2,imesh,isd)
      real au(1),u(1),helm1(1),helm2(1)
      integer imesh, isd
      write(6,'(a,2i4)') 'nek_axhelm', imesh, isd
      return
      end

      subroutine nek_dssum(u,nx,ny,nz)
      real u(1)
      integer nx, ny, nz
      write(6,'(a,3i4)') 'nek_dssum', nx, ny, nz
      return
      end
