c     Stand-ins for the Nek5000 routines the nekamd shims call (test
c     driver only): exitti prints Nek's EXIT line and stops with status
c     3; fgslib_gs_op reports the handle it was given; get_fast_bc
c     returns the face codes the driver stored per element.
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
