c     Test driver of the nekamd Fortran shims against stand-in commons.
c     Modes (first argument):
c       sizes            c_sizeof of the bind(c) types
c       zwgll NP         GLL points and weights through nkh_zwgll
c       setvert IN OUT   glo_num through nkh_setvert3d (IN: nel, vertex)
c       dssum0           dssum before any set-up: Nek's own gs handle
c       nogpu            nekamd_setup without a GPU: the EXIT line
c       solve IN OUT     a case through the shims (hmholtz, axhelm and
c                        dssum after set-up)
c       route            axisymmetric and param(42) = 1 cases go to the
c                        reference's routines (no library call)
      program drive
      use nekamd
      include 'SIZE'
      include 'TOTAL'
      common /ivrtx/ vertex ((2**ldim)*lelt)
      integer*8 vertex
      common /fastmd/ ifdfrm(lelt), iffast(lelt), ifh2, ifsolv
      logical ifdfrm, iffast, ifh2, ifsolv
      integer ifbc
      common /tfbc/ ifbc(6,lelt)
      common /iterhm/ niterhm

      type(nkh_params) prm
      type(nkh_result) res
      character*256 mode, arg1, arg2
      real z(16), w(16), tli
      real, allocatable :: u(:), rhs(:), h1(:), h2(:), vmask(:)
      real, allocatable :: vmult(:), au(:), ds(:)
      integer*8, allocatable :: glo(:)
      integer*8 ngv
      integer np, nel, ierr, n, k, e, maxit, isplit, p42, idef(lelt)
      character*4 name

      call get_command_argument(1, mode)
      if (mode.eq.'sizes') then
         write(6,'(a,2i6)') 'sizes', c_sizeof(prm), c_sizeof(res)

      elseif (mode.eq.'zwgll') then
         call get_command_argument(2, arg1)
         read(arg1,*) np
         ierr = nkh_zwgll(z, w, np)
         call nekamd_check(ierr, 'zwgll')
         do k=1,np
            write(6,'(a,i4,2es26.17)') 'gll', k, z(k), w(k)
         enddo

      elseif (mode.eq.'setvert') then
         call get_command_argument(2, arg1)
         call get_command_argument(3, arg2)
         open(10, file=arg1, access='stream', form='unformatted')
         read(10) nel
         read(10) (vertex(k), k=1,8*nel)
         close(10)
         n = lx1*ly1*lz1*nel
         allocate(glo(n))
         ierr = nkh_setvert3d(glo, ngv, lx1, nel, vertex)
         call nekamd_check(ierr, 'setvert3d')
         open(11, file=arg2, access='stream', form='unformatted',
     $        status='replace')
         write(11) ngv
         write(11) glo
         close(11)

      elseif (mode.eq.'dssum0') then
c        setup_topo's order: vmult (ifield 1), tmult (ifield 2), then a
c        passive scalar that shares the velocity's handle
         gsh_fld(1) = 11
         gsh_fld(2) = 12
         gsh_fld(3) = 11
         ifldmhd = 99
         allocate(u(lx1*ly1*lz1))
         u = 1.0
         do k=1,3
            ifield = k
            call dssum(u, lx1, ly1, lz1)
         enddo

      elseif (mode.eq.'route') then
         allocate(u(8), rhs(8), h1(8), h2(8))
         allocate(vmask(8), vmult(8), au(8))
         param = 0.0
         ifaxis = .true.
         call hmholtz('VELX', u, rhs, h1, h2, vmask, vmult, 1, tli,
     $                10, 1)
         call axhelm(au, u, h1, h2, 1, 1)
         call hsolve('TEMP', u, rhs, h1, h2, vmask, vmult, 2, tli,
     $               10, 1, au, np, au)
         ifaxis = .false.
         param(42) = 1.0
         call hmholtz('PRES', u, rhs, h1, h2, vmask, vmult, 1, tli,
     $                10, 1)
         write(6,'(a)') 'routed'

      elseif (mode.eq.'nogpu') then
         nelv = 1
         nelt = 1
         call nekamd_setup(1)

      elseif (mode.eq.'solve') then
         call get_command_argument(2, arg1)
         call get_command_argument(3, arg2)
         open(10, file=arg1, access='stream', form='unformatted')
         read(10) nel, isplit, p42, maxit
         read(10) tli
         read(10) name
         n = lx1*ly1*lz1*nel
         nelv = nel
         nelt = nel
         read(10) (vertex(k), k=1,8*nel)
         read(10) (idef(k), k=1,nel)
         read(10) ((ifbc(k,e), k=1,6), e=1,nel)
         read(10) (xm1(k,1,1,1), k=1,n), (ym1(k,1,1,1), k=1,n),
     $            (zm1(k,1,1,1), k=1,n)
         read(10) (g1m1(k,1,1,1), k=1,n), (g2m1(k,1,1,1), k=1,n),
     $            (g3m1(k,1,1,1), k=1,n), (g4m1(k,1,1,1), k=1,n),
     $            (g5m1(k,1,1,1), k=1,n), (g6m1(k,1,1,1), k=1,n)
         read(10) (bm1(k,1,1,1), k=1,n)
         allocate(u(n), rhs(n), h1(n), h2(n), vmask(n), vmult(n))
         allocate(au(n), ds(n))
         read(10) rhs, h1, h2, vmask, vmult, ds
         close(10)
         do e=1,nel
            ifdfrm(e) = idef(e).ne.0
         enddo
         ifsplit = isplit.ne.0
         istep = 1
         ifield = 1
         param = 0.0
         param(22) = 1.0e-7
         param(42) = p42
         param(21) = tli
         restol = 1.0e-7
         ifprojfld = .false.
         ifmvbd = .false.
         gsh_fld(1) = 11
         gsh_fld(2) = 11
         ifldmhd = 99
         u = 0.0
         call hmholtz(name, u, rhs, h1, h2, vmask, vmult, 1, tli,
     $                maxit, 1)
c        the operator and the sum once the element set is resident
         call axhelm(au, u, h1, h2, 1, 1)
         call dssum(ds, lx1, ly1, lz1)
         open(11, file=arg2, access='stream', form='unformatted',
     $        status='replace')
         write(11) niterhm
         write(11) u, au, ds
         close(11)
         write(6,'(a,i6)') 'solved niterhm', niterhm
      endif
      end
