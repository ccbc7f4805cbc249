c     Drop-in replacements for Nek5000's Helmholtz/Poisson solve path,
c     linked as a USR object ahead of libnek5000.a (makenek.inc:304-307
c     allows the multiple definitions).  Each keeps the reference's
c     signature and reads the commons the reference reads; a nonzero
c     return code of the library becomes exitti (comm_mpi.f:550-637).
c
c       hmholtz  hmholtz.f:2-69     (cggo, hmh_flex_cg, hmh_gmres behind it)
c       hsolve   navier4.f:562-634  (project1, hmhzpf, project2)
c       axhelm   hmholtz.f:72-259
c       dssum    dssum.f:33-98
c
c     nekamd_setup uploads an element set once, on its first solve: Nek's
c     geometric factors (after setdef/sfastax, with IFDFRM), mass matrix,
c     vertex ids and, for the PnPn pressure, the multigrid face codes.
c
c     Cases the library does not build -- 2-D meshes, axisymmetric ones
c     (ifaxis, hmholtz.f:128-168) and the pressure with param(42) = 1 --
c     go to the reference's own routines, linked under the names
c     nek_hmholtz, nek_hsolve, nek_axhelm and nek_dssum (the reference's
c     objects with their symbols renamed; the link recipe is in the
c     integration notes).

c-----------------------------------------------------------------------
      subroutine nekamd_params(prm)
c     the knobs cggo/hsolve read from Nek's commons
      use nekamd
      include 'SIZE'
      include 'TOTAL'
      type(nkh_params) prm
      prm%istep     = istep
      prm%param18   = int(param(18))
      prm%param22   = param(22)
      prm%restol    = restol(ifield)
      prm%precision = 0
      prm%param42   = int(param(42))
      prm%param21   = param(21)
      prm%param40   = int(param(40))
      prm%param93   = int(param(93))
      prm%param94   = int(param(94))
      prm%param95   = int(param(95))
      prm%ifprojfld = 0
      if (ifprojfld(ifield)) prm%ifprojfld = 1
      prm%ifield    = ifield
      prm%ldimt_proj = ldimt_proj
      prm%ifmvbd    = 0
      if (ifmvbd) prm%ifmvbd = 1
      return
      end

c-----------------------------------------------------------------------
      subroutine nekamd_setup(imsh)
c     one-time upload of element set imsh (1: nelv, 2: nelt)
      use nekamd
      include 'SIZE'
      include 'TOTAL'
      common /ivrtx/ vertex ((2**ldim)*lelt)
      integer*8 vertex
      common /fastmd/ ifdfrm(lelt), iffast(lelt), ifh2, ifsolv
      logical ifdfrm, iffast, ifh2, ifsolv
      integer*8, allocatable :: glo_num(:)
      integer, allocatable :: fbc(:,:), idef(:)
      integer*8 ngv
      integer imsh, nel, e, ierr, n, ifsave, two
      integer lbr, rbr, lbs, rbs, lbt, rbt

      if (ldim.ne.3) call exitti('nekamd: 3-D only, ldim = $', ldim)
      nel = nelv
      if (imsh.eq.2) nel = nelt
      n   = lx1*ly1*lz1*nel
      if (nkh_mesh(1).lt.0 .and. nkh_mesh(2).lt.0) then
         ierr = nkh_init(0)
         call nekamd_check(ierr, 'nkh_init')
      endif
      ierr = nkh_mesh_create(lx1, nel, nkh_mesh(imsh))
      call nekamd_check(ierr, 'nkh_mesh_create')
      ierr = nkh_mesh_set_geom(nkh_mesh(imsh), g1m1, g2m1, g3m1,
     $                         g4m1, g5m1, g6m1, bm1)
      call nekamd_check(ierr, 'nkh_mesh_set_geom')
c     undeformed elements (param(59) = 0): their cross terms are dropped
c     as the reference's axhelm and setprec drop them
      allocate(idef(nel))
      do e=1,nel
         idef(e) = 0
         if (ifdfrm(e)) idef(e) = 1
      enddo
      ierr = nkh_mesh_set_deformed(nkh_mesh(imsh), idef)
      deallocate(idef)
      call nekamd_check(ierr, 'nkh_mesh_set_deformed')
      allocate(glo_num(n))
      ierr = nkh_setvert3d(glo_num, ngv, lx1, nel, vertex)
      call nekamd_check(ierr, 'nkh_setvert3d')
      ierr = nkh_mesh_gs_setup(nkh_mesh(imsh), glo_num)
      deallocate(glo_num)
      call nekamd_check(ierr, 'nkh_mesh_gs_setup')
      if (ifsplit .and. imsh.eq.1) then
c        PnPn pressure: h1mg with get_fast_bc's face codes of the
c        velocity boundary conditions, as h1mg_setup_mask reads them
         allocate(fbc(6,nel))
         ifsave = ifield
         ifield = 1
         two    = 2
         ierr   = 0
         do e=1,nel
            call get_fast_bc(lbr,rbr,lbs,rbs,lbt,rbt,e,two,ierr)
            fbc(1,e) = lbr
            fbc(2,e) = rbr
            fbc(3,e) = lbs
            fbc(4,e) = rbs
            fbc(5,e) = lbt
            fbc(6,e) = rbt
         enddo
         ifield = ifsave
         ierr = nkh_mg_setup(nkh_mesh(1), fbc, vertex, xm1, ym1, zm1)
         deallocate(fbc)
         call nekamd_check(ierr, 'nkh_mg_setup')
      endif
      return
      end

c-----------------------------------------------------------------------
      subroutine hmholtz(name,u,rhs,h1,h2,mask,mult,imsh,tli,maxit,isd)
      use nekamd
      include 'SIZE'
      include 'TOTAL'
      character*4 name
      real u(1),rhs(1),h1(1),h2(1),mask(1),mult(1)
      type(nkh_params) prm
      type(nkh_result) res
      integer ierr
      common /iterhm/ niterhm
      if (ldim.ne.3 .or. ifaxis .or.
     $    (name.eq.'PRES' .and. int(param(42)).eq.1)) then
         call nek_hmholtz(name,u,rhs,h1,h2,mask,mult,imsh,tli,maxit,isd)
         return
      endif
      if (nkh_mesh(imsh).lt.0) call nekamd_setup(imsh)
      call nekamd_params(prm)
      ierr = nkh_hmholtz(nkh_mesh(imsh), name//c_null_char, u, rhs,
     $                   h1, h2, mask, mult, tli, maxit, isd, prm, res,
     $                   c_null_ptr)
      call nekamd_check(ierr, 'hmholtz')
      niterhm = res%niter
      return
      end

c-----------------------------------------------------------------------
      subroutine hsolve(name,u,r,h1,h2,vmk,vml,imsh,tol,maxit,isd,
     $                  approx,napprox,bi)
c     the projection space lives on the GPU, one per approx array of the
c     caller (approx and napprox themselves are not read or written)
      use nekamd
      include 'SIZE'
      include 'TOTAL'
      character*4 name
      real u(1),r(1),h1(1),h2(1),vmk(1),vml(1),bi(1),approx(1)
      integer napprox(1)
      type(nkh_params) prm
      type(nkh_result) res
      integer ierr, k, p
      integer(c_intptr_t) key
      common /iterhm/ niterhm
      if (ldim.ne.3 .or. ifaxis .or.
     $    (name.eq.'PRES' .and. int(param(42)).eq.1)) then
         call nek_hsolve(name,u,r,h1,h2,vmk,vml,imsh,tol,maxit,isd,
     $                   approx,napprox,bi)
         return
      endif
      if (nkh_mesh(imsh).lt.0) call nekamd_setup(imsh)
      key = loc(approx)
      p = -1
      do k=1,nkh_maxproj
         if (nkh_projkey(k).eq.key) p = nkh_projid(k)
      enddo
      if (p.lt.0) then
         do k=1,nkh_maxproj
            if (nkh_projkey(k).eq.0 .and. p.lt.0) then
               ierr = nkh_proj_create(nkh_mesh(imsh), mxprev, p)
               call nekamd_check(ierr, 'nkh_proj_create')
               nkh_projkey(k) = key
               nkh_projid(k)  = p
            endif
         enddo
         if (p.lt.0) call exitti('nekamd hsolve: spaces > $',
     $                           nkh_maxproj)
      endif
      call nekamd_params(prm)
      ierr = nkh_hsolve(nkh_mesh(imsh), name//c_null_char, u, r, h1,
     $                  h2, vmk, vml, tol, maxit, isd, p, prm, res,
     $                  c_null_ptr)
      call nekamd_check(ierr, 'hsolve')
      niterhm = res%niter
      return
      end

c-----------------------------------------------------------------------
      subroutine axhelm(au,u,helm1,helm2,imesh,isd)
      use nekamd
      include 'SIZE'
      include 'TOTAL'
      real au(1),u(1),helm1(1),helm2(1)
      integer imesh, isd, ierr
      if (ldim.ne.3 .or. ifaxis) then
         call nek_axhelm(au,u,helm1,helm2,imesh,isd)
         return
      endif
      if (nkh_mesh(imesh).lt.0) call nekamd_setup(imesh)
      ierr = nkh_axhelm(nkh_mesh(imesh), au, u, helm1, helm2, isd)
      call nekamd_check(ierr, 'axhelm')
      return
      end

c-----------------------------------------------------------------------
      subroutine dssum(u,nx,ny,nz)
c     the element set follows gsh_fld(ifield) as in the reference; until
c     the library holds that set (setup_topo's vmult/tmult come first)
c     and for other orders, Nek's own gather-scatter handle does the sum
      use nekamd
      include 'SIZE'
      include 'TOTAL'
      real u(1)
      integer nx, ny, nz, ifldt, imsh, ierr
      if (ldim.ne.3) then
         call nek_dssum(u,nx,ny,nz)
         return
      endif
      ifldt = ifield
      if (ifldt.eq.ifldmhd) ifldt = 1
      imsh = 2
      if (gsh_fld(ifldt).eq.gsh_fld(1)) imsh = 1
      if (nx.ne.lx1 .or. nkh_mesh(imsh).lt.0) then
         if (gsh_fld(ifldt).ge.0)
     $      call fgslib_gs_op(gsh_fld(ifldt),u,1,1,0)
         return
      endif
      ierr = nkh_dssum(nkh_mesh(imsh), u)
      call nekamd_check(ierr, 'dssum')
      return
      end
