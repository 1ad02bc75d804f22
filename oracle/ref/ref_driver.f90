! ref_driver.f90 -- drives the reference freeCappuccino routines, compiled
! from the reference sources by oracle/ref/Makefile, to write golden vectors.
! TEST INFRASTRUCTURE ONLY: run in this container by tests/golden/make_golden.py.
!
! Usage (run inside a case directory that holds polyMesh/):
!   ref_driver <native|foam> <jobfile>
! Every file name in the job file is relative to the case directory.
! Job lines (list-directed; names without '/'):
!   dump                                      -> geometry and CSR arrays
!   dump_periodic                             -> owner after face_mapping and the
!              slot maps over inner + periodic faces (sparse_matrix.f90:241-296)
!   laplacian  mu phi su0 a_out su_out        (laplacian.f90)
!   grad_gauss phi out                        (gradients.f90:1607)
!   grad_lsq   variant phi out                variant: lsq | lsq_dm | lsq_qr
!   grad_fvx   phi gin out2 outc              fvxGradient.f90: the two-sweep
!              corrected Gauss grad_gauss (:1549-1645) into out2, and
!              grad_gauss_corrected (:1664-1758) started from gin into outc
!              (gradients AoS (3,numCells))
!   solve      solver itr_max tol_abs tol_rel a fi0 rhs fi_out info_out
!   convdiff   flmass dcoef phi grad gam scheme bckind a_out su_out sp_out
!   convdiff_cell flmass vis phi grad gam scheme bckind a_out su_out sp_out
!              (as convdiff with facefluxscalar, scalar_fluxes.f90:344-464, on
!               the inner faces: the cell array vis(numCells) interpolated to
!               the face; boundary faces take vis of their owner)
!   nonorth    gamma grad su0 su_out         (facefluxsc with zero mass flux:
!              the explicit non-orthogonal part only, scalar_fluxes.f90:99-109,139)
!   update_boundary phi phi_out               (updateBoundary.f90)
!   temperature T To Too den vis flmass wallfrac viscos pranl sigt urf gds dt
!               tscheme scheme solver itr_max tol_abs tol_rel gradmethod
!               T_out a_out su_out sp_out grad_out info_out
!              (calculate_temperature, temperature.f90:35-385, on the module
!               globals set from the files; tscheme steady|bdf|bdf2|cn,
!               gradmethod gauss|lsq|lsq_dm|lsq_qr, wallfrac = bcDFraction(7,:))
!   solve_explicit solver n nnz itr_max tol_abs tol_rel ia ja diag a fi0 rhs fi_out info_out
!              (explicit-array entry points iccg/dpcg/bicgstab/GaussSeidel on a
!               CSR system of its own, 1-based int32 ia/ja/diag files)
!   pcorr_assemble u v w p dpdxi den apu apv apw flmass const_mflux flomas
!                  a_out su_out flmass_out u_out v_out w_out
!              (calcp_simple.f90:69-228 for an incompressible case without
!               pressure patches: facefluxmass2 on the inner faces,
!               adjustMassFlow unless const_mflux, the inlet/outlet sources
!               and facefluxmass2_periodic; faceflux_mass.f90:175-252,
!               313-372, 833-912.  calcp_simple.f90 itself does not compile
!               at this revision -- lSolverP, maxiterP, tolAbsP, tolRelP and
!               urfp are undeclared at :324,419 -- so its steps are composed
!               here from the compiled routines, in its statement order)
!   pcorr_correct pp flmass a u v w p apu apv apw urfp prefcell
!                 pp_out flmass_out u_out v_out w_out p_out dpdxi_out su_out sv_out sw_out
!              (calcp_simple.f90:333-419 for one corrector: flux correction
!               on inner and periodic faces, ppref = pp(pRefCell),
!               gradp_and_sources (nablap.f90:19-209, pscheme 'linear', bpres
!               stages), velocity and pressure correction, then
!               updateVelocityAtBoundary velocity.f90:1184-1279)
!   momentum u v w uo vo wo uoo voo p den vis visw flmass astale viscos dt tscheme scheme gds
!            urf1 urf2 urf3 maxiter tolrel
!            a_out su_out sv_out sw_out spu_out spv_out sp_out apu_out apv_out apw_out
!            u_out v_out w_out p_out dpdxi_out dudxi_out
!              (calcuvw, velocity.f90:50-750, on module globals set from the files:
!               woo = uoo/2; a holds astale before the call -- calcuvw never clears it and
!               the u row sums read its diagonal; tscheme steady|bdf|bdf2|cn;
!               bicgstab with maxiter iterations, 0 leaves u, v, w unsolved;
!               no buoyancy, Lorentz force, PISO or constant mass flux)
! Binary files are raw little-endian stream files (real(8) / integer(4)).
program ref_driver
  use types
  use parameters
  use geometry
  use sparse_matrix
  use gradients
  use fvxGradient, only: fvx_gauss => grad_gauss, fvx_gauss_corrected => grad_gauss_corrected
  use linear_solvers
  use scalar_fluxes
  implicit none

  character(len=256) :: fmt_arg, jobfile
  character(len=32) :: cmd
  integer :: ju, ios

  call get_command_argument(1, fmt_arg)
  call get_command_argument(2, jobfile)
  if (trim(fmt_arg) == 'foam') then
    mesh_format = 'foamMesh'
  else
    mesh_format = 'nativeMesh'
  endif
  ltest = .false.

  call read_mesh
  call create_CSR_matrix

  allocate(res(numCells), reso(numCells), pk(numCells), zk(numCells), d(numCells), &
           uk(numCells), vk(numCells))

  open(newunit=ju, file=trim(jobfile), status='old', action='read')
  do
    read(ju, *, iostat=ios) cmd
    if (ios /= 0) exit
    backspace(ju)
    select case (trim(cmd))
    case ('dump')
      read(ju, *) cmd
      call job_dump()
    case ('dump_periodic')
      read(ju, *) cmd
      call job_dump_periodic()
    case ('laplacian')
      call job_laplacian(ju)
    case ('grad_gauss')
      call job_grad_gauss(ju)
    case ('grad_lsq')
      call job_grad_lsq(ju)
    case ('grad_fvx')
      call job_grad_fvx(ju)
    case ('solve')
      call job_solve(ju)
    case ('convdiff')
      call job_convdiff(ju, .false.)
    case ('convdiff_cell')
      call job_convdiff(ju, .true.)
    case ('nonorth')
      call job_nonorth(ju)
    case ('update_boundary')
      call job_update_boundary(ju)
    case ('temperature')
      call job_temperature(ju)
    case ('solve_explicit')
      call job_solve_explicit(ju)
    case ('ilu')
      call job_ilu(ju)
    case ('pcorr_assemble')
      call job_pcorr_assemble(ju)
    case ('pcorr_correct')
      call job_pcorr_correct(ju)
    case ('momentum')
      call job_momentum(ju)
    case default
      read(ju, *) cmd
      write(*,*) 'ref_driver: unknown job ', trim(cmd)
      stop 2
    end select
  end do
  close(ju)

contains

  subroutine wr8(fname, v, nv)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: nv
    real(dp), intent(in) :: v(nv)
    integer :: u
    open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
    write(u) v
    close(u)
  end subroutine

  subroutine wi4(fname, v, nv)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: nv
    integer, intent(in) :: v(nv)
    integer :: u
    open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
    write(u) v
    close(u)
  end subroutine

  subroutine rd8(fname, v, nv)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: nv
    real(dp), intent(out) :: v(nv)
    integer :: u
    open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='old')
    read(u) v
    close(u)
  end subroutine

  subroutine job_dump()
    integer :: u
    call wr8('g_xc.bin', xc, numCells)
    call wr8('g_yc.bin', yc, numCells)
    call wr8('g_zc.bin', zc, numCells)
    call wr8('g_vol.bin', vol, numCells)
    call wr8('g_arx.bin', arx, numFaces)
    call wr8('g_ary.bin', ary, numFaces)
    call wr8('g_arz.bin', arz, numFaces)
    call wr8('g_xf.bin', xf, numFaces)
    call wr8('g_yf.bin', yf, numFaces)
    call wr8('g_zf.bin', zf, numFaces)
    call wr8('g_facint.bin', facint, numInnerFaces)
    call wr8('g_Df.bin', Df, numInnerFaces)
    if (nwal > 0) then
      call wr8('g_dnw.bin', dnw, nwal)
      call wr8('g_srdw.bin', srdw, nwal)
    endif
    if (nsym > 0) then
      call wr8('g_dns.bin', dns, nsym)
      call wr8('g_srds.bin', srds, nsym)
    endif
    call wi4('c_ia.bin', ia, numCells+1)
    call wi4('c_ja.bin', ja, nnz)
    call wi4('c_diag.bin', diag, numCells)
    call wi4('c_slot_pn.bin', icell_jcell_csr_index, numInnerFaces)
    call wi4('c_slot_np.bin', jcell_icell_csr_index, numInnerFaces)
    open(newunit=u, file='sizes.txt', status='replace')
    write(u,*) numNodes, numCells, numInnerFaces, numBoundaryFaces, numFaces, nnz, nwal, nsym
    close(u)
  end subroutine

  subroutine job_dump_periodic()
    integer :: u
    call wi4('p_owner.bin', owner, numFaces)
    call wi4('p_slot_pn.bin', icell_jcell_csr_index, numInnerFaces+numPeriodic)
    call wi4('p_slot_np.bin', jcell_icell_csr_index, numInnerFaces+numPeriodic)
    open(newunit=u, file='psizes.txt', status='replace')
    write(u,*) numPeriodic, nper
    close(u)
  end subroutine

  subroutine job_grad_fvx(ju)
    integer, intent(in) :: ju
    character(len=64) :: c, fphi, fgin, fo2, foc
    real(dp), allocatable :: phi(:), g(:), dx(:), dy(:), dz(:)
    integer :: i
    read(ju, *) c, fphi, fgin, fo2, foc
    allocate(phi(numTotal), g(3*numCells), dx(numCells), dy(numCells), dz(numCells))
    call rd8(fphi, phi, numTotal)
    call fvx_gauss(phi, dx, dy, dz)
    do i = 1, numCells
      g(3*i-2) = dx(i)
      g(3*i-1) = dy(i)
      g(3*i) = dz(i)
    end do
    call wr8(fo2, g, 3*numCells)
    call rd8(fgin, g, 3*numCells)
    do i = 1, numCells
      dx(i) = g(3*i-2)
      dy(i) = g(3*i-1)
      dz(i) = g(3*i)
    end do
    call fvx_gauss_corrected(phi, dx, dy, dz)
    do i = 1, numCells
      g(3*i-2) = dx(i)
      g(3*i-1) = dy(i)
      g(3*i) = dz(i)
    end do
    call wr8(foc, g, 3*numCells)
    deallocate(phi, g, dx, dy, dz)
  end subroutine

  subroutine job_laplacian(ju)
    integer, intent(in) :: ju
    character(len=64) :: c, fmu, fphi, fsu0, fa, fsu
    real(dp), allocatable :: mu(:), phi(:)
    read(ju, *) c, fmu, fphi, fsu0, fa, fsu
    allocate(mu(numCells), phi(numTotal))
    call rd8(fmu, mu, numCells)
    call rd8(fphi, phi, numTotal)
    call rd8(fsu0, su, numCells)
    call laplacian(mu, phi)
    call wr8(fa, a, nnz)
    call wr8(fsu, su, numCells)
    deallocate(mu, phi)
  end subroutine

  subroutine job_grad_gauss(ju)
    integer, intent(in) :: ju
    character(len=64) :: c, fphi, fout
    real(dp), allocatable :: phi(:), g(:,:)
    read(ju, *) c, fphi, fout
    allocate(phi(numTotal), g(3,numTotal))
    call rd8(fphi, phi, numTotal)
    call grad_gauss(phi, g)
    call wr8(fout, g, 3*numTotal)
    deallocate(phi, g)
  end subroutine

  subroutine job_grad_lsq(ju)
    integer, intent(in) :: ju
    character(len=64) :: c, variant, fphi, fout
    real(dp), allocatable :: phi(:), g(:,:)
    read(ju, *) c, variant, fphi, fout
    lstsq = .false.; lstsq_qr = .false.; lstsq_dm = .false.
    gauss = .false.; node_gauss = .false.
    limiter = 'none'
    select case (trim(variant))
    case ('lsq')
      lstsq = .true.
    case ('lsq_dm')
      lstsq_dm = .true.
    case ('lsq_qr')
      lstsq_qr = .true.
    end select
    call create_lsq_grad_matrix
    allocate(phi(numTotal), g(3,numTotal))
    call rd8(fphi, phi, numTotal)
    call grad(phi, g)
    call wr8(fout, g, 3*numTotal)
    deallocate(phi, g)
  end subroutine

  subroutine job_solve(ju)
    integer, intent(in) :: ju
    character(len=64) :: c, solver, fa, ffi, frhs, fout, finfo
    integer :: itr_max, u
    real(dp) :: tol_abs, tol_rel, res0
    real(dp), allocatable :: fi(:), rhs(:)
    read(ju, *) c, solver, itr_max, tol_abs, tol_rel, fa, ffi, frhs, fout, finfo
    allocate(fi(numTotal), rhs(numCells))
    call rd8(fa, a, nnz)
    fi = 0.0_dp
    call rd8(ffi, fi, numCells)
    call rd8(frhs, rhs, numCells)
    res = 0.0_dp; reso = 0.0_dp; pk = 0.0_dp; zk = 0.0_dp; d = 0.0_dp
    uk = 0.0_dp; vk = 0.0_dp
    res0 = 0.0_dp
    call csrsolve(trim(solver), fi, rhs, res0, itr_max, tol_abs, tol_rel, 'phi')
    call wr8(fout, fi, numCells)
    open(newunit=u, file=trim(finfo), status='replace')
    write(u,'(es25.17)') res0
    close(u)
    deallocate(fi, rhs)
  end subroutine

  subroutine job_solve_explicit(ju)
    integer, intent(in) :: ju
    character(len=64) :: c, solver, fia, fja, fdg, fa, ffi, frhs, fout, finfo
    integer :: n, nz, itr_max, u
    real(dp) :: tol_abs, tol_rel, resor
    integer, allocatable :: eia(:), eja(:), edg(:)
    real(dp), allocatable :: ea(:), fi(:), rhs(:)
    read(ju, *) c, solver, n, nz, itr_max, tol_abs, tol_rel, fia, fja, fdg, fa, ffi, frhs, fout, finfo
    allocate(eia(n+1), eja(nz), edg(n), ea(nz), fi(n), rhs(n))
    open(newunit=u, file=trim(fia), access='stream', form='unformatted', status='old'); read(u) eia; close(u)
    open(newunit=u, file=trim(fja), access='stream', form='unformatted', status='old'); read(u) eja; close(u)
    open(newunit=u, file=trim(fdg), access='stream', form='unformatted', status='old'); read(u) edg; close(u)
    call rd8(fa, ea, nz)
    call rd8(ffi, fi, n)
    call rd8(frhs, rhs, n)
    deallocate(res, reso, pk, zk, d, uk, vk)
    allocate(res(n), reso(n), pk(n), zk(n), d(n), uk(n), vk(n))
    res = 0.0_dp; reso = 0.0_dp; pk = 0.0_dp; zk = 0.0_dp; d = 0.0_dp; uk = 0.0_dp; vk = 0.0_dp
    resor = 0.0_dp
    select case (trim(solver))
    case ('iccg')
      call iccg(n, nz, eia, eja, ea, edg, fi, rhs, resor, itr_max, tol_abs, tol_rel, 'x', .false.)
    case ('dpcg')
      call dpcg(n, nz, eia, eja, ea, edg, fi, rhs, resor, itr_max, tol_abs, tol_rel, 'x', .false.)
    case ('bicgstab')
      call bicgstab(n, nz, eia, eja, ea, edg, fi, rhs, resor, itr_max, tol_abs, tol_rel, 'x', .false.)
    case ('gauss-seidel')
      call GaussSeidel(n, nz, eia, eja, ea, edg, fi, rhs, resor, itr_max, tol_abs, tol_rel, 'x', .false.)
    case ('pmgmres')
      call pmgmres_ilu(n, nz, eia, eja, ea, edg, fi, rhs, resor, itr_max, 4, tol_abs, tol_rel, 'x', .false.)
    end select
    call wr8(fout, fi, n)
    open(newunit=u, file=trim(finfo), status='replace')
    write(u,'(es25.17)') resor
    close(u)
    deallocate(res, reso, pk, zk, d, uk, vk)
    allocate(res(numCells), reso(numCells), pk(numCells), zk(numCells), d(numCells), &
             uk(numCells), vk(numCells))
    deallocate(eia, eja, edg, ea, fi, rhs)
  end subroutine

  ! ILU(0) factors (ilu_cr), one preconditioner application (lus_cr) and
  ! the transposed product (atx_cr) of an explicit CSR matrix.
  subroutine job_ilu(ju)
    integer, intent(in) :: ju
    character(len=64) :: c, fia, fja, fdg, fa, fr, fl, fz, fw
    integer :: n, nz, u
    integer, allocatable :: eia(:), eja(:), edg(:)
    real(dp), allocatable :: ea(:), el(:), r(:), z(:), w(:)
    read(ju, *) c, n, nz, fia, fja, fdg, fa, fr, fl, fz, fw
    allocate(eia(n+1), eja(nz), edg(n), ea(nz), el(nz), r(n), z(n), w(n))
    open(newunit=u, file=trim(fia), access='stream', form='unformatted', status='old'); read(u) eia; close(u)
    open(newunit=u, file=trim(fja), access='stream', form='unformatted', status='old'); read(u) eja; close(u)
    open(newunit=u, file=trim(fdg), access='stream', form='unformatted', status='old'); read(u) edg; close(u)
    call rd8(fa, ea, nz)
    call rd8(fr, r, n)
    call ilu_cr(n, nz, eia, eja, ea, edg, el)
    call lus_cr(n, nz, eia, eja, el, edg, r, z)
    call atx_cr(n, nz, eia, eja, ea, r, w)
    call wr8(fl, el, nz)
    call wr8(fz, z, n)
    call wr8(fw, w, n)
    deallocate(eia, eja, edg, ea, el, r, z, w)
  end subroutine

  ! Steady convection-diffusion assembly with the reference flux routines,
  ! composed as calculate_temperature does (temperature.f90:118-366).
  ! bckind per boundary face: 0 Dirichlet wall, 1 zero-gradient, 2 flux bc.
  subroutine job_convdiff(ju, cellgam)
    integer, intent(in) :: ju
    logical, intent(in) :: cellgam
    character(len=64) :: c, ffm, fdc, fphi, fgr, fbk, fa, fsu, fsp
    character(len=30) :: scheme
    real(dp) :: gam, cap, can, suadd, dw, off
    real(dp), allocatable :: fm(:), dc(:), phi(:), gr(:,:), bk(:), spl(:), vis(:)
    integer :: i, ijp, ijn, k, iface, ijb, inp
    read(ju, *) c, ffm, fdc, fphi, fgr, gam, scheme, fbk, fa, fsu, fsp
    allocate(fm(numFaces), dc(numFaces), phi(numTotal), gr(3,numTotal), bk(numBoundaryFaces), &
             spl(numCells), vis(numCells))
    call rd8(ffm, fm, numFaces)
    if (cellgam) then
      call rd8(fdc, vis, numCells)
      dc = 0.0_dp
      do iface = numInnerFaces + 1, numFaces
        dc(iface) = vis(owner(iface))
      enddo
    else
      call rd8(fdc, dc, numFaces)
    endif
    call rd8(fphi, phi, numTotal)
    call rd8(fgr, gr, 3*numTotal)
    call rd8(fbk, bk, numBoundaryFaces)
    a = 0.0_dp; su = 0.0_dp; spl = 0.0_dp
    do i = 1, numInnerFaces
      ijp = owner(i)
      ijn = neighbour(i)
      if (cellgam) then
        call facefluxscalar(i, ijp, ijn, xf(i), yf(i), zf(i), arx(i), ary(i), arz(i), &
                            fm(i), facint(i), gam, scheme, phi, gr, vis, cap, can, suadd)
      else
        call facefluxsc(i, ijp, ijn, xf(i), yf(i), zf(i), arx(i), ary(i), arz(i), &
                        fm(i), facint(i), gam, scheme, phi, gr, dc(i), cap, can, suadd)
      endif
      k = icell_jcell_csr_index(i); a(k) = can
      k = jcell_icell_csr_index(i); a(k) = cap
      k = diag(ijp); a(k) = a(k) - can
      k = diag(ijn); a(k) = a(k) - cap
      su(ijp) = su(ijp) + suadd
      su(ijn) = su(ijn) - suadd
    enddo
    do i = 1, numBoundaryFaces
      iface = numInnerFaces + i
      ijp = owner(iface)
      ijb = numCells + i
      if (nint(bk(i)) == 2) then
        call facefluxsc(ijp, xf(iface), yf(iface), zf(iface), arx(iface), ary(iface), arz(iface), &
                        fm(iface), gr, dc(iface), cap, can, suadd)
        spl(ijp) = spl(ijp) - can
        su(ijp) = su(ijp) - can*phi(ijb) + suadd
      elseif (nint(bk(i)) == 0) then
        dw = dc(iface) * ( sqrt(arx(iface)**2+ary(iface)**2+arz(iface)**2) / &
             ( (xf(iface)-xc(ijp))*(arx(iface)/sqrt(arx(iface)**2+ary(iface)**2+arz(iface)**2)) + &
               (yf(iface)-yc(ijp))*(ary(iface)/sqrt(arx(iface)**2+ary(iface)**2+arz(iface)**2)) + &
               (zf(iface)-zc(ijp))*(arz(iface)/sqrt(arx(iface)**2+ary(iface)**2+arz(iface)**2)) ) )
        spl(ijp) = spl(ijp) + dw
        su(ijp) = su(ijp) + dw*phi(ijb)
      else
        phi(ijb) = phi(ijp)
      endif
    enddo
    do inp = 1, numCells
      off = sum( a(ia(inp) : ia(inp+1)-1) ) - a(diag(inp))
      a(diag(inp)) = spl(inp) - off
    enddo
    call wr8(fa, a, nnz)
    call wr8(fsu, su, numCells)
    call wr8(fsp, spl, numCells)
    deallocate(fm, dc, phi, gr, bk, spl, vis)
  end subroutine

  ! The explicit non-orthogonal correction of the inner faces, taken from
  ! the reference's facefluxsc with zero mass flux (then fcfie = fcfii = 0
  ! and suadd = fdfie exactly), scattered as temperature.f90:152-155 does.
  subroutine job_nonorth(ju)
    integer, intent(in) :: ju
    character(len=64) :: c, fdc, fgr, fsu0, fsu
    character(len=30) :: scheme
    real(dp) :: cap, can, suadd
    real(dp), allocatable :: dc(:), phi(:), gr(:,:)
    integer :: i, ijp, ijn
    read(ju, *) c, fdc, fgr, fsu0, fsu
    allocate(dc(numFaces), phi(numTotal), gr(3,numTotal))
    call rd8(fdc, dc, numFaces)
    call rd8(fgr, gr, 3*numTotal)
    call rd8(fsu0, su, numCells)
    phi = 0.0_dp
    scheme = 'cds'
    do i = 1, numInnerFaces
      ijp = owner(i)
      ijn = neighbour(i)
      call facefluxsc(i, ijp, ijn, xf(i), yf(i), zf(i), arx(i), ary(i), arz(i), &
                      0.0_dp, facint(i), 1.0_dp, scheme, phi, gr, dc(i), cap, can, suadd)
      su(ijp) = su(ijp) + suadd
      su(ijn) = su(ijn) - suadd
    enddo
    call wr8(fsu, su, numCells)
    deallocate(dc, phi, gr)
  end subroutine

  subroutine job_update_boundary(ju)
    integer, intent(in) :: ju
    character(len=64) :: c, fphi, fout
    real(dp), allocatable :: phi(:)
    read(ju, *) c, fphi, fout
    allocate(phi(numTotal))
    call rd8(fphi, phi, numTotal)
    call updateBoundary(phi)
    call wr8(fout, phi, numTotal)
    deallocate(phi)
  end subroutine

  ! One call of the reference's calculate_temperature: every module global it
  ! reads is set from the job line, then the assembled system (a, su, sp),
  ! the updated field, its gradient and the solver's res0 are written.
  subroutine job_temperature(ju)
    use variables, only: T, To, Too, dTdxi, den, vis, flmass
    use temperature
    integer, intent(in) :: ju
    character(len=64) :: c, fT, fTo, fToo, fden, fvis, ffm, fwf, fTout, fa, fsu, fsp, fgr, finfo
    character(len=30) :: tsch, schm, slv, gmeth
    real(dp) :: vsc, prn, sgt, urf, gds, dtt, tabs, trel
    integer :: itmax, uu
    real(dp), allocatable :: wf(:)
    read(ju, *) c, fT, fTo, fToo, fden, fvis, ffm, fwf, vsc, prn, sgt, urf, gds, dtt, tsch, schm, &
                slv, itmax, tabs, trel, gmeth, fTout, fa, fsu, fsp, fgr, finfo
    if (allocated(T)) deallocate(T, To, Too, dTdxi, den, vis, flmass)
    allocate(T(numTotal), To(numTotal), Too(numTotal), dTdxi(3,numTotal), den(numTotal), &
             vis(numTotal), flmass(numFaces), wf(numBoundaries))
    To = 0.0_dp; Too = 0.0_dp; den = 0.0_dp; vis = 0.0_dp; dTdxi = 0.0_dp
    call rd8(fT, T, numTotal)
    call rd8(fTo, To, numCells)
    call rd8(fToo, Too, numCells)
    call rd8(fden, den, numCells)
    call rd8(fvis, vis, numCells)
    call rd8(ffm, flmass, numFaces)
    call rd8(fwf, wf, numBoundaries)
    bcDFraction(7, 1:numBoundaries) = wf
    viscos = vsc
    pranl = prn
    sigt = sgt
    urfT = urf
    gdsT = gds
    timestep = dtt
    ltransient = trim(tsch) /= 'steady'
    bdf = trim(tsch) == 'bdf'
    bdf2 = trim(tsch) == 'bdf2'
    bdf3 = .false.
    cn = trim(tsch) == 'cn'
    lturb = .false.
    cSchemeT = schm
    lSolverT = slv
    maxiterT = itmax
    tolAbsT = tabs
    tolRelT = trel
    lstsq = trim(gmeth) == 'lsq'
    lstsq_dm = trim(gmeth) == 'lsq_dm'
    lstsq_qr = trim(gmeth) == 'lsq_qr'
    gauss = trim(gmeth) == 'gauss'
    node_gauss = .false.
    limiter = 'none'
    if (lstsq .or. lstsq_dm .or. lstsq_qr) call create_lsq_grad_matrix
    res = 0.0_dp; reso = 0.0_dp; pk = 0.0_dp; zk = 0.0_dp; d = 0.0_dp; uk = 0.0_dp; vk = 0.0_dp
    resor = 0.0_dp
    call calculate_temperature
    call wr8(fTout, T, numTotal)
    call wr8(fa, a, nnz)
    call wr8(fsu, su, numCells)
    call wr8(fsp, sp, numCells)
    call wr8(fgr, dTdxi, 3*numTotal)
    open(newunit=uu, file=trim(finfo), status='replace')
    write(uu,'(es25.17)') resor(7)
    close(uu)
    deallocate(wf)
  end subroutine

  subroutine pcorr_alloc()
    use variables, only: u, v, w, p, pp, den, dPdxi, flmass
    if (allocated(u)) deallocate(u, v, w, p, pp, den, dPdxi, flmass)
    allocate(u(numTotal), v(numTotal), w(numTotal), p(numTotal), pp(numTotal), den(numTotal), &
             dPdxi(3,numTotal), flmass(numFaces))
    u = 0.0_dp; v = 0.0_dp; w = 0.0_dp; p = 0.0_dp; pp = 0.0_dp; den = 0.0_dp
    dPdxi = 0.0_dp; flmass = 0.0_dp
  end subroutine

  subroutine job_pcorr_assemble(ju)
    use variables, only: u, v, w, p, den, dPdxi, flmass
    use faceflux_mass
    integer, intent(in) :: ju
    character(len=64) :: c, fu, fv, fw, fp, fdp, fden, fapu, fapv, fapw, ffm, fa, fsu, ffo, &
                         fuo, fvo, fwo
    integer :: cmf, i, k, ib, iface, ijp, ijn, iPer, l, ifc, iftwin
    real(dp) :: flo, cap, can
    read(ju, *) c, fu, fv, fw, fp, fdp, fden, fapu, fapv, fapw, ffm, cmf, flo, fa, fsu, ffo, &
                fuo, fvo, fwo
    call pcorr_alloc()
    call rd8(fu, u, numTotal)
    call rd8(fv, v, numTotal)
    call rd8(fw, w, numTotal)
    call rd8(fp, p, numTotal)
    call rd8(fdp, dPdxi, 3*numTotal)
    call rd8(fden, den, numTotal)
    call rd8(fapu, apu, numCells)
    call rd8(fapv, apv, numCells)
    call rd8(fapw, apw, numCells)
    call rd8(ffm, flmass, numFaces)
    const_mflux = cmf /= 0
    flomas = flo
    a = 0.0_dp
    su = 0.0_dp
    do i = 1, numInnerFaces
      ijp = owner(i)
      ijn = neighbour(i)
      call facefluxmass2(i, ijp, ijn, arx(i), ary(i), arz(i), facint(i), cap, can, flmass(i))
      k = icell_jcell_csr_index(i); a(k) = cap
      k = jcell_icell_csr_index(i); a(k) = cap
      k = diag(ijp); a(k) = a(k) - cap
      k = diag(ijn); a(k) = a(k) - cap
      su(ijp) = su(ijp) - flmass(i)
      su(ijn) = su(ijn) + flmass(i)
    enddo
    if (.not. const_mflux) call adjustMassFlow
    iPer = 0
    l = numInnerFaces
    do ib = 1, numBoundaries
      if (bctype(ib) == 'inlet' .or. bctype(ib) == 'outlet') then
        do i = 1, nfaces(ib)
          iface = startFace(ib) + i
          ijp = owner(iface)
          su(ijp) = su(ijp) - flmass(iface)
        enddo
      elseif (bctype(ib) == 'periodic') then
        iPer = iPer + 1
        do i = 1, nfaces(ib)
          ifc = startFace(ib) + i
          ijp = owner(ifc)
          iftwin = startFaceTwin(iPer) + i
          ijn = owner(iftwin)
          call facefluxmass2_periodic(i, ijp, ijn, xf(ifc), yf(ifc), zf(ifc), arx(ifc), ary(ifc), &
                                      arz(ifc), half, cap, can, flmass(ifc))
          l = l + 1
          k = icell_jcell_csr_index(l); a(k) = cap
          k = jcell_icell_csr_index(l); a(k) = cap
          k = diag(ijp); a(k) = a(k) - cap
          k = diag(ijn); a(k) = a(k) - cap
          su(ijp) = su(ijp) - flmass(ifc)
          su(ijn) = su(ijn) + flmass(ifc)
        enddo
      endif
    enddo
    call wr8(fa, a, nnz)
    call wr8(fsu, su, numCells)
    call wr8(ffo, flmass, numFaces)
    call wr8(fuo, u, numTotal)
    call wr8(fvo, v, numTotal)
    call wr8(fwo, w, numTotal)
  end subroutine

  subroutine job_pcorr_correct(ju)
    use variables, only: u, v, w, p, pp, dPdxi, flmass
    use nablap
    use velocity, only: updateVelocityAtBoundary
    integer, intent(in) :: ju
    character(len=64) :: c, fpp, ffm, fa, fu, fv, fw, fp, fapu, fapv, fapw, fppo, ffmo, fuo, fvo, &
                         fwo, fpo, fdpo, fsu, fsv, fsw
    integer :: iref, ib, i, iface, ijp, ijn, iPer, l, iftwin, k
    real(dp) :: urfpp, ppref
    read(ju, *) c, fpp, ffm, fa, fu, fv, fw, fp, fapu, fapv, fapw, urfpp, iref, fppo, ffmo, fuo, &
                fvo, fwo, fpo, fdpo, fsu, fsv, fsw
    call pcorr_alloc()
    call rd8(fpp, pp, numTotal)
    call rd8(ffm, flmass, numFaces)
    call rd8(fa, a, nnz)
    call rd8(fu, u, numTotal)
    call rd8(fv, v, numTotal)
    call rd8(fw, w, numTotal)
    call rd8(fp, p, numTotal)
    call rd8(fapu, apu, numCells)
    call rd8(fapv, apv, numCells)
    call rd8(fapw, apw, numCells)
    pRefCell = iref
    do iface = 1, numInnerFaces
      ijp = owner(iface)
      ijn = neighbour(iface)
      k = icell_jcell_csr_index(iface)
      flmass(iface) = flmass(iface) + a(k) * (pp(ijn)-pp(ijp))
    enddo
    iPer = 0
    l = numInnerFaces
    do ib = 1, numBoundaries
      if (bctype(ib) == 'periodic') then
        iPer = iPer + 1
        do i = 1, nfaces(ib)
          iface = startFace(ib) + i
          ijp = owner(iface)
          iftwin = startFaceTwin(iPer) + i
          ijn = owner(iftwin)
          l = l + 1
          k = icell_jcell_csr_index(l)
          flmass(iface) = flmass(iface) + a(k) * (pp(ijn)-pp(ijp))
          flmass(iftwin) = flmass(iface)
        enddo
      endif
    enddo
    ppref = pp(pRefCell)
    call gradp_and_sources(pp)
    u(1:numCells) = u(1:numCells) + su*apu
    v(1:numCells) = v(1:numCells) + sv*apv
    w(1:numCells) = w(1:numCells) + sw*apw
    p(1:numCells) = p(1:numCells) + urfpp*(pp(1:numCells)-ppref)
    call updateVelocityAtBoundary
    call wr8(fppo, pp, numTotal)
    call wr8(ffmo, flmass, numFaces)
    call wr8(fuo, u, numTotal)
    call wr8(fvo, v, numTotal)
    call wr8(fwo, w, numTotal)
    call wr8(fpo, p, numTotal)
    call wr8(fdpo, dPdxi, 3*numTotal)
    call wr8(fsu, su, numCells)
    call wr8(fsv, sv, numCells)
    call wr8(fsw, sw, numCells)
  end subroutine

  subroutine job_momentum(ju)
    use variables, only: u, v, w, uo, vo, wo, uoo, voo, woo, p, den, vis, visw, flmass, &
                         dUdxi, dVdxi, dWdxi, dPdxi
    use velocity, only: calcuvw, urfU, gdsU, cSchemeU, lSolverU, maxiterU, tolAbsU, tolRelU
    use mhd, only: calcEpot
    integer, intent(in) :: ju
    character(len=64) :: c, fu, fv, fw, fuo, fvo, fwo, fuoo, fvoo, fp, fden, fvis, fvisw, ffm, fas
    character(len=64) :: fa, fsu, fsv, fsw, fspu, fspv, fsp, fapu, fapv, fapw, fuo2, fvo2, fwo2
    character(len=64) :: fpo, fdpo, fduo
    character(len=30) :: tsch, schm
    real(dp) :: vsc, dtt, gds, u1, u2, u3, trel
    integer :: itmax
    read(ju, *) c, fu, fv, fw, fuo, fvo, fwo, fuoo, fvoo, fp, fden, fvis, fvisw, ffm, fas, vsc, dtt, &
                tsch, schm, gds, u1, u2, u3, itmax, trel, fa, fsu, fsv, fsw, fspu, fspv, fsp, fapu, &
                fapv, fapw, fuo2, fvo2, fwo2, fpo, fdpo, fduo
    if (allocated(u)) deallocate(u, v, w, p, den, flmass)
    if (allocated(dPdxi)) deallocate(dPdxi)
    if (allocated(uo)) deallocate(uo, vo, wo, uoo, voo, woo)
    if (allocated(vis)) deallocate(vis)
    if (allocated(visw)) deallocate(visw)
    if (allocated(dUdxi)) deallocate(dUdxi, dVdxi, dWdxi)
    allocate(u(numTotal), v(numTotal), w(numTotal), p(numTotal), den(numTotal), flmass(numFaces), &
             dPdxi(3,numTotal), uo(numTotal), vo(numTotal), wo(numTotal), uoo(numTotal), &
             voo(numTotal), woo(numTotal), vis(numTotal), visw(max(nwal,1)), dUdxi(3,numTotal), &
             dVdxi(3,numTotal), dWdxi(3,numTotal))
    uo = 0.0_dp; vo = 0.0_dp; wo = 0.0_dp; uoo = 0.0_dp; voo = 0.0_dp; woo = 0.0_dp
    dPdxi = 0.0_dp; dUdxi = 0.0_dp; dVdxi = 0.0_dp; dWdxi = 0.0_dp; visw = 0.0_dp
    call rd8(fu, u, numTotal)
    call rd8(fv, v, numTotal)
    call rd8(fw, w, numTotal)
    call rd8(fuo, uo, numCells)
    call rd8(fvo, vo, numCells)
    call rd8(fwo, wo, numCells)
    call rd8(fuoo, uoo, numCells)
    call rd8(fvoo, voo, numCells)
    woo = 0.5_dp*uoo
    call rd8(fp, p, numTotal)
    call rd8(fden, den, numTotal)
    call rd8(fvis, vis, numTotal)
    if (nwal > 0) call rd8(fvisw, visw, nwal)
    call rd8(ffm, flmass, numFaces)
    call rd8(fas, a, nnz)
    viscos = vsc
    timestep = dtt
    ltransient = trim(tsch) /= 'steady'
    bdf = trim(tsch) == 'bdf'
    bdf2 = trim(tsch) == 'bdf2'
    bdf3 = .false.
    cn = trim(tsch) == 'cn'
    piso = .false.
    lbuoy = .false.
    const_mflux = .false.
    calcEpot = .false.
    lstsq = .false.; lstsq_dm = .false.; lstsq_qr = .false.; gauss = .true.
    node_gauss = .false.
    limiter = 'none'
    urfU = (/ u1, u2, u3 /)
    gdsU = gds
    cSchemeU = schm
    lSolverU = 'bicgstab'
    maxiterU = itmax
    tolAbsU = 0.0_dp
    tolRelU = trel
    su = 0.0_dp; sv = 0.0_dp; sw = 0.0_dp
    res = 0.0_dp; reso = 0.0_dp; pk = 0.0_dp; zk = 0.0_dp; d = 0.0_dp; uk = 0.0_dp; vk = 0.0_dp
    resor = 0.0_dp
    call calcuvw
    call wr8(fa, a, nnz)
    call wr8(fsu, su, numCells)
    call wr8(fsv, sv, numCells)
    call wr8(fsw, sw, numCells)
    call wr8(fspu, spu, numCells)
    call wr8(fspv, spv, numCells)
    call wr8(fsp, sp, numCells)
    call wr8(fapu, apu, numCells)
    call wr8(fapv, apv, numCells)
    call wr8(fapw, apw, numCells)
    call wr8(fuo2, u, numTotal)
    call wr8(fvo2, v, numTotal)
    call wr8(fwo2, w, numTotal)
    call wr8(fpo, p, numTotal)
    call wr8(fdpo, dPdxi, 3*numTotal)
    call wr8(fduo, dUdxi, 3*numTotal)
  end subroutine

end program ref_driver
