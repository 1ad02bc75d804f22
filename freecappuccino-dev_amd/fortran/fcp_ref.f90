! fcp_ref.f90 -- the reference's procedural module API on top of libfcp.
!
! Same routine names, argument lists and module-global conventions as the
! reference so that callers keep working unchanged:
!   create_CSR_matrix          sparse_matrix.f90:86   (fills nnz, ia, ja, diag, a, su,
!                                                      icell_jcell_csr_index, jcell_icell_csr_index)
!   laplacian(mu, phi)         laplacian.f90:1        (writes a, su)
!   grad_gauss(u, dudxi)       gradients.f90:1607
!   csrsolve(solver, fi, rhs, res0, itr_max, tol_abs, tol_rel, chvar)
!                              linear_solvers.f90:40
!   iccg / dpcg / bicgstab / GaussSeidel(n, nnz, ia, ja, a, diag, fi, rhs, resor, itr_max,
!                          tol_abs, tol_rel, chvar, verbose)  linear_solvers.f90:364,206,548,96
!   pmgmres_ilu(n, nz_num, ia, ja, a, ua, x, rhs, resor, itr_max, mr, tol_abs, tol_rel,
!               chvar, verbose)                               linear_solvers.f90:789
!   spsolve(solver, fi, rhs, res0, itr_max, tol_abs, tol_rel, chvar)
!                              test/test_linear_solvers_spsolve.f90:103 (csrsolve's body on
!                              the module arrays; also 'bicgstab_ilu' and 'pmgmres_ilu')
!   read_mesh([case_dir])      geometry.f90:90 (mesh_format 'nativeMesh' or 'foamMesh')
!   grad(phi, dPhidxi) / grad(U, V, W, dUdxi, dVdxi, dWdxi)   gradients.f90:106-214
!   create_lsq_grad_matrix, grad_lsq, grad_lsq_dm, grad_lsq_qr gradients.f90:72,782,1334,1063
!   updateBoundary(phi)        updateBoundary.f90:1
!   fvm_laplacian / fvm_div / fvm_ddt   face-Gamma Laplacian, facefluxsc assembly
!                              (temperature.f90:118-366), time terms (:82-102)
!   calculate_scalar(...)      calculate_temperature (temperature.f90:35-385) for a
!                              scalar with explicit parameters
!   calcuvw(...)               velocity.f90:50-750 (assembly + the three solves with a
!                              diagonal per component) with explicit arguments
!   calcp_simple(...)          calcp_simple.f90:69-419 (incompressible, no pressure
!                              patches, one corrector) with explicit arguments
!   fvx_grad_gauss / fvx_grad_gauss_corrected   fvxGradient.f90:1549,1664
!   fvm_div_cell               facefluxscalar assembly (scalar_fluxes.f90:344-464)
! The mesh is handed over with read_mesh (from disk) or fcp_read_mesh (the
! arrays read_mesh produces, geometry.f90:118-406); geometry is computed on
! the GPU and returned in the module arrays below (geometry.f90:19-83 names).
module fcp_ref
  use iso_c_binding
  use fcp_capi
  implicit none

  integer, parameter :: dp = kind(1.0d0)

  type(c_ptr), save :: ctx = c_null_ptr
  logical, save :: ltest = .false.

  ! geometry module globals (subset)
  integer :: numCells = 0, numInnerFaces = 0, numBoundaryFaces = 0, numFaces = 0, numTotal = 0
  real(dp), allocatable :: xc(:), yc(:), zc(:), vol(:)
  real(dp), allocatable :: arx(:), ary(:), arz(:), xf(:), yf(:), zf(:), facint(:), Df(:)

  ! patches of the resident mesh (geometry.f90:64-80; types as FCP_BC_* codes)
  integer :: numBoundaries = 0
  integer, allocatable :: nfaces(:), startFace(:), bctype_code(:)
  character(len=10), save :: mesh_format = 'nativeMesh'  ! or 'foamMesh' (geometry.f90:92)

  ! sparse_matrix module globals (sparse_matrix.f90:10-26)
  integer :: nnz = 0
  integer, allocatable :: ia(:), ja(:), diag(:)
  integer, allocatable :: icell_jcell_csr_index(:), jcell_icell_csr_index(:)
  real(dp), allocatable, target :: a(:), su(:), sp(:)

  ! gradients module globals (gradients.f90:9-10); lsq_compat keeps the
  ! reference's least-squares arithmetic, defects included
  logical, save :: lstsq = .false., lstsq_qr = .false., lstsq_dm = .false., gauss = .true.
  logical, save :: node_gauss = .false.
  character(len=20), save :: limiter = 'none'
  logical, save :: lsq_compat = .true.

  interface grad
    module procedure grad_scalar_field
    module procedure grad_vector_field
  end interface

contains

  subroutine fcp_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    if (rc < 0) then
      call c_f_pointer(fcp_last_error(ctx), msg, [512])
      write(*,'(3a,i0)') '  fcp error in ', what, ': code ', rc
      do i = 1, 512
        if (msg(i) == c_null_char) exit
        write(*,'(a)', advance='no') msg(i)
      end do
      write(*,*)
      stop 1
    endif
  end subroutine

  subroutine fcp_read_mesh(nNodes, nCells, nInner, nBnd, nPatches, xyz, face_nnodes, node, owner, &
                           neighbour, nfaces, startFace, bctype, device)
    integer, intent(in) :: nNodes, nCells, nInner, nBnd, nPatches
    real(dp), intent(in) :: xyz(3, nNodes)
    integer, intent(in) :: face_nnodes(nInner+nBnd), node(FCP_NOMAX, nInner+nBnd)
    integer, intent(in) :: owner(nInner+nBnd), neighbour(max(nInner,1))
    integer, intent(in) :: nfaces(nPatches), startFace(nPatches), bctype(nPatches)
    integer, intent(in), optional :: device
    integer :: dev
    real(dp), allocatable :: dnb(:), srdb(:)
    dev = 0
    if (present(device)) dev = device
    if (.not. c_associated(ctx)) call fcp_check(fcp_init(dev, ctx), 'fcp_init')
    numCells = nCells; numInnerFaces = nInner; numBoundaryFaces = nBnd
    numFaces = nInner + nBnd; numTotal = nCells + nBnd
    call fcp_check(fcp_mesh_upload(ctx, nNodes, nCells, nInner, nBnd, nPatches, xyz, face_nnodes, node, &
                                   owner, neighbour, nfaces, startFace, bctype), 'fcp_mesh_upload')
    if (allocated(xc)) deallocate(xc, yc, zc, vol, arx, ary, arz, xf, yf, zf, facint, Df)
    allocate(xc(nCells), yc(nCells), zc(nCells), vol(nCells))
    allocate(arx(numFaces), ary(numFaces), arz(numFaces), xf(numFaces), yf(numFaces), zf(numFaces))
    allocate(facint(max(nInner,1)), Df(max(nInner,1)), dnb(max(nBnd,1)), srdb(max(nBnd,1)))
    call fcp_check(fcp_geometry(ctx, xc, yc, zc, vol, arx, ary, arz, xf, yf, zf, facint, Df, &
                                dnb, srdb), 'fcp_geometry')
    deallocate(dnb, srdb)
  end subroutine

  ! read_mesh (geometry.f90:90-116): the case's polyMesh files, parsed and
  ! uploaded by the library; geometry into the module arrays.
  subroutine read_mesh(case_dir, device)
    character(len=*), intent(in), optional :: case_dir
    integer, intent(in), optional :: device
    character(len=1024) :: cdir
    integer(c_int) :: sizes(5), fmt, dev
    real(dp), allocatable :: dnb(:), srdb(:)
    cdir = '.'
    if (present(case_dir)) cdir = case_dir
    dev = 0
    if (present(device)) dev = device
    fmt = 0
    if (trim(mesh_format) == 'foamMesh') fmt = 1
    if (.not. c_associated(ctx)) call fcp_check(fcp_init(dev, ctx), 'fcp_init')
    call fcp_check(fcp_read_polymesh(ctx, trim(cdir), len_trim(cdir), fmt, sizes), 'fcp_read_polymesh')
    numCells = sizes(2); numInnerFaces = sizes(3); numBoundaryFaces = sizes(4)
    numFaces = numInnerFaces + numBoundaryFaces; numTotal = numCells + numBoundaryFaces
    numBoundaries = sizes(5)
    if (allocated(nfaces)) deallocate(nfaces, startFace, bctype_code)
    allocate(nfaces(max(numBoundaries,1)), startFace(max(numBoundaries,1)), &
             bctype_code(max(numBoundaries,1)))
    call fcp_check(fcp_mesh_patches(ctx, nfaces, startFace, bctype_code), 'fcp_mesh_patches')
    if (allocated(xc)) deallocate(xc, yc, zc, vol, arx, ary, arz, xf, yf, zf, facint, Df)
    allocate(xc(numCells), yc(numCells), zc(numCells), vol(numCells))
    allocate(arx(numFaces), ary(numFaces), arz(numFaces), xf(numFaces), yf(numFaces), zf(numFaces))
    allocate(facint(max(numInnerFaces,1)), Df(max(numInnerFaces,1)), dnb(max(numBoundaryFaces,1)), &
             srdb(max(numBoundaryFaces,1)))
    call fcp_check(fcp_geometry(ctx, xc, yc, zc, vol, arx, ary, arz, xf, yf, zf, facint, Df, &
                                dnb, srdb), 'fcp_geometry')
    deallocate(dnb, srdb)
  end subroutine

  subroutine create_CSR_matrix
    integer(c_int) :: n
    if (allocated(ia)) deallocate(ia, ja, diag, icell_jcell_csr_index, jcell_icell_csr_index, a, su, sp)
    nnz = 2*numInnerFaces + numCells                      ! sparse_matrix.f90:110
    allocate(ia(numCells+1), ja(nnz), diag(numCells), a(nnz), su(numCells), sp(numCells))
    allocate(icell_jcell_csr_index(max(numInnerFaces,1)), jcell_icell_csr_index(max(numInnerFaces,1)))
    call fcp_check(fcp_csr_build(ctx, n, ia, ja, diag, icell_jcell_csr_index, jcell_icell_csr_index), &
                   'fcp_csr_build')
    nnz = n
    a = 0.0_dp
    su = 0.0_dp
    sp = 0.0_dp
  end subroutine

  subroutine laplacian(mu, phi)
    real(dp), intent(in) :: mu(numCells), phi(numTotal)
    call fcp_check(fcp_laplacian(ctx, mu, phi, a, su), 'fcp_laplacian')
  end subroutine

  subroutine grad_gauss(u, dudxi)
    real(dp), intent(in) :: u(numTotal)
    real(dp), intent(inout) :: dudxi(3, numTotal)
    call fcp_check(fcp_grad_gauss(ctx, u, dudxi), 'fcp_grad_gauss')
  end subroutine

  ! --- gradients (gradients.f90:72-214) ----------------------------------
  integer(c_int) function grad_method()
    grad_method = FCP_GM_GAUSS
    if (lstsq) then
      grad_method = FCP_GM_LSQ
    elseif (lstsq_qr) then
      grad_method = FCP_GM_LSQ_QR
    elseif (lstsq_dm) then
      grad_method = FCP_GM_LSQ_DM
    endif
  end function

  subroutine create_lsq_grad_matrix
    integer(c_int) :: m, cm
    m = grad_method()
    if (m == FCP_GM_GAUSS) return
    cm = 0
    if (lsq_compat) cm = 1
    call fcp_check(fcp_create_lsq_grad_matrix(ctx, m, cm, c_null_ptr), 'fcp_create_lsq_grad_matrix')
  end subroutine

  subroutine grad_scalar_field(phi, dPhidxi)
    real(dp), intent(in) :: phi(numTotal)
    real(dp), intent(inout) :: dPhidxi(3, numTotal)
    if (node_gauss .or. (trim(limiter) /= 'none' .and. trim(limiter) /= 'no-limit')) then
      write(*,'(a)') '  grad: node-based Gauss and gradient limiters are not provided'
      stop 1
    endif
    call fcp_check(fcp_grad(ctx, grad_method(), phi, dPhidxi), 'fcp_grad')
  end subroutine

  subroutine grad_vector_field(U, V, W, dUdxi, dVdxi, dWdxi)
    real(dp), intent(in) :: U(numTotal), V(numTotal), W(numTotal)
    real(dp), intent(inout) :: dUdxi(3, numTotal), dVdxi(3, numTotal), dWdxi(3, numTotal)
    call grad_scalar_field(U, dUdxi)
    call grad_scalar_field(V, dVdxi)
    call grad_scalar_field(W, dWdxi)
  end subroutine

  subroutine grad_lsq(Phi, dPhidxi)
    real(dp), intent(in) :: Phi(numTotal)
    real(dp), intent(inout) :: dPhidxi(3, numTotal)
    call fcp_check(fcp_grad(ctx, FCP_GM_LSQ, Phi, dPhidxi), 'fcp_grad')
  end subroutine

  subroutine grad_lsq_dm(Phi, dPhidxi)
    real(dp), intent(in) :: Phi(numTotal)
    real(dp), intent(inout) :: dPhidxi(3, numTotal)
    call fcp_check(fcp_grad(ctx, FCP_GM_LSQ_DM, Phi, dPhidxi), 'fcp_grad')
  end subroutine

  subroutine grad_lsq_qr(Phi, dPhidxi)
    real(dp), intent(in) :: Phi(numTotal)
    real(dp), intent(inout) :: dPhidxi(3, numTotal)
    call fcp_check(fcp_grad(ctx, FCP_GM_LSQ_QR, Phi, dPhidxi), 'fcp_grad')
  end subroutine

  ! fvxGradient.f90's grad_gauss (:1549-1645) and grad_gauss_corrected
  ! (:1664-1758), with their separate component arrays of numCells; a caller
  ! of that module renames them on use (use fcp_ref, only: grad_gauss =>
  ! fvx_grad_gauss).  The corrected form starts from dUdx, dUdy, dUdz on entry.
  subroutine fvx_grad_gauss(u, dUdx, dUdy, dUdz)
    real(dp), intent(in) :: u(numTotal)
    real(dp), intent(inout) :: dUdx(numCells), dUdy(numCells), dUdz(numCells)
    call fvx_grad(FCP_GM_GAUSS_FVX, u, dUdx, dUdy, dUdz)
  end subroutine

  subroutine fvx_grad_gauss_corrected(u, dUdx, dUdy, dUdz)
    real(dp), intent(in) :: u(numTotal)
    real(dp), intent(inout) :: dUdx(numCells), dUdy(numCells), dUdz(numCells)
    call fvx_grad(FCP_GM_GAUSS_CORRECTED, u, dUdx, dUdy, dUdz)
  end subroutine

  subroutine fvx_grad(method, u, dUdx, dUdy, dUdz)
    integer(c_int), intent(in) :: method
    real(dp), intent(in) :: u(numTotal)
    real(dp), intent(inout) :: dUdx(numCells), dUdy(numCells), dUdz(numCells)
    real(dp), allocatable :: g(:, :)
    allocate(g(3, numTotal))
    g = 0.0_dp
    g(1, 1:numCells) = dUdx
    g(2, 1:numCells) = dUdy
    g(3, 1:numCells) = dUdz
    call fcp_check(fcp_grad(ctx, method, u, g), 'fcp_grad')
    dUdx = g(1, 1:numCells)
    dUdy = g(2, 1:numCells)
    dUdz = g(3, 1:numCells)
    deallocate(g)
  end subroutine

  subroutine updateBoundary(phi)
    real(dp), intent(inout) :: phi(numTotal)
    call fcp_check(fcp_update_boundary(ctx, phi), 'fcp_update_boundary')
  end subroutine

  ! --- assembly ------------------------------------------------------------
  ! Face-diffusivity Laplacian: bc_kind 0 Dirichlet, 1 zero-gradient per
  ! boundary face; writes a and adds the boundary terms to su.
  subroutine fvm_laplacian(mu_f, bc_kind, phi)
    real(dp), intent(in) :: mu_f(numFaces), phi(numTotal)
    integer(c_signed_char), intent(in) :: bc_kind(numBoundaryFaces)
    call fcp_check(fcp_laplacian_face(ctx, mu_f, bc_kind, phi, a, su), 'fcp_laplacian_face')
  end subroutine

  ! facefluxsc over every face with the boundary loop by bc_kind and the
  ! under-relaxed diagonal (temperature.f90:118-366); su, sp accumulate.
  subroutine fvm_div(flmass, dcoef, phi, dphidxi, gam, cScheme, bc_kind, urf)
    real(dp), intent(in) :: flmass(numFaces), dcoef(numFaces), dphidxi(3, numTotal), gam, urf
    real(dp), intent(inout) :: phi(numTotal)
    character(len=*), intent(in) :: cScheme
    integer(c_signed_char), intent(in) :: bc_kind(numBoundaryFaces)
    integer(c_int) :: sch
    sch = 1
    if (trim(cScheme) == 'cds') sch = 0
    call fcp_check(fcp_convdiff(ctx, flmass, dcoef, phi, dphidxi, gam, sch, bc_kind, urf, &
                                c_loc(a), su, sp), 'fcp_convdiff')
  end subroutine

  ! fvm_div with facefluxscalar on the inner faces (scalar_fluxes.f90:
  ! 344-464): the diffusivity per cell; boundary faces take the owner's.
  subroutine fvm_div_cell(flmass, diffcoef, phi, dphidxi, gam, cScheme, bc_kind, urf)
    real(dp), intent(in) :: flmass(numFaces), diffcoef(numCells), dphidxi(3, numTotal), gam, urf
    real(dp), intent(inout) :: phi(numTotal)
    character(len=*), intent(in) :: cScheme
    integer(c_signed_char), intent(in) :: bc_kind(numBoundaryFaces)
    integer(c_int) :: sch
    sch = 1
    if (trim(cScheme) == 'cds') sch = 0
    call fcp_check(fcp_convdiff_cell(ctx, flmass, diffcoef, phi, dphidxi, gam, sch, bc_kind, &
                                     urf, c_loc(a), su, sp), 'fcp_convdiff_cell')
  end subroutine

  ! calcuvw (velocity.f90:50-750) with the reference's module globals passed
  ! explicitly: the momentum assembly (fcp_momentum_assemble) on the resident
  ! matrix, then the three components solved together with their own
  ! diagonals (fcp_csrsolve_multi_diag); apu, apv, apw receive 1/a_P for
  ! calcp_simple, resor(3) each component's res0.  uo..woo are read for the
  ! transient schemes only.
  subroutine calcuvw(u, v, w, uo, vo, wo, uoo, voo, woo, p, dPdxi, den, vis, visw, flmass, &
                     viscos, timestep, tscheme, cSchemeU, gdsU, urfU, apu, apv, apw, lSolverU, &
                     maxiterU, tolAbsU, tolRelU, resor)
    real(dp), intent(inout) :: u(numTotal), v(numTotal), w(numTotal), p(numTotal)
    real(dp), intent(inout), target :: dPdxi(3, numTotal)
    real(dp), intent(in) :: uo(numCells), vo(numCells), wo(numCells)
    real(dp), intent(in) :: uoo(numCells), voo(numCells), woo(numCells)
    real(dp), intent(in) :: den(numTotal), vis(numTotal), visw(*), flmass(numFaces)
    real(dp), intent(in) :: viscos, timestep, gdsU, urfU(3), tolAbsU, tolRelU
    character(len=*), intent(in) :: tscheme, cSchemeU, lSolverU
    integer, intent(in) :: maxiterU
    real(dp), intent(out), target :: apu(numCells), apv(numCells), apw(numCells)
    real(dp), intent(out) :: resor(3)
    real(dp), allocatable, target :: rhs(:, :), dg(:, :), ap3(:, :), fi(:, :)
    type(fcp_solve_opts) :: opts
    integer(c_int) :: ts, sch, rc
    select case (trim(tscheme))
    case ('bdf')
      ts = 1
    case ('bdf2')
      ts = 2
    case ('cn')
      ts = 3
    case default
      ts = 0
    end select
    sch = 1
    if (trim(cSchemeU) == 'cds') sch = 0
    allocate(rhs(numCells, 3), dg(numCells, 3), ap3(numCells, 3), fi(numCells, 3))
    call fcp_check(fcp_momentum_assemble(ctx, u, v, w, uo, vo, wo, uoo, voo, woo, p, den, vis, &
                                         visw, flmass, viscos, timestep, ts, sch, gdsU, urfU, &
                                         c_loc(dPdxi), c_null_ptr, rhs, dg, c_loc(ap3), &
                                         c_null_ptr, c_null_ptr, c_loc(a)), 'fcp_momentum_assemble')
    apu = ap3(:, 1)
    apv = ap3(:, 2)
    apw = ap3(:, 3)
    fi(:, 1) = u(1:numCells)
    fi(:, 2) = v(1:numCells)
    fi(:, 3) = w(1:numCells)
    call fcp_solve_opts_default(opts)
    rc = fcp_csrsolve_multi_diag(ctx, lSolverU, len_trim(lSolverU), 3, fi, numCells, rhs, dg, &
                                 resor, maxiterU, tolAbsU, tolRelU, 0, opts)
    if (rc < 0) call fcp_check(rc, 'fcp_csrsolve_multi_diag')
    u(1:numCells) = fi(:, 1)
    v(1:numCells) = fi(:, 2)
    w(1:numCells) = fi(:, 3)
    deallocate(rhs, dg, ap3, fi)
  end subroutine

  ! calcp_simple (calcp_simple.f90:69-419) for an incompressible case
  ! without pressure patches and one corrector, with the reference's module
  ! globals passed explicitly: the pressure-correction assembly into a, su
  ! (facefluxmass2, adjustMassFlow unless const_mflux), csrsolve(solver) for
  ! pp, then the flux, velocity and pressure correction (gradp_and_sources,
  ! updateVelocityAtBoundary).  resor receives csrsolve's res0.
  subroutine calcp_simple(u, v, w, p, pp, dPdxi, den, apu, apv, apw, flmass, const_mflux, &
                          flomas, urfp, pRefCell, solver, itr_max, tol_abs, tol_rel, resor)
    real(dp), intent(inout) :: u(numTotal), v(numTotal), w(numTotal), p(numTotal), pp(numTotal)
    real(dp), intent(inout) :: dPdxi(3, numTotal), flmass(numFaces)
    real(dp), intent(in) :: den(numTotal), apu(numCells), apv(numCells), apw(numCells)
    logical, intent(in) :: const_mflux
    real(dp), intent(in) :: flomas, urfp, tol_abs, tol_rel
    integer, intent(in) :: pRefCell, itr_max
    character(len=*), intent(in) :: solver
    real(dp), intent(out) :: resor
    integer(c_int) :: cm
    cm = 0
    if (const_mflux) cm = 1
    call fcp_check(fcp_pressure_assemble(ctx, u, v, w, p, dPdxi, den, apu, apv, apw, flmass, cm, &
                                         flomas, c_loc(a), su), 'fcp_pressure_assemble')
    call csrsolve(solver, pp, su, resor, itr_max, tol_abs, tol_rel, 'p')
    call fcp_check(fcp_pressure_correct(ctx, pp, flmass, u, v, w, p, dPdxi, apu, apv, apw, urfp, &
                                        pRefCell, c_loc(su), c_null_ptr, c_null_ptr), &
                   'fcp_pressure_correct')
  end subroutine

  ! Time terms of temperature.f90:82-102 into su, sp (scheme 1 BDF1 / the
  ! volume part of Crank-Nicolson, 2 BDF2).
  subroutine fvm_ddt(scheme, den, dt, phi_o, phi_oo)
    integer, intent(in) :: scheme
    real(dp), intent(in) :: den(numCells), dt, phi_o(numCells)
    real(dp), intent(in), target, optional :: phi_oo(numCells)
    type(c_ptr) :: poo
    poo = c_null_ptr
    if (present(phi_oo)) poo = c_loc(phi_oo)
    call fcp_check(fcp_ddt(ctx, scheme, den, dt, phi_o, poo, su, sp), 'fcp_ddt')
  end subroutine

  ! calculate_temperature (temperature.f90:35-385) for a scalar phi with the
  ! parameters the reference takes from its modules: tscheme 'steady' |
  ! 'bdf' | 'bdf2' | 'cn', cScheme 'cds' | 'linearUpwind', wall_frac =
  ! bcDFraction(field,:) per patch; the gradient follows the module flags.
  ! Leaves the system in a, su, sp and prints the range line as the
  ! reference does.
  subroutine calculate_scalar(phi, phi_o, phi_oo, den, vis, flmass, wall_frac, viscos, pranl, &
                              sigt, urf, gds, dt, tscheme, cScheme, solver, itr_max, tol_abs, &
                              tol_rel, resor, chvar)
    real(dp), intent(inout) :: phi(numTotal)
    real(dp), intent(in), target :: phi_o(numCells), phi_oo(numCells), den(numCells)
    real(dp), intent(in) :: vis(numCells), flmass(numFaces), wall_frac(numBoundaries)
    real(dp), intent(in) :: viscos, pranl, sigt, urf, gds, dt, tol_abs, tol_rel
    character(len=*), intent(in) :: tscheme, cScheme, solver, chvar
    integer, intent(in) :: itr_max
    real(dp), intent(out) :: resor
    type(fcp_solve_opts) :: opts
    integer(c_int) :: ts, sch, rc
    real(dp) :: fimin, fimax
    select case (trim(tscheme))
    case ('bdf')
      ts = 1
    case ('bdf2')
      ts = 2
    case ('cn')
      ts = 3
    case default
      ts = 0
    end select
    sch = 1
    if (trim(cScheme) == 'cds') sch = 0
    call fcp_solve_opts_default(opts)
    rc = fcp_transport(ctx, phi, c_loc(phi_o), c_loc(phi_oo), c_loc(den), vis, flmass, wall_frac, &
                       viscos, pranl, sigt, urf, gds, dt, ts, sch, grad_method(), trim(solver), &
                       len_trim(solver), itr_max, tol_abs, tol_rel, 1, c_null_ptr, c_loc(a), &
                       c_loc(su), c_loc(sp), resor, fimin, fimax, opts)
    call fcp_check(rc, 'fcp_transport')
    write(6,'(2x,es11.4,3a,es11.4)') fimin, ' <= ', trim(chvar), ' <= ', fimax
  end subroutine


  subroutine csrsolve(solver, fi, rhs, res0, itr_max, tol_abs, tol_rel, chvar)
    character(len=*), intent(in) :: solver, chvar
    real(dp), intent(inout) :: fi(numTotal)
    real(dp), intent(in) :: rhs(numCells)
    real(dp), intent(out) :: res0
    integer, intent(in) :: itr_max
    real(dp), intent(in) :: tol_abs, tol_rel
    type(fcp_solve_opts) :: opts
    integer(c_int) :: rc
    call fcp_solve_opts_default(opts)
    call fcp_check(fcp_matrix_set(ctx, a), 'fcp_matrix_set')
    rc = fcp_csrsolve(ctx, solver, len_trim(solver), fi, rhs, res0, itr_max, tol_abs, tol_rel, &
                      chvar, len_trim(chvar), 1, opts)
    call fcp_check(rc, 'fcp_csrsolve')
  end subroutine

  ! LIS_linear_solvers.f90:38-200 under the reference's names: the LIS
  ! option string selects the library's solver; lis_spsolve solves the module
  ! matrix (numCells, ia, ja, a) for fi and prints LIS's report line.
  subroutine lis_create_solver(solver_options)
    character(len=*), intent(in) :: solver_options
    call fcp_check(fcp_lis_create_solver(ctx, solver_options, len_trim(solver_options)), &
                   'fcp_lis_create_solver')
  end subroutine

  subroutine lis_spsolve(fi, rhs, chvar)
    real(dp), intent(inout) :: fi(numCells)
    real(dp), intent(in) :: rhs(numCells)
    character(len=*), intent(in) :: chvar
    real(dp) :: res0, resid
    integer(c_int) :: iter
    call fcp_check(fcp_matrix_set(ctx, a), 'fcp_matrix_set')
    call fcp_check(fcp_lis_spsolve(ctx, fi, rhs, chvar, len_trim(chvar), res0, resid, iter), &
                   'fcp_lis_spsolve')
  end subroutine

  subroutine lis_destroy_solver
    call fcp_check(fcp_lis_destroy_solver(ctx), 'fcp_lis_destroy_solver')
  end subroutine

  subroutine iccg(n, nz, ia_, ja_, a_, diag_, fi, rhs, resor, itr_max, tol_abs, tol_rel, chvar, verbose)
    integer, intent(in) :: n, nz, itr_max
    integer, intent(in) :: ia_(n+1), ja_(nz), diag_(n)
    real(dp), intent(in) :: a_(nz), rhs(n), tol_abs, tol_rel
    real(dp), intent(inout) :: fi(n)
    real(dp), intent(out) :: resor
    character(len=*), intent(in) :: chvar
    logical, intent(in) :: verbose
    integer(c_int) :: rc, vb
    vb = 0
    if (verbose) vb = 2
    rc = fcp_iccg(n, nz, ia_, ja_, a_, diag_, fi, rhs, resor, itr_max, tol_abs, tol_rel, vb)
    if (rc < 0) then
      write(*,'(3a,i0)') '  iccg (', trim(chvar), ') failed: ', rc
      stop 1
    endif
  end subroutine

  subroutine dpcg(n, nz, ia_, ja_, a_, diag_, fi, rhs, resor, itr_max, tol_abs, tol_rel, chvar, verbose)
    integer, intent(in) :: n, nz, itr_max
    integer, intent(in) :: ia_(n+1), ja_(nz), diag_(n)
    real(dp), intent(in) :: a_(nz), rhs(n), tol_abs, tol_rel
    real(dp), intent(inout) :: fi(n)
    real(dp), intent(out) :: resor
    character(len=*), intent(in) :: chvar
    logical, intent(in) :: verbose
    integer(c_int) :: rc, vb
    vb = 0
    if (verbose) vb = 2
    rc = fcp_dpcg(n, nz, ia_, ja_, a_, diag_, fi, rhs, resor, itr_max, tol_abs, tol_rel, vb)
    if (rc < 0) then
      write(*,'(3a,i0)') '  dpcg (', trim(chvar), ') failed: ', rc
      stop 1
    endif
  end subroutine

  subroutine bicgstab(n, nz, ia_, ja_, a_, diag_, fi, rhs, resor, itr_max, tol_abs, tol_rel, chvar, verbose)
    integer, intent(in) :: n, nz, itr_max
    integer, intent(in) :: ia_(n+1), ja_(nz), diag_(n)
    real(dp), intent(in) :: a_(nz), rhs(n), tol_abs, tol_rel
    real(dp), intent(inout) :: fi(n)
    real(dp), intent(out) :: resor
    character(len=*), intent(in) :: chvar
    logical, intent(in) :: verbose
    integer(c_int) :: rc, vb
    vb = 0
    if (verbose) vb = 2
    rc = fcp_bicgstab(n, nz, ia_, ja_, a_, diag_, fi, rhs, resor, itr_max, tol_abs, tol_rel, vb)
    if (rc < 0) then
      write(*,'(3a,i0)') '  bicgstab (', trim(chvar), ') failed: ', rc
      stop 1
    endif
  end subroutine

  subroutine GaussSeidel(n, nz, ia_, ja_, a_, diag_, fi, rhs, resor, itr_max, tol_abs, tol_rel, chvar, &
                         verbose)
    integer, intent(in) :: n, nz, itr_max
    integer, intent(in) :: ia_(n+1), ja_(nz), diag_(n)
    real(dp), intent(in) :: a_(nz), rhs(n), tol_abs, tol_rel
    real(dp), intent(inout) :: fi(n)
    real(dp), intent(out) :: resor
    character(len=*), intent(in) :: chvar
    logical, intent(in) :: verbose
    integer(c_int) :: rc, vb
    vb = 0
    if (verbose) vb = 2
    rc = fcp_gauss_seidel(n, nz, ia_, ja_, a_, diag_, fi, rhs, resor, itr_max, tol_abs, tol_rel, vb)
    if (rc < 0) then
      write(*,'(3a,i0)') '  Gauss-Seidel (', trim(chvar), ') failed: ', rc
      stop 1
    endif
  end subroutine

  subroutine pmgmres_ilu(n, nz_num, ia_, ja_, a_, ua, x, rhs, resor, itr_max, mr, tol_abs, &
                         tol_rel, chvar, verbose)
    integer, intent(in) :: n, nz_num, itr_max, mr
    integer, intent(in) :: ia_(n+1), ja_(nz_num), ua(n)
    real(dp), intent(in) :: a_(nz_num), rhs(n), tol_abs, tol_rel
    real(dp), intent(inout) :: x(n)
    real(dp), intent(out) :: resor
    character(len=*), intent(in) :: chvar
    logical, intent(in) :: verbose
    integer(c_int) :: rc, vb
    vb = 0
    if (verbose) vb = 2
    rc = fcp_pmgmres(n, nz_num, ia_, ja_, a_, ua, x, rhs, resor, itr_max, mr, tol_abs, tol_rel, vb)
    if (rc < 0) then
      write(*,'(3a,i0)') '  PMGMRES_ILU (', trim(chvar), ') failed: ', rc
      stop 1
    endif
  end subroutine

  ! The body of csrsolve (linear_solvers.f90:62-88) on the module arrays
  ! numCells, nnz, ia, ja, a, diag, under the name the reference's solver
  ! test calls it by; 'bicgstab_ilu' and 'pmgmres_ilu' name the same solvers.
  subroutine spsolve(solver, fi, rhs, res0, itr_max, tol_abs, tol_rel, chvar)
    character(len=*), intent(in) :: solver, chvar
    real(dp), intent(inout) :: fi(*)
    real(dp), intent(in) :: rhs(numCells)
    real(dp), intent(out) :: res0
    integer, intent(in) :: itr_max
    real(dp), intent(in) :: tol_abs, tol_rel
    select case (trim(solver))
    case ('gauss-seidel')
      call GaussSeidel(numCells, nnz, ia, ja, a, diag, fi(1:numCells), rhs, res0, itr_max, &
                       tol_abs, tol_rel, chvar, ltest)
    case ('dpcg')
      call dpcg(numCells, nnz, ia, ja, a, diag, fi(1:numCells), rhs, res0, itr_max, tol_abs, &
                tol_rel, chvar, ltest)
    case ('iccg')
      call iccg(numCells, nnz, ia, ja, a, diag, fi(1:numCells), rhs, res0, itr_max, tol_abs, &
                tol_rel, chvar, ltest)
    case ('bicgstab', 'bicgstab_ilu')
      call bicgstab(numCells, nnz, ia, ja, a, diag, fi(1:numCells), rhs, res0, itr_max, tol_abs, &
                    tol_rel, chvar, ltest)
    case ('pmgmres', 'pmgmres_ilu')
      call pmgmres_ilu(numCells, nnz, ia, ja, a, diag, fi(1:numCells), rhs, res0, itr_max, 4, &
                       tol_abs, tol_rel, chvar, ltest)
    case default
      write(*,'(3a)') '  spsolve: unknown solver "', trim(solver), '"'
      stop 1
    end select
  end subroutine


  subroutine fcp_ref_finalize
    if (c_associated(ctx)) call fcp_finalize(ctx)
    ctx = c_null_ptr
  end subroutine

end module fcp_ref
