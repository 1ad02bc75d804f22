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
! The mesh is handed over once with fcp_read_mesh (the arrays read_mesh
! produces, geometry.f90:118-406); geometry is computed on the GPU and
! returned in the module arrays below (geometry.f90:19-83 names).
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

  ! sparse_matrix module globals (sparse_matrix.f90:10-26)
  integer :: nnz = 0
  integer, allocatable :: ia(:), ja(:), diag(:)
  integer, allocatable :: icell_jcell_csr_index(:), jcell_icell_csr_index(:)
  real(dp), allocatable :: a(:), su(:)

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

  subroutine create_CSR_matrix
    integer(c_int) :: n
    if (allocated(ia)) deallocate(ia, ja, diag, icell_jcell_csr_index, jcell_icell_csr_index, a, su)
    nnz = 2*numInnerFaces + numCells                      ! sparse_matrix.f90:110
    allocate(ia(numCells+1), ja(nnz), diag(numCells), a(nnz), su(numCells))
    allocate(icell_jcell_csr_index(max(numInnerFaces,1)), jcell_icell_csr_index(max(numInnerFaces,1)))
    call fcp_check(fcp_csr_build(ctx, n, ia, ja, diag, icell_jcell_csr_index, jcell_icell_csr_index), &
                   'fcp_csr_build')
    nnz = n
    a = 0.0_dp
    su = 0.0_dp
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

  subroutine fcp_ref_finalize
    if (c_associated(ctx)) call fcp_finalize(ctx)
    ctx = c_null_ptr
  end subroutine

end module fcp_ref
