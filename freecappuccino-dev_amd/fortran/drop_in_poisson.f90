! drop_in_poisson.f90 -- a reference-style Fortran caller of the MI355X path.
!
! Builds the unit-cube N^3 hexahedral mesh in the native polyMesh layout
! (cell id i+N*j+N*N*k, inner faces sorted by owner then neighbour, boundary
! patches x-, x+, y-, y+, z-, z+), hands it over, and runs the reference call
! sequence of a Poisson solve (applications/Poisson/poisson.f90:52-90):
! create_CSR_matrix -> laplacian(mu=-1, phi) -> iccg via csrsolve, plus a
! Gauss gradient and the explicit-array iccg entry point.
! Usage: drop_in_poisson N outfile
program drop_in_poisson
  use fcp_ref
  implicit none
  integer :: N, M, nNodes, nCells, nInner, nBnd, nFacesAll, i, j, k, c, f, p, u
  real(dp), allocatable :: xyz(:,:), mu(:), phi(:), fi(:), fi2(:), rhs(:), g(:,:)
  integer, allocatable :: fnn(:), node(:,:), owner(:), neighbour(:)
  integer :: pnf(6), pStart(6), bctype(6)
  character(len=256) :: arg, outfile
  real(dp) :: res0, resor

  call get_command_argument(1, arg); read(arg, *) N
  call get_command_argument(2, outfile)
  M = N + 1
  nNodes = M**3
  nCells = N**3
  nInner = 3*N*N*(N-1)
  nBnd = 6*N*N
  nFacesAll = nInner + nBnd
  allocate(xyz(3,nNodes), fnn(nFacesAll), node(FCP_NOMAX,nFacesAll), owner(nFacesAll), neighbour(nInner))
  do k = 0, N
    do j = 0, N
      do i = 0, N
        xyz(:, 1 + i + M*j + M*M*k) = [real(i,dp)/N, real(j,dp)/N, real(k,dp)/N]
      end do
    end do
  end do
  fnn = 4
  node = 0
  f = 0
  do c = 0, nCells - 1                       ! inner faces: +x, +y, +z of each owner
    i = mod(c, N); j = mod(c/N, N); k = c/(N*N)
    if (i < N-1) then
      f = f + 1; owner(f) = c + 1; neighbour(f) = c + 2
      node(1:4,f) = [nid(i+1,j,k), nid(i+1,j+1,k), nid(i+1,j+1,k+1), nid(i+1,j,k+1)]
    endif
    if (j < N-1) then
      f = f + 1; owner(f) = c + 1; neighbour(f) = c + 1 + N
      node(1:4,f) = [nid(i,j+1,k), nid(i,j+1,k+1), nid(i+1,j+1,k+1), nid(i+1,j+1,k)]
    endif
    if (k < N-1) then
      f = f + 1; owner(f) = c + 1; neighbour(f) = c + 1 + N*N
      node(1:4,f) = [nid(i,j,k+1), nid(i+1,j,k+1), nid(i+1,j+1,k+1), nid(i,j+1,k+1)]
    endif
  end do
  do p = 1, 6                                ! boundary patches, faces by owner
    pStart(p) = f
    pnf(p) = N*N
    bctype(p) = 1
    do c = 0, nCells - 1
      i = mod(c, N); j = mod(c/N, N); k = c/(N*N)
      select case (p)
      case (1); if (i /= 0) cycle
        f = f + 1; node(1:4,f) = [nid(0,j,k+1), nid(0,j+1,k+1), nid(0,j+1,k), nid(0,j,k)]
      case (2); if (i /= N-1) cycle
        f = f + 1; node(1:4,f) = [nid(N,j,k), nid(N,j+1,k), nid(N,j+1,k+1), nid(N,j,k+1)]
      case (3); if (j /= 0) cycle
        f = f + 1; node(1:4,f) = [nid(i+1,0,k), nid(i+1,0,k+1), nid(i,0,k+1), nid(i,0,k)]
      case (4); if (j /= N-1) cycle
        f = f + 1; node(1:4,f) = [nid(i,N,k), nid(i,N,k+1), nid(i+1,N,k+1), nid(i+1,N,k)]
      case (5); if (k /= 0) cycle
        f = f + 1; node(1:4,f) = [nid(i,j+1,0), nid(i+1,j+1,0), nid(i+1,j,0), nid(i,j,0)]
      case (6); if (k /= N-1) cycle
        f = f + 1; node(1:4,f) = [nid(i,j,N), nid(i+1,j,N), nid(i+1,j+1,N), nid(i,j+1,N)]
      end select
      owner(f) = c + 1
    end do
  end do

  call fcp_read_mesh(nNodes, nCells, nInner, nBnd, 6, xyz, fnn, node, owner, neighbour, &
                     pnf, pStart, bctype)
  call create_CSR_matrix

  ! poisson.f90:62-75: source, zero Dirichlet values, mu = -1
  allocate(mu(nCells), phi(numTotal), fi(numTotal), fi2(nCells), rhs(nCells), g(3,numTotal))
  do c = 1, nCells
    su(c) = sin(0.37_dp*c) * vol(c)
  end do
  phi = 0.0_dp
  mu = -1.0_dp
  call laplacian(mu, phi)
  rhs = su
  fi = 0.0_dp
  call csrsolve('iccg', fi, rhs, res0, 1000, 0.0_dp, 1.0d-10, 'p')
  fi2 = 0.0_dp
  call iccg(nCells, nnz, ia, ja, a, diag, fi2, rhs, resor, 1000, 0.0_dp, 1.0d-10, 'p', .false.)
  phi(1:nCells) = fi(1:nCells)
  call grad_gauss(phi, g)

  open(newunit=u, file=trim(outfile), access='stream', form='unformatted', status='replace')
  write(u) nCells, nnz
  write(u) vol, a, su, fi(1:nCells), fi2, g
  write(u) res0, resor
  close(u)
  write(*,'(a,i0,a,i0)') 'drop_in_poisson ok: cells=', nCells, ' nnz=', nnz
  call fcp_ref_finalize

contains

  integer function nid(ii, jj, kk)
    integer, intent(in) :: ii, jj, kk
    nid = 1 + ii + M*jj + M*M*kk
  end function

end program drop_in_poisson
