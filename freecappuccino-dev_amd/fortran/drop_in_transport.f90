! drop_in_transport.f90 -- a reference-style Fortran caller of the transport
! path through fcp_ref: read_mesh from disk, a least-squares gradient
! (create_lsq_grad_matrix + grad), the face-diffusivity Laplacian
! (fvm_laplacian), a BDF2 transport system from the building blocks
! (fvm_ddt + fvm_div) solved with spsolve, and one calculate_scalar step
! (calculate_temperature: BDF2, linearUpwind, BiCGStab, urf 0.9).
! Usage: drop_in_transport case_dir outfile
! Inputs: raw little-endian files in case_dir -- REAL(8) T (numTotal), To,
! Too, den, vis (numCells), flm, muf, dco (numFaces), wf (numBoundaries);
! INTEGER(1) bkl, bkc (numBoundaryFaces): boundary kinds of the Laplacian and
! of the transport system.
program drop_in_transport
  use fcp_ref
  implicit none
  character(len=1024) :: cdir, outfile
  real(dp), allocatable :: T(:), To(:), Too(:), den(:), vis(:), flm(:), muf(:), dco(:), wf(:)
  real(dp), allocatable :: g(:,:), w(:), x(:), T2(:), alap(:), sulap(:)
  integer(c_signed_char), allocatable :: bkl(:), bkc(:)
  real(dp) :: res0, resor
  integer :: u

  call get_command_argument(1, cdir)
  call get_command_argument(2, outfile)
  call read_mesh(trim(cdir))
  call create_CSR_matrix
  allocate(T(numTotal), To(numCells), Too(numCells), den(numCells), vis(numCells))
  allocate(flm(numFaces), muf(numFaces), dco(numFaces), wf(numBoundaries))
  allocate(bkl(numBoundaryFaces), bkc(numBoundaryFaces), g(3, numTotal), w(numTotal))
  allocate(x(numTotal), T2(numTotal), alap(nnz), sulap(numCells))
  call rd8('T.bin', T)
  call rd8('To.bin', To)
  call rd8('Too.bin', Too)
  call rd8('den.bin', den)
  call rd8('vis.bin', vis)
  call rd8('flm.bin', flm)
  call rd8('muf.bin', muf)
  call rd8('dco.bin', dco)
  call rd8('wf.bin', wf)
  call rdi1('bkl.bin', bkl)
  call rdi1('bkc.bin', bkc)

  ! least-squares gradient (distance-weighted, the reference's arithmetic)
  gauss = .false.
  lstsq_dm = .true.
  call create_lsq_grad_matrix
  call grad(T, g)

  ! face-diffusivity Laplacian with zero-gradient faces
  su = 0.0_dp
  call fvm_laplacian(muf, bkl, T)
  alap = a
  sulap = su

  ! BDF2 transport system from the building blocks, solved with spsolve
  su = 0.0_dp
  sp = 0.0_dp
  w = T
  call fvm_ddt(2, den, 1.0d-2, To, Too)
  call fvm_div(flm, dco, w, g, 1.0d0, 'linearUpwind', bkc, 0.9d0)
  x = 0.0_dp
  call spsolve('bicgstab', x, su, res0, 50, 0.0d0, 1.0d-10, 'phi')

  ! one step of calculate_temperature with the Gauss gradient
  lstsq_dm = .false.
  gauss = .true.
  T2 = T
  call calculate_scalar(T2, To, Too, den, vis, flm, wf, 1.0d-3, 0.7d0, 1.0d0, 0.9d0, 1.0d0, &
                        1.0d-2, 'bdf2', 'linearUpwind', 'bicgstab', 10, 1.0d-13, 0.02d0, resor, &
                        'Temperature')

  open(newunit=u, file=trim(outfile), access='stream', form='unformatted', status='replace')
  write(u) numCells, nnz, numTotal
  write(u) g, alap, sulap, w, x(1:numCells), T2, a, su, sp, resor
  close(u)
  write(*,'(a)') 'drop_in_transport ok'
  call fcp_ref_finalize

contains

  subroutine rd8(name, v)
    character(len=*), intent(in) :: name
    real(dp), intent(out) :: v(:)
    integer :: uu
    open(newunit=uu, file=trim(cdir)//'/'//name, access='stream', form='unformatted', status='old')
    read(uu) v
    close(uu)
  end subroutine

  subroutine rdi1(name, v)
    character(len=*), intent(in) :: name
    integer(c_signed_char), intent(out) :: v(:)
    integer :: uu
    open(newunit=uu, file=trim(cdir)//'/'//name, access='stream', form='unformatted', status='old')
    read(uu) v
    close(uu)
  end subroutine

end program drop_in_transport
