! drop_in_spsolve.f90 -- the reference's linear-solver test
! (test/test_linear_solvers_spsolve.f90) on top of fcp_ref: the two 5x5
! systems (a nonsymmetric one and a symmetric positive definite one, entries
! as default-REAL literals) solved through spsolve on the module arrays with
! the test's solver names, tolerances and iteration limits.  The test leaves
! x undefined before its first call and carries it from solve to solve; here
! every solve starts from zero.  The solutions are printed as the test prints
! them and written to a raw REAL(8) file for comparison.
! Usage: drop_in_spsolve outfile
program drop_in_spsolve
  use fcp_ref
  implicit none
  integer :: i, k, u
  integer :: itr_max
  real(dp) :: b(5), x(5), xa(5), res0, tol_abs, tol_rel
  real(dp) :: sol(5, 5)
  character(len=256) :: outfile
  character(len=12) :: names(5)

  call get_command_argument(1, outfile)
  itr_max = 5
  tol_rel = 1e-7
  tol_abs = 1e-10
  numCells = 5
  numTotal = 5
  nnz = 25
  allocate(ia(6), ja(25), diag(5), a(25))
  ja = (/ 1,2,3,4,5, 1,2,3,4,5, 1,2,3,4,5, 1,2,3,4,5, 1,2,3,4,5 /)
  ia = (/ 1, 6, 11, 16, 21, 26 /)
  diag = (/ 1, 7, 13, 19, 25 /)
  names = [character(len=12) :: 'bicgstab_ilu', 'pmgmres_ilu', 'iccg', 'dpcg', 'gauss-seidel']

  ! Test 1 - real nonsymmetric 5x5 system
  a = (/ 6.80, -6.05, -0.45,  8.32, -9.67, &
        -2.11, -3.30,  2.58,  2.71, -5.14, &
         5.66,  5.36, -2.70,  4.35, -7.26, &
         5.97, -4.44,  0.27, -7.17,  6.08, &
         8.23,  1.08,  9.04,  2.14, -6.87 /)
  b = (/ 4.02, 6.19, -8.22, -7.57, -3.03 /)
  xa = (/ -0.80, -0.70, 0.59, 1.32, 0.57 /)
  do k = 1, 2
    x = 0.0_dp
    call spsolve(trim(names(k)), x, b, res0, itr_max, tol_abs, tol_rel, 'x')
    sol(:, k) = x
    call report(trim(names(k)))
  end do

  ! Test 2 - symmetric positive definite 5x5 system
  a = (/ 3.14,  0.17, -0.90,  1.65, -0.72, &
         0.17,  0.79,  0.83, -0.65,  0.28, &
        -0.90,  0.83,  4.53, -3.70,  1.60, &
         1.65, -0.65, -3.70,  5.32, -1.37, &
        -0.72,  0.28,  1.60, -1.37,  1.98 /)
  b = (/ -7.29, 9.25, 5.99, -1.94, -8.30 /)
  xa = (/ -6.02, 15.62, 3.02, 3.25, -8.78 /)
  do k = 3, 5
    x = 0.0_dp
    if (k == 5) then
      call spsolve(trim(names(k)), x, b, res0, 100, tol_abs, tol_rel, 'x')
    else
      call spsolve(trim(names(k)), x, b, res0, itr_max, tol_abs, tol_rel, 'x')
    endif
    sol(:, k) = x
    call report(trim(names(k)))
  end do

  open(newunit=u, file=trim(outfile), access='stream', form='unformatted', status='replace')
  write(u) sol
  close(u)
  write(*,'(a)') 'drop_in_spsolve ok'
  call fcp_ref_finalize

contains

  subroutine report(name)
    character(len=*), intent(in) :: name
    write(*,'(a)') ' '
    write(*,'(19x,2a)') ' solver: ', name
    write(*,'(19x,a,5x,a)') 'Solution: ', 'Analytic solution:'
    do i = 1, 5
      write(*,'(19x,f5.2,10x,f5.2)') x(i), xa(i)
    end do
  end subroutine

end program drop_in_spsolve
