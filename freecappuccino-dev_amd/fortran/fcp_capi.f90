! fcp_capi.f90 -- ISO_C_BINDING interfaces of libfcp (include/fcp.h).
!
! Every entry point is bound as declared in the C header: INTEGER is
! integer(c_int), REAL(8) is real(c_double), arrays are passed by address,
! scalars by value where the C prototype takes a value.  Index arrays are
! 1-based, field arrays have length numTotal = numCells + numBoundaryFaces,
! gradients are dPhidxi(3,numTotal) -- the reference's conventions
! (geometry.f90:282-289, gradients.f90:1620).
module fcp_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: FCP_NOMAX = 6
  integer(c_int), parameter :: FCP_PREC_FP64 = 0, FCP_PREC_MIXED = 1
  integer(c_int), parameter :: FCP_CRIT_REF_L1 = 0, FCP_CRIT_TRUE_L2 = 1
  ! gradient methods (FCP_GRAD_* of the C header; Fortran names are case-blind,
  ! so FCP_GRAD_GAUSS would clash with fcp_grad_gauss)
  integer(c_int), parameter :: FCP_GM_GAUSS = 0, FCP_GM_LSQ = 1, FCP_GM_LSQ_DM = 2, &
                               FCP_GM_LSQ_QR = 3, FCP_GM_GAUSS_FVX = 4, &
                               FCP_GM_GAUSS_CORRECTED = 5

  type, bind(C) :: fcp_solve_opts
    integer(c_int) :: precision
    integer(c_int) :: criterion
    real(c_double) :: inner_tol
    integer(c_int) :: max_outer
    integer(c_int) :: check_every
    integer(c_int) :: iterations
    integer(c_int) :: outer_iterations
    real(c_double) :: res0
    real(c_double) :: resl
    real(c_double) :: resor
    real(c_double) :: rel_res_l2
    real(c_double) :: time_ms
  end type

  interface
    integer(c_int) function fcp_init(device, ctx) bind(C, name='fcp_init')
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr) :: ctx
    end function

    subroutine fcp_finalize(ctx) bind(C, name='fcp_finalize')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine

    type(c_ptr) function fcp_last_error(ctx) bind(C, name='fcp_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    subroutine fcp_solve_opts_default(opts) bind(C, name='fcp_solve_opts_default')
      import :: fcp_solve_opts
      type(fcp_solve_opts) :: opts
    end subroutine

    integer(c_int) function fcp_mesh_upload(ctx, nNodes, nCells, nInnerFaces, nBoundaryFaces, &
        nBoundaries, xyz, face_nnodes, node, owner, neighbour, nfaces, startFace, bctype) &
        bind(C, name='fcp_mesh_upload')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nNodes, nCells, nInnerFaces, nBoundaryFaces, nBoundaries
      real(c_double) :: xyz(*)
      integer(c_int) :: face_nnodes(*), node(*), owner(*), neighbour(*), nfaces(*), startFace(*), bctype(*)
    end function

    integer(c_int) function fcp_mesh_periodic(ctx, nPeriodic, startFaceTwin) &
        bind(C, name='fcp_mesh_periodic')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nPeriodic
      integer(c_int) :: startFaceTwin(*)
    end function

    integer(c_int) function fcp_geometry(ctx, xc, yc, zc, vol, arx, ary, arz, xf, yf, zf, &
        facint, Df, dnb, srdb) bind(C, name='fcp_geometry')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: xc(*), yc(*), zc(*), vol(*), arx(*), ary(*), arz(*), xf(*), yf(*), zf(*)
      real(c_double) :: facint(*), Df(*), dnb(*), srdb(*)
    end function

    integer(c_int) function fcp_csr_build(ctx, nnz, ia, ja, diag, slot_pn, slot_np) &
        bind(C, name='fcp_csr_build')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: nnz
      integer(c_int) :: ia(*), ja(*), diag(*), slot_pn(*), slot_np(*)
    end function

    integer(c_int) function fcp_levels(ctx, upper, lev, nlev) bind(C, name='fcp_levels')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: upper
      integer(c_int) :: lev(*), nlev
    end function

    integer(c_int) function fcp_laplacian(ctx, mu, phi, a, su) bind(C, name='fcp_laplacian')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: mu(*), phi(*), a(*), su(*)
    end function

    integer(c_int) function fcp_laplacian_face(ctx, mu_f, bc_kind, phi, a, su) &
        bind(C, name='fcp_laplacian_face')
      import :: c_int, c_double, c_ptr, c_signed_char
      type(c_ptr), value :: ctx
      real(c_double) :: mu_f(*), phi(*), a(*), su(*)
      integer(c_signed_char) :: bc_kind(*)
    end function

    integer(c_int) function fcp_nonorth_correction(ctx, gamma_f, grad, su) &
        bind(C, name='fcp_nonorth_correction')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: gamma_f(*), grad(*), su(*)
    end function

    integer(c_int) function fcp_read_polymesh(ctx, case_dir, case_len, format, sizes) &
        bind(C, name='fcp_read_polymesh')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char) :: case_dir(*)
      integer(c_int), value :: case_len, format
      integer(c_int) :: sizes(5)
    end function

    integer(c_int) function fcp_mesh_patches(ctx, nfaces, startFace, bctype) &
        bind(C, name='fcp_mesh_patches')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: nfaces(*), startFace(*), bctype(*)
    end function

    integer(c_int) function fcp_convdiff(ctx, flmass, dcoef, phi, grad, gam, scheme, bc_kind, urf, &
        a, su, sp) bind(C, name='fcp_convdiff')
      import :: c_int, c_double, c_ptr, c_signed_char
      type(c_ptr), value :: ctx
      real(c_double) :: flmass(*), dcoef(*), phi(*), grad(*), su(*), sp(*)
      real(c_double), value :: gam, urf
      integer(c_int), value :: scheme
      integer(c_signed_char) :: bc_kind(*)
      type(c_ptr), value :: a  ! c_null_ptr or c_loc of a(nnz)
    end function

    integer(c_int) function fcp_convdiff_cell(ctx, flmass, diffcoef, phi, grad, gam, scheme, &
        bc_kind, urf, a, su, sp) bind(C, name='fcp_convdiff_cell')
      import :: c_int, c_double, c_ptr, c_signed_char
      type(c_ptr), value :: ctx
      real(c_double) :: flmass(*), diffcoef(*), phi(*), grad(*), su(*), sp(*)
      real(c_double), value :: gam, urf
      integer(c_int), value :: scheme
      integer(c_signed_char) :: bc_kind(*)
      type(c_ptr), value :: a  ! c_null_ptr or c_loc of a(nnz)
    end function

    integer(c_int) function fcp_momentum_assemble(ctx, u, v, w, uo, vo, wo, uoo, voo, woo, p, &
        den, vis, visw, flmass, viscos, dt, time_scheme, conv_scheme, gds, urf, dpdxi, grad, &
        rhs, diagv, ap, spv, src, a) bind(C, name='fcp_momentum_assemble')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: u(*), v(*), w(*), uo(*), vo(*), wo(*), uoo(*), voo(*), woo(*), p(*)
      real(c_double) :: den(*), vis(*), visw(*), flmass(*), urf(*), rhs(*), diagv(*)
      real(c_double), value :: viscos, dt, gds
      integer(c_int), value :: time_scheme, conv_scheme
      type(c_ptr), value :: dpdxi, grad, ap, spv, src, a  ! c_null_ptr or c_loc of arrays
    end function

    integer(c_int) function fcp_pressure_assemble(ctx, u, v, w, p, dpdxi, den, apu, apv, apw, &
        flmass, const_mflux, flomas, a, su) bind(C, name='fcp_pressure_assemble')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: u(*), v(*), w(*), p(*), dpdxi(*), den(*), apu(*), apv(*), apw(*)
      real(c_double) :: flmass(*), su(*)
      integer(c_int), value :: const_mflux
      real(c_double), value :: flomas
      type(c_ptr), value :: a  ! c_null_ptr or c_loc of a(nnz)
    end function

    integer(c_int) function fcp_pressure_correct(ctx, pp, flmass, u, v, w, p, dpdxi, apu, apv, &
        apw, urfp, pref_cell, su, sv, sw) bind(C, name='fcp_pressure_correct')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: pp(*), flmass(*), u(*), v(*), w(*), p(*), dpdxi(*), apu(*), apv(*), apw(*)
      real(c_double), value :: urfp
      integer(c_int), value :: pref_cell
      type(c_ptr), value :: su, sv, sw  ! c_null_ptr or c_loc of arrays (numCells)
    end function

    integer(c_int) function fcp_ddt(ctx, scheme, den, dt, phi_o, phi_oo, su, sp) &
        bind(C, name='fcp_ddt')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: scheme
      real(c_double) :: den(*), phi_o(*), su(*), sp(*)
      real(c_double), value :: dt
      type(c_ptr), value :: phi_oo  ! c_null_ptr unless scheme 2
    end function

    integer(c_int) function fcp_transport(ctx, phi, phi_o, phi_oo, den, vis, flmass, wall_frac, &
        viscos, pranl, sigt, urf, gds, dt, time_scheme, conv_scheme, grad_method, solver, &
        solver_len, itr_max, tol_abs, tol_rel, clip, grad, a, su, sp, resor, fimin, fimax, opts) &
        bind(C, name='fcp_transport')
      import :: c_int, c_double, c_ptr, c_char, fcp_solve_opts
      type(c_ptr), value :: ctx
      real(c_double) :: phi(*), vis(*), flmass(*), wall_frac(*)
      type(c_ptr), value :: phi_o, phi_oo, den  ! c_null_ptr when not needed
      real(c_double), value :: viscos, pranl, sigt, urf, gds, dt, tol_abs, tol_rel
      integer(c_int), value :: time_scheme, conv_scheme, grad_method, solver_len, itr_max, clip
      character(kind=c_char) :: solver(*)
      type(c_ptr), value :: grad, a, su, sp  ! optional outputs
      real(c_double) :: resor, fimin, fimax
      type(fcp_solve_opts) :: opts
    end function

    integer(c_int) function fcp_matrix_transpose(ctx) bind(C, name='fcp_matrix_transpose')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function fcp_convdiff_backward(ctx, flmass, dcoef, phi_old, grad_old, gam, &
        scheme, bc_kind, lam, phi_new, dcoef_bar, phi_old_bar, grad_old_bar) &
        bind(C, name='fcp_convdiff_backward')
      import :: c_int, c_double, c_ptr, c_signed_char
      type(c_ptr), value :: ctx
      real(c_double) :: flmass(*), dcoef(*), phi_old(*), grad_old(*), lam(*), phi_new(*)
      real(c_double) :: dcoef_bar(*), phi_old_bar(*), grad_old_bar(*)
      real(c_double), value :: gam
      integer(c_int), value :: scheme
      integer(c_signed_char) :: bc_kind(*)
    end function

    integer(c_int) function fcp_laplacian_face_backward(ctx, mu_f, bc_kind, phi_bc, gamma_f, &
        grad, lam, phi, mu_f_bar, phi_bc_bar, gamma_bar, grad_bar) &
        bind(C, name='fcp_laplacian_face_backward')
      import :: c_int, c_double, c_ptr, c_signed_char
      type(c_ptr), value :: ctx
      real(c_double) :: mu_f(*), phi_bc(*), gamma_f(*), grad(*), lam(*), phi(*)
      real(c_double) :: mu_f_bar(*), phi_bc_bar(*), gamma_bar(*), grad_bar(*)
      integer(c_signed_char) :: bc_kind(*)
    end function

    integer(c_int) function fcp_grad_gauss_backward(ctx, grad_bar, phi_bar) &
        bind(C, name='fcp_grad_gauss_backward')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: grad_bar(*), phi_bar(*)
    end function

    integer(c_int) function fcp_rcm(ctx, pp_iters, perm) bind(C, name='fcp_rcm')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: pp_iters
      integer(c_int) :: perm(*)
    end function

    integer(c_int) function fcp_create_lsq_grad_matrix(ctx, method, compat, dmat) &
        bind(C, name='fcp_create_lsq_grad_matrix')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: method, compat
      type(c_ptr), value :: dmat  ! c_null_ptr or c_loc of a real(c_double) array
    end function

    integer(c_int) function fcp_grad(ctx, method, phi, dphidxi) bind(C, name='fcp_grad')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: method
      real(c_double) :: phi(*), dphidxi(*)
    end function

    integer(c_int) function fcp_grad_gauss(ctx, u, dudxi) bind(C, name='fcp_grad_gauss')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: u(*), dudxi(*)
    end function

    integer(c_int) function fcp_update_boundary(ctx, phi) bind(C, name='fcp_update_boundary')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: phi(*)
    end function

    integer(c_int) function fcp_matrix_set(ctx, a) bind(C, name='fcp_matrix_set')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: a(*)
    end function

    integer(c_int) function fcp_spmv(ctx, x, y) bind(C, name='fcp_spmv')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: x(*), y(*)
    end function

    integer(c_int) function fcp_precond_apply(ctx, kind, r, z, d) bind(C, name='fcp_precond_apply')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind
      real(c_double) :: r(*), z(*), d(*)
    end function

    integer(c_int) function fcp_spmv_transpose(ctx, x, y) bind(C, name='fcp_spmv_transpose')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: x(*), y(*)
    end function

    integer(c_int) function fcp_sweep_tiles(ctx, tile) bind(C, name='fcp_sweep_tiles')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: tile(*)
    end function

    integer(c_int) function fcp_autotune(ctx, solver, solver_len, tile_out, cached) &
        bind(C, name='fcp_autotune')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char) :: solver(*)
      integer(c_int), value :: solver_len
      integer(c_int) :: tile_out, cached
    end function

    integer(c_int) function fcp_ilu_factor(ctx, l) bind(C, name='fcp_ilu_factor')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: l(*)
    end function

    integer(c_int) function fcp_csrsolve(ctx, solver, solver_len, fi, rhs, res0, itr_max, &
        tol_abs, tol_rel, chvar, chvar_len, verbose, opts) bind(C, name='fcp_csrsolve')
      import :: c_int, c_double, c_ptr, c_char, fcp_solve_opts
      type(c_ptr), value :: ctx
      character(kind=c_char) :: solver(*), chvar(*)
      integer(c_int), value :: solver_len, itr_max, chvar_len, verbose
      real(c_double) :: fi(*), rhs(*), res0
      real(c_double), value :: tol_abs, tol_rel
      type(fcp_solve_opts) :: opts
    end function

    integer(c_int) function fcp_lis_create_solver(ctx, options, options_len) &
        bind(C, name='fcp_lis_create_solver')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char) :: options(*)
      integer(c_int), value :: options_len
    end function

    integer(c_int) function fcp_lis_spsolve(ctx, fi, rhs, chvar, chvar_len, res0, resid, iter) &
        bind(C, name='fcp_lis_spsolve')
      import :: c_int, c_double, c_ptr, c_char
      type(c_ptr), value :: ctx
      real(c_double) :: fi(*), rhs(*), res0, resid
      character(kind=c_char) :: chvar(*)
      integer(c_int), value :: chvar_len
      integer(c_int) :: iter
    end function

    integer(c_int) function fcp_lis_destroy_solver(ctx) bind(C, name='fcp_lis_destroy_solver')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function fcp_csrsolve_multi(ctx, solver, solver_len, nrhs, fi, ldfi, rhs, res0, &
        itr_max, tol_abs, tol_rel, verbose, opts) bind(C, name='fcp_csrsolve_multi')
      import :: c_int, c_double, c_ptr, c_char, fcp_solve_opts
      type(c_ptr), value :: ctx
      character(kind=c_char) :: solver(*)
      integer(c_int), value :: solver_len, nrhs, ldfi, itr_max, verbose
      real(c_double) :: fi(*), rhs(*), res0(*)
      real(c_double), value :: tol_abs, tol_rel
      type(fcp_solve_opts) :: opts
    end function

    integer(c_int) function fcp_csrsolve_multi_diag(ctx, solver, solver_len, nrhs, fi, ldfi, rhs, &
        diagv, res0, itr_max, tol_abs, tol_rel, verbose, opts) bind(C, name='fcp_csrsolve_multi_diag')
      import :: c_int, c_double, c_ptr, c_char, fcp_solve_opts
      type(c_ptr), value :: ctx
      character(kind=c_char) :: solver(*)
      integer(c_int), value :: solver_len, nrhs, ldfi, itr_max, verbose
      real(c_double) :: fi(*), rhs(*), diagv(*), res0(*)
      real(c_double), value :: tol_abs, tol_rel
      type(fcp_solve_opts) :: opts
    end function

    integer(c_int) function fcp_iccg(n, nnz, ia, ja, a, diag, fi, rhs, resor, itr_max, &
        tol_abs, tol_rel, verbose) bind(C, name='fcp_iccg')
      import :: c_int, c_double
      integer(c_int), value :: n, nnz, itr_max, verbose
      integer(c_int) :: ia(*), ja(*), diag(*)
      real(c_double) :: a(*), fi(*), rhs(*), resor
      real(c_double), value :: tol_abs, tol_rel
    end function

    integer(c_int) function fcp_dpcg(n, nnz, ia, ja, a, diag, fi, rhs, resor, itr_max, &
        tol_abs, tol_rel, verbose) bind(C, name='fcp_dpcg')
      import :: c_int, c_double
      integer(c_int), value :: n, nnz, itr_max, verbose
      integer(c_int) :: ia(*), ja(*), diag(*)
      real(c_double) :: a(*), fi(*), rhs(*), resor
      real(c_double), value :: tol_abs, tol_rel
    end function

    integer(c_int) function fcp_bicgstab(n, nnz, ia, ja, a, diag, fi, rhs, resor, itr_max, &
        tol_abs, tol_rel, verbose) bind(C, name='fcp_bicgstab')
      import :: c_int, c_double
      integer(c_int), value :: n, nnz, itr_max, verbose
      integer(c_int) :: ia(*), ja(*), diag(*)
      real(c_double) :: a(*), fi(*), rhs(*), resor
      real(c_double), value :: tol_abs, tol_rel
    end function

    integer(c_int) function fcp_gauss_seidel(n, nnz, ia, ja, a, diag, fi, rhs, resor, itr_max, &
        tol_abs, tol_rel, verbose) bind(C, name='fcp_gauss_seidel')
      import :: c_int, c_double
      integer(c_int), value :: n, nnz, itr_max, verbose
      integer(c_int) :: ia(*), ja(*), diag(*)
      real(c_double) :: a(*), fi(*), rhs(*), resor
      real(c_double), value :: tol_abs, tol_rel
    end function

    integer(c_int) function fcp_pmgmres(n, nnz, ia, ja, a, diag, fi, rhs, resor, itr_max, mr, &
        tol_abs, tol_rel, verbose) bind(C, name='fcp_pmgmres')
      import :: c_int, c_double
      integer(c_int), value :: n, nnz, itr_max, mr, verbose
      integer(c_int) :: ia(*), ja(*), diag(*)
      real(c_double) :: a(*), fi(*), rhs(*), resor
      real(c_double), value :: tol_abs, tol_rel
    end function

    ! device arrays only (c_loc of device memory, e.g. from hipMalloc)
    integer(c_int) function fcp_dpcg_noalloc(n, nnz, ia, ja, a, diag, fi, rhs, resor, itr_max, &
        tol_abs, tol_rel, stream) bind(C, name='fcp_dpcg_noalloc')
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: n, nnz, itr_max
      type(c_ptr), value :: ia, ja, a, diag, fi, rhs, stream
      real(c_double) :: resor
      real(c_double), value :: tol_abs, tol_rel
    end function

    integer(c_int) function fcp_tune_set(ctx, param, value) bind(C, name='fcp_tune_set')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: param, value
    end function

    integer(c_int) function fcp_profile(ctx, enable) bind(C, name='fcp_profile')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
    end function

    integer(c_int) function fcp_profile_report(ctx, buf, len) bind(C, name='fcp_profile_report')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char) :: buf(*)
      integer(c_int), value :: len
    end function

    type(c_ptr) function fcp_stream(ctx) bind(C, name='fcp_stream')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface

end module fcp_capi
