/*
 * fcp.h -- C-ABI of libfcp, the MI355X (gfx950) finite-volume
 * assemble-and-solve path.
 *
 * Drop-in boundary for freeCappuccino's Fortran module API.  The reference
 * interface each entry point replaces is cited (file:line, relative to the
 * reference repository root).  Conventions follow the reference so a
 * Fortran ISO_C_BINDING module can bind every symbol directly:
 *   - default INTEGER = int32, REAL(8) = double, all arguments by pointer or
 *     value as declared here (no hidden string lengths: lengths are explicit);
 *   - every index array crossing the boundary is 1-based (owner, neighbour,
 *     face vertex lists, ia/ja/diag, face->slot maps);
 *   - field arrays have length numTotal = nCells + nBoundaryFaces with the
 *     boundary-face values at positions nCells+1.. (geometry.f90:282-289);
 *   - gradients are dPhidxi(3,numTotal), i.e. xyz-contiguous per cell.
 * Array arguments may be host pointers (staged through device memory) or
 * device pointers from hipMalloc (used in place, no copies).
 *
 * Return codes: 0 ok; > 0 solver did not converge (value = iterations);
 * < 0 error (FCP_ERR_*), message via fcp_last_error().
 *
 * Threading: one context per host thread; calls on one context must not
 * overlap.  Every call is ordered on the context's HIP stream and returns
 * after its results are complete.
 */
#ifndef FCP_H
#define FCP_H

#ifdef __cplusplus
extern "C" {
#endif

#define FCP_NOMAX 6 /* max vertices per face, geometry.f90:40 */

#define FCP_OK 0
#define FCP_ERR_ARG (-1)
#define FCP_ERR_HIP (-2)
#define FCP_ERR_STATE (-3)
#define FCP_ERR_UNSUPPORTED (-4)
#define FCP_ERR_NONFINITE (-5)
#define FCP_ERR_BREAKDOWN (-6)
#define FCP_ERR_SWEEP (-7)

/* Boundary-type codes of polyMesh/boundary (geometry.f90:234-261). */
#define FCP_BC_WALL 1
#define FCP_BC_SYMMETRY 2
#define FCP_BC_INLET 3
#define FCP_BC_OUTLET 4
#define FCP_BC_PRESSURE 5
#define FCP_BC_EMPTY 6
#define FCP_BC_PERIODIC 7
#define FCP_BC_WALL_QFLUX 8 /* 'wallQFlux': prescribed heat-flux wall (temperature.f90:231) */

/* Precision of a solve. */
#define FCP_PREC_FP64 0  /* reference arithmetic, FP64 throughout            */
#define FCP_PREC_MIXED 1 /* FP32 inner Krylov + FP64 iterative refinement    */
#define FCP_PREC_FP64_TILE 2 /* iccg only: the refinement driver of
                                FCP_PREC_MIXED with FP64 inner iterations
                                and the same two-level tile preconditioner
                                (tile-local DIC + coarse correction), the
                                like-for-like FP64 baseline of the mixed
                                path; FCP_ERR_UNSUPPORTED where that
                                preconditioner does not apply (no reference
                                counterpart) */

/* Stopping criterion of a solve. */
#define FCP_CRIT_REF_L1 0  /* reference test: ||r||_1/||r0||_1 < tol_rel or
                              ||r||_1 < tol_abs, linear_solvers.f90:512-532 */
#define FCP_CRIT_TRUE_L2 1 /* ||b - A x||_2 / ||b||_2 <= tol_rel (FP64)       */

typedef struct fcp_ctx fcp_ctx;

typedef struct fcp_solve_opts {
  /* inputs */
  int precision;      /* FCP_PREC_*                                        */
  int criterion;      /* FCP_CRIT_*                                        */
  double inner_tol;   /* mixed precision: relative L2 reduction per inner
                         FP32 solve (e.g. 1e-4)                            */
  int max_outer;      /* mixed precision: max refinement steps             */
  int check_every;    /* iterations between host convergence checks (>=1)  */
  /* outputs */
  int iterations;     /* total (inner) Krylov iterations                   */
  int outer_iterations;
  double res0;        /* initial L1 residual                               */
  double resl;        /* final L1 residual (recursive)                     */
  double resor;       /* reference report value res0/factor (:522-526)     */
  double rel_res_l2;  /* final true ||b-Ax||_2/||b||_2                     */
  double time_ms;     /* device time of the solve                          */
} fcp_solve_opts;

/* --- context ------------------------------------------------------------ */
int fcp_init(int device, fcp_ctx **ctx);
void fcp_finalize(fcp_ctx *ctx);
const char *fcp_last_error(const fcp_ctx *ctx);
/* Default solve options (FP64, reference criterion). */
void fcp_solve_opts_default(fcp_solve_opts *opts);

/* --- mesh and geometry -------------------------------------------------- */
/* Replaces read_mesh_native's data hand-over (geometry.f90:118-406): the
 * caller passes the arrays the reader produced.  xyz is [3*nNodes] (x,y,z per
 * node); node is [FCP_NOMAX*nFaces] (node(nomax,numFaces), geometry.f90:386);
 * startFace as in polyMesh/boundary (0-based offset of the patch's first
 * face); bctype FCP_BC_* codes.  Resident in HBM until the next upload. */
int fcp_mesh_upload(fcp_ctx *ctx, int nNodes, int nCells, int nInnerFaces,
                    int nBoundaryFaces, int nBoundaries, const double *xyz,
                    const int *nnodes, const int *node, const int *owner,
                    const int *neighbour, const int *nfaces, const int *startFace,
                    const int *bctype);

/* The twin of every FCP_BC_PERIODIC patch, in patch order: startFaceTwin as
 * the periodic line of polyMesh/boundary gives it (geometry.f90:251-258);
 * the twin is a whole patch of the same size (typically FCP_BC_EMPTY), not
 * periodic and the twin of no other patch.  Call after fcp_mesh_upload;
 * fcp_geometry then maps the twin faces as read_mesh does (face_mapping,
 * geometry.f90:1848-1997: owners and twin-face centres and area vectors
 * permuted), fcp_csr_build adds the periodic pairs (sparse_matrix.f90:
 * 141-171) and their slots after the inner faces' (:262-293), the
 * Laplacian leaves their entries zero, fcp_update_boundary averages them
 * (updateBoundary.f90:46-68) and fcp_convdiff / fcp_transport couple them
 * (boundary kind 5, facefluxsc_periodic, scalar_fluxes.f90:147-251).  A
 * pair inside one cell or doubling an existing entry is FCP_ERR_UNSUPPORTED
 * at fcp_csr_build. */
int fcp_mesh_periodic(fcp_ctx *ctx, int nPeriodic, const int *startFaceTwin);

/* read_mesh (geometry.f90:90-116) from disk: reads <case_dir>/polyMesh in the
 * native format (format 0, read_mesh_native :118-406: 1-based text files
 * size, boundary, points, faces, owner, neighbour) or the OpenFOAM format
 * (format 1, read_mesh_openfoam :846-1300: 0-based lists under FoamFile
 * headers, numCells = largest owner/neighbour + 1), both with the boundary
 * table "name type nFaces startFace" after a header line, and uploads the
 * mesh as fcp_mesh_upload does.  Boundary types other than the FCP_BC_ names
 * get code 0 (no treatment).  sizes (optional) receives numNodes, numCells,
 * numInnerFaces, numBoundaryFaces, numBoundaries. */
int fcp_read_polymesh(fcp_ctx *ctx, const char *case_dir, int len, int format, int *sizes);

/* Patches of the resident mesh: nfaces, startFace, bctype (FCP_BC_* codes),
 * each [numBoundaries] and optional. */
int fcp_mesh_patches(fcp_ctx *ctx, int *nfaces, int *startFace, int *bctype);

/* FP64 geometry on the GPU, same arithmetic as geometry.f90:416-754
 * (interpolation_coeff_variant = 2).  Any output pointer may be NULL.
 * dnb/srdb: normal distance and |S|/dn for every boundary face (the
 * reference's dnw/srdw, dns/srds are their wall/symmetry subsets). */
int fcp_geometry(fcp_ctx *ctx, double *xc, double *yc, double *zc, double *vol,
                 double *arx, double *ary, double *arz, double *xf, double *yf,
                 double *zf, double *facint, double *Df, double *dnb,
                 double *srdb);

/* create_CSR_matrix (sparse_matrix.f90:86-296): CSR pattern, diagonal
 * positions and face->slot maps icell_jcell_csr_index /
 * jcell_icell_csr_index, built on the GPU and kept resident together with
 * the level schedules of the strict-lower/upper parts.  Outputs (1-based,
 * optional): ia[n+1], ja[nnz], diag[n], slot_pn[Fi + numPeriodic],
 * slot_np[Fi + numPeriodic] (the periodic pairs' slots after the inner
 * faces', numPeriodic = faces on periodic patches; see fcp_mesh_periodic). */
int fcp_csr_build(fcp_ctx *ctx, int *nnz, int *ia, int *ja, int *diag,
                  int *slot_pn, int *slot_np);

/* Level sets of the strict-lower (upper=0) / strict-upper (upper=1) DAG:
 * lev[i] (0-based level numbers), returns number of levels via *nlev. */
int fcp_levels(fcp_ctx *ctx, int upper, int *lev, int *nlev);

/* --- assembly ----------------------------------------------------------- */
/* laplacian(mu,phi) (laplacian.f90:1-151): a = 0, inner faces
 * a(PN)=a(NP)=Gamma_f*D_f with Gamma_f=(1-lambda)mu_P+lambda*mu_N, diagonal
 * minus the row sum, every boundary face Dirichlet: a(PP) -= mu_P|S|/dn,
 * su -= mu_P|S|/dn*phi_b.  The matrix stays resident as the solve operator;
 * a (optional) receives a copy.  su is updated in place (not zeroed). */
int fcp_laplacian(fcp_ctx *ctx, const double *mu, const double *phi, double *a,
                  double *su);

/* Face-diffusivity form of laplacian (extension, no reference counterpart
 * for the per-face coefficient): mu_f[nFaces], bc_kind[nBoundaryFaces]
 * 0 = Dirichlet (as laplacian.f90:93-96 with mu_f of the face),
 * 1 = zero-gradient (no matrix/source term). */
int fcp_laplacian_face(fcp_ctx *ctx, const double *mu_f,
                       const unsigned char *bc_kind, const double *phi,
                       double *a, double *su);

/* Explicit non-orthogonal diffusion correction of facefluxsc
 * (scalar_fluxes.f90:99-109,139): per inner face
 * s = gamma_f*(grad_f . (S - D_f d_PN)), grad_f interpolated with lambda;
 * su_P += s, su_N -= s (temperature.f90:153-154). */
int fcp_nonorth_correction(fcp_ctx *ctx, const double *gamma_f,
                           const double *grad, double *su);

/* Transport-equation assembly as calculate_temperature composes it
 * (temperature.f90:118-366): per inner face facefluxsc (scalar_fluxes.f90:
 * 32-141) -- implicit upwind convection + diffusion gamma*D_f, explicit
 * non-orthogonal diffusion correction and deferred correction
 * gam*(F*phi_f^HO - F*phi^UD) with face value scheme 0 = 'cds',
 * 1 = 'linearUpwind' (interpolation.f90:116-157, 318-356); per boundary face
 * bc_kind 0 = Dirichlet wall (:206-210), 1 = zero-gradient wall (phi_b =
 * phi_P, :224), 2 = inlet/outlet/pressure flux face (facefluxsc_boundary
 * :256-337, :182-190), 3 = no term (symmetry, empty), 4 = prescribed normal
 * gradient (wallQFlux, :231-262: su += dcoef |S|/dn (grad_b . n), grad_b the
 * face's entry of grad), 5 = periodic face (:262-313, facefluxsc_periodic
 * with the twin face's owner; dcoef of the periodic face; its twin takes
 * no term of its own when given 3); then a_ii = sp - (row sum - a_ii) and under-relaxation
 * a_ii /= urf, su += (1-urf) a_ii phi (:350-366).  flmass/dcoef [nFaces],
 * phi [numTotal] (boundary values of zero-gradient faces updated), grad
 * [3*numTotal]; su, sp [nCells] accumulate (pass them cleared, or holding
 * the time terms of fcp_ddt).  The matrix stays resident; a optional copy. */
int fcp_convdiff(fcp_ctx *ctx, const double *flmass, const double *dcoef, double *phi,
                 const double *grad, double gam, int scheme,
                 const unsigned char *bc_kind, double urf, double *a, double *su,
                 double *sp);

/* fcp_convdiff with facefluxscalar (scalar_fluxes.f90:344-464) on the inner
 * faces: the diffusivity diffcoef[numCells] is given per cell and each inner
 * face takes game = diffcoef(P)*(1-lambda) + diffcoef(N)*lambda (:384-392),
 * used as facefluxsc uses its dcoef.  The reference has no boundary form of
 * facefluxscalar; here a boundary face (periodic ones included) takes
 * diffcoef of its owner.  Other arguments and outputs as fcp_convdiff. */
int fcp_convdiff_cell(fcp_ctx *ctx, const double *flmass, const double *diffcoef, double *phi,
                      const double *grad, double gam, int scheme,
                      const unsigned char *bc_kind, double urf, double *a, double *su,
                      double *sp);

/* The pressure-correction step of SIMPLE (calcp_simple.f90:69-419) for an
 * incompressible flow on a mesh without 'pressure' patches
 * (FCP_ERR_UNSUPPORTED), pressure interpolated 'linear', one corrector
 * (npcor = 1); the caller solves between the two calls (fcp_csrsolve on the
 * resident matrix, or lis_spsolve).
 *
 * fcp_pressure_assemble (:69-228): per inner face facefluxmass2
 * (faceflux_mass.f90:175-252) -- coefficient -den_f (V/a_P)_f D_f into both
 * off-diagonals and subtracted from both diagonals, Rhie-Chow mass flux into
 * flmass and -/+ flmass into su; unless const_mflux, adjustMassFlow (:833-912:
 * outlet u, v, w copied from the owner, outlet fluxes den u.S summed in
 * face order and scaled with u, v, w by flomas/(sum + small)); su -= flmass
 * on inlet and outlet faces; periodic pairs by facefluxmass2_periodic
 * (:313-372, which reads Df of inner face i for periodic face i).
 * u, v, w, p, den [numTotal]; dpdxi [3*numTotal]; apu, apv, apw [numCells]
 * (1/a_P of the momentum rows); flmass [numFaces] in/out; a (optional) [nnz]
 * and su [numCells] overwritten.  The matrix stays resident.
 *
 * fcp_pressure_correct (:333-419): flmass += a_PN (pp_N - pp_P) on inner and
 * periodic faces (the twin face takes the periodic face's flux); ppref =
 * pp(pref_cell) (1-based); gradp_and_sources(pp) (nablap.f90:19-209:
 * bpres stage 1 copies pp to the boundary faces, stage 2 extrapolates it to
 * wall faces with the stage-1 gradient; dpdxi's cell rows receive the
 * gradient of pp, su/sv/sw = -sum_f pp_f S_f); u += su apu, v += sv apv,
 * w += sw apw, p += urfp (pp - ppref) on the cells; then
 * updateVelocityAtBoundary (velocity.f90:1184-1279: empty and periodic
 * faces take the owner's velocity, symmetry faces u_P - (u_P.S) S).  pp
 * (boundary values updated), flmass, u, v, w, p, dpdxi in/out; su, sv, sw
 * [numCells] optional outputs.  Uses the resident matrix of the assembly. */
int fcp_pressure_assemble(fcp_ctx *ctx, double *u, double *v, double *w, const double *p,
                          const double *dpdxi, const double *den, const double *apu,
                          const double *apv, const double *apw, double *flmass, int const_mflux,
                          double flomas, double *a, double *su);
int fcp_pressure_correct(fcp_ctx *ctx, double *pp, double *flmass, double *u, double *v,
                         double *w, double *p, double *dpdxi, const double *apu,
                         const double *apv, const double *apw, double urfp, int pref_cell,
                         double *su, double *sv, double *sw);

/* The momentum assembly of calcuvw (velocity.f90:50-750) before its three
 * solves, for steady, BDF1, BDF2 and Crank-Nicolson time schemes
 * (time_scheme 0/1/2/3), conv_scheme 0 'cds' / 1 'linearUpwind', Gauss
 * gradients, no buoyancy, Lorentz force, PISO or constant mass flux:
 * updateVelocityAtBoundary, the gradients of u, v, w, gradp_and_sources(p),
 * the time terms, facefluxuvw per inner face (with sngrad's transposed
 * stress terms, :754-878), per boundary face by patch type inlet / outlet /
 * pressure (facefluxuvw_bnd :882-1034), symmetry (:354-391, which reads the
 * viscosity at index numCells+1), periodic (facefluxuvw_periodic
 * :1038-1179) and wall (visw of the face's place among the wall faces, in
 * patch order), the Crank-Nicolson block and, per component, the diagonal
 * a_P = sp - (sum(row) - a_row_diag) with its relaxation factor urf[3].  As in
 * the reference the row sums of u read whatever the resident matrix held on
 * its diagonal before the call, and v's and w's a zeroed diagonal.
 * u, v, w, p [numTotal] (boundary values updated); uo, vo, wo (transient),
 * uoo, voo, woo (BDF2) [numCells]; den, vis [numTotal]; visw [wall faces];
 * flmass [numFaces].  Outputs: the resident matrix's off-diagonals (shared)
 * and w's diagonal, as the reference leaves a (a optional copy);
 * rhs and diagv [3*numCells]: the right-hand side and relaxed diagonal of
 * u, v and w (column-major), ready for fcp_csrsolve_multi_diag; optional
 * dpdxi [3*numTotal] (cell rows: grad p), grad [9*numTotal] (grad u, v, w),
 * ap [3*numCells] (1/a_P: apu, apv, apw), spv [3*numCells] (the diagonal
 * sources spu, spv, sp) and src [3*numCells] (su, sv, sw before the
 * relaxation terms). */
int fcp_momentum_assemble(fcp_ctx *ctx, double *u, double *v, double *w, const double *uo,
                          const double *vo, const double *wo, const double *uoo,
                          const double *voo, const double *woo, double *p, const double *den,
                          const double *vis, const double *visw, const double *flmass,
                          double viscos, double dt, int time_scheme, int conv_scheme, double gds,
                          const double *urf, double *dpdxi, double *grad, double *rhs,
                          double *diagv, double *ap, double *spv, double *src, double *a);

/* Time terms (temperature.f90:82-102): scheme 1 BDF1 (and the volume part of
 * Crank-Nicolson), 2 BDF2; apotime = den*vol/dt; su/sp updated in place. */
int fcp_ddt(fcp_ctx *ctx, int scheme, const double *den, double dt,
            const double *phi_o, const double *phi_oo, double *su, double *sp);

/* calculate_temperature (temperature.f90:35-385) as one call, for any scalar
 * transported like the temperature: grad(phi) with grad_method (FCP_GRAD_*;
 * least-squares matrices created beforehand), the face diffusion
 * coefficients from the cell viscosity vis (viscos*prl + ((vis_P + (vis_N -
 * vis_P) lambda) - viscos)*prtr inside, prl = 1/pranl, prtr = 1/sigt, and
 * the boundary forms of :178,203,237), the time terms (time_scheme 0 steady,
 * 1 BDF1, 2 BDF2, 3 Crank-Nicolson with its coefficient block :325-347),
 * facefluxsc with conv_scheme (0 'cds', 1 'linearUpwind') and deferred
 * correction gds, the boundary loop by patch type (inlet/outlet/pressure flux
 * faces; 'wall' patches isothermal when wall_frac[ib] = bcDFraction(7,ib) is
 * 1, adiabatic when 0; wallQFlux; symmetry/empty without a term; periodic
 * patches are FCP_ERR_UNSUPPORTED), the diagonal with under-relaxation urf,
 * csrsolve(solver) from the current cell values, updateBoundary, and -- when
 * clip is 1 -- the clip of negative fields to `small` (:381).  Wall terms read
 * |S|/dn through the reference's own wall counter (an isothermal wall that
 * follows an adiabatic wall takes the earlier wall's faces' values).
 * phi[numTotal] in/out; phi_o, phi_oo, den, vis [numCells] (phi_o, den for
 * transient schemes, phi_oo for BDF2); flmass[numFaces]; wall_frac
 * [nBoundaries]; grad (optional, [3*numTotal]) receives dphi/dxi and a, su,
 * sp (optional; [nnz], [numCells], [numCells]) the assembled system, which
 * the reference leaves in its module globals; resor the solver's resor
 * (csrsolve's res0); fimin/fimax the cell range before the clip.  The
 * assembled matrix stays resident.  Returns the solve's code. */
int fcp_transport(fcp_ctx *ctx, double *phi, const double *phi_o, const double *phi_oo,
                  const double *den, const double *vis, const double *flmass,
                  const double *wall_frac, double viscos, double pranl, double sigt, double urf,
                  double gds, double dt, int time_scheme, int conv_scheme, int grad_method,
                  const char *solver, int solver_len, int itr_max, double tol_abs, double tol_rel,
                  int clip, double *grad, double *a, double *su, double *sp, double *resor,
                  double *fimin, double *fimax, fcp_solve_opts *opts);

/* grad_gauss (gradients.f90:1607-1715). dudxi[3*numTotal], boundary rows 0. */
int fcp_grad_gauss(fcp_ctx *ctx, const double *u, double *dudxi);

/* --- discrete adjoint ----------------------------------------------------
 * No counterpart in the reference (its only transpose primitive is atx_cr,
 * linear_solvers.f90:1079-1171).  For a system A(c) phi = b(c, phi_old,
 * grad_old) assembled by the entry points above and an adjoint lam solving
 * A^T lam = dJ/dphi, these accumulate (+=) the reverse-mode derivatives of J;
 * every output is written by one thread (deterministic). */

/* A <- A^T for the resident matrix (face-based swap of the slot_pn / slot_np
 * coefficients); call twice to restore A.  A following csrsolve then solves
 * the transposed system. */
int fcp_matrix_transpose(fcp_ctx *ctx);

/* Reverse of fcp_convdiff (urf = 1 only): inputs as given to the forward
 * call (phi_old = its phi[numTotal], grad_old = its grad), lam[numCells],
 * phi_new[numCells] = solution of the assembled system.  Accumulates
 * dcoef_bar[nFaces], phi_old_bar[numTotal], grad_old_bar[3*numTotal]. */
int fcp_convdiff_backward(fcp_ctx *ctx, const double *flmass, const double *dcoef,
                          const double *phi_old, const double *grad_old, double gam,
                          int scheme, const unsigned char *bc_kind, const double *lam,
                          const double *phi_new, double *dcoef_bar, double *phi_old_bar,
                          double *grad_old_bar);

/* Reverse of fcp_laplacian_face (mu_f, bc_kind, phi) followed by
 * fcp_nonorth_correction (gamma_f, grad) (gamma_f and grad may be NULL when
 * no correction was applied).  Accumulates mu_f_bar[nFaces],
 * phi_bc_bar[numTotal] (boundary entries), gamma_bar[nFaces],
 * grad_bar[3*numTotal]. */
int fcp_laplacian_face_backward(fcp_ctx *ctx, const double *mu_f,
                                const unsigned char *bc_kind, const double *phi_bc,
                                const double *gamma_f, const double *grad, const double *lam,
                                const double *phi, double *mu_f_bar, double *phi_bc_bar,
                                double *gamma_bar, double *grad_bar);

/* Reverse of fcp_grad_gauss: phi_bar[numTotal] += G^T grad_bar. */
int fcp_grad_gauss_backward(fcp_ctx *ctx, const double *grad_bar, double *phi_bar);

/* Reverse Cuthill-McKee renumbering of the cell graph of the resident mesh
 * (adjacency of the CSR pattern), computed on the GPU; no counterpart in the
 * reference.  Each connected component starts from its minimum-(degree, id)
 * cell refined by at most pp_iters George-Liu pseudo-peripheral steps;
 * Cuthill-McKee appends unvisited neighbours in (degree, id) order; the
 * order is reversed.  perm[numCells] receives the new 1-based number of
 * every cell (bit-identical to the CPU rule of the same definition). */
int fcp_rcm(fcp_ctx *ctx, int pp_iters, int *perm);

/* Gradient methods of grad_scalar_field (gradients.f90:106-163). */
#define FCP_GRAD_GAUSS 0
#define FCP_GRAD_LSQ 1     /* lstsq    : grad_lsq    (:782-893)   */
#define FCP_GRAD_LSQ_DM 2  /* lstsq_dm : grad_lsq_dm (:1334-1486) */
#define FCP_GRAD_LSQ_QR 3  /* lstsq_qr : grad_lsq_qr (:1063-1152) */
/* fvxGradient.f90's Gauss gradients, face values corrected by an old
 * gradient (gradco :1761-1817): */
#define FCP_GRAD_GAUSS_FVX 4        /* grad_gauss, two sweeps (:1549-1645)   */
#define FCP_GRAD_GAUSS_CORRECTED 5  /* grad_gauss_corrected (:1664-1758)     */

/* create_lsq_grad_matrix (gradients.f90:72-101): builds and keeps the
 * per-cell least-squares matrices of `method` (Dmat(9,numCells) for lsq and
 * lsq_dm, D(3,6,numCells) for lsq_qr, which needs cells of <= 6 faces).
 * compat = 1 reproduces the reference arithmetic including its defects
 * (y-row cofactors at :887/:1480, boundary weight xf(i) at :1459); 0 computes
 * the corrected gradient.  dmat (may be NULL) receives the matrices. */
int fcp_create_lsq_grad_matrix(fcp_ctx *ctx, int method, int compat, double *dmat);

/* grad(phi, dPhidxi) for a scalar field (gradients.f90:106-163, no limiter):
 * FCP_GRAD_GAUSS, or the least-squares method whose matrices were created;
 * or fvxGradient.f90's FCP_GRAD_GAUSS_FVX or FCP_GRAD_GAUSS_CORRECTED, for
 * which dPhidxi is intent(inout) as in the reference: its cell rows on entry
 * are the old gradient the one corrected sweep starts from.
 * dPhidxi[3*numTotal], boundary rows 0 on return. */
int fcp_grad(fcp_ctx *ctx, int method, const double *phi, double *dphidxi);

/* updateBoundary (updateBoundary.f90:1-74): zero-gradient copy for
 * outlet/symmetry/pressure/empty patches. */
int fcp_update_boundary(fcp_ctx *ctx, double *phi);

/* --- linear algebra ----------------------------------------------------- */
/* Replace the resident matrix values (CSR order of fcp_csr_build). */
int fcp_matrix_set(fcp_ctx *ctx, const double *a);

/* y = A x with the resident matrix (ax_cr, linear_solvers.f90:1173;
 * matvec of iccg :493-498).  x, y of length nCells. */
int fcp_spmv(fcp_ctx *ctx, const double *x, double *y);

/* y = A^T x with the resident matrix (atx_cr, linear_solvers.f90:1079-1171),
 * summed in the reference's row order.  Needs a symmetric sparsity pattern
 * (every finite-volume matrix has one), else FCP_ERR_UNSUPPORTED. */
int fcp_spmv_transpose(fcp_ctx *ctx, const double *x, double *y);

/* One application z = M^{-1} r of the DIC (kind 0, iccg :439-475), DILU
 * (kind 1, bicgstab :613-683) or ILU(0) (kind 2, ilu_cr + lus_cr
 * :1267-1477, the preconditioner of pmgmres_ilu) preconditioner of the
 * resident matrix, with the reference's operation order; d (optional, [n])
 * receives the inverted pivots of kinds 0 and 1 and is ignored for kind 2
 * (fcp_ilu_factor returns those factors).  Kinds 3 and 4 apply, in FP64,
 * the tile-local DIC / DILU that preconditions the FP32 inner solves of
 * FCP_PREC_MIXED (no reference counterpart: the DIC / DILU recurrences with
 * the couplings between tiles dropped; fcp_sweep_tiles gives the tiles).
 * Kinds 5 and 6 apply the same in FP32 with the tile-order kernel of the
 * inner solves (input rounded to FP32); kind 7 is kind 5 plus the coarse
 * correction P0 A0^-1 P0^T r of FCP_TUNE_INNER_COARSE (FCP_ERR_UNSUPPORTED
 * where it does not apply).  Kinds 8 and 9 are kinds 5 and 7 in FP64 (the
 * tile-order kernel of FCP_PREC_FP64_TILE).  r, z of length nCells. */
int fcp_precond_apply(fcp_ctx *ctx, int kind, const double *r, double *z, double *d);

/* Tile (0-based block number) of every cell in the tile-local
 * preconditioner of the FP32 inner solves; tile[nCells]. */
int fcp_sweep_tiles(fcp_ctx *ctx, int *tile);

/* ILU(0) factors of the resident matrix as ilu_cr leaves them
 * (linear_solvers.f90:1267-1385): l[nnz] in CSR order, multipliers left of
 * the diagonal, U from it on.  FCP_ERR_BREAKDOWN on a zero pivot (the
 * reference stops there). */
int fcp_ilu_factor(fcp_ctx *ctx, double *l);

/* csrsolve (linear_solvers.f90:40-91) on the resident matrix: solver is
 * "iccg" | "dpcg" | "bicgstab" | "gauss-seidel" | "pmgmres"; fi[numTotal]
 * (only 1..n used), rhs[n], res0 receives the reference's report value
 * (resor).  opts may be NULL (FP64, reference criterion).  verbose >= 1
 * prints the reference's report line, verbose >= 2 per-iteration residuals.
 * "pmgmres" is GMRES(4) with ILU(0) (pmgmres_ilu :789-1077, restart fixed
 * at 4 as :85 does; itr_max counts restart cycles), FP64 only.  With
 * FCP_CRIT_REF_L1 it reproduces the reference, including its stopping test
 * resl/resor <= resor*tol_rel and its normalisation of the preconditioned
 * residual by the unpreconditioned norm (:970-978), which can stop early
 * on a wrong solution; FCP_CRIT_TRUE_L2 normalises by the preconditioned
 * norm and stops on the true relative L2 residual at a restart. */
int fcp_csrsolve(fcp_ctx *ctx, const char *solver, int solver_len, double *fi,
                 const double *rhs, double *res0, int itr_max, double tol_abs,
                 double tol_rel, const char *chvar, int chvar_len, int verbose,
                 fcp_solve_opts *opts);

/* The LIS interface (LIS_linear_solvers.f90:38-200) without the LIS library:
 * fcp_lis_create_solver (lis_create_solver :38-90) takes the LIS option
 * string -- -i cg|bicgstab|gmres|gs (or LIS's numbers 1, 4, 9, 11), -p
 * none|jacobi|ilu (0, 1, 2), -tol (1e-12), -maxiter (1000), -restart (40),
 * -initx_zeros true|false (true: LIS starts from zero), -conv_cond 0|1
 * (||r||_2 <= tol ||r_0||_2 | <= tol ||b||_2), -f double; -print, -scale,
 * -storage*, -adds*, -omp_num_threads are accepted and have no effect --
 * and maps the method to this library's solver of the same Krylov method
 * and preconditioner family: cg+ilu -> iccg (DIC), cg+jacobi -> dpcg,
 * bicgstab+ilu -> bicgstab (DILU), gs+none -> gauss-seidel, gmres+ilu ->
 * GMRES(restart) with ILU(0); other pairs and options are
 * FCP_ERR_UNSUPPORTED.  fcp_lis_spsolve (lis_spsolve :92-173) solves the
 * resident system for fi[n] (in: the current field, out: the solution),
 * prints its report line (initial residual = L1 of rhs - A fi, final =
 * that times the relative residual norm), and returns 0 when converged,
 * the iteration count when not, < 0 on error; res0 (L1 initial residual),
 * resid (relative residual norm of the convergence condition) and iter are
 * optional outputs.  No LIS counterpart of the iteration count exists:
 * it is the iteration count of the mapped solver. */
int fcp_lis_create_solver(fcp_ctx *ctx, const char *options, int len);
int fcp_lis_spsolve(fcp_ctx *ctx, double *fi, const double *rhs, const char *chvar,
                    int chvar_len, double *res0, double *resid, int *iter);
int fcp_lis_destroy_solver(fcp_ctx *ctx);

/* Batched csrsolve: nrhs (1..4) systems sharing the resident matrix, as the
 * three momentum components of velocity.f90:638-748 share one operator
 * (there each is a separate csrsolve call).  fi is fi(ldfi, nrhs) and rhs is
 * rhs(numCells, nrhs), column-major (Fortran order), ldfi >= numCells; only
 * rows 1..numCells of each column are touched.  solver "iccg" | "bicgstab".
 * Each column runs its own recurrence, bit-equal to fcp_csrsolve on that
 * column; SpMV and the preconditioner sweeps read the operator once for all
 * columns.  res0[nrhs] (may be NULL) receives each column's res0; opts
 * reports the largest iteration count and the worst true relative residual.
 * Returns 0, the iteration count when a column did not converge, or < 0. */
int fcp_csrsolve_multi(fcp_ctx *ctx, const char *solver, int solver_len, int nrhs,
                       double *fi, int ldfi, const double *rhs, double *res0,
                       int itr_max, double tol_abs, double tol_rel, int verbose,
                       fcp_solve_opts *opts);

/* fcp_csrsolve_multi with a diagonal of its own per column and the resident
 * matrix's off-diagonals shared, as the momentum components of
 * velocity.f90:611-748 share theirs: diagv(numCells, nrhs), column-major,
 * holds each column's diagonal entries (a_ii after under-relaxation, :617-
 * 627, :673-681, :729-737); the resident matrix's own diagonal is not used.
 * Each column is bit-equal to fcp_csrsolve on the resident matrix with its
 * diagonal replaced by that column's (DIC / DILU pivots per column, set up
 * from that matrix); the batched SpMV and sweeps read the off-diagonals
 * once for all columns.  FP64 only (FCP_ERR_UNSUPPORTED for other
 * precisions, and where the batched sweeps do not fit the mesh). */
int fcp_csrsolve_multi_diag(fcp_ctx *ctx, const char *solver, int solver_len, int nrhs,
                            double *fi, int ldfi, const double *rhs, const double *diagv,
                            double *res0, int itr_max, double tol_abs, double tol_rel,
                            int verbose, fcp_solve_opts *opts);

/* Explicit-array entry points with the reference signatures
 * iccg/dpcg/bicgstab/GaussSeidel(n,nnz,ia,ja,a,diag,fi,rhs,resor,itr_max,
 * tol_abs,tol_rel,chvar,verbose) (linear_solvers.f90:364,206,548,96).
 * ia/ja/diag 1-based.  They run on a library-owned context on device 0. */
int fcp_iccg(int n, int nnz, const int *ia, const int *ja, const double *a,
             const int *diag, double *fi, const double *rhs, double *resor,
             int itr_max, double tol_abs, double tol_rel, int verbose);
int fcp_dpcg(int n, int nnz, const int *ia, const int *ja, const double *a,
             const int *diag, double *fi, const double *rhs, double *resor,
             int itr_max, double tol_abs, double tol_rel, int verbose);
int fcp_bicgstab(int n, int nnz, const int *ia, const int *ja, const double *a,
                 const int *diag, double *fi, const double *rhs, double *resor,
                 int itr_max, double tol_abs, double tol_rel, int verbose);
int fcp_gauss_seidel(int n, int nnz, const int *ia, const int *ja, const double *a,
                     const int *diag, double *fi, const double *rhs, double *resor,
                     int itr_max, double tol_abs, double tol_rel, int verbose);
/* pmgmres_ilu(n,nz_num,ia,ja,a,ua,x,rhs,resor,itr_max,mr,tol_abs,tol_rel,
 * chvar,verbose) (linear_solvers.f90:789): restart length mr (1..32, < n),
 * itr_max restart cycles, reference stopping test. */
int fcp_pmgmres(int n, int nnz, const int *ia, const int *ja, const double *a,
                const int *diag, double *fi, const double *rhs, double *resor,
                int itr_max, int mr, double tol_abs, double tol_rel, int verbose);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the Jacobi-preconditioned CG of fcp_dpcg (dpcg's
 * recurrence and L1 stopping test) on DEVICE arrays of the current device,
 * ia/ja/diag 1-based as there, on the given stream (NULL: the null stream).
 * Its work vectors and reduction state live in a fixed device array of the
 * library (about 6 MB, held by every device it runs on until the process
 * ends), so it allocates nothing and takes at most FCP_NOALLOC_MAX_ROWS rows;
 * larger systems, host arrays, arrays shorter than their stated sizes and
 * index arrays that are not a CSR matrix with its diagonal are rejected with
 * FCP_ERR_ARG before any solve kernel runs (message: fcp_last_error(NULL)).
 * Every call on a device shares that device's array: calls that overlap on
 * one device race.  resor (host) receives the reference's resor.  Returns 0,
 * the iteration count when the stopping test was not met, or < 0. */
#define FCP_NOALLOC_MAX_ROWS (1 << 18)
int fcp_dpcg_noalloc(int n, int nnz, const int *ia, const int *ja, const double *a,
                     const int *diag, double *fi, const double *rhs, double *resor,
                     int itr_max, double tol_abs, double tol_rel, void *stream);

/* --- launch-shape tuning (autotuner hook) --------------------------------- */
#define FCP_TUNE_SWEEP_WAVES 1   /* resident waves of a sweep kernel (0 = auto) */
#define FCP_TUNE_SWEEP_SLEEP 2   /* s_sleep argument between dependency polls   */
#define FCP_TUNE_SWEEP_KIND 3    /* 0 row-block LDS sweeps, 1 level-task sweeps */
#define FCP_TUNE_SWEEP_ROWS 4    /* rows per block of the row-block sweeps      */
#define FCP_TUNE_SWEEP_TILE 5    /* rows per geometric tile of the pipelined
                                    sweeps (0 = auto, 1 = contiguous blocks)   */
#define FCP_TUNE_INNER_PRECOND 6 /* preconditioner of the FP32 inner solves of
                                    FCP_PREC_MIXED: 1 (default) tile-local
                                    DIC/DILU (the reference recurrences
                                    restricted to ~4096-cell tiles, which all
                                    run at once), 0 the global DIC/DILU of the
                                    FP64 solve.  FP64 solves and the FP64
                                    refinement always use the global one.      */
#define FCP_TUNE_INNER_TILE 7    /* rows per tile of the tile-local
                                    preconditioner (0 = auto: 4096)            */
#define FCP_TUNE_GRAPH 8         /* 1: inner FP32 iterations of a mixed-
                                    precision PCG are replayed from a captured
                                    HIP graph (two iterations per replay);
                                    0 (default): launched one kernel at a
                                    time (as fast on MI355X: the gap between
                                    dependent kernels is on the device side).
                                    Profiling launches them one at a time.     */
#define FCP_TUNE_INNER_COARSE 9  /* 1 (default): the tile-local DIC of the
                                    FP32 inner PCG is completed by a coarse
                                    correction with one unknown per tile
                                    (z += P0 A0^-1 P0^T r, A0 = P0^T A P0, P0
                                    the tile indicator; up to 2048 tiles),
                                    which cuts the inner iterations several
                                    times on large meshes; 0: tile-local only */
#define FCP_TUNE_SPMV 10         /* variant of the standalone FP64 SELL SpMV
                                    (fcp_spmv): -1 (default) automatic: 4 when
                                    the operator and two vectors exceed the
                                    256 MB Infinity Cache, else 0; 0 one
                                    row per thread, 1 the same with
                                    non-temporal matrix loads, 2 two rows per
                                    thread with non-temporal matrix loads, 3
                                    two rows per thread, 4 as 1 with
                                    non-temporal y stores; each row ends at
                                    its first padding slot (column -1), so
                                    no variant reads the row pointers      */
#define FCP_TUNE_TILE_PACK 11    /* tile-order sweeps: 1 (default) the packed
                                    operand layout (three 10-bit dependency
                                    distances in each coefficient vector's
                                    fourth component) where every position
                                    has at most 3 in-tile dependencies per
                                    sweep; 0 the 4-slot layout.  Both give
                                    the same results bit for bit          */
#define FCP_TUNE_COARSE_FUSE 12  /* inner PCG with the coarse correction: 1
                                    (default) the tile kernel adds it itself
                                    after a grid barrier where every tile's
                                    workgroup fits on the device at once; 0
                                    a separate kernel.  The two differ only
                                    in the summation order of one dot
                                    product                                */
#define FCP_TUNE_ASM_BLOCKS 13  /* face Laplacian, Gauss gradient and
                                    non-orthogonal correction: 1 (default)
                                    blocks of 256 cells compute each face's
                                    term once into LDS (the faces their
                                    cells own, and those reaching them from
                                    other blocks) and write their rows of a
                                    from LDS, where the inner faces are
                                    sorted by owner and a block's faces fit;
                                    0 every cell reads its faces from HBM.
                                    Same results bit for bit               */
/* Sets a launch-shape parameter of the context; results do not depend on it
 * (FCP_TUNE_INNER_PRECOND and FCP_TUNE_INNER_COARSE change the path of a
 * mixed-precision solve, not its stopping test). */
int fcp_tune_set(fcp_ctx *ctx, int param, int value);

/* Autotuner (no counterpart in the reference; the launch-shape search of the
 * mixed-precision path): for solver "iccg" or "bicgstab", times the mixed
 * FP32-inner solve of A x = A*1 on the resident matrix to a true relative
 * residual of 1e-6 with tile-local preconditioner tiles of 1024, 2048, 4096
 * and 8192 cells (those with at least four tiles), keeps the fastest in the
 * context (FCP_TUNE_INNER_TILE) and caches it for the rest of the process
 * under the CSR pattern (cell count, entry count, a 64-bit signature of ia
 * and ja) and the solver, so later contexts on the same mesh take it
 * without trials.  tile_out (optional) receives the tile rows chosen (0:
 * the default), cached (optional) 1 when the cache supplied it. */
int fcp_autotune(fcp_ctx *ctx, const char *solver, int solver_len, int *tile_out, int *cached);

/* --- instrumentation ------------------------------------------------------ */
/* Per-kernel device timing accumulated on the context's stream with HIP
 * events (name -> total ms, launches).  enable=1 starts, 0 stops. */
int fcp_profile(fcp_ctx *ctx, int enable);
/* Writes "name launches total_ms\n" lines into buf (truncated to len). */
int fcp_profile_report(fcp_ctx *ctx, char *buf, int len);

/* Context stream (hipStream_t) for callers that enqueue their own work. */
void *fcp_stream(fcp_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
