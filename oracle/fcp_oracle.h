/*
 * fcp_oracle.h -- CPU restatement of freeCappuccino's finite-volume
 * assemble-and-solve path.  TEST INFRASTRUCTURE ONLY: this library is the
 * checker for the HIP path (tests, __graft_entry__.smoke and the CPU
 * baseline leg of bench.py).  Nothing in the product links or calls it.
 *
 * Conventions: every index is 0-based.  Field arrays have length
 * numTotal = nCells + nBoundaryFaces with boundary-face values at
 * nCells.. (geometry.f90:282-289).  Gradients are AoS [3*numTotal]
 * (dPhidxi(3,numTotal), gradients.f90:1620).  Each function follows the
 * reference's order of floating-point operations; file:line citations are
 * relative to the reference repository root.
 */
#ifndef FCP_ORACLE_H
#define FCP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NOMAX 6

typedef struct {
  int n_nodes, n_cells, n_inner, n_faces, n_patches;
  const double *points;     /* [n_nodes*3] xyz AoS                        */
  const int *nnodes;        /* [n_faces]                                  */
  const int *node;          /* [n_faces*ORC_NOMAX] vertex ids             */
  const int *owner;         /* [n_faces]                                  */
  const int *neighbour;     /* [n_inner]                                  */
  const int *nfaces;        /* [n_patches]                                */
  const int *start_face;    /* [n_patches] id of the patch's first face    */
  const int *bctype;        /* [n_patches] codes: 1 wall 2 symmetry 3 inlet
                               4 outlet 5 pressure 6 empty 7 periodic
                               8 wallQFlux                                 */
  const int *twin_start;    /* [periodic patches] start face of each
                               periodic patch's twin, in patch order
                               (startFaceTwin, geometry.f90:257); NULL
                               without periodic patches                    */
} orc_mesh;

typedef struct {
  double *xc, *yc, *zc, *vol;               /* [n_cells]  */
  double *arx, *ary, *arz, *xf, *yf, *zf;   /* [n_faces]  */
  double *facint, *Df;                      /* [n_inner]  */
  double *dnb, *srdb;                       /* [n_bnd]: normal distance face
                                               centre -> owner centre and
                                               |S|/dn for every boundary face */
} orc_geom;

/* geometry.f90:416-754 (native reader, interpolation variant 2). */
int orc_geometry(const orc_mesh *m, orc_geom *g);

/* face_mapping (geometry.f90:1848-1997), which read_mesh runs after the
 * geometry when there are periodic patches: for face i of the k-th periodic
 * patch, the twin face whose owner centre is nearest to the face centre (L1
 * distance, first minimum) becomes twin face i -- its owner, centre and area
 * vector move there.  owner[n_faces] receives the mapped owners (m->owner
 * holds the mesh's own); g's face arrays are permuted in place. */
void orc_face_mapping(const orc_mesh *m, orc_geom *g, int *owner);

/* Faces on periodic patches (numPeriodic, geometry.f90:253). */
int orc_n_periodic(const orc_mesh *m);

/* sparse_matrix.f90:86-296: CSR pattern, diagonal and face->slot maps with
 * the periodic pairs (owner of periodic face i, owner of twin face i; m's
 * owners after face mapping).  ia[n+1] (ia[0]=0), ja[nnz], diag[n],
 * slot_pn / slot_np [F_i + numPeriodic] (inner faces, then periodic faces
 * in patch order).  Returns nnz (>0), or -2 where a periodic pair repeats a
 * matrix entry (a pair inside one cell or one already coupled). */
int orc_csr_nnz(const orc_mesh *m);
int orc_csr_build(const orc_mesh *m, int *ia, int *ja, int *diag,
                  int *slot_pn, int *slot_np);

typedef struct {
  int n, nnz;
  const int *ia, *ja, *diag;
} orc_csr;

/* laplacian.f90:1-151 (cell mu, all boundary faces Dirichlet). a is zeroed,
 * su is updated in place (not zeroed), exactly as the reference. */
void orc_laplacian(const orc_mesh *m, const orc_geom *g, const orc_csr *c,
                   const int *slot_pn, const int *slot_np,
                   const double *mu, const double *phi, double *a, double *su);

/* Face-diffusivity extension of laplacian.f90: mu_f[n_faces] per face and
 * bc_kind[n_bnd] (0 = Dirichlet, 1 = zero-gradient: no contribution).
 * With mu_f = (1-lambda)*mu_P + lambda*mu_N on inner faces, mu_f = mu_P on
 * boundary faces and every boundary Dirichlet it reproduces orc_laplacian
 * bit for bit. */
void orc_laplacian_face(const orc_mesh *m, const orc_geom *g, const orc_csr *c,
                        const int *slot_pn, const int *slot_np,
                        const double *mu_f, const unsigned char *bc_kind,
                        const double *phi, double *a, double *su);

/* gradients.f90:1607-1715 grad_gauss (dudxi[3*numTotal], boundary rows 0). */
void orc_grad_gauss(const orc_mesh *m, const orc_geom *g, const double *u,
                    double *dudxi);

/* fvxGradient.f90 Gauss gradients with face values corrected by an old
 * gradient (gradco :1761-1817).  gold == NULL: grad_gauss (:1549-1645), two
 * sweeps from a zero old gradient; else grad_gauss_corrected (:1664-1758),
 * one sweep from gold[3*n].  dudxi[3*numTotal], boundary rows 0. */
int orc_grad_gauss_fvx(const orc_mesh *m, const orc_geom *g, const double *u,
                       const double *gold, double *dudxi);

/* LSQ gradients: variant 0 = lsq (gradients.f90:660-893), 1 = lsq_dm
 * (:1157-1486), 2 = lsq_qr with modified Gram-Schmidt (:900-1152).
 * compat=1 reproduces the reference arithmetic including its defects
 * (y-row cofactors at :887/:1480, xf(i) for xf(iface) at :1459);
 * compat=0 computes the corrected least-squares gradient.
 * dmat: [9*n] for variants 0/1, [18*n] (3x6 per cell) for variant 2. */
int orc_lsq_matrix(const orc_mesh *m, const orc_geom *g, int variant, int compat,
                   double *dmat);
int orc_grad_lsq(const orc_mesh *m, const orc_geom *g, int variant, int compat,
                 const double *dmat, const double *phi, double *dphidxi);

/* updateBoundary.f90:1-74: zero-gradient copy for outlet/symmetry/pressure/
 * empty patches and, for periodic ones, the average of the owner and twin
 * owner values written to both faces, patch by patch in order. */
void orc_update_boundary(const orc_mesh *m, double *phi);

/* Linear solvers on CSR (linear_solvers.f90).  fi[n] in/out, rhs[n].
 * hist (may be NULL) receives resl after each iteration (capacity itr_max).
 * Return value: iterations used.  *resor: reference report value
 * res0/factor;  *res0_out: initial L1 residual; *resl_out: final L1. */
typedef struct {
  double resor, res0, resl, factor;
  int iters;
} orc_solve_info;

int orc_iccg(const orc_csr *c, const double *a, double *fi, const double *rhs,
             int itr_max, double tol_abs, double tol_rel, double *hist,
             orc_solve_info *info);
/* orc_iccg with the stopping test ||r||_2 <= tol_rel ||r0||_2 on the
 * recursively updated residual instead of the reference's L1 ratio. */
int orc_iccg_l2(const orc_csr *c, const double *a, double *fi, const double *rhs,
                int itr_max, double tol_abs, double tol_rel, double *hist,
                orc_solve_info *info);
int orc_dpcg(const orc_csr *c, const double *a, double *fi, const double *rhs,
             int itr_max, double tol_abs, double tol_rel, double *hist,
             orc_solve_info *info);
int orc_bicgstab(const orc_csr *c, const double *a, double *fi, const double *rhs,
                 int itr_max, double tol_abs, double tol_rel, double *hist,
                 orc_solve_info *info);
int orc_gauss_seidel(const orc_csr *c, const double *a, double *fi,
                     const double *rhs, int itr_max, double tol_abs,
                     double tol_rel, double *hist, orc_solve_info *info);

/* DIC diagonal (linear_solvers.f90:439-445) and DILU diagonal (:613-624),
 * both stored inverted (d = 1/pivot). */
void orc_dic_diag(const orc_csr *c, const double *a, double *d);
void orc_dilu_diag(const orc_csr *c, const double *a, double *d);
/* One application z = M^{-1} r of the DIC/DILU preconditioner as the
 * reference's sweeps do it (:458-475). */
void orc_dic_apply(const orc_csr *c, const double *a, const double *d,
                   const double *r, double *z);

/* y = A x, row by row in CSR order (:493-498). */
void orc_spmv(const orc_csr *c, const double *a, const double *x, double *y);

/* Level sets of the strict-lower DAG: lev[i] = 0 if row i has no strict-lower
 * entry, else 1 + max lev[ja[k]] over ia[i] <= k < diag[i].  upper=1 uses the
 * strict-upper DAG processed from the last row.  Returns number of levels. */
int orc_levels(const orc_csr *c, int upper, int *lev);

/* Reverse Cuthill-McKee ordering of the matrix graph (rows adjacent through
 * off-diagonal entries; degree = off-diagonal count).  The reference has no
 * renumbering; this is the specification the GPU implementation follows:
 *  - each connected component starts from the unvisited vertex of minimum
 *    (degree, id), refined by at most pp_iters George-Liu steps: BFS from
 *    the current root r, take x = min (degree, id) of the last BFS level,
 *    and move to x when its eccentricity is larger than r's, else stop;
 *  - Cuthill-McKee from the root: a FIFO queue; every dequeued vertex
 *    appends its unvisited neighbours in (degree, id) order;
 *  - the final order is reversed: new_of_old[order[k]] = n-1-k.
 * Returns 0. */
int orc_rcm(const orc_csr *c, int pp_iters, int *new_of_old);

/* Convection-diffusion assembly as calculate_temperature composes it
 * (temperature.f90:75-366 with facefluxsc scalar_fluxes.f90:32-337 and
 * face_value interpolation.f90:28-157,318-356), steady, no ddt.
 * flmass[n_faces], dcoef[n_faces] (diffusivity per face), phi[numTotal],
 * grad[3*numTotal] (AoS), gam = deferred-correction blend, scheme 0 = cds,
 * 1 = linearUpwind.  bc_kind per boundary face: 0 = Dirichlet wall
 * (:206-210), 1 = zero-gradient wall (:224, copies phi_P to phi_b),
 * 2 = inlet/outlet/pressure flux boundary (:182-190).  urf: under-relaxation
 * (1 = none).  Outputs a[nnz] (zeroed first), su[n], sp[n] (zeroed first).
 * phi boundary entries of zero-gradient faces are updated in place. */
/* The assembly of calculate_temperature (temperature.f90:75-366): time terms
 * (time_scheme 0 steady, 1 BDF1, 2 BDF2, 3 Crank-Nicolson), facefluxsc over
 * the inner faces, the boundary loop by bc_kind (0 Dirichlet wall, 1
 * zero-gradient wall, 2 inlet/outlet/pressure flux face, 3 no term, 4
 * prescribed normal gradient, 5 periodic face: facefluxsc_periodic with the
 * twin face's owner, temperature.f90:262-313), the Crank-Nicolson coefficient block, then
 * the diagonal and under-relaxation.  srd_map (NULL: identity) gives, per
 * boundary face, the boundary face whose |S|/dn the reference reads through
 * its wall counter.  a, su, sp are overwritten.  Returns 0 or -1. */
int orc_transport_assemble(const orc_mesh *m, const orc_geom *g, const orc_csr *c,
                           const int *slot_pn, const int *slot_np, const double *flmass,
                           const double *dcoef, double *phi, const double *grad, double gam,
                           int scheme, const unsigned char *bc_kind, const int *srd_map,
                           double urf, int time_scheme, const double *den, double dt,
                           const double *phi_o, const double *phi_oo, double *a, double *su,
                           double *sp);
/* Face diffusion coefficients of calculate_temperature from the cell
 * viscosity (temperature.f90:125,178,203,237). */
void orc_transport_dcoef(const orc_mesh *m, const orc_geom *g, const double *vis,
                         const unsigned char *bc_kind, double viscos, double pranl, double sigt,
                         double *dcoef);
/* facefluxscalar's face diffusivity (scalar_fluxes.f90:384-392): inner face
 * game = diffcoef(P)*(1-lambda) + diffcoef(N)*lambda; a boundary face takes
 * diffcoef of its owner.  With it, orc_convdiff is facefluxscalar's
 * assembly: facefluxscalar is facefluxsc with dcoef = game. */
void orc_face_gamma_cells(const orc_mesh *m, const orc_geom *g, const double *diffcoef,
                          double *dcoef);
void orc_convdiff(const orc_mesh *m, const orc_geom *g, const orc_csr *c,
                  const int *slot_pn, const int *slot_np,
                  const double *flmass, const double *dcoef,
                  double *phi, const double *grad, double gam, int scheme,
                  const unsigned char *bc_kind, double urf,
                  double *a, double *su, double *sp);

/* Explicit non-orthogonal diffusion correction of facefluxsc
 * (scalar_fluxes.f90:99-109) accumulated as calculate_temperature does
 * (temperature.f90:153-154): su_P += s, su_N -= s, per inner face in face
 * order, gamma per face, grad AoS [3*numTotal]. su updated in place. */
void orc_nonorth_correction(const orc_mesh *m, const orc_geom *g, const double *gamma,
                            const double *grad, double *su);

/* y = A^T x by the row-order scatter of atx_cr (linear_solvers.f90:1079-1171):
 * y = 0, then for every row i in ascending order y[ja[k]] += a[k]*x[i]. */
void orc_atx(const orc_csr *c, const double *a, const double *x, double *y);

/* ILU(0) factors of ilu_cr (linear_solvers.f90:1267-1385) into l[nnz]: the
 * strict lower part holds the multipliers, the rest U; every pivot is
 * inverted at the end of its row and inverted back at the end (:1378,1382),
 * as the reference does.  Rows must be sorted by column.  Returns 0, or
 * -(i+1) when row i has no diagonal or a zero pivot (the reference stops). */
int orc_ilu0(const orc_csr *c, const double *a, double *l);

/* z = (LU)^{-1} r of lus_cr (:1386-1477, sweeps :1457-1470): unit-lower forward sweep, then
 * the backward sweep divided by the pivot.  z may alias r. */
void orc_ilu_apply(const orc_csr *c, const double *l, const double *r, double *z);

/* Restarted GMRES(mr) with the ILU(0) preconditioner, pmgmres_ilu
 * (:789-1077; csrsolve passes mr = 4 at :85).  itr_max counts restart
 * cycles.  hist (may be NULL, capacity itr_max*mr) receives resl after every
 * inner step.  info: resor = res0 = L1 norm of the first residual, resl =
 * the last estimate |g(k+1)|, iters = inner steps (itr_used).  Stopping
 * test as the reference writes it: resl/resor <= resor*tol_rel or
 * resl <= tol_abs.  One guard departs from the reference: a zero initial
 * residual returns at once (the reference divides by it and fills x with
 * NaN).  Returns the inner steps, or the orc_ilu0 error. */
int orc_pmgmres_ilu(const orc_csr *c, const double *a, double *x, const double *rhs,
                    int itr_max, int mr, double tol_abs, double tol_rel, double *hist,
                    orc_solve_info *info);

/* Reference constants (parameters.f90:6-7 widened from default REAL). */
double orc_small(void);
double orc_s0_init(void);

/* Pressure-correction step of SIMPLE, calcp_simple.f90:69-419, for an
 * incompressible case without pressure patches (pscheme 'linear', one
 * corrector).  orc_pcorr_assemble: facefluxmass2 on the inner faces
 * (faceflux_mass.f90:175-252), adjustMassFlow unless const_mflux (:833-912,
 * outlet values of u, v, w and outlet fluxes rescaled), inlet/outlet sources
 * and facefluxmass2_periodic (:313-372); a[nnz] and su[n] overwritten,
 * flmass[F] updated.  Returns -1 for a pressure patch.
 * orc_pcorr_correct: from the solved correction pp[numTotal], the mass-flux
 * correction (:333-376), ppref = pp[pref] (0-based), gradp_and_sources
 * (nablap.f90:19-209 with bpres) into su/sv/sw[n] and dpdxi's cell rows,
 * u,v,w += s*ap, p += urfp (pp - ppref), then updateVelocityAtBoundary
 * (velocity.f90:1184-1279). */
int orc_pcorr_assemble(const orc_mesh *m, const orc_geom *g, const orc_csr *c,
                       const int *slot_pn, const int *slot_np, double *u, double *v, double *w,
                       const double *pr, const double *dpdxi, const double *den,
                       const double *apu, const double *apv, const double *apw, double *flmass,
                       int const_mflux, double flomas, double *a, double *su);
void orc_pcorr_correct(const orc_mesh *m, const orc_geom *g, const int *slot_pn,
                       const double *a, double *pp, double *flmass, double *u, double *v,
                       double *w, double *pr, const double *apu, const double *apv,
                       const double *apw, double urfp, int pref, double *dpdxi, double *su,
                       double *sv, double *sw);

/* The assembly of calcuvw (velocity.f90:50-750) before its three solves,
 * steady / BDF1 / BDF2 / Crank-Nicolson (time_scheme 0/1/2/3), no
 * buoyancy, Lorentz force, PISO or constant mass flux, Gauss gradients:
 * updateVelocityAtBoundary, grad(u, v, w), gradp_and_sources(p), time terms,
 * facefluxuvw on the inner faces, the boundary loop (inlet/outlet/pressure
 * facefluxuvw_bnd, symmetry, periodic facefluxuvw_periodic, wall with
 * visw[wall face] in wall order), CN and the diagonal of every component
 * with its URF.  a[nnz] on entry holds the matrix's previous content (the
 * reference never clears it, and the u row sums read its diagonal); on
 * return the shared off-diagonals and w's diagonal.  rhs, diag3, ap3, sp3
 * [3n] (u, v, w blocks): the right-hand sides, URF-scaled diagonals, 1/a_P
 * and the diagonal sources; sv_out/sw_out the v and w sources before their
 * URF terms.  u, v, w, pr: boundary values updated; dpdxi, dudxi, dvdxi,
 * dwdxi [3*numTotal]. */
int orc_momentum(const orc_mesh *m, const orc_geom *g, const orc_csr *c, const int *slot_pn,
                 const int *slot_np, double *u, double *v, double *w, const double *uo,
                 const double *vo, const double *wo, const double *uoo, const double *voo,
                 const double *woo, double *pr, const double *den, const double *vis,
                 const double *visw, const double *flmass, double viscos, double dt,
                 int time_scheme, int scheme, double gds, const double *urf, double *a,
                 double *dpdxi, double *dudxi, double *dvdxi, double *dwdxi, double *rhs,
                 double *diag3, double *ap3, double *sp3, double *sv_out, double *sw_out);

#ifdef __cplusplus
}
#endif
#endif
