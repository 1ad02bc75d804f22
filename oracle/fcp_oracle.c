/*
 * fcp_oracle.c -- CPU restatement of the reference finite-volume path.
 *
 * TEST INFRASTRUCTURE ONLY (see fcp_oracle.h).  Built with
 * -O2 -ffp-contract=off and no fast-math so that every floating-point
 * operation happens in the order the reference Fortran performs it; the
 * reference (flang, x86-64 baseline ISA) has no FMA contraction either.
 *
 * Literal values: the reference writes several constants as default-REAL
 * literals that are widened to REAL(8) (parameters.f90:6 `small = 1e-20`,
 * linear_solvers.f90:276,447 `s0 = 1.e20`).  They are reproduced with the
 * float->double widening below.
 */
#include "fcp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SMALL ((double)1e-20f)  /* parameters.f90:6, gradients.f90:18 */
#define S0INIT ((double)1e20f)  /* linear_solvers.f90:276,447          */
static const double THIRD = 1.0 / 3.0; /* geometry.f90:151 third = 1./3._dp */

double orc_small(void) { return SMALL; }
double orc_s0_init(void) { return S0INIT; }

/* ------------------------------------------------------------------------ */
/* Geometry: geometry.f90:416-754                                           */
/* ------------------------------------------------------------------------ */
int orc_geometry(const orc_mesh *m, orc_geom *g) {
  const int n = m->n_cells, F = m->n_faces, Fi = m->n_inner;
  const double *P = m->points;
  double *r8tmp = (double *)calloc((size_t)n, sizeof(double));
  if (!r8tmp) return -1;
  /* F7: the reference relies on zero-filled fresh allocations. */
  memset(g->xc, 0, sizeof(double) * n);
  memset(g->yc, 0, sizeof(double) * n);
  memset(g->zc, 0, sizeof(double) * n);
  memset(g->vol, 0, sizeof(double) * n);
  memset(g->arx, 0, sizeof(double) * F);
  memset(g->ary, 0, sizeof(double) * F);
  memset(g->arz, 0, sizeof(double) * F);
  memset(g->xf, 0, sizeof(double) * F);
  memset(g->yf, 0, sizeof(double) * F);
  memset(g->zf, 0, sizeof(double) * F);

  for (int f = 0; f < F; ++f) {
    const int *nd = m->node + (size_t)f * ORC_NOMAX;
    const int inp = m->owner[f];
    double areasum = 0.0;
    const double x1 = P[3 * nd[0]], y1 = P[3 * nd[0] + 1], z1 = P[3 * nd[0] + 2];
    for (int i = 0; i < m->nnodes[f] - 2; ++i) {
      const double *pb = P + 3 * nd[i + 1];
      const double *pc = P + 3 * nd[i + 2];
      const double px = pb[0] - x1, py = pb[1] - y1, pz = pb[2] - z1;
      const double qx = pc[0] - x1, qy = pc[1] - y1, qz = pc[2] - z1;
      /* triangle_area_vector geometry.f90:1721-1749 */
      const double nx = 0.5 * (py * qz - pz * qy);
      const double ny = 0.5 * (pz * qx - px * qz);
      const double nz = 0.5 * (px * qy - py * qx);
      g->arx[f] += nx;
      g->ary[f] += ny;
      g->arz[f] += nz;
      const double cx = THIRD * ((pc[0] + pb[0]) + x1);
      const double cy = THIRD * ((pc[1] + pb[1]) + y1);
      const double cz = THIRD * ((pc[2] + pb[2]) + z1);
      const double are = sqrt((nx * nx + ny * ny) + nz * nz);
      g->xf[f] += are * cx;
      g->yf[f] += are * cy;
      g->zf[f] += are * cz;
      areasum += are;
      double riSi = (cx * nx + cy * ny) + cz * nz;
      g->vol[inp] += THIRD * riSi;
      g->xc[inp] += 0.75 * riSi * cx;
      g->yc[inp] += 0.75 * riSi * cy;
      g->zc[inp] += 0.75 * riSi * cz;
      r8tmp[inp] += riSi;
      if (f < Fi) {
        const int inn = m->neighbour[f];
        riSi = -((cx * nx + cy * ny) + cz * nz);
        g->vol[inn] += THIRD * riSi;
        g->xc[inn] += 0.75 * riSi * cx;
        g->yc[inn] += 0.75 * riSi * cy;
        g->zc[inn] += 0.75 * riSi * cz;
        r8tmp[inn] += riSi;
      }
    }
    g->xf[f] = g->xf[f] / areasum;
    g->yf[f] = g->yf[f] / areasum;
    g->zf[f] = g->zf[f] / areasum;
  }
  for (int c = 0; c < n; ++c) {
    g->xc[c] = g->xc[c] / r8tmp[c];
    g->yc[c] = g->yc[c] / r8tmp[c];
    g->zc[c] = g->zc[c] / r8tmp[c];
  }
  free(r8tmp);

  /* Interpolation factor, variant 2 (geometry.f90:46, 581-606). */
  for (int f = 0; f < Fi; ++f) {
    const int p = m->owner[f], q = m->neighbour[f];
    double dx = g->xf[f] - g->xc[p], dy = g->yf[f] - g->yc[p], dz = g->zf[f] - g->zc[p];
    const double djp = sqrt((dx * dx + dy * dy) + dz * dz);
    dx = g->xf[f] - g->xc[q];
    dy = g->yf[f] - g->yc[q];
    dz = g->zf[f] - g->zc[q];
    const double djn = sqrt((dx * dx + dy * dy) + dz * dz);
    g->facint[f] = djp / (djp + djn);
  }
  /* Df = Sf.Sf/(Sf.dpn) (geometry.f90:648-664). */
  for (int f = 0; f < Fi; ++f) {
    const int p = m->owner[f], q = m->neighbour[f];
    const double xpn = g->xc[q] - g->xc[p], ypn = g->yc[q] - g->yc[p], zpn = g->zc[q] - g->zc[p];
    const double are = (g->arx[f] * g->arx[f] + g->ary[f] * g->ary[f]) + g->arz[f] * g->arz[f];
    g->Df[f] = are / ((g->arx[f] * xpn + g->ary[f] * ypn) + g->arz[f] * zpn);
  }
  /* Boundary normal distance and |S|/dn (geometry.f90:701-754), for every
   * boundary face (the reference keeps them for wall/symmetry only). */
  for (int f = Fi; f < F; ++f) {
    const int p = m->owner[f];
    const double are = sqrt((g->arx[f] * g->arx[f] + g->ary[f] * g->ary[f]) + g->arz[f] * g->arz[f]);
    const double nxf = g->arx[f] / are, nyf = g->ary[f] / are, nzf = g->arz[f] / are;
    const double dn = ((g->xf[f] - g->xc[p]) * nxf + (g->yf[f] - g->yc[p]) * nyf) +
                      (g->zf[f] - g->zc[p]) * nzf;
    g->dnb[f - Fi] = dn;
    g->srdb[f - Fi] = are / dn;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* CSR: sparse_matrix.f90:86-296, utils.f90:96-189                          */
/* ------------------------------------------------------------------------ */
int orc_n_periodic(const orc_mesh *m) {
  int np = 0;
  for (int ib = 0; ib < m->n_patches; ++ib)
    if (m->bctype[ib] == 7) np += m->nfaces[ib];
  return np;
}

/* Periodic pair k (patch order): the owner of periodic face i and of twin
 * face i (sparse_matrix.f90:141-171); returns the number of pairs. */
static int periodic_pairs(const orc_mesh *m, int *pp, int *pn) {
  int k = 0, iper = 0;
  for (int ib = 0; ib < m->n_patches; ++ib) {
    if (m->bctype[ib] != 7) continue;
    const int t0 = m->twin_start[iper++];
    for (int i = 0; i < m->nfaces[ib]; ++i, ++k) {
      if (pp) pp[k] = m->owner[m->start_face[ib] + i];
      if (pn) pn[k] = m->owner[t0 + i];
    }
  }
  return k;
}

/* Boundary face f on a periodic patch: its index in the patch, the twin's
 * start face and its periodic ordinal (patch order); 0 when f is elsewhere. */
static int periodic_face(const orc_mesh *m, int f, int *i, int *t0, int *k) {
  int base = 0, iper = 0;
  for (int ib = 0; ib < m->n_patches; ++ib) {
    if (m->bctype[ib] != 7) continue;
    const int s0 = m->start_face[ib];
    if (f >= s0 && f < s0 + m->nfaces[ib]) {
      *i = f - s0;
      *t0 = m->twin_start[iper];
      *k = base + (f - s0);
      return 1;
    }
    base += m->nfaces[ib];
    ++iper;
  }
  return 0;
}

void orc_face_mapping(const orc_mesh *m, orc_geom *g, int *owner) {
  memcpy(owner, m->owner, sizeof(int) * (size_t)m->n_faces);
  int iper = 0;
  for (int ib = 0; ib < m->n_patches; ++ib) {
    if (m->bctype[ib] != 7) continue;
    const int nf = m->nfaces[ib], s0 = m->start_face[ib], t0 = m->twin_start[iper++];
    int *tmp = (int *)malloc(sizeof(int) * (size_t)(nf > 0 ? nf : 1));
    int *itmp = (int *)malloc(sizeof(int) * (size_t)(nf > 0 ? nf : 1));
    double *v = (double *)malloc(sizeof(double) * (size_t)(nf > 0 ? nf : 1));
    for (int i = 0; i < nf; ++i) { /* geometry.f90:1910-1935 */
      const int iface = s0 + i;
      double mindist = (double)1e30f; /* mindist = 1e+30, a default-real literal */
      tmp[i] = owner[t0 + i];
      itmp[i] = i;
      for (int it = 0; it < nf; ++it) {
        const int ijn = owner[t0 + it];
        const double dist = (fabs(g->xf[iface] - g->xc[ijn]) + fabs(g->yf[iface] - g->yc[ijn])) +
                            fabs(g->zf[iface] - g->zc[ijn]);
        if (dist < mindist) {
          mindist = dist;
          tmp[i] = ijn;
          itmp[i] = it;
        }
      }
    }
    for (int i = 0; i < nf; ++i) owner[t0 + i] = tmp[i]; /* :1939-1941 */
    double *arrs[6] = {g->xf, g->yf, g->zf, g->arx, g->ary, g->arz}; /* :1943-1985 */
    for (int a = 0; a < 6; ++a) {
      for (int i = 0; i < nf; ++i) v[i] = arrs[a][t0 + itmp[i]];
      for (int i = 0; i < nf; ++i) arrs[a][t0 + i] = v[i];
    }
    free(tmp);
    free(itmp);
    free(v);
  }
}

int orc_csr_nnz(const orc_mesh *m) {
  return 2 * (m->n_inner + orc_n_periodic(m)) + m->n_cells; /* sparse_matrix.f90:110 */
}

int orc_csr_build(const orc_mesh *m, int *ia, int *ja, int *diag, int *slot_pn,
                  int *slot_np) {
  const int n = m->n_cells, Fi = m->n_inner, np = orc_n_periodic(m);
  /* The reference sorts the COO list (diagonal + (P,N) + (N,P) + the
   * periodic pairs both ways) lexically (utils.f90:443) -- the sorted
   * multiset is unique, so per-row counting followed by an in-row sort
   * yields the same arrays. */
  int *cnt = (int *)calloc((size_t)n + 1, sizeof(int));
  int *pp = (int *)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
  int *pn = (int *)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
  if (!cnt || !pp || !pn) return -1;
  periodic_pairs(m, pp, pn);
  for (int i = 0; i < n; ++i) cnt[i] = 1;
  for (int f = 0; f < Fi; ++f) {
    cnt[m->owner[f]]++;
    cnt[m->neighbour[f]]++;
  }
  for (int k = 0; k < np; ++k) {
    cnt[pp[k]]++;
    cnt[pn[k]]++;
  }
  ia[0] = 0;
  for (int i = 0; i < n; ++i) ia[i + 1] = ia[i] + cnt[i];
  for (int i = 0; i < n; ++i) {
    ja[ia[i]] = i;
    cnt[i] = ia[i] + 1;
  }
  for (int f = 0; f < Fi; ++f) {
    const int p = m->owner[f], q = m->neighbour[f];
    ja[cnt[p]++] = q;
    ja[cnt[q]++] = p;
  }
  for (int k = 0; k < np; ++k) {
    ja[cnt[pp[k]]++] = pn[k];
    ja[cnt[pn[k]]++] = pp[k];
  }
  for (int i = 0; i < n; ++i) { /* insertion sort each row */
    for (int k = ia[i] + 1; k < ia[i + 1]; ++k) {
      const int v = ja[k];
      int l = k - 1;
      while (l >= ia[i] && ja[l] > v) {
        ja[l + 1] = ja[l];
        --l;
      }
      ja[l + 1] = v;
    }
  }
  for (int i = 0; i < n; ++i) /* a repeated entry: not supported */
    for (int k = ia[i] + 1; k < ia[i + 1]; ++k)
      if (ja[k] == ja[k - 1]) {
        free(cnt);
        free(pp);
        free(pn);
        return -2;
      }
  /* find_main_diag_element_positions utils.f90:149-168: the id-th entry with
   * row==col in sorted order is the diagonal of row id. */
  for (int i = 0; i < n; ++i) {
    for (int k = ia[i]; k < ia[i + 1]; ++k)
      if (ja[k] == i) {
        diag[i] = k;
        break;
      }
  }
  /* csr_to_k utils.f90:96-146: first match in the row. */
  for (int f = 0; f < Fi; ++f) {
    const int p = m->owner[f], q = m->neighbour[f];
    int kpn = -1, knp = -1;
    for (int k = ia[p]; k < ia[p + 1]; ++k)
      if (ja[k] == q) { kpn = k; break; }
    for (int k = ia[q]; k < ia[q + 1]; ++k)
      if (ja[k] == p) { knp = k; break; }
    slot_pn[f] = kpn;
    slot_np[f] = knp;
  }
  for (int k = 0; k < np; ++k) { /* sparse_matrix.f90:262-293 */
    const int p = pp[k], q = pn[k];
    int kpn = -1, knp = -1;
    for (int j = ia[p]; j < ia[p + 1]; ++j)
      if (ja[j] == q) { kpn = j; break; }
    for (int j = ia[q]; j < ia[q + 1]; ++j)
      if (ja[j] == p) { knp = j; break; }
    slot_pn[Fi + k] = kpn;
    slot_np[Fi + k] = knp;
  }
  free(cnt);
  free(pp);
  free(pn);
  return ia[n];
}

/* ------------------------------------------------------------------------ */
/* Laplacian: laplacian.f90:1-151                                           */
/* ------------------------------------------------------------------------ */
void orc_laplacian(const orc_mesh *m, const orc_geom *g, const orc_csr *c,
                   const int *slot_pn, const int *slot_np, const double *mu,
                   const double *phi, double *a, double *su) {
  const int n = m->n_cells, Fi = m->n_inner;
  memset(a, 0, sizeof(double) * c->nnz);
  for (int f = 0; f < Fi; ++f) {
    const int p = m->owner[f], q = m->neighbour[f];
    const double fxn = g->facint[f], fxp = 1.0 - g->facint[f];
    const double xpn = g->xc[q] - g->xc[p], ypn = g->yc[q] - g->yc[p], zpn = g->zc[q] - g->zc[p];
    const double ax = g->arx[f], ay = g->ary[f], az = g->arz[f];
    const double smdpn = ((ax * ax + ay * ay) + az * az) / ((ax * xpn + ay * ypn) + az * zpn);
    const double cap = (fxp * mu[p] + fxn * mu[q]) * smdpn;
    const double can = cap;
    a[slot_pn[f]] = can;
    a[slot_np[f]] = cap;
    a[c->diag[p]] = a[c->diag[p]] - can;
    a[c->diag[q]] = a[c->diag[q]] - cap;
  }
  for (int ib = 0; ib < m->n_patches; ++ib) {
    for (int i = 0; i < m->nfaces[ib]; ++i) {
      const int f = m->start_face[ib] + i;
      const int p = m->owner[f];
      const int ijb = n + (m->start_face[ib] - Fi) + i;
      const double ax = g->arx[f], ay = g->ary[f], az = g->arz[f];
      const double are = sqrt((ax * ax + ay * ay) + az * az);
      const double nxf = ax / are, nyf = ay / are, nzf = az / are;
      const double dfn = ((g->xf[f] - g->xc[p]) * nxf + (g->yf[f] - g->yc[p]) * nyf) +
                         (g->zf[f] - g->zc[p]) * nzf;
      const double dcoef = mu[p] * are / dfn;
      a[c->diag[p]] = a[c->diag[p]] - dcoef;
      su[p] = su[p] - dcoef * phi[ijb];
    }
  }
}

void orc_laplacian_face(const orc_mesh *m, const orc_geom *g, const orc_csr *c,
                        const int *slot_pn, const int *slot_np, const double *mu_f,
                        const unsigned char *bc_kind, const double *phi, double *a,
                        double *su) {
  const int n = m->n_cells, Fi = m->n_inner;
  memset(a, 0, sizeof(double) * c->nnz);
  for (int f = 0; f < Fi; ++f) {
    const int p = m->owner[f], q = m->neighbour[f];
    const double cap = mu_f[f] * g->Df[f];
    a[slot_pn[f]] = cap;
    a[slot_np[f]] = cap;
    a[c->diag[p]] = a[c->diag[p]] - cap;
    a[c->diag[q]] = a[c->diag[q]] - cap;
  }
  for (int f = Fi; f < m->n_faces; ++f) {
    if (bc_kind[f - Fi] != 0) continue;
    const int p = m->owner[f];
    const double ax = g->arx[f], ay = g->ary[f], az = g->arz[f];
    const double are = sqrt((ax * ax + ay * ay) + az * az);
    const double nxf = ax / are, nyf = ay / are, nzf = az / are;
    const double dfn = ((g->xf[f] - g->xc[p]) * nxf + (g->yf[f] - g->yc[p]) * nyf) +
                       (g->zf[f] - g->zc[p]) * nzf;
    const double dcoef = mu_f[f] * are / dfn;
    a[c->diag[p]] = a[c->diag[p]] - dcoef;
    su[p] = su[p] - dcoef * phi[n + (f - Fi)];
  }
}

/* ------------------------------------------------------------------------ */
/* Gauss gradient: gradients.f90:1607-1715                                  */
/* ------------------------------------------------------------------------ */
void orc_grad_gauss(const orc_mesh *m, const orc_geom *g, const double *u,
                    double *dudxi) {
  const int n = m->n_cells, Fi = m->n_inner, F = m->n_faces;
  memset(dudxi, 0, sizeof(double) * 3 * (size_t)(n + F - Fi));
  for (int f = 0; f < Fi; ++f) {
    const int p = m->owner[f], q = m->neighbour[f];
    const double fie = u[p] + (u[q] - u[p]) * g->facint[f];
    const double dfxe = fie * g->arx[f], dfye = fie * g->ary[f], dfze = fie * g->arz[f];
    dudxi[3 * p] += dfxe;
    dudxi[3 * p + 1] += dfye;
    dudxi[3 * p + 2] += dfze;
    dudxi[3 * q] -= dfxe;
    dudxi[3 * q + 1] -= dfye;
    dudxi[3 * q + 2] -= dfze;
  }
  for (int f = Fi; f < F; ++f) {
    const int p = m->owner[f];
    const double fi = u[n + (f - Fi)];
    dudxi[3 * p] = dudxi[3 * p] + fi * g->arx[f];
    dudxi[3 * p + 1] = dudxi[3 * p + 1] + fi * g->ary[f];
    dudxi[3 * p + 2] = dudxi[3 * p + 2] + fi * g->arz[f];
  }
  for (int c = 0; c < n; ++c) {
    const double volr = 1.0 / g->vol[c];
    dudxi[3 * c] *= volr;
    dudxi[3 * c + 1] *= volr;
    dudxi[3 * c + 2] *= volr;
  }
}

/* One Gauss sweep of fvxGradient.f90 with face values corrected by the old
 * gradient dfo[3*n] (gradco :1761-1817), boundary faces by gradbc
 * (:1819-1838), then the 1/vol scaling (:1623-1631 / :1746-1754). */
static void fvx_sweep(const orc_mesh *m, const orc_geom *g, const double *u,
                      const double *dfo, double *out) {
  const int n = m->n_cells, Fi = m->n_inner, F = m->n_faces;
  memset(out, 0, sizeof(double) * 3 * (size_t)(n + F - Fi));
  for (int f = 0; f < Fi; ++f) {
    const int p = m->owner[f], q = m->neighbour[f];
    const double fxn = g->facint[f];
    const double fxp = 1.0 - fxn;
    const double xi = g->xc[p] * fxp + g->xc[q] * fxn;
    const double yi = g->yc[p] * fxp + g->yc[q] * fxn;
    const double zi = g->zc[p] * fxp + g->zc[q] * fxn;
    const double dfxi = dfo[3 * p] * fxp + dfo[3 * q] * fxn;
    const double dfyi = dfo[3 * p + 1] * fxp + dfo[3 * q + 1] * fxn;
    const double dfzi = dfo[3 * p + 2] * fxp + dfo[3 * q + 2] * fxn;
    const double fie = (((u[p] * fxp + u[q] * fxn) + dfxi * (g->xf[f] - xi)) +
                        dfyi * (g->yf[f] - yi)) + dfzi * (g->zf[f] - zi);
    const double dfxe = fie * g->arx[f], dfye = fie * g->ary[f], dfze = fie * g->arz[f];
    out[3 * p] = out[3 * p] + dfxe;
    out[3 * p + 1] = out[3 * p + 1] + dfye;
    out[3 * p + 2] = out[3 * p + 2] + dfze;
    out[3 * q] = out[3 * q] - dfxe;
    out[3 * q + 1] = out[3 * q + 1] - dfye;
    out[3 * q + 2] = out[3 * q + 2] - dfze;
  }
  for (int f = Fi; f < F; ++f) {
    const int p = m->owner[f];
    const double fi = u[n + (f - Fi)];
    out[3 * p] = out[3 * p] + fi * g->arx[f];
    out[3 * p + 1] = out[3 * p + 1] + fi * g->ary[f];
    out[3 * p + 2] = out[3 * p + 2] + fi * g->arz[f];
  }
  for (int c = 0; c < n; ++c) {
    const double volr = 1.0 / g->vol[c];
    out[3 * c] = out[3 * c] * volr;
    out[3 * c + 1] = out[3 * c + 1] * volr;
    out[3 * c + 2] = out[3 * c + 2] * volr;
  }
}

int orc_grad_gauss_fvx(const orc_mesh *m, const orc_geom *g, const double *u,
                       const double *gold, double *dudxi) {
  const int n = m->n_cells;
  double *dfo = (double *)calloc(3 * (size_t)n + 3, sizeof(double));
  if (!dfo) return -1;
  if (gold) {
    /* grad_gauss_corrected :1700-1703: old gradient = the given one */
    memcpy(dfo, gold, sizeof(double) * 3 * (size_t)n);
    fvx_sweep(m, g, u, dfo, dudxi);
  } else {
    /* grad_gauss :1593-1641: lc = 1 from a zero old gradient, lc = 2 from
     * the first sweep's result */
    fvx_sweep(m, g, u, dfo, dudxi);
    memcpy(dfo, dudxi, sizeof(double) * 3 * (size_t)n);
    fvx_sweep(m, g, u, dfo, dudxi);
  }
  free(dfo);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Least-squares gradients                                                  */
/* ------------------------------------------------------------------------ */
static void lsq_invert_sym(double *D) {
  /* gradients.f90:745-777 (and :1296-1326): closed-form adjugate with +small */
  const double d11 = D[0], d12 = D[1], d13 = D[2], d22 = D[3], d23 = D[4], d33 = D[5];
  const double d21 = d12, d31 = d13, d32 = d23;
  const double tmp = 1.0 / ((((((d11 * d22 * d33 - d11 * d23 * d32) - d12 * d21 * d33) +
                                d12 * d23 * d31) + d13 * d21 * d32) - d13 * d22 * d31) + SMALL);
  D[0] = (d22 * d33 - d23 * d32) * tmp;
  D[1] = (d21 * d33 - d23 * d31) * tmp;
  D[2] = (d21 * d32 - d22 * d31) * tmp;
  D[3] = (d11 * d33 - d13 * d31) * tmp;
  D[4] = (d12 * d33 - d13 * d32) * tmp;
  D[5] = (d11 * d32 - d12 * d31) * tmp;
  D[6] = (d12 * d23 - d13 * d22) * tmp;
  D[7] = (d11 * d23 - d13 * d21) * tmp;
  D[8] = (d11 * d22 - d12 * d21) * tmp;
}

static void lsq_accum(double *D, double w, double dx, double dy, double dz, int weighted) {
  /* Dmat(1..6) += [w*]dx*dx, dx*dy, dx*dz, dy*dy, dy*dz, dz*dz in the
   * reference's storage slots 1,2,3,4,5,6 = xx,xy,xz,yy,yz,zz. */
  if (weighted) {
    D[0] += w * dx * dx;
    D[3] += w * dy * dy;
    D[5] += w * dz * dz;
    D[1] += w * dx * dy;
    D[2] += w * dx * dz;
    D[4] += w * dy * dz;
  } else {
    D[0] += dx * dx;
    D[3] += dy * dy;
    D[5] += dz * dz;
    D[1] += dx * dy;
    D[2] += dx * dz;
    D[4] += dy * dz;
  }
}

/* Thin QR of an l x 3 matrix by modified Gram-Schmidt (matrix.f90:366-419),
 * returning R^{-1} Q^T (3 x l, column-major per neighbour) in out[3*6]. */
static int mgs_rinv_qt(const double *A /* l x 3, row-major */, int l, double *out) {
  double q[6][3], r[3][3] = {{0}};
  for (int i = 0; i < l; ++i)
    for (int j = 0; j < 3; ++j) q[i][j] = A[3 * i + j];
  for (int j = 0; j < 3; ++j) {
    double z = 0.0;
    for (int i = 0; i < l; ++i) z = z + q[i][j] * q[i][j];
    r[j][j] = sqrt(z);
    for (int i = 0; i < l; ++i) q[i][j] = q[i][j] / r[j][j];
    for (int k = j + 1; k < 3; ++k) {
      z = 0.0;
      for (int i = 0; i < l; ++i) z = z + q[i][j] * q[i][k];
      r[j][k] = z;
      for (int i = 0; i < l; ++i) q[i][k] = q[i][k] - r[j][k] * q[i][j];
    }
  }
  /* R^{-1} for upper-triangular R */
  double ri[3][3] = {{0}};
  ri[0][0] = 1.0 / r[0][0];
  ri[1][1] = 1.0 / r[1][1];
  ri[2][2] = 1.0 / r[2][2];
  ri[0][1] = -r[0][1] / (r[0][0] * r[1][1]);
  ri[1][2] = -r[1][2] / (r[1][1] * r[2][2]);
  ri[0][2] = (r[0][1] * r[1][2] - r[0][2] * r[1][1]) / (r[0][0] * r[1][1] * r[2][2]);
  memset(out, 0, sizeof(double) * 18);
  for (int k = 0; k < l; ++k)
    for (int i = 0; i < 3; ++i) {
      double s = 0.0;
      for (int j = 0; j < 3; ++j) s += ri[i][j] * q[k][j];
      out[3 * k + i] = s;
    }
  return 0;
}

int orc_lsq_matrix(const orc_mesh *m, const orc_geom *g, int variant, int compat,
                   double *dmat) {
  const int n = m->n_cells, Fi = m->n_inner, F = m->n_faces;
  (void)compat;
  if (variant == 0 || variant == 1) {
    const int weighted = (variant == 1);
    memset(dmat, 0, sizeof(double) * 9 * (size_t)n);
    for (int f = 0; f < Fi; ++f) {
      const int p = m->owner[f], q = m->neighbour[f];
      const double dx = g->xc[q] - g->xc[p], dy = g->yc[q] - g->yc[p], dz = g->zc[q] - g->zc[p];
      double w = 0.0;
      if (weighted) w = 1.0 / ((dx * dx + dy * dy) + dz * dz);
      /* reference interleaves owner/neighbour per component; each Dmat
       * entry still receives its face contributions in face order. */
      lsq_accum(dmat + 9 * (size_t)p, w, dx, dy, dz, weighted);
      lsq_accum(dmat + 9 * (size_t)q, w, dx, dy, dz, weighted);
    }
    for (int f = Fi; f < F; ++f) {
      const int p = m->owner[f];
      const double dx = g->xf[f] - g->xc[p], dy = g->yf[f] - g->yc[p], dz = g->zf[f] - g->zc[p];
      double w = 0.0;
      if (weighted) w = 1.0 / ((dx * dx + dy * dy) + dz * dz);
      lsq_accum(dmat + 9 * (size_t)p, w, dx, dy, dz, weighted);
    }
    for (int c = 0; c < n; ++c) lsq_invert_sym(dmat + 9 * (size_t)c);
    return 0;
  }
  if (variant == 2) {
    /* neighbour slots in face-visit order (gradients.f90:966-990) */
    double *A = (double *)calloc((size_t)n * 18, sizeof(double));
    int *cnt = (int *)calloc((size_t)n, sizeof(int));
    if (!A || !cnt) return -1;
    for (int f = 0; f < Fi; ++f) {
      const int p = m->owner[f], q = m->neighbour[f];
      if (cnt[p] >= 6 || cnt[q] >= 6) { free(A); free(cnt); return -3; }
      double *ap = A + 18 * (size_t)p + 3 * cnt[p]++;
      ap[0] = g->xc[q] - g->xc[p]; ap[1] = g->yc[q] - g->yc[p]; ap[2] = g->zc[q] - g->zc[p];
      double *aq = A + 18 * (size_t)q + 3 * cnt[q]++;
      aq[0] = g->xc[p] - g->xc[q]; aq[1] = g->yc[p] - g->yc[q]; aq[2] = g->zc[p] - g->zc[q];
    }
    for (int f = Fi; f < F; ++f) {
      const int p = m->owner[f];
      if (cnt[p] >= 6) { free(A); free(cnt); return -3; }
      double *ap = A + 18 * (size_t)p + 3 * cnt[p]++;
      ap[0] = g->xf[f] - g->xc[p]; ap[1] = g->yf[f] - g->yc[p]; ap[2] = g->zf[f] - g->zc[p];
    }
    for (int c = 0; c < n; ++c) mgs_rinv_qt(A + 18 * (size_t)c, cnt[c], dmat + 18 * (size_t)c);
    free(A);
    free(cnt);
    return 0;
  }
  return -4;
}

int orc_grad_lsq(const orc_mesh *m, const orc_geom *g, int variant, int compat,
                 const double *dmat, const double *phi, double *dphidxi) {
  const int n = m->n_cells, Fi = m->n_inner, F = m->n_faces;
  const size_t ntot = (size_t)n + (size_t)(F - Fi);
  memset(dphidxi, 0, sizeof(double) * 3 * ntot);
  if (variant == 0 || variant == 1) {
    for (int f = 0; f < Fi; ++f) {
      const int p = m->owner[f], q = m->neighbour[f];
      double Dx, Dy, Dz;
      if (variant == 0) {
        const double delta = phi[q] - phi[p];
        Dx = (g->xc[q] - g->xc[p]) * delta;
        Dy = (g->yc[q] - g->yc[p]) * delta;
        Dz = (g->zc[q] - g->zc[p]) * delta;
      } else {
        const double dx = g->xc[q] - g->xc[p], dy = g->yc[q] - g->yc[p], dz = g->zc[q] - g->zc[p];
        const double w = (phi[q] - phi[p]) / ((dx * dx + dy * dy) + dz * dz);
        Dx = w * dx;
        Dy = w * dy;
        Dz = w * dz;
      }
      dphidxi[3 * p] += Dx; dphidxi[3 * p + 1] += Dy; dphidxi[3 * p + 2] += Dz;
      dphidxi[3 * q] += Dx; dphidxi[3 * q + 1] += Dy; dphidxi[3 * q + 2] += Dz;
    }
    for (int f = Fi; f < F; ++f) {
      const int p = m->owner[f];
      const int ib = f - Fi;
      const double delta = phi[n + ib] - phi[p];
      double Dx, Dy, Dz;
      if (variant == 0) {
        Dx = (g->xf[f] - g->xc[p]) * delta;
        Dy = (g->yf[f] - g->yc[p]) * delta;
        Dz = (g->zf[f] - g->zc[p]) * delta;
      } else {
        /* gradients.f90:1459 weights with xf(i) (i = boundary counter) */
        const int fw = compat ? ib : f;
        const double wx = g->xf[fw] - g->xc[p], wy = g->yf[fw] - g->yc[p], wz = g->zf[fw] - g->zc[p];
        const double w = delta / ((wx * wx + wy * wy) + wz * wz);
        Dx = w * (g->xf[f] - g->xc[p]);
        Dy = w * (g->yf[f] - g->yc[p]);
        Dz = w * (g->zf[f] - g->zc[p]);
      }
      dphidxi[3 * p] += Dx; dphidxi[3 * p + 1] += Dy; dphidxi[3 * p + 2] += Dz;
    }
    for (int c = 0; c < n; ++c) {
      const double *D = dmat + 9 * (size_t)c;
      const double b1 = dphidxi[3 * c], b2 = dphidxi[3 * c + 1], b3 = dphidxi[3 * c + 2];
      dphidxi[3 * c] = (b1 * D[0] - b2 * D[1]) + b3 * D[2];
      if (compat)
        dphidxi[3 * c + 1] = (b1 * D[3] - b2 * D[4]) - b3 * D[5];
      else
        dphidxi[3 * c + 1] = (-(b1 * D[4]) + b2 * D[3]) - b3 * D[5];
      dphidxi[3 * c + 2] = (b1 * D[6] - b2 * D[7]) + b3 * D[8];
    }
    return 0;
  }
  if (variant == 2) {
    double *b = (double *)calloc((size_t)n * 6, sizeof(double));
    int *cnt = (int *)calloc((size_t)n, sizeof(int));
    if (!b || !cnt) return -1;
    for (int f = 0; f < Fi; ++f) {
      const int p = m->owner[f], q = m->neighbour[f];
      b[6 * (size_t)p + cnt[p]++] = phi[q] - phi[p];
      b[6 * (size_t)q + cnt[q]++] = phi[p] - phi[q];
    }
    for (int f = Fi; f < F; ++f) {
      const int p = m->owner[f];
      b[6 * (size_t)p + cnt[p]++] = phi[n + (f - Fi)] - phi[p];
    }
    for (int c = 0; c < n; ++c) {
      const double *D = dmat + 18 * (size_t)c;
      for (int i = 0; i < 3; ++i) {
        double s = 0.0; /* sum(D(i,1:l)*b(1:l)) */
        for (int k = 0; k < cnt[c]; ++k) s = s + D[3 * k + i] * b[6 * (size_t)c + k];
        dphidxi[3 * c + i] = s;
      }
    }
    free(b);
    free(cnt);
    return 0;
  }
  return -4;
}

void orc_update_boundary(const orc_mesh *m, double *phi) {
  const int n = m->n_cells, Fi = m->n_inner;
  int iper = 0;
  for (int ib = 0; ib < m->n_patches; ++ib) {
    const int t = m->bctype[ib];
    if (t == 4 || t == 2 || t == 5 || t == 6) {
      for (int i = 0; i < m->nfaces[ib]; ++i) {
        const int f = m->start_face[ib] + i;
        phi[n + (f - Fi)] = phi[m->owner[f]];
      }
    } else if (t == 7) { /* updateBoundary.f90:46-68 */
      const int t0 = m->twin_start[iper++];
      for (int i = 0; i < m->nfaces[ib]; ++i) {
        const int f = m->start_face[ib] + i;
        const int ijb = n + (f - Fi);
        phi[ijb] = 0.5 * (phi[m->owner[f]] + phi[m->owner[t0 + i]]);
        phi[n + (t0 - Fi) + i] = phi[ijb]; /* the twin takes the same value */
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Linear algebra helpers                                                   */
/* ------------------------------------------------------------------------ */
void orc_spmv(const orc_csr *c, const double *a, const double *x, double *y) {
  for (int i = 0; i < c->n; ++i) {
    double s = 0.0;
    for (int k = c->ia[i]; k < c->ia[i + 1]; ++k) s = s + a[k] * x[c->ja[k]];
    y[i] = s;
  }
}

static void residual(const orc_csr *c, const double *a, const double *fi,
                     const double *rhs, double *res) {
  for (int i = 0; i < c->n; ++i) {
    double r = rhs[i];
    for (int k = c->ia[i]; k < c->ia[i + 1]; ++k) r = r - a[k] * fi[c->ja[k]];
    res[i] = r;
  }
}

static double l1(const double *v, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s = s + fabs(v[i]);
  return s;
}

static double dot(const double *u, const double *v, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s = s + u[i] * v[i];
  return s;
}

static double factor_of(const orc_csr *c, const double *a, const double *fi) {
  double s = 0.0;
  for (int i = 0; i < c->n; ++i) s = s + fabs(a[c->diag[i]] * fi[i]);
  return s + SMALL;
}

void orc_dic_diag(const orc_csr *c, const double *a, double *d) {
  for (int i = 0; i < c->n; ++i) {
    double di = a[c->diag[i]];
    for (int k = c->ia[i]; k < c->diag[i]; ++k) di = di - a[k] * a[k] * d[c->ja[k]];
    d[i] = 1.0 / di;
  }
}

void orc_dilu_diag(const orc_csr *c, const double *a, double *d) {
  const int n = c->n;
  for (int i = 0; i < n; ++i) {
    double di = a[c->diag[i]];
    for (int k = c->ia[i]; k < c->diag[i]; ++k) {
      const int j = c->ja[k];
      int l;
      for (l = c->diag[j]; l < c->ia[j + 1]; ++l)
        if (c->ja[l] == i) break;
      /* a miss leaves l one past the row (linear_solvers.f90:616-620) */
      di = di - a[k] * d[j] * (l < c->nnz ? a[l] : 0.0);
    }
    d[i] = 1.0 / di;
  }
}

void orc_dic_apply(const orc_csr *c, const double *a, const double *d,
                   const double *r, double *z) {
  const int n = c->n;
  for (int i = 0; i < n; ++i) {
    double zi = r[i];
    for (int k = c->ia[i]; k < c->diag[i]; ++k) zi = zi - a[k] * z[c->ja[k]];
    z[i] = zi * d[i];
  }
  for (int i = 0; i < n; ++i) z[i] = z[i] / (d[i] + SMALL);
  for (int i = n - 1; i >= 0; --i) {
    double zi = z[i];
    for (int k = c->diag[i] + 1; k < c->ia[i + 1]; ++k) zi = zi - a[k] * z[c->ja[k]];
    z[i] = zi * d[i];
  }
}

/* ------------------------------------------------------------------------ */
/* Solvers: linear_solvers.f90                                              */
/* ------------------------------------------------------------------------ */
/* crit 0: the reference's L1 test rsm < tol_rel (:528-532); crit 1: the
 * recursively updated residual's L2 norm against tol_rel * ||r0||_2 (the
 * stopping test of the mixed-precision GPU solves, used only for the CPU
 * baseline timing). */
static int iccg_impl(const orc_csr *c, const double *a, double *fi, const double *rhs,
                     int itr_max, double tol_abs, double tol_rel, double *hist,
                     orc_solve_info *info, int crit) {
  const int n = c->n;
  double *res = (double *)calloc((size_t)n, sizeof(double));
  double *pk = (double *)calloc((size_t)n, sizeof(double));
  double *zk = (double *)calloc((size_t)n, sizeof(double));
  double *d = (double *)calloc((size_t)n, sizeof(double));
  double factor = 0.0, resor = 0.0, resl = 0.0;
  int itr_used = 0;
  residual(c, a, fi, rhs, res);                 /* :417-422 */
  const double res0 = l1(res, n);               /* :425     */
  const double l2r0 = sqrt(dot(res, res, n));
  if (res0 < tol_abs) {                         /* :428-432 */
    resor = res0;
    resl = res0;
    goto done;
  }
  orc_dic_diag(c, a, d);                        /* :439-445 */
  double s0 = S0INIT;                           /* :447     */
  for (int l = 1; l <= itr_max; ++l) {
    orc_dic_apply(c, a, d, res, zk);            /* :458-475 */
    const double sk = dot(res, zk, n);          /* :478     */
    const double bet = sk / s0;
    for (int i = 0; i < n; ++i) pk[i] = zk[i] + bet * pk[i];
    orc_spmv(c, a, pk, zk);                     /* :493-498 */
    const double pkapk = dot(pk, zk, n);
    const double alf = sk / pkapk;
    for (int i = 0; i < n; ++i) fi[i] = fi[i] + alf * pk[i];
    for (int i = 0; i < n; ++i) res[i] = res[i] - alf * zk[i];
    resl = l1(res, n);
    s0 = sk;
    itr_used++;
    if (hist) hist[l - 1] = resl;
    if (l == 1) {
      factor = factor_of(c, a, fi);
      resor = res0 / factor;
    }
    const double rsm = resl / (res0 + SMALL);
    if (crit == 0 && (rsm < tol_rel || resl < tol_abs)) break;
    if (crit == 1 && (sqrt(dot(res, res, n)) <= tol_rel * l2r0 || resl < tol_abs)) break;
  }
done:
  if (info) {
    info->resor = resor; info->res0 = res0; info->resl = resl;
    info->factor = factor; info->iters = itr_used;
  }
  free(res); free(pk); free(zk); free(d);
  return itr_used;
}

int orc_iccg(const orc_csr *c, const double *a, double *fi, const double *rhs,
             int itr_max, double tol_abs, double tol_rel, double *hist,
             orc_solve_info *info) {
  return iccg_impl(c, a, fi, rhs, itr_max, tol_abs, tol_rel, hist, info, 0);
}

int orc_iccg_l2(const orc_csr *c, const double *a, double *fi, const double *rhs,
                int itr_max, double tol_abs, double tol_rel, double *hist,
                orc_solve_info *info) {
  return iccg_impl(c, a, fi, rhs, itr_max, tol_abs, tol_rel, hist, info, 1);
}

int orc_dpcg(const orc_csr *c, const double *a, double *fi, const double *rhs,
             int itr_max, double tol_abs, double tol_rel, double *hist,
             orc_solve_info *info) {
  const int n = c->n;
  double *res = (double *)calloc((size_t)n, sizeof(double));
  double *pk = (double *)calloc((size_t)n, sizeof(double));
  double *zk = (double *)calloc((size_t)n, sizeof(double));
  double factor = 0.0, resor = 0.0, resl = 0.0;
  int itr_used = 0;
  residual(c, a, fi, rhs, res);
  const double res0 = l1(res, n);
  if (res0 < tol_abs) {
    resor = res0;
    resl = res0;
    goto done;
  }
  double s0 = S0INIT;
  for (int l = 1; l <= itr_max; ++l) {
    for (int i = 0; i < n; ++i) zk[i] = res[i] / a[c->diag[i]];
    const double sk = dot(res, zk, n);
    const double bet = sk / s0;
    for (int i = 0; i < n; ++i) pk[i] = zk[i] + bet * pk[i];
    orc_spmv(c, a, pk, zk);
    const double pkapk = dot(pk, zk, n);
    const double alf = sk / pkapk;
    for (int i = 0; i < n; ++i) fi[i] = fi[i] + alf * pk[i];
    for (int i = 0; i < n; ++i) res[i] = res[i] - alf * zk[i];
    resl = l1(res, n);
    s0 = sk;
    itr_used++;
    if (hist) hist[l - 1] = resl;
    if (l == 1) {
      factor = factor_of(c, a, fi);
      resor = res0 / factor;
    }
    const double rsm = resl / (res0 + SMALL);
    if (rsm < tol_rel || resl < tol_abs) break;
  }
done:
  if (info) {
    info->resor = resor; info->res0 = res0; info->resl = resl;
    info->factor = factor; info->iters = itr_used;
  }
  free(res); free(pk); free(zk);
  return itr_used;
}

int orc_bicgstab(const orc_csr *c, const double *a, double *fi, const double *rhs,
                 int itr_max, double tol_abs, double tol_rel, double *hist,
                 orc_solve_info *info) {
  const int n = c->n;
  double *res = (double *)calloc((size_t)n, sizeof(double));
  double *reso = (double *)calloc((size_t)n, sizeof(double));
  double *pk = (double *)calloc((size_t)n, sizeof(double));
  double *uk = (double *)calloc((size_t)n, sizeof(double));
  double *zk = (double *)calloc((size_t)n, sizeof(double));
  double *vk = (double *)calloc((size_t)n, sizeof(double));
  double *d = (double *)calloc((size_t)n, sizeof(double));
  double factor = 0.0, resor = 0.0, resl = 0.0;
  int itr_used = 0;
  residual(c, a, fi, rhs, res);                 /* :571-576 */
  const double res0 = l1(res, n);
  if (res0 < tol_abs) {
    resor = res0;
    resl = res0;
    goto done;
  }
  orc_dilu_diag(c, a, d);                       /* :613-624 */
  memcpy(reso, res, sizeof(double) * n);
  double alf = 1.0, beto = 1.0, gam = 1.0;
  for (int l = 1; l <= itr_max; ++l) {
    const double bet = dot(res, reso, n);       /* :651     */
    const double om = bet * gam / (alf * beto + SMALL);
    beto = bet;
    for (int i = 0; i < n; ++i) pk[i] = res[i] + om * (pk[i] - alf * uk[i]);
    orc_dic_apply(c, a, d, pk, zk);             /* :664-683 */
    orc_spmv(c, a, zk, uk);                     /* :689-694 */
    const double ukreso = dot(uk, reso, n);
    gam = bet / ukreso;
    for (int i = 0; i < n; ++i) fi[i] = fi[i] + gam * zk[i];
    for (int i = 0; i < n; ++i) res[i] = res[i] - gam * uk[i];
    orc_dic_apply(c, a, d, res, zk);            /* :714-732 */
    orc_spmv(c, a, zk, vk);                     /* :737-742 */
    alf = dot(vk, res, n) / (dot(vk, vk, n) + SMALL);
    for (int i = 0; i < n; ++i) fi[i] = fi[i] + alf * zk[i];
    for (int i = 0; i < n; ++i) res[i] = res[i] - alf * vk[i];
    resl = l1(res, n);
    itr_used++;
    if (hist) hist[l - 1] = resl;
    if (l == 1) {
      factor = factor_of(c, a, fi);
      resor = res0 / factor;
    }
    const double rsm = resl / (res0 + SMALL);
    if (rsm < tol_rel || resl < tol_abs) break;
  }
done:
  if (info) {
    info->resor = resor; info->res0 = res0; info->resl = resl;
    info->factor = factor; info->iters = itr_used;
  }
  free(res); free(reso); free(pk); free(uk); free(zk); free(vk); free(d);
  return itr_used;
}

int orc_gauss_seidel(const orc_csr *c, const double *a, double *fi,
                     const double *rhs, int itr_max, double tol_abs,
                     double tol_rel, double *hist, orc_solve_info *info) {
  const int n = c->n;
  double *res = (double *)calloc((size_t)n, sizeof(double));
  double factor = 0.0, resor = 0.0, resl = 0.0, res0 = 0.0;
  int itr_used = 0;
  for (int l = 1; l <= itr_max; ++l) {
    for (int i = 0; i < n; ++i) {              /* :147-153 */
      double r = rhs[i];
      for (int k = c->ia[i]; k < c->ia[i + 1]; ++k) r = r - a[k] * fi[c->ja[k]];
      res[i] = r;
      fi[i] = fi[i] + res[i] / (a[c->diag[i]] + SMALL);
    }
    if (l == 1) {
      res0 = l1(res, n);
      if (res0 < tol_abs) {
        resor = res0;
        resl = res0;
        itr_used = 1;
        break;
      }
    }
    resl = l1(res, n);
    itr_used++;
    if (hist) hist[l - 1] = resl;
    if (l == 1) {
      factor = factor_of(c, a, fi);
      resor = res0 / factor;
    }
    const double rsm = resl / (res0 + SMALL);
    if (rsm < tol_rel || resl < tol_abs) break;
  }
  if (info) {
    info->resor = resor; info->res0 = res0; info->resl = resl;
    info->factor = factor; info->iters = itr_used;
  }
  free(res);
  return itr_used;
}

/* ------------------------------------------------------------------------ */
/* GMRES(m) with ILU(0): linear_solvers.f90:789-1555                        */
/* ------------------------------------------------------------------------ */
void orc_atx(const orc_csr *c, const double *a, const double *x, double *y) {
  const int n = c->n;
  for (int i = 0; i < n; ++i) y[i] = 0.0;                       /* :1162 */
  for (int i = 0; i < n; ++i)                                    /* :1164-1168 */
    for (int k = c->ia[i]; k < c->ia[i + 1]; ++k) y[c->ja[k]] = y[c->ja[k]] + a[k] * x[i];
}

int orc_ilu0(const orc_csr *c, const double *a, double *l) {
  const int n = c->n, nnz = c->nnz;
  int *iw = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *ua = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  if (!iw || !ua) {
    free(iw);
    free(ua);
    return -1;
  }
  memcpy(l, a, sizeof(double) * (size_t)nnz);                    /* :1332 */
  memcpy(ua, c->diag, sizeof(int) * (size_t)n);
  /* iw(1:n) = -1 per row (:1338) is kept as "entries of row i only": set
   * before the row and cleared after it, which reads the same values. */
  for (int i = 0; i < n; ++i) iw[i] = -1;
  int rc = 0;
  for (int i = 0; i < n && rc == 0; ++i) {
    for (int k = c->ia[i]; k < c->ia[i + 1]; ++k) iw[c->ja[k]] = k;  /* :1340-1342 */
    int j = c->ia[i], jrow = -1;
    for (; j < c->ia[i + 1]; ++j) {                               /* :1344-1357 */
      jrow = c->ja[j];
      if (i <= jrow) break;
      const double tl = l[j] * l[ua[jrow]];
      l[j] = tl;
      for (int jj = ua[jrow] + 1; jj < c->ia[jrow + 1]; ++jj) {
        const int jw = iw[c->ja[jj]];
        if (jw != -1) l[jw] = l[jw] - tl * l[jj];
      }
    }
    ua[i] = j;                                                    /* :1359 */
    if (jrow != i || l[j] == 0.0) rc = -(i + 1);                  /* :1361-1376 */
    else l[j] = 1.0 / l[j];                                       /* :1378 */
    for (int k = c->ia[i]; k < c->ia[i + 1]; ++k) iw[c->ja[k]] = -1;
  }
  if (rc == 0)
    for (int i = 0; i < n; ++i) l[ua[i]] = 1.0 / l[ua[i]];        /* :1382 */
  free(iw);
  free(ua);
  return rc;
}

void orc_ilu_apply(const orc_csr *c, const double *l, const double *r, double *z) {
  const int n = c->n;
  double *w = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  if (!w) return;
  for (int i = 0; i < n; ++i) w[i] = r[i];                        /* :1453 */
  for (int i = 1; i < n; ++i)                                     /* :1457-1461 */
    for (int j = c->ia[i]; j < c->diag[i]; ++j) w[i] = w[i] - l[j] * w[c->ja[j]];
  for (int i = n - 1; i >= 0; --i) {                              /* :1465-1470 */
    for (int j = c->diag[i] + 1; j < c->ia[i + 1]; ++j) w[i] = w[i] - l[j] * w[c->ja[j]];
    w[i] = w[i] / l[c->diag[i]];
  }
  for (int i = 0; i < n; ++i) z[i] = w[i];                        /* :1474 */
  free(w);
}

/* mult_givens (:1548-1552) on g(k), g(k+1), 0-based k. */
static void givens(double cs, double sn, int k, double *g) {
  const double g1 = cs * g[k] - sn * g[k + 1];
  const double g2 = sn * g[k] + cs * g[k + 1];
  g[k] = g1;
  g[k + 1] = g2;
}

int orc_pmgmres_ilu(const orc_csr *c, const double *a, double *x, const double *rhs,
                    int itr_max, int mr, double tol_abs, double tol_rel, double *hist,
                    orc_solve_info *info) {
  const int n = c->n;
  const int ld = mr + 1;
  const double delta = 1.0e-3;                                    /* :903 */
  double *r = (double *)calloc((size_t)n, sizeof(double));
  double *v = (double *)calloc((size_t)n * (size_t)(mr + 1), sizeof(double));
  double *l = (double *)calloc((size_t)c->nnz, sizeof(double));
  double *h = (double *)calloc((size_t)ld * (size_t)mr, sizeof(double));
  double *cs = (double *)calloc((size_t)ld, sizeof(double));
  double *sn = (double *)calloc((size_t)ld, sizeof(double));
  double *g = (double *)calloc((size_t)ld, sizeof(double));
  double *y = (double *)calloc((size_t)ld, sizeof(double));
  int itr_used = 0, k_copy = 0, rc;
  double rho_tol = 0.0, resor = 0.0, resl = 0.0;
#define H(i, k) h[(i) + ld * (k)]
  rc = orc_ilu0(c, a, l);                                         /* :947 */
  if (rc < 0) goto done;
  for (int itr = 1; itr <= itr_max; ++itr) {
    orc_spmv(c, a, x, r);                                         /* :959-960 */
    for (int i = 0; i < n; ++i) r[i] = rhs[i] - r[i];
    if (itr == 1) {                                               /* :962-968 */
      resor = l1(r, n);
      rho_tol = resor * tol_rel;
    }
    const double rho = sqrt(dot(r, r, n));                        /* :970 */
    if (rho == 0.0) break;  /* guard: see the header */
    orc_ilu_apply(c, l, r, r);                                    /* :976 */
    for (int i = 0; i < n; ++i) v[i] = r[i] / rho;                /* :978 */
    g[0] = rho;                                                   /* :980-983 */
    for (int i = 1; i <= mr; ++i) g[i] = 0.0;
    for (int i = 0; i < ld * mr; ++i) h[i] = 0.0;
    for (int k = 0; k < mr; ++k) {                                /* :985-1043 */
      k_copy = k + 1;
      double *vk = v + (size_t)k * n, *w = v + (size_t)(k + 1) * n;
      orc_spmv(c, a, vk, w);
      orc_ilu_apply(c, l, w, w);
      const double av = sqrt(dot(w, w, n));
      for (int j = 0; j <= k; ++j) {
        const double *vj = v + (size_t)j * n;
        H(j, k) = dot(w, vj, n);
        for (int i = 0; i < n; ++i) w[i] = w[i] - vj[i] * H(j, k);
      }
      H(k + 1, k) = sqrt(dot(w, w, n));
      if (av + delta * H(k + 1, k) == av) {                       /* :1002-1009 */
        for (int j = 0; j <= k; ++j) {
          const double *vj = v + (size_t)j * n;
          const double htmp = dot(w, vj, n);
          H(j, k) = H(j, k) + htmp;
          for (int i = 0; i < n; ++i) w[i] = w[i] - htmp * vj[i];
        }
        H(k + 1, k) = sqrt(dot(w, w, n));
      }
      if (H(k + 1, k) != 0.0)                                     /* :1011-1013 */
        for (int i = 0; i < n; ++i) w[i] = w[i] / H(k + 1, k);
      if (k > 0) {                                                /* :1015-1021 */
        for (int i = 0; i <= k + 1; ++i) y[i] = H(i, k);
        for (int j = 0; j < k; ++j) givens(cs[j], sn[j], j, y);
        for (int i = 0; i <= k + 1; ++i) H(i, k) = y[i];
      }
      const double mu = sqrt(H(k, k) * H(k, k) + H(k + 1, k) * H(k + 1, k));
      cs[k] = H(k, k) / mu;                                       /* :1023-1029 */
      sn[k] = -H(k + 1, k) / mu;
      H(k, k) = cs[k] * H(k, k) - sn[k] * H(k + 1, k);
      H(k + 1, k) = 0.0;
      givens(cs[k], sn[k], k, g);
      resl = fabs(g[k + 1]);                                      /* :1031-1033 */
      if (hist) hist[itr_used] = resl;
      itr_used++;
      if (resl / resor <= rho_tol || resl <= tol_abs) break;      /* :1039-1041 */
    }
    const int m = k_copy - 1;                                     /* :1045-1051 */
    y[m] = g[m] / H(m, m);
    for (int i = m - 1; i >= 0; --i) {
      double s = 0.0;
      for (int q = i + 1; q <= m; ++q) s = s + H(i, q) * y[q];
      y[i] = (g[i] - s) / H(i, i);
    }
    for (int i = 0; i < n; ++i) {                                 /* :1053-1055 */
      double s = 0.0;
      for (int q = 0; q <= m; ++q) s = s + v[(size_t)q * n + i] * y[q];
      x[i] = x[i] + s;
    }
    if (resl / resor <= rho_tol || resl <= tol_abs) break;        /* :1057-1059 */
  }
#undef H
  rc = itr_used;
done:
  if (info) {
    info->resor = resor; info->res0 = resor; info->resl = resl;
    info->factor = 1.0; info->iters = itr_used;
  }
  free(r); free(v); free(l); free(h); free(cs); free(sn); free(g); free(y);
  return rc;
}

int orc_levels(const orc_csr *c, int upper, int *lev) {
  const int n = c->n;
  int nlev = 0;
  if (!upper) {
    for (int i = 0; i < n; ++i) {
      int l = 0;
      for (int k = c->ia[i]; k < c->diag[i]; ++k)
        if (lev[c->ja[k]] + 1 > l) l = lev[c->ja[k]] + 1;
      lev[i] = l;
      if (l + 1 > nlev) nlev = l + 1;
    }
  } else {
    for (int i = n - 1; i >= 0; --i) {
      int l = 0;
      for (int k = c->diag[i] + 1; k < c->ia[i + 1]; ++k)
        if (lev[c->ja[k]] + 1 > l) l = lev[c->ja[k]] + 1;
      lev[i] = l;
      if (l + 1 > nlev) nlev = l + 1;
    }
  }
  return nlev;
}

/* ------------------------------------------------------------------------ */
/* Convection-diffusion assembly (temperature.f90 + scalar_fluxes.f90)      */
/* ------------------------------------------------------------------------ */
static double face_value(int scheme, int ijp, int ijn, double xf, double yf, double zf,
                         double lambda, const double *u, const double *gr,
                         const orc_geom *g) {
  if (scheme == 0) /* face_value_cds interpolation.f90:116-157 */
    return u[ijp] + (u[ijn] - u[ijp]) * lambda;
  /* face_value_2nd_upwind interpolation.f90:318-356 */
  const double gradfidr = (gr[3 * ijp] * (xf - g->xc[ijp]) + gr[3 * ijp + 1] * (yf - g->yc[ijp])) +
                          gr[3 * ijp + 2] * (zf - g->zc[ijp]);
  return u[ijp] + gradfidr;
}

int orc_transport_assemble(const orc_mesh *m, const orc_geom *g, const orc_csr *c,
                           const int *slot_pn, const int *slot_np, const double *flmass,
                           const double *dcoef, double *phi, const double *grad, double gam,
                           int scheme, const unsigned char *bc_kind, const int *srd_map,
                           double urf, int time_scheme, const double *den, double dt,
                           const double *phi_o, const double *phi_oo, double *a, double *su,
                           double *sp) {
  const int n = m->n_cells, Fi = m->n_inner, F = m->n_faces;
  memset(a, 0, sizeof(double) * c->nnz);
  memset(su, 0, sizeof(double) * n);
  memset(sp, 0, sizeof(double) * n);
  if (time_scheme < 0 || time_scheme > 3) return -1;
  if (time_scheme != 0) { /* temperature.f90:82-102 */
    for (int i = 0; i < n; ++i) {
      const double apotime = den[i] * g->vol[i] / dt;
      if (time_scheme == 2) {
        su[i] = su[i] + apotime * (2 * phi_o[i] - 0.5 * phi_oo[i]);
        sp[i] = sp[i] + 1.5 * apotime;
      } else {
        su[i] = su[i] + apotime * phi_o[i];
        sp[i] = sp[i] + apotime;
      }
    }
  }
  for (int f = 0; f < Fi; ++f) { /* temperature.f90:118-157 */
    const int p = m->owner[f], q = m->neighbour[f];
    const double fxn = g->facint[f], fxp = 1.0 - g->facint[f];
    const double xpn = g->xc[q] - g->xc[p], ypn = g->yc[q] - g->yc[p], zpn = g->zc[q] - g->zc[p];
    const double fm = flmass[f], dc = dcoef[f];
    const double de = dc * g->Df[f];
    const double ce = fm < 0.0 ? fm : 0.0;
    const double cp = fm > 0.0 ? fm : 0.0;
    const double can = -de + ce;
    const double cap = -de - cp;
    double dfixi = grad[3 * p] * fxp + grad[3 * q] * fxn;
    double dfiyi = grad[3 * p + 1] * fxp + grad[3 * q + 1] * fxn;
    double dfizi = grad[3 * p + 2] * fxp + grad[3 * q + 2] * fxn;
    dfixi = dfixi * (g->arx[f] - g->Df[f] * xpn);
    dfiyi = dfiyi * (g->ary[f] - g->Df[f] * ypn);
    dfizi = dfizi * (g->arz[f] - g->Df[f] * zpn);
    const double fdfie = dc * ((dfixi + dfiyi) + dfizi);
    double fii;
    if (fm >= 0.0)
      fii = face_value(scheme, p, q, g->xf[f], g->yf[f], g->zf[f], fxp, phi, grad, g);
    else
      fii = face_value(scheme, q, p, g->xf[f], g->yf[f], g->zf[f], fxn, phi, grad, g);
    double fcfie = fm * fii;
    const double fcfii = ce * phi[q] + cp * phi[p];
    fcfie = gam * (fcfie - fcfii);
    const double suadd = -fcfie + fdfie;
    a[slot_pn[f]] = can;
    a[slot_np[f]] = cap;
    a[c->diag[p]] = a[c->diag[p]] - can;
    a[c->diag[q]] = a[c->diag[q]] - cap;
    su[p] = su[p] + suadd;
    su[q] = su[q] - suadd;
  }
  for (int f = Fi; f < F; ++f) { /* boundary loop temperature.f90:168-320 */
    const int p = m->owner[f];
    const int ijb = n + (f - Fi);
    const int kind = bc_kind[f - Fi];
    /* the reference indexes srdw by a counter of the faces it treats as
       walls (temperature.f90:201-202,235-236) */
    const int srd = srd_map ? srd_map[f - Fi] : f - Fi;
    if (kind == 2) { /* facefluxsc_boundary scalar_fluxes.f90:256-337 */
      const double xpn = g->xf[f] - g->xc[p], ypn = g->yf[f] - g->yc[p], zpn = g->zf[f] - g->zc[p];
      const double ax = g->arx[f], ay = g->ary[f], az = g->arz[f];
      const double Dfi = ((ax * ax + ay * ay) + az * az) / ((xpn * ax + ypn * ay) + zpn * az);
      const double de = dcoef[f] * Dfi;
      const double fm = flmass[f];
      const double ce = fm < 0.0 ? fm : 0.0;
      const double cp = fm > 0.0 ? fm : 0.0;
      const double can = -de + ce;
      (void)cp;
      double dfixi = grad[3 * p] * (ax - Dfi * xpn);
      double dfiyi = grad[3 * p + 1] * (ay - Dfi * ypn);
      double dfizi = grad[3 * p + 2] * (az - Dfi * zpn);
      const double suadd = dcoef[f] * ((dfixi + dfiyi) + dfizi);
      sp[p] = sp[p] - can;
      su[p] = su[p] - can * phi[ijb] + suadd;
    } else if (kind == 0) { /* Dirichlet wall :206-210 */
      const double dw = dcoef[f] * g->srdb[srd];
      sp[p] = sp[p] + dw;
      su[p] = su[p] + dw * phi[ijb];
    } else if (kind == 1) { /* zero-gradient wall :224 */
      phi[ijb] = phi[p];
    } else if (kind == 5) { /* periodic face :262-313, facefluxsc_periodic
                               scalar_fluxes.f90:147-251; the reference
                               takes Df and facint of inner face i */
      int i, t0, k;
      if (!periodic_face(m, f, &i, &t0, &k)) return -1;
      const int q = m->owner[t0 + i];
      const double fxn = 0.5, fxp = fxn;
      const double xpn = 2 * (g->xf[f] - g->xc[p]), ypn = 2 * (g->yf[f] - g->yc[p]),
                   zpn = 2 * (g->zf[f] - g->zc[p]);
      const double dc = dcoef[f], fm = flmass[f];
      const double de = dc * g->Df[i];
      const double ce = fm < 0.0 ? fm : 0.0;
      const double cp = fm > 0.0 ? fm : 0.0;
      const double can = -de + ce;
      const double cap = -de - cp;
      double dfixi = grad[3 * p] * fxp + grad[3 * q] * fxn;
      double dfiyi = grad[3 * p + 1] * fxp + grad[3 * q + 1] * fxn;
      double dfizi = grad[3 * p + 2] * fxp + grad[3 * q + 2] * fxn;
      dfixi = dfixi * (g->arx[f] - g->Df[i] * xpn);
      dfiyi = dfiyi * (g->ary[f] - g->Df[i] * ypn);
      dfizi = dfizi * (g->arz[f] - g->Df[i] * zpn);
      const double fdfie = dc * ((dfixi + dfiyi) + dfizi);
      const double fii = fm >= 0.0 ? face_value(0, p, q, 0, 0, 0, fxp, phi, grad, g)
                                   : face_value(0, q, p, 0, 0, 0, fxn, phi, grad, g);
      double fcfie = fm * fii;
      const double fcfii = ce * phi[q] + cp * phi[p];
      fcfie = gam * (fcfie - fcfii);
      const double suadd = -fcfie + fdfie;
      a[slot_pn[Fi + k]] = can;
      a[slot_np[Fi + k]] = cap;
      a[c->diag[p]] = a[c->diag[p]] - can;
      a[c->diag[q]] = a[c->diag[q]] - cap;
      su[p] = su[p] + suadd;
      su[q] = su[q] - suadd;
    } else if (kind == 4) { /* prescribed normal gradient :231-262 */
      const double dw = dcoef[f] * g->srdb[srd];
      const double ax = g->arx[f], ay = g->ary[f], az = g->arz[f];
      const double are = sqrt((ax * ax + ay * ay) + az * az);
      const double nxf = ax / are, nyf = ay / are, nzf = az / are;
      const double dtn = (grad[3 * ijb] * nxf + grad[3 * ijb + 1] * nyf) + grad[3 * ijb + 2] * nzf;
      su[p] = su[p] + dw * dtn;
    } /* kind 3: no term (symmetry, empty) */
  }
  if (time_scheme == 3) { /* Crank-Nicolson :325-347 */
    for (int k = 0; k < c->nnz; ++k) a[k] = 0.5 * a[k];
    for (int f = 0; f < Fi; ++f) {
      const int p = m->owner[f], q = m->neighbour[f];
      su[p] = su[p] - a[slot_pn[f]] * phi_o[q];
      su[q] = su[q] - a[slot_np[f]] * phi_o[p];
    }
    for (int i = 0; i < n; ++i) {
      const double apotime = den[i] * g->vol[i] / dt;
      double s = 0.0;
      for (int k = c->ia[i]; k < c->ia[i + 1]; ++k) s = s + a[k];
      const double off = s - a[c->diag[i]];
      su[i] = su[i] + (apotime + off) * phi_o[i];
      sp[i] = sp[i] + apotime;
    }
  }
  const double urfr = 1.0 / urf, urfm = 1.0 - urf; /* :350-366 */
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int k = c->ia[i]; k < c->ia[i + 1]; ++k) s = s + a[k];
    const double off = s - a[c->diag[i]];
    a[c->diag[i]] = sp[i] - off;
    a[c->diag[i]] = a[c->diag[i]] * urfr;
    su[i] = su[i] + urfm * a[c->diag[i]] * phi[i];
  }
  return 0;
}

void orc_face_gamma_cells(const orc_mesh *m, const orc_geom *g, const double *diffcoef,
                          double *dcoef) {
  const int Fi = m->n_inner, F = m->n_faces;
  for (int f = 0; f < Fi; ++f) {
    const double fxn = g->facint[f];
    const double fxp = 1.0 - fxn;
    dcoef[f] = diffcoef[m->owner[f]] * fxp + diffcoef[m->neighbour[f]] * fxn;
  }
  for (int f = Fi; f < F; ++f) dcoef[f] = diffcoef[m->owner[f]];
}

void orc_convdiff(const orc_mesh *m, const orc_geom *g, const orc_csr *c,
                  const int *slot_pn, const int *slot_np, const double *flmass,
                  const double *dcoef, double *phi, const double *grad, double gam,
                  int scheme, const unsigned char *bc_kind, double urf, double *a,
                  double *su, double *sp) {
  orc_transport_assemble(m, g, c, slot_pn, slot_np, flmass, dcoef, phi, grad, gam, scheme,
                         bc_kind, NULL, urf, 0, NULL, 1.0, NULL, NULL, a, su, sp);
}

/* Face diffusion coefficients of calculate_temperature: inner faces
 * viscos*prl + ((vis_P + (vis_N - vis_P) lambda) - viscos)*prtr (:125),
 * inlet/outlet/pressure faces viscos*prl + (vis_P - viscos)/sigt (:178),
 * wall faces viscos/pranl + (vis_P - viscos)/sigt (:203, :237); other
 * boundary faces 0.  prl = 1/pranl, prtr = 1/sigt (:66-68). */
void orc_transport_dcoef(const orc_mesh *m, const orc_geom *g, const double *vis,
                         const unsigned char *bc_kind, double viscos, double pranl, double sigt,
                         double *dcoef) {
  const double prl = 1.0 / pranl, prtr = 1.0 / sigt;
  for (int f = 0; f < m->n_inner; ++f) {
    const int p = m->owner[f], q = m->neighbour[f];
    dcoef[f] = viscos * prl + ((vis[p] + (vis[q] - vis[p]) * g->facint[f]) - viscos) * prtr;
  }
  for (int f = m->n_inner; f < m->n_faces; ++f) {
    const int p = m->owner[f];
    const int kind = bc_kind[f - m->n_inner];
    int i, t0, k;
    if (kind == 2)
      dcoef[f] = viscos * prl + (vis[p] - viscos) / sigt;
    else if (kind == 0 || kind == 4)
      dcoef[f] = viscos / pranl + (vis[p] - viscos) / sigt;
    else if (kind == 5 && periodic_face(m, f, &i, &t0, &k)) {
      /* temperature.f90:276: the interpolation factor of inner face i */
      const int q = m->owner[t0 + i];
      dcoef[f] = viscos * prl + ((vis[p] + (vis[q] - vis[p]) * g->facint[i]) - viscos) * prtr;
    } else
      dcoef[f] = 0.0;
  }
}

void orc_nonorth_correction(const orc_mesh *m, const orc_geom *g, const double *gamma,
                            const double *grad, double *su) {
  for (int f = 0; f < m->n_inner; ++f) {
    const int p = m->owner[f], q = m->neighbour[f];
    const double fxn = g->facint[f], fxp = 1.0 - g->facint[f];
    const double xpn = g->xc[q] - g->xc[p], ypn = g->yc[q] - g->yc[p], zpn = g->zc[q] - g->zc[p];
    double dfixi = grad[3 * p] * fxp + grad[3 * q] * fxn;
    double dfiyi = grad[3 * p + 1] * fxp + grad[3 * q + 1] * fxn;
    double dfizi = grad[3 * p + 2] * fxp + grad[3 * q + 2] * fxn;
    dfixi = dfixi * (g->arx[f] - g->Df[f] * xpn);
    dfiyi = dfiyi * (g->ary[f] - g->Df[f] * ypn);
    dfizi = dfizi * (g->arz[f] - g->Df[f] * zpn);
    const double fdfie = gamma[f] * ((dfixi + dfiyi) + dfizi);
    su[p] = su[p] + fdfie;
    su[q] = su[q] - fdfie;
  }
}

/* ------------------------------------------------------------------------ */
/* Reverse Cuthill-McKee (no reference counterpart; specification in the    */
/* header)                                                                  */
/* ------------------------------------------------------------------------ */
static int rcm_less(const int *deg, int a, int b) {
  return deg[a] < deg[b] || (deg[a] == deg[b] && a < b);
}

/* BFS over vertices with done[v] == 0 from r; returns the eccentricity and
 * the minimum (degree, id) vertex of the last level in *last_min. */
static int rcm_bfs(const orc_csr *c, const int *deg, const char *done, int r, int *dist,
                   int *queue, int *last_min) {
  const int n = c->n;
  for (int i = 0; i < n; ++i) dist[i] = -1;
  int head = 0, tail = 0;
  queue[tail++] = r;
  dist[r] = 0;
  int ecc = 0;
  while (head < tail) {
    const int v = queue[head++];
    for (int k = c->ia[v]; k < c->ia[v + 1]; ++k) {
      const int u = c->ja[k];
      if (u == v || done[u] || dist[u] >= 0) continue;
      dist[u] = dist[v] + 1;
      if (dist[u] > ecc) ecc = dist[u];
      queue[tail++] = u;
    }
  }
  int best = -1;
  for (int q = 0; q < tail; ++q) {
    const int v = queue[q];
    if (dist[v] == ecc && (best < 0 || rcm_less(deg, v, best))) best = v;
  }
  *last_min = best;
  return ecc;
}

static void rcm_sort_by_degree(const int *deg, int *v, int m) {
  for (int i = 1; i < m; ++i) { /* insertion sort: lists are short */
    const int x = v[i];
    int j = i - 1;
    while (j >= 0 && rcm_less(deg, x, v[j])) {
      v[j + 1] = v[j];
      --j;
    }
    v[j + 1] = x;
  }
}

int orc_rcm(const orc_csr *c, int pp_iters, int *new_of_old) {
  const int n = c->n;
  int *deg = (int *)malloc(sizeof(int) * (size_t)n);
  int *dist = (int *)malloc(sizeof(int) * (size_t)n);
  int *queue = (int *)malloc(sizeof(int) * (size_t)n);
  int *order = (int *)malloc(sizeof(int) * (size_t)n);
  char *done = (char *)calloc((size_t)n, 1);
  int *kids = (int *)malloc(sizeof(int) * (size_t)(n > 64 ? n : 64));
  if (!deg || !dist || !queue || !order || !done || !kids) return -1;
  for (int i = 0; i < n; ++i) {
    int d = 0;
    for (int k = c->ia[i]; k < c->ia[i + 1]; ++k) d += (c->ja[k] != i);
    deg[i] = d;
  }
  int pos = 0;
  while (pos < n) {
    int s = -1;
    for (int i = 0; i < n; ++i)
      if (!done[i] && (s < 0 || rcm_less(deg, i, s))) s = i;
    int r = s, last;
    int ecc = rcm_bfs(c, deg, done, r, dist, queue, &last);
    for (int it = 0; it < pp_iters; ++it) {
      int last2;
      const int x = last;
      const int ecc2 = rcm_bfs(c, deg, done, x, dist, queue, &last2);
      if (ecc2 > ecc) {
        r = x;
        ecc = ecc2;
        last = last2;
      } else {
        break;
      }
    }
    int head = pos;
    order[pos++] = r;
    done[r] = 1;
    while (head < pos) {
      const int v = order[head++];
      int m = 0;
      for (int k = c->ia[v]; k < c->ia[v + 1]; ++k) {
        const int u = c->ja[k];
        if (u == v || done[u]) continue;
        done[u] = 1;
        kids[m++] = u;
      }
      rcm_sort_by_degree(deg, kids, m);
      for (int q = 0; q < m; ++q) order[pos++] = kids[q];
    }
  }
  for (int k = 0; k < n; ++k) new_of_old[order[k]] = n - 1 - k;
  free(deg);
  free(dist);
  free(queue);
  free(order);
  free(done);
  free(kids);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Pressure correction of SIMPLE: calcp_simple.f90:69-419 (incompressible,  */
/* no pressure patches, one corrector, pscheme 'linear')                    */
/* ------------------------------------------------------------------------ */
/* Rhie-Chow mass flux and pressure-correction coefficient of an inner face:
 * facefluxmass2, faceflux_mass.f90:175-252. */
static double rc_inner(const orc_geom *g, int f, int p, int q, const double *u, const double *v,
                       const double *w, const double *pr, const double *dp, const double *den,
                       const double *apu, double *fm) {
  const double lam = g->facint[f];
  const double fxn = lam, fxp = 1.0 - lam;
  const double xpn = g->xc[q] - g->xc[p], ypn = g->yc[q] - g->yc[p], zpn = g->zc[q] - g->zc[p];
  const double dene = den[p] * fxp + den[q] * fxn;
  const double Kj = g->vol[p] * apu[p] * fxp + g->vol[q] * apu[q] * fxn;
  const double cap = -(dene * Kj * g->Df[f]);
  const double ui = u[p] + (u[q] - u[p]) * lam; /* face_value_cds, interpolation.f90:116-157 */
  const double vi = v[p] + (v[q] - v[p]) * lam;
  const double wi = w[p] + (w[q] - w[p]) * lam;
  const double dpxi = (dp[3 * q] * fxp + dp[3 * p] * fxn) * xpn;
  const double dpyi = (dp[3 * q + 1] * fxp + dp[3 * p + 1] * fxn) * ypn;
  const double dpzi = (dp[3 * q + 2] * fxp + dp[3 * p + 2] * fxn) * zpn;
  *fm = dene * ((ui * g->arx[f] + vi * g->ary[f]) + wi * g->arz[f]) +
        cap * ((((pr[q] - pr[p]) - dpxi) - dpyi) - dpzi);
  return cap;
}

/* facefluxmass2_periodic, faceflux_mass.f90:313-372: face f of a periodic
 * patch (patch-local index i, whose Df(i) -- an inner face's -- it reads),
 * p its owner, q the twin face's owner, lambda = 1/2, velocities weighted by
 * the other cell's 1/a_P. */
static double rc_periodic(const orc_geom *g, int f, int i, int p, int q, const double *u,
                          const double *v, const double *w, const double *pr, const double *dp,
                          const double *den, const double *apu, const double *apv,
                          const double *apw, double *fm) {
  const double fxn = 0.5, fxp = 1.0 - 0.5;
  const double xpn = 2.0 * (g->xf[f] - g->xc[p]), ypn = 2.0 * (g->yf[f] - g->yc[p]),
               zpn = 2.0 * (g->zf[f] - g->zc[p]);
  const double dene = den[p] * fxp + den[q] * fxn;
  double Kj = g->vol[p] * apu[p] * fxp + g->vol[q] * apu[q] * fxn;
  const double cap = -(dene * Kj * g->Df[i]);
  Kj = (apu[p] + apu[q]) + SMALL;
  const double ui = (u[p] * apu[q] + u[q] * apu[p]) / Kj;
  const double vi = (v[p] * apv[q] + v[q] * apv[p]) / Kj;
  const double wi = (w[p] * apw[q] + w[q] * apw[p]) / Kj;
  const double dpxi = (dp[3 * q] * fxp + dp[3 * p] * fxn) * xpn;
  const double dpyi = (dp[3 * q + 1] * fxp + dp[3 * p + 1] * fxn) * ypn;
  const double dpzi = (dp[3 * q + 2] * fxp + dp[3 * p + 2] * fxn) * zpn;
  *fm = dene * ((ui * g->arx[f] + vi * g->ary[f]) + wi * g->arz[f]) +
        cap * ((((pr[q] - pr[p]) - dpxi) - dpyi) - dpzi);
  return cap;
}

int orc_pcorr_assemble(const orc_mesh *m, const orc_geom *g, const orc_csr *c,
                       const int *slot_pn, const int *slot_np, double *u, double *v, double *w,
                       const double *pr, const double *dpdxi, const double *den,
                       const double *apu, const double *apv, const double *apw, double *flmass,
                       int const_mflux, double flomas, double *a, double *su) {
  const int n = m->n_cells, Fi = m->n_inner;
  for (int ib = 0; ib < m->n_patches; ++ib) {
    if (m->bctype[ib] == 5) return -1; /* pressure patches: not restated */
    if (m->bctype[ib] == 7 && m->nfaces[ib] > Fi) return -1;
  }
  memset(a, 0, sizeof(double) * (size_t)c->nnz);
  memset(su, 0, sizeof(double) * (size_t)n);
  for (int f = 0; f < Fi; ++f) { /* :78-119 */
    const int p = m->owner[f], q = m->neighbour[f];
    const double cap = rc_inner(g, f, p, q, u, v, w, pr, dpdxi, den, apu, &flmass[f]);
    a[slot_pn[f]] = cap;
    a[slot_np[f]] = cap;
    a[c->diag[p]] = a[c->diag[p]] - cap;
    a[c->diag[q]] = a[c->diag[q]] - cap;
    su[p] = su[p] - flmass[f];
    su[q] = su[q] + flmass[f];
  }
  if (!const_mflux) { /* adjustMassFlow, faceflux_mass.f90:833-912 */
    double flowo = 0.0;
    for (int ib = 0; ib < m->n_patches; ++ib) {
      if (m->bctype[ib] != 4) continue;
      for (int i = 0; i < m->nfaces[ib]; ++i) {
        const int f = m->start_face[ib] + i, p = m->owner[f], b = n + (f - Fi);
        u[b] = u[p];
        v[b] = v[p];
        w[b] = w[p];
        flmass[f] = den[p] * ((u[b] * g->arx[f] + v[b] * g->ary[f]) + w[b] * g->arz[f]);
        flowo = flowo + flmass[f];
      }
    }
    const double fac = flomas / (flowo + SMALL);
    for (int ib = 0; ib < m->n_patches; ++ib) {
      if (m->bctype[ib] != 4) continue;
      for (int i = 0; i < m->nfaces[ib]; ++i) {
        const int f = m->start_face[ib] + i, b = n + (f - Fi);
        flmass[f] = flmass[f] * fac;
        u[b] = u[b] * fac;
        v[b] = v[b] * fac;
        w[b] = w[b] * fac;
      }
    }
  }
  int iper = 0, l = Fi;
  for (int ib = 0; ib < m->n_patches; ++ib) { /* :129-228 */
    const int t = m->bctype[ib];
    if (t == 3 || t == 4) {
      for (int i = 0; i < m->nfaces[ib]; ++i) {
        const int f = m->start_face[ib] + i, p = m->owner[f];
        su[p] = su[p] - flmass[f];
      }
    } else if (t == 7) {
      const int t0 = m->twin_start[iper++];
      for (int i = 0; i < m->nfaces[ib]; ++i, ++l) {
        const int f = m->start_face[ib] + i, p = m->owner[f], q = m->owner[t0 + i];
        const double cap =
            rc_periodic(g, f, i, p, q, u, v, w, pr, dpdxi, den, apu, apv, apw, &flmass[f]);
        a[slot_pn[l]] = cap;
        a[slot_np[l]] = cap;
        a[c->diag[p]] = a[c->diag[p]] - cap;
        a[c->diag[q]] = a[c->diag[q]] - cap;
        su[p] = su[p] - flmass[f];
        su[q] = su[q] + flmass[f];
      }
    }
  }
  return 0;
}

/* bpres (bpres.f90): stage 1 copies the owner's value to every non-pressure
 * boundary face, stage 2 extrapolates with the gradient to wall faces. */
static void orc_bpres(const orc_mesh *m, const orc_geom *g, double *pp, const double *dp,
                      int stage) {
  const int n = m->n_cells, Fi = m->n_inner;
  for (int ib = 0; ib < m->n_patches; ++ib) {
    const int t = m->bctype[ib];
    if (stage == 1 ? t == 5 : t != 1) continue;
    for (int i = 0; i < m->nfaces[ib]; ++i) {
      const int f = m->start_face[ib] + i, p = m->owner[f], b = n + (f - Fi);
      if (stage == 1) {
        pp[b] = pp[p];
      } else {
        const double xpb = g->xf[f] - g->xc[p], ypb = g->yf[f] - g->yc[p],
                     zpb = g->zf[f] - g->zc[p];
        pp[b] = ((pp[p] + dp[3 * p] * xpb) + dp[3 * p + 1] * ypb) + dp[3 * p + 2] * zpb;
      }
    }
  }
}

/* gradp_and_sources(p), nablap.f90:19-209 with pscheme 'linear': su/sv/sw =
 * -sum_f p_f S_f, p's boundary values by bpres's two stages, dp's cell rows
 * the gradient of p. */
static void gradp_sources(const orc_mesh *m, const orc_geom *g, double *pp, double *dpdxi,
                          double *su, double *sv, double *sw) {
  const int n = m->n_cells, Fi = m->n_inner, F = m->n_faces;
  memset(su, 0, sizeof(double) * (size_t)n);
  memset(sv, 0, sizeof(double) * (size_t)n);
  memset(sw, 0, sizeof(double) * (size_t)n);
  for (int f = 0; f < Fi; ++f) {
    const int p = m->owner[f], q = m->neighbour[f];
    const double pf = pp[p] + (pp[q] - pp[p]) * g->facint[f];
    const double dfxe = pf * g->arx[f], dfye = pf * g->ary[f], dfze = pf * g->arz[f];
    su[p] = su[p] - dfxe;
    sv[p] = sv[p] - dfye;
    sw[p] = sw[p] - dfze;
    su[q] = su[q] + dfxe;
    sv[q] = sv[q] + dfye;
    sw[q] = sw[q] + dfze;
  }
  for (int stage = 1; stage <= 2; ++stage) {
    orc_bpres(m, g, pp, dpdxi, stage);
    for (int c = 0; c < n; ++c) {
      dpdxi[3 * c] = -su[c];
      dpdxi[3 * c + 1] = -sv[c];
      dpdxi[3 * c + 2] = -sw[c];
    }
    for (int f = Fi; f < F; ++f) {
      const int p = m->owner[f], b = n + (f - Fi);
      dpdxi[3 * p] = dpdxi[3 * p] + pp[b] * g->arx[f];
      dpdxi[3 * p + 1] = dpdxi[3 * p + 1] + pp[b] * g->ary[f];
      dpdxi[3 * p + 2] = dpdxi[3 * p + 2] + pp[b] * g->arz[f];
    }
    for (int c = 0; c < n; ++c) {
      const double volr = 1.0 / g->vol[c];
      dpdxi[3 * c] = dpdxi[3 * c] * volr;
      dpdxi[3 * c + 1] = dpdxi[3 * c + 1] * volr;
      dpdxi[3 * c + 2] = dpdxi[3 * c + 2] * volr;
    }
  }
  for (int f = Fi; f < F; ++f) {
    const int p = m->owner[f], b = n + (f - Fi);
    su[p] = su[p] - pp[b] * g->arx[f];
    sv[p] = sv[p] - pp[b] * g->ary[f];
    sw[p] = sw[p] - pp[b] * g->arz[f];
  }
}

/* updateVelocityAtBoundary, velocity.f90:1184-1279 */
static void vel_bnd(const orc_mesh *m, const orc_geom *g, double *u, double *v, double *w) {
  const int n = m->n_cells, Fi = m->n_inner;
  for (int ib = 0; ib < m->n_patches; ++ib) {
    const int t = m->bctype[ib];
    if (t != 6 && t != 7 && t != 2) continue;
    for (int i = 0; i < m->nfaces[ib]; ++i) {
      const int f = m->start_face[ib] + i, p = m->owner[f], b = n + (f - Fi);
      if (t == 2) {
        const double Unmag = (u[p] * g->arx[f] + v[p] * g->ary[f]) + w[p] * g->arz[f];
        u[b] = u[p] - Unmag * g->arx[f];
        v[b] = v[p] - Unmag * g->ary[f];
        w[b] = w[p] - Unmag * g->arz[f];
      } else {
        u[b] = u[p];
        v[b] = v[p];
        w[b] = w[p];
      }
    }
  }
}

void orc_pcorr_correct(const orc_mesh *m, const orc_geom *g, const int *slot_pn,
                       const double *a, double *pp, double *flmass, double *u, double *v,
                       double *w, double *pr, const double *apu, const double *apv,
                       const double *apw, double urfp, int pref, double *dpdxi, double *su,
                       double *sv, double *sw) {
  const int n = m->n_cells, Fi = m->n_inner;
  for (int f = 0; f < Fi; ++f) { /* :333-345 */
    const int p = m->owner[f], q = m->neighbour[f];
    flmass[f] = flmass[f] + a[slot_pn[f]] * (pp[q] - pp[p]);
  }
  int iper = 0, l = Fi;
  for (int ib = 0; ib < m->n_patches; ++ib) { /* :348-376 */
    if (m->bctype[ib] != 7) continue;
    const int t0 = m->twin_start[iper++];
    for (int i = 0; i < m->nfaces[ib]; ++i, ++l) {
      const int f = m->start_face[ib] + i, p = m->owner[f], q = m->owner[t0 + i];
      flmass[f] = flmass[f] + a[slot_pn[l]] * (pp[q] - pp[p]);
      flmass[t0 + i] = flmass[f];
    }
  }
  const double ppref = pp[pref]; /* :395 */
  gradp_sources(m, g, pp, dpdxi, su, sv, sw);
  for (int c = 0; c < n; ++c) { /* :401-404 */
    u[c] = u[c] + su[c] * apu[c];
    v[c] = v[c] + sv[c] * apv[c];
    w[c] = w[c] + sw[c] * apw[c];
    pr[c] = pr[c] + urfp * (pp[c] - ppref);
  }
  vel_bnd(m, g, u, v, w);
}

/* ------------------------------------------------------------------------ */
/* Momentum assembly of calcuvw: velocity.f90:50-750                        */
/* ------------------------------------------------------------------------ */
/* sngrad (gradients.f90:1720-1781): the face-interpolated gradient, its
 * corrected form d_ii and the non-orthogonal part d_i (S - D_f d). */
static void orc_sngrad(const orc_geom *g, int f, int p, int q, const double *phi,
                       const double *gr, double *di, double *dii) {
  const double lam = g->facint[f];
  const double fxn = lam, fxp = 1.0 - lam;
  const double xpn = g->xc[q] - g->xc[p], ypn = g->yc[q] - g->yc[p], zpn = g->zc[q] - g->zc[p];
  const double ax = g->arx[f], ay = g->ary[f], az = g->arz[f];
  const double vole = (xpn * ax + ypn * ay) + zpn * az;
  double dx = gr[3 * p] * fxp + gr[3 * q] * fxn;
  double dy = gr[3 * p + 1] * fxp + gr[3 * q + 1] * fxn;
  double dz = gr[3 * p + 2] * fxp + gr[3 * q + 2] * fxn;
  const double e = (((phi[q] - phi[p]) - dx * xpn) - dy * ypn) - dz * zpn;
  dii[0] = dx + ax / vole * e;
  dii[1] = dy + ay / vole * e;
  dii[2] = dz + az / vole * e;
  di[0] = dx * (ax - g->Df[f] * xpn);
  di[1] = dy * (ay - g->Df[f] * ypn);
  di[2] = dz * (az - g->Df[f] * zpn);
}

/* The corrected boundary-face gradient of facefluxuvw_bnd (velocity.f90:
 * 916-1037): owner gradient, corrected by the face value. */
static void orc_snbnd(const double *gr, int p, int b, const double *phi, double xpn, double ypn,
                      double zpn, double ax, double ay, double az, double vole, double Dfi,
                      double *di, double *dii) {
  const double dx = gr[3 * p], dy = gr[3 * p + 1], dz = gr[3 * p + 2];
  const double e = (((phi[b] - phi[p]) - dx * xpn) - dy * ypn) - dz * zpn;
  dii[0] = dx + ax / vole * e;
  dii[1] = dy + ay / vole * e;
  dii[2] = dz + az / vole * e;
  di[0] = dx * (ax - Dfi * xpn);
  di[1] = dy * (ay - Dfi * ypn);
  di[2] = dz * (az - Dfi * zpn);
}

int orc_momentum(const orc_mesh *m, const orc_geom *g, const orc_csr *c, const int *slot_pn,
                 const int *slot_np, double *u, double *v, double *w, const double *uo,
                 const double *vo, const double *wo, const double *uoo, const double *voo,
                 const double *woo, double *pr, const double *den, const double *vis,
                 const double *visw, const double *flmass, double viscos, double dt,
                 int time_scheme, int scheme, double gds, const double *urf, double *a,
                 double *dpdxi, double *dudxi, double *dvdxi, double *dwdxi, double *rhs,
                 double *diag3, double *ap3, double *sp3, double *sv_out, double *sw_out) {
  const int n = m->n_cells, Fi = m->n_inner;
  for (int ib = 0; ib < m->n_patches; ++ib)
    if (m->bctype[ib] == 7 && m->nfaces[ib] > Fi) return -1;
  double *su = (double *)calloc(3 * (size_t)n, sizeof(double));
  if (!su) return -1;
  double *sv = su + n, *sw = su + 2 * (size_t)n;
  double *spu = sp3, *spv = sp3 + n, *sp = sp3 + 2 * (size_t)n;
  memset(sp3, 0, sizeof(double) * 3 * (size_t)n);
  const double *U[3] = {u, v, w}, *G[3] = {dudxi, dvdxi, dwdxi};
  vel_bnd(m, g, u, v, w); /* :166 */
  orc_grad_gauss(m, g, u, dudxi);
  orc_grad_gauss(m, g, v, dvdxi);
  orc_grad_gauss(m, g, w, dwdxi);
  gradp_sources(m, g, pr, dpdxi, su, sv, sw); /* :173 */
  if (time_scheme > 0) /* :226-279 */
    for (int i = 0; i < n; ++i) {
      const double apotime = den[i] * g->vol[i] / dt;
      if (time_scheme == 1 || time_scheme == 3) {
        su[i] = su[i] + apotime * uo[i];
        sv[i] = sv[i] + apotime * vo[i];
        sw[i] = sw[i] + apotime * wo[i];
        spu[i] = spu[i] + apotime;
        spv[i] = spv[i] + apotime;
        sp[i] = sp[i] + apotime;
      } else {
        su[i] = su[i] + apotime * (2.0 * uo[i] - 0.5 * uoo[i]);
        sv[i] = sv[i] + apotime * (2.0 * vo[i] - 0.5 * voo[i]);
        sw[i] = sw[i] + apotime * (2.0 * wo[i] - 0.5 * woo[i]);
        spu[i] = spu[i] + 1.5 * apotime;
        spv[i] = spv[i] + 1.5 * apotime;
        sp[i] = sp[i] + 1.5 * apotime;
      }
    }
  for (int f = 0; f < Fi; ++f) { /* facefluxuvw :754-878, scatter :286-314 */
    const int p = m->owner[f], q = m->neighbour[f];
    const double lam = g->facint[f], fxn = lam, fxp = 1.0 - lam;
    const double fm = flmass[f];
    const double game = vis[p] + (vis[q] - vis[p]) * lam;
    const double de = game * g->Df[f];
    const double ce = fm < 0.0 ? fm : 0.0, cp = fm > 0.0 ? fm : 0.0;
    const double can = -de + ce, cap = -de - cp;
    double di[3][3], dii[3][3];
    for (int r = 0; r < 3; ++r) orc_sngrad(g, f, p, q, U[r], G[r], di[r], dii[r]);
    const double ax = g->arx[f], ay = g->ary[f], az = g->arz[f];
    double sadd[3];
    for (int r = 0; r < 3; ++r) {
      /* fdue = game (du_ii_x ax + dv_ii_x ay + dw_ii_x az), transposed */
      double fde = game * ((dii[0][r] * ax + dii[1][r] * ay) + dii[2][r] * az);
      const double fdi = game * ((di[r][0] + di[r][1]) + di[r][2]);
      fde = fde + fdi;
      const double fuds = cp * U[r][p] + ce * U[r][q];
      const double ue = fm >= 0.0 ? face_value(scheme, p, q, g->xf[f], g->yf[f], g->zf[f], fxp,
                                               U[r], G[r], g)
                                  : face_value(scheme, q, p, g->xf[f], g->yf[f], g->zf[f], fxn,
                                               U[r], G[r], g);
      const double fhigh = fm * ue;
      sadd[r] = -(gds * (fhigh - fuds)) + fde;
    }
    a[slot_pn[f]] = can;
    a[slot_np[f]] = cap;
    su[p] = su[p] + sadd[0];
    sv[p] = sv[p] + sadd[1];
    sw[p] = sw[p] + sadd[2];
    su[q] = su[q] - sadd[0];
    sv[q] = sv[q] - sadd[1];
    sw[q] = sw[q] - sadd[2];
  }
  int iper = 0, iwall = 0, l = Fi;
  for (int ib = 0; ib < m->n_patches; ++ib) { /* :326-482 */
    const int t = m->bctype[ib];
    for (int i = 0; i < m->nfaces[ib]; ++i) {
      const int f = m->start_face[ib] + i, p = m->owner[f], b = n + (f - Fi);
      const double ax = g->arx[f], ay = g->ary[f], az = g->arz[f];
      if (t == 3 || t == 4 || t == 5) { /* facefluxuvw_bnd :882-1034 */
        const double xpn = g->xf[f] - g->xc[p], ypn = g->yf[f] - g->yc[p],
                     zpn = g->zf[f] - g->zc[p];
        const double vole = (xpn * ax + ypn * ay) + zpn * az;
        const double Dfi = ((ax * ax + ay * ay) + az * az) / vole;
        const double game = vis[b];
        const double de = game * Dfi;
        const double fm = flmass[f];
        const double cb = -de + (fm < 0.0 ? fm : 0.0);
        double di[3][3], dii[3][3];
        for (int r = 0; r < 3; ++r)
          orc_snbnd(G[r], p, b, U[r], xpn, ypn, zpn, ax, ay, az, vole, Dfi, di[r], dii[r]);
        double sadd[3];
        for (int r = 0; r < 3; ++r) {
          const double fde = game * ((dii[0][r] * ax + dii[1][r] * ay) + dii[2][r] * az);
          const double fdi = game * ((di[r][0] + di[r][1]) + di[r][2]);
          sadd[r] = fde + fdi;
        }
        spu[p] = spu[p] - cb;
        spv[p] = spv[p] - cb;
        sp[p] = sp[p] - cb;
        su[p] = su[p] - cb * u[b] + sadd[0];
        sv[p] = sv[p] - cb * v[b] + sadd[1];
        sw[p] = sw[p] - cb * w[b] + sadd[2];
      } else if (t == 2) { /* symmetry :354-391 (vis(inp) with inp = numCells+1) */
        const double are = sqrt((ax * ax + ay * ay) + az * az);
        const double arer = 1.0 / are;
        const double nxf = ax * arer, nyf = ay * arer, nzf = az * arer;
        const double dpb = ((g->xf[f] - g->xc[p]) * nxf + (g->yf[f] - g->yc[p]) * nyf) +
                           (g->zf[f] - g->zc[p]) * nzf;
        const double cf = 2.0 * vis[n] * are / dpb;
        su[p] = su[p] - cf * nxf * (nyf * v[p] + nzf * w[p]);
        sv[p] = sv[p] - cf * nyf * (nxf * u[p] + nzf * w[p]);
        sw[p] = sw[p] - cf * nzf * (nxf * u[p] + nyf * v[p]);
        spu[p] = spu[p] + cf * (nxf * nxf);
        spv[p] = spv[p] + cf * (nyf * nyf);
        sp[p] = sp[p] + cf * (nzf * nzf);
      } else if (t == 7) { /* facefluxuvw_periodic :1038-1179 */
        const int q = m->owner[m->twin_start[iper] + i];
        const double fxn = 0.5, fxp = fxn;
        const double xpn = 2.0 * (g->xf[f] - g->xc[p]), ypn = 2.0 * (g->yf[f] - g->yc[p]),
                     zpn = 2.0 * (g->zf[f] - g->zc[p]);
        const double dpn = sqrt((xpn * xpn + ypn * ypn) + zpn * zpn);
        const double are = sqrt((ax * ax + ay * ay) + az * az);
        const double game = vis[p] * fxp + vis[q] * fxn;
        const double de = game * ((ax * ax + ay * ay) + az * az) / ((xpn * ax + ypn * ay) + zpn * az);
        const double fm = flmass[f];
        const double ce = fm < 0.0 ? fm : 0.0, cp = fm > 0.0 ? fm : 0.0;
        const double can = -de + ce, cap = -de - cp;
        double d[3][3];
        for (int r = 0; r < 3; ++r)
          for (int k = 0; k < 3; ++k) d[r][k] = G[r][3 * p + k] * fxp + G[r][3 * q + k] * fxn;
        /* d[0] = (duxi, duyi, duzi), d[1] = dv.., d[2] = dw.. */
        const double fde[3] = {
            game * (((d[0][0] + d[0][0]) * ax + (d[0][1] + d[1][0]) * ay) + (d[0][2] + d[2][0]) * az),
            game * (((d[0][1] + d[1][0]) * ax + (d[1][1] + d[1][1]) * ay) + (d[1][2] + d[2][1]) * az),
            game * (((d[0][2] + d[2][0]) * ax + (d[2][1] + d[1][2]) * ay) + (d[2][2] + d[2][2]) * az)};
        double sadd[3];
        for (int r = 0; r < 3; ++r) {
          const double fdi = game * are / dpn * ((d[r][0] * xpn + d[r][1] * ypn) + d[r][2] * zpn);
          const double fuds = cp * U[r][p] + ce * U[r][q];
          const double ue = fm >= 0.0 ? U[r][p] + (U[r][q] - U[r][p]) * fxp
                                      : U[r][q] + (U[r][p] - U[r][q]) * fxn;
          const double fhigh = fm * ue;
          sadd[r] = (-(gds * (fhigh - fuds)) + fde[r]) - fdi;
        }
        a[slot_pn[l]] = can;
        a[slot_np[l]] = cap;
        ++l;
        su[p] = su[p] + sadd[0];
        sv[p] = sv[p] + sadd[1];
        sw[p] = sw[p] + sadd[2];
        su[q] = su[q] - sadd[0];
        sv[q] = sv[q] - sadd[1];
        sw[q] = sw[q] - sadd[2];
      } else if (t == 1) { /* wall :437-482 */
        const double viss = viscos > visw[iwall] ? viscos : visw[iwall];
        ++iwall;
        const double are = sqrt((ax * ax + ay * ay) + az * az);
        const double arer = 1.0 / are;
        const double nxf = ax * arer, nyf = ay * arer, nzf = az * arer;
        const double dpb = ((g->xf[f] - g->xc[p]) * nxf + (g->yf[f] - g->yc[p]) * nyf) +
                           (g->zf[f] - g->zc[p]) * nzf;
        const double vsol = viss * are / dpb;
        const double upb = u[p] - u[b], vpb = v[p] - v[b], wpb = w[p] - w[b];
        spu[p] = spu[p] + vsol * (1.0 - nxf * nxf);
        spv[p] = spv[p] + vsol * (1.0 - nyf * nyf);
        sp[p] = sp[p] + vsol * (1.0 - nzf * nzf);
        su[p] = su[p] + vsol * ((u[b] * (1.0 - nxf * nxf) + vpb * nyf * nxf) + wpb * nzf * nxf);
        sv[p] = sv[p] + vsol * ((upb * nxf * nyf + v[b] * (1.0 - nyf * nyf)) + wpb * nzf * nyf);
        sw[p] = sw[p] + vsol * ((upb * nxf * nzf + vpb * nyf * nzf) + w[b] * (1.0 - nzf * nzf));
      }
    }
    if (t == 7) ++iper;
  }
  /* CN (:582-606) and the three diagonals (:608-746), component by
   * component; the row sums read whatever the diagonal entry holds */
  const double *O[3] = {uo, vo, wo};
  double *S[3] = {su, sv, sw}, *SP[3] = {spu, spv, sp};
  if (time_scheme == 3)
    for (int k = 0; k < c->nnz; ++k) a[k] = 0.5 * a[k];
  for (int r = 0; r < 3; ++r) {
    if (time_scheme == 3) {
      for (int f = 0; f < Fi; ++f) {
        const int p = m->owner[f], q = m->neighbour[f];
        S[r][p] = S[r][p] - a[slot_pn[f]] * O[r][q];
        S[r][q] = S[r][q] - a[slot_np[f]] * O[r][p];
      }
      for (int i = 0; i < n; ++i) {
        const double apotime = den[i] * g->vol[i] / dt;
        double s = 0.0;
        for (int k = c->ia[i]; k < c->ia[i + 1]; ++k) s = s + a[k];
        const double off = s - a[c->diag[i]];
        S[r][i] = S[r][i] + (apotime + off) * O[r][i];
        SP[r][i] = SP[r][i] + apotime;
      }
    }
    if (r > 0)
      for (int i = 0; i < n; ++i) a[c->diag[i]] = 0.0;
    const double urfr = 1.0 / urf[r], urfm = 1.0 - urf[r];
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int k = c->ia[i]; k < c->ia[i + 1]; ++k) s = s + a[k];
      const double off = s - a[c->diag[i]];
      double ad = SP[r][i] - off;
      ap3[r * (size_t)n + i] = 1.0 / (ad + SMALL);
      ad = ad * urfr;
      a[c->diag[i]] = ad;
      diag3[r * (size_t)n + i] = ad;
      rhs[r * (size_t)n + i] = S[r][i] + urfm * ad * U[r][i];
    }
  }
  memcpy(sv_out, sv, sizeof(double) * (size_t)n);
  memcpy(sw_out, sw, sizeof(double) * (size_t)n);
  free(su);
  return 0;
}
