// fcp_adjoint.hip -- discrete adjoint building blocks of the assembly path.
//
// For one assembled system A(c) phi = b(c, phi_old, grad_old) and an adjoint
// vector lam solving A^T lam = dJ/dphi, the reverse-mode contributions are
//   dJ/dc       += -lam^T (dA/dc) phi + lam^T (db/dc)
//   dJ/dphi_old += (db/dphi_old)^T lam,   dJ/dgrad_old += (db/dgrad_old)^T lam
// and the Gauss gradient is linear, grad = G phi, so dJ/dphi += G^T dJ/dgrad.
// The kernels below state these transposes face by face for the assembly
// kernels of fcp_assemble.hip (facefluxsc with linear-upwind or central
// deferred correction, urf = 1; the face-diffusivity Laplacian with the
// explicit non-orthogonal correction; grad_gauss).  Every output element is
// written by exactly one thread (a cell gathers its faces; a face's own
// derivative is written by its owner cell), so results are deterministic.
// A^T is the face-based coefficient swap: a(slot_pn) <-> a(slot_np).
#include <hip/hip_runtime.h>

#include <cmath>

#include "fcp_internal.h"

namespace {

constexpr int BLK = 256;

struct AGeom {
  const double *xc, *yc, *zc, *vol, *arx, *ary, *arz, *xf, *yf, *zf, *facint, *Df, *srdb;
};

__global__ void k_transpose(int nnz, const int *__restrict__ tslot, const double *__restrict__ a,
                            double *__restrict__ at) {
  const int k = blockIdx.x * BLK + threadIdx.x;
  if (k < nnz) at[k] = a[tslot[k]];
}

// Reverse of k_convdiff (urf = 1).  lam[n], phi_new[n] (solution of the
// system), phi_old[numTotal] and grad_old[3*numTotal] (inputs of the
// assembly).  Accumulates dcoef_bar[F], phi_old_bar[numTotal],
// grad_old_bar[3*numTotal].
__global__ void k_convdiff_bwd(int n, int Fi, const int *__restrict__ cf_ptr,
                               const int *__restrict__ cf, const int *__restrict__ owner,
                               const int *__restrict__ neighbour, AGeom g,
                               const double *__restrict__ flmass, const double *__restrict__ dcoef,
                               const double *__restrict__ phi_old,
                               const double *__restrict__ gr, double gam, int scheme,
                               const unsigned char *__restrict__ bc_kind,
                               const double *__restrict__ lam, const double *__restrict__ phi_new,
                               double *__restrict__ dcoef_bar, double *__restrict__ phi_old_bar,
                               double *__restrict__ grad_old_bar) {
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= n) return;
  double pb = 0.0, gbx = 0.0, gby = 0.0, gbz = 0.0;
  const double lc = lam[c];
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int f = cf[k];
    if (f < Fi) {
      const int p = owner[f], q = neighbour[f];
      const double wf = lam[p] - lam[q];  // su_P += suadd, su_N -= suadd
      const double fm = flmass[f], dc = dcoef[f];
      const double fxn = g.facint[f], fxp = 1.0 - fxn;
      const double Dff = g.Df[f];
      const double xpn = g.xc[q] - g.xc[p], ypn = g.yc[q] - g.yc[p], zpn = g.zc[q] - g.zc[p];
      const double tx = g.arx[f] - Dff * xpn, ty = g.ary[f] - Dff * ypn, tz = g.arz[f] - Dff * zpn;
      const double ce = fm < 0.0 ? fm : 0.0, cp = fm > 0.0 ? fm : 0.0;
      const int u = fm >= 0.0 ? p : q, d = fm >= 0.0 ? q : p;
      const double lamw = fm >= 0.0 ? fxp : fxn;
      // d suadd / d phi_old: -gam*(F dfii/dphi - ce [q] - cp [p])
      double dphi_c = 0.0;
      if (scheme == 0) {
        if (c == u) dphi_c += -gam * fm * (1.0 - lamw);
        if (c == d) dphi_c += -gam * fm * lamw;
      } else {
        if (c == u) dphi_c += -gam * fm;
      }
      if (c == q) dphi_c += gam * ce;
      if (c == p) dphi_c += gam * cp;
      pb += wf * dphi_c;
      // d suadd / d grad_old: non-orthogonal term and linear-upwind face value
      const double wgt = (c == p) ? fxp : fxn;
      gbx += wf * dc * wgt * tx;
      gby += wf * dc * wgt * ty;
      gbz += wf * dc * wgt * tz;
      if (scheme == 1 && c == u) {
        gbx += wf * (-gam * fm * (g.xf[f] - g.xc[u]));
        gby += wf * (-gam * fm * (g.yf[f] - g.yc[u]));
        gbz += wf * (-gam * fm * (g.zf[f] - g.zc[u]));
      }
      if (c == p) {
        // matrix: dA/ddcoef = -D at (P,N),(N,P); +D at (P,P),(N,N)
        const double gp_x = gr[3 * p] * fxp + gr[3 * q] * fxn;
        const double gp_y = gr[3 * p + 1] * fxp + gr[3 * q + 1] * fxn;
        const double gp_z = gr[3 * p + 2] * fxp + gr[3 * q + 2] * fxn;
        const double rhs = (gp_x * tx + gp_y * ty) + gp_z * tz;
        const double mat = -Dff * (lam[p] - lam[q]) * (phi_new[p] - phi_new[q]);
        dcoef_bar[f] += mat + wf * rhs;
      }
    } else {
      const int b = f - Fi;
      const int kind = bc_kind[b];
      if (kind == 0) {  // Dirichlet wall: sp += dc*srdb, su += dc*srdb*phi_b
        const double sr = g.srdb[b];
        dcoef_bar[f] += -lc * phi_new[c] * sr + lc * sr * phi_old[n + b];
        phi_old_bar[n + b] += lc * dcoef[f] * sr;
      } else if (kind == 2) {  // in/outlet: sp -= can, su += -can*phi_b + suadd
        const double xpn = g.xf[f] - g.xc[c], ypn = g.yf[f] - g.yc[c], zpn = g.zf[f] - g.zc[c];
        const double ax = g.arx[f], ay = g.ary[f], az = g.arz[f];
        const double Dfi = ((ax * ax + ay * ay) + az * az) / ((xpn * ax + ypn * ay) + zpn * az);
        const double tx = ax - Dfi * xpn, ty = ay - Dfi * ypn, tz = az - Dfi * zpn;
        const double fm = flmass[f];
        const double can = -dcoef[f] * Dfi + (fm < 0.0 ? fm : 0.0);
        const double gt = (gr[3 * c] * tx + gr[3 * c + 1] * ty) + gr[3 * c + 2] * tz;
        dcoef_bar[f] += -lc * phi_new[c] * Dfi + lc * (Dfi * phi_old[n + b] + gt);
        phi_old_bar[n + b] += lc * (-can);
        gbx += lc * dcoef[f] * tx;
        gby += lc * dcoef[f] * ty;
        gbz += lc * dcoef[f] * tz;
      }
    }
  }
  phi_old_bar[c] += pb;
  grad_old_bar[3 * (size_t)c] += gbx;
  grad_old_bar[3 * (size_t)c + 1] += gby;
  grad_old_bar[3 * (size_t)c + 2] += gbz;
}

// Reverse of k_laplacian (face form) followed by k_nonorth: accumulates
// mu_f_bar[F], phi_bc_bar[numTotal] (boundary entries), gamma_bar[F] and
// grad_bar[3*numTotal] of the non-orthogonal correction.
__global__ void k_laplacian_bwd(int n, int Fi, const int *__restrict__ cf_ptr,
                                const int *__restrict__ cf, const int *__restrict__ owner,
                                const int *__restrict__ neighbour, AGeom g,
                                const double *__restrict__ mu_f,
                                const unsigned char *__restrict__ bc_kind,
                                const double *__restrict__ phi_bc, const double *__restrict__ gamma,
                                const double *__restrict__ gr, const double *__restrict__ lam,
                                const double *__restrict__ phi, double *__restrict__ mu_f_bar,
                                double *__restrict__ phi_bc_bar, double *__restrict__ gamma_bar,
                                double *__restrict__ grad_bar) {
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= n) return;
  const double lc = lam[c];
  double gbx = 0.0, gby = 0.0, gbz = 0.0;
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int f = cf[k];
    if (f < Fi) {
      const int p = owner[f], q = neighbour[f];
      const double fxn = g.facint[f], fxp = 1.0 - fxn, Dff = g.Df[f];
      const double xpn = g.xc[q] - g.xc[p], ypn = g.yc[q] - g.yc[p], zpn = g.zc[q] - g.zc[p];
      const double tx = g.arx[f] - Dff * xpn, ty = g.ary[f] - Dff * ypn, tz = g.arz[f] - Dff * zpn;
      const double wf = lam[p] - lam[q];
      if (gamma) {
        const double wgt = (c == p) ? fxp : fxn;
        gbx += wf * gamma[f] * wgt * tx;
        gby += wf * gamma[f] * wgt * ty;
        gbz += wf * gamma[f] * wgt * tz;
      }
      if (c == p) {
        // a(PN) = a(NP) = mu_f D, a(PP) and a(NN) -= mu_f D
        mu_f_bar[f] += Dff * (lam[p] - lam[q]) * (phi[p] - phi[q]);
        if (gamma) {
          const double gx = gr[3 * p] * fxp + gr[3 * q] * fxn;
          const double gy = gr[3 * p + 1] * fxp + gr[3 * q + 1] * fxn;
          const double gz = gr[3 * p + 2] * fxp + gr[3 * q + 2] * fxn;
          gamma_bar[f] += wf * ((gx * tx + gy * ty) + gz * tz);
        }
      }
    } else {
      const int b = f - Fi;
      if (bc_kind && bc_kind[b] != 0) continue;
      const double ax = g.arx[f], ay = g.ary[f], az = g.arz[f];
      const double are = sqrt((ax * ax + ay * ay) + az * az);
      const double dfn = ((g.xf[f] - g.xc[c]) * (ax / are) + (g.yf[f] - g.yc[c]) * (ay / are)) +
                         (g.zf[f] - g.zc[c]) * (az / are);
      const double co = are / dfn;
      // a(PP) -= mu_f co ; su -= mu_f co phi_b
      mu_f_bar[f] += lc * phi[c] * co - lc * co * phi_bc[n + b];
      phi_bc_bar[n + b] += -lc * mu_f[f] * co;
    }
  }
  if (gamma) {
    grad_bar[3 * (size_t)c] += gbx;
    grad_bar[3 * (size_t)c + 1] += gby;
    grad_bar[3 * (size_t)c + 2] += gbz;
  }
}

// Reverse of k_grad_gauss: phi_bar[numTotal] += G^T grad_bar.
__global__ void k_grad_gauss_bwd(int n, int Fi, const int *__restrict__ cf_ptr,
                                 const int *__restrict__ cf, const int *__restrict__ owner,
                                 const int *__restrict__ neighbour, AGeom g,
                                 const double *__restrict__ gb, double *__restrict__ phi_bar) {
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= n) return;
  double acc = 0.0;
  const double vc = 1.0 / g.vol[c];
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int f = cf[k];
    if (f < Fi) {
      const int p = owner[f], q = neighbour[f];
      const double vp = 1.0 / g.vol[p], vq = 1.0 / g.vol[q];
      // phi_f enters grad_P with +S/V_P and grad_N with -S/V_N
      const double fb = ((gb[3 * p] * vp - gb[3 * q] * vq) * g.arx[f] +
                         (gb[3 * p + 1] * vp - gb[3 * q + 1] * vq) * g.ary[f]) +
                        (gb[3 * p + 2] * vp - gb[3 * q + 2] * vq) * g.arz[f];
      const double w = g.facint[f];
      acc += (c == p) ? (1.0 - w) * fb : w * fb;
    } else {
      const int b = f - Fi;
      phi_bar[n + b] += ((gb[3 * c] * g.arx[f] + gb[3 * c + 1] * g.ary[f]) +
                         gb[3 * c + 2] * g.arz[f]) * vc;
    }
  }
  phi_bar[c] += acc;
}

AGeom ageom(fcp_ctx *ctx) {
  return AGeom{ctx->xc.p,  ctx->yc.p,  ctx->zc.p,  ctx->vol.p,    ctx->arx.p, ctx->ary.p, ctx->arz.p,
               ctx->xf.p,  ctx->yf.p,  ctx->zf.p,  ctx->facint.p, ctx->Df.p,  ctx->srdb.p};
}

int grid(long long m) { return (int)((m + BLK - 1) / BLK); }

}  // namespace

int fcp_launch_transpose(fcp_ctx *ctx) {
  FCP_TRY(fcp_coarse_wait(ctx));
  const int nnz = ctx->nnz;
  DBuf<double> at;
  FCP_HIP(ctx, at.alloc(nnz));
  k_transpose<<<grid(nnz), BLK, 0, ctx->stream>>>(nnz, ctx->tslot.p, ctx->a.p, at.p);
  FCP_HIP(ctx, hipGetLastError());
  std::swap(ctx->a.p, at.p);
  std::swap(ctx->a.n, at.n);
  ctx->a_version++;
  return 0;
}

int fcp_launch_convdiff_bwd(fcp_ctx *ctx, const double *flmass, const double *dcoef,
                            const double *phi_old, const double *grad_old, double gam, int scheme,
                            const unsigned char *bc_kind, const double *lam,
                            const double *phi_new, double *dcoef_bar, double *phi_old_bar,
                            double *grad_old_bar) {
  ProfScope ps(ctx, "convdiff_bwd");
  k_convdiff_bwd<<<grid(ctx->n), BLK, 0, ctx->stream>>>(
      ctx->n, ctx->Fi, ctx->cf_ptr.p, ctx->cf_face.p, ctx->owner.p, ctx->neighbour.p, ageom(ctx),
      flmass, dcoef, phi_old, grad_old, gam, scheme, bc_kind, lam, phi_new, dcoef_bar,
      phi_old_bar, grad_old_bar);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

int fcp_launch_laplacian_bwd(fcp_ctx *ctx, const double *mu_f, const unsigned char *bc_kind,
                             const double *phi_bc, const double *gamma, const double *grad,
                             const double *lam, const double *phi, double *mu_f_bar,
                             double *phi_bc_bar, double *gamma_bar, double *grad_bar) {
  ProfScope ps(ctx, "laplacian_bwd");
  k_laplacian_bwd<<<grid(ctx->n), BLK, 0, ctx->stream>>>(
      ctx->n, ctx->Fi, ctx->cf_ptr.p, ctx->cf_face.p, ctx->owner.p, ctx->neighbour.p, ageom(ctx),
      mu_f, bc_kind, phi_bc, gamma, grad, lam, phi, mu_f_bar, phi_bc_bar, gamma_bar, grad_bar);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

int fcp_launch_grad_gauss_bwd(fcp_ctx *ctx, const double *grad_bar, double *phi_bar) {
  ProfScope ps(ctx, "grad_gauss_bwd");
  k_grad_gauss_bwd<<<grid(ctx->n), BLK, 0, ctx->stream>>>(ctx->n, ctx->Fi, ctx->cf_ptr.p,
                                                           ctx->cf_face.p, ctx->owner.p,
                                                           ctx->neighbour.p, ageom(ctx), grad_bar,
                                                           phi_bar);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}
