// fcp_lsq.hip -- least-squares cell gradients (gradients.f90:660-1486).
//
// The reference builds the per-cell systems by scattering over faces (inner
// faces to both cells, then boundary faces to their owner); here every cell
// gathers its faces in ascending face id -- the order in which the scatter
// reaches it -- so every accumulated entry sees the same additions in the
// same order and the results are bit-equal.
//   method 1 lsq     : D = sum d d^T, closed-form inverse (+small in the
//                      determinant), b = sum d (phi_N - phi_P)
//   method 2 lsq_dm  : weights 1/|d|^2 in D and b
//   method 3 lsq_qr  : thin QR of the <=6 neighbour offsets by modified
//                      Gram-Schmidt, R^{-1} Q^T stored (3 x 6 per cell)
// compat = 1 reproduces the reference arithmetic including its two defects
// (y-row cofactors at :887/:1480; xf(i) with a boundary counter at :1459);
// compat = 0 computes the corrected least-squares gradient.
#include <hip/hip_runtime.h>

#include <cmath>

#include "fcp_internal.h"

namespace {

constexpr int BLK = 256;

struct LsqGeom {
  const double *xc, *yc, *zc, *xf, *yf, *zf;
};

__device__ __forceinline__ void offset_of(const LsqGeom &g, int c, int f, int Fi,
                                          const int *__restrict__ owner,
                                          const int *__restrict__ neighbour, double &dx,
                                          double &dy, double &dz) {
  if (f < Fi) {
    const int p = owner[f], q = neighbour[f];
    dx = g.xc[q] - g.xc[p];
    dy = g.yc[q] - g.yc[p];
    dz = g.zc[q] - g.zc[p];
  } else {
    dx = g.xf[f] - g.xc[c];
    dy = g.yf[f] - g.yc[c];
    dz = g.zf[f] - g.zc[c];
  }
}

// create_matrix_lsq (:660-779) and create_matrix_lsq_dm (:1157-1328)
__global__ void k_lsq_dmat(int n, int Fi, int weighted, const int *__restrict__ cf_ptr,
                           const int *__restrict__ cf, const int *__restrict__ owner,
                           const int *__restrict__ neighbour, LsqGeom g, double *__restrict__ dmat) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double d11 = 0.0, d12 = 0.0, d13 = 0.0, d22 = 0.0, d23 = 0.0, d33 = 0.0;
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int f = cf[k];
    double dx, dy, dz;
    offset_of(g, c, f, Fi, owner, neighbour, dx, dy, dz);
    if (weighted) {
      const double w = 1.0 / ((dx * dx + dy * dy) + dz * dz);
      d11 += w * dx * dx;
      d22 += w * dy * dy;
      d33 += w * dz * dz;
      d12 += w * dx * dy;
      d13 += w * dx * dz;
      d23 += w * dy * dz;
    } else {
      d11 += dx * dx;
      d22 += dy * dy;
      d33 += dz * dz;
      d12 += dx * dy;
      d13 += dx * dz;
      d23 += dy * dz;
    }
  }
  // closed-form adjugate (:745-777, :1296-1326)
  const double d21 = d12, d31 = d13, d32 = d23;
  const double tmp = 1.0 / ((((((d11 * d22 * d33 - d11 * d23 * d32) - d12 * d21 * d33) +
                                d12 * d23 * d31) + d13 * d21 * d32) - d13 * d22 * d31) +
                            FCP_SMALL);
  double *D = dmat + 9 * (size_t)c;
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

// create_matrix_lsq_qr (:900-1056): rows = neighbour offsets in face-visit
// order (:966-990), MGS (matrix.f90:366-419), R^{-1} Q^T into D(3,6,cell).
// Cells with more than 6 faces are flagged in err.
__global__ void k_lsq_qr(int n, int Fi, const int *__restrict__ cf_ptr, const int *__restrict__ cf,
                         const int *__restrict__ owner, const int *__restrict__ neighbour,
                         LsqGeom g, double *__restrict__ dmat, int *__restrict__ err) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double q[6][3];
  int l = 0;
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int f = cf[k];
    if (l >= 6) {
      atomicExch(err, FCP_ERR_UNSUPPORTED);
      break;
    }
    double dx, dy, dz;
    if (f < Fi) {
      const int p = owner[f], o = (c == p) ? neighbour[f] : p;
      dx = g.xc[o] - g.xc[c];
      dy = g.yc[o] - g.yc[c];
      dz = g.zc[o] - g.zc[c];
    } else {
      dx = g.xf[f] - g.xc[c];
      dy = g.yf[f] - g.yc[c];
      dz = g.zf[f] - g.zc[c];
    }
    q[l][0] = dx;
    q[l][1] = dy;
    q[l][2] = dz;
    ++l;
  }
  for (int i = l; i < 6; ++i) q[i][0] = q[i][1] = q[i][2] = 0.0;
  double r[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double z = 0.0;
    for (int i = 0; i < l; ++i) z = z + q[i][j] * q[i][j];
    r[j][j] = sqrt(z);
    for (int i = 0; i < l; ++i) q[i][j] = q[i][j] / r[j][j];
#pragma unroll
    for (int k = j + 1; k < 3; ++k) {
      z = 0.0;
      for (int i = 0; i < l; ++i) z = z + q[i][j] * q[i][k];
      r[j][k] = z;
      for (int i = 0; i < l; ++i) q[i][k] = q[i][k] - r[j][k] * q[i][j];
    }
  }
  double ri[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  ri[0][0] = 1.0 / r[0][0];
  ri[1][1] = 1.0 / r[1][1];
  ri[2][2] = 1.0 / r[2][2];
  ri[0][1] = -r[0][1] / (r[0][0] * r[1][1]);
  ri[1][2] = -r[1][2] / (r[1][1] * r[2][2]);
  ri[0][2] = (r[0][1] * r[1][2] - r[0][2] * r[1][1]) / (r[0][0] * r[1][1] * r[2][2]);
  double *D = dmat + 18 * (size_t)c;
#pragma unroll
  for (int k = 0; k < 6; ++k)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double s = 0.0;
      if (k < l)
#pragma unroll
        for (int j = 0; j < 3; ++j) s += ri[i][j] * q[k][j];
      D[3 * k + i] = s;
    }
}

// grad_lsq (:782-893) / grad_lsq_dm (:1334-1486): right-hand side gathered in
// face order, then the stored inverse applied.
__global__ void k_grad_lsq(int n, int Fi, int ntot, int weighted, int compat,
                           const int *__restrict__ cf_ptr, const int *__restrict__ cf,
                           const int *__restrict__ owner, const int *__restrict__ neighbour,
                           LsqGeom g, const double *__restrict__ dmat,
                           const double *__restrict__ phi, double *__restrict__ dphi) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ntot) return;
  if (c >= n) {
    dphi[3 * (size_t)c] = 0.0;
    dphi[3 * (size_t)c + 1] = 0.0;
    dphi[3 * (size_t)c + 2] = 0.0;
    return;
  }
  double b1 = 0.0, b2 = 0.0, b3 = 0.0;
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int f = cf[k];
    double Dx, Dy, Dz;
    if (f < Fi) {
      const int p = owner[f], q = neighbour[f];
      const double dx = g.xc[q] - g.xc[p], dy = g.yc[q] - g.yc[p], dz = g.zc[q] - g.zc[p];
      if (!weighted) {
        const double delta = phi[q] - phi[p];
        Dx = dx * delta;
        Dy = dy * delta;
        Dz = dz * delta;
      } else {
        const double w = (phi[q] - phi[p]) / ((dx * dx + dy * dy) + dz * dz);
        Dx = w * dx;
        Dy = w * dy;
        Dz = w * dz;
      }
    } else {
      const int ib = f - Fi;
      const double delta = phi[n + ib] - phi[c];
      const double dx = g.xf[f] - g.xc[c], dy = g.yf[f] - g.yc[c], dz = g.zf[f] - g.zc[c];
      if (!weighted) {
        Dx = dx * delta;
        Dy = dy * delta;
        Dz = dz * delta;
      } else {
        const int fw = compat ? ib : f;  // :1459 indexes xf with the boundary counter
        const double wx = g.xf[fw] - g.xc[c], wy = g.yf[fw] - g.yc[c], wz = g.zf[fw] - g.zc[c];
        const double w = delta / ((wx * wx + wy * wy) + wz * wz);
        Dx = w * dx;
        Dy = w * dy;
        Dz = w * dz;
      }
    }
    b1 += Dx;
    b2 += Dy;
    b3 += Dz;
  }
  const double *D = dmat + 9 * (size_t)c;
  dphi[3 * (size_t)c] = (b1 * D[0] - b2 * D[1]) + b3 * D[2];
  if (compat)
    dphi[3 * (size_t)c + 1] = (b1 * D[3] - b2 * D[4]) - b3 * D[5];
  else
    dphi[3 * (size_t)c + 1] = (-(b1 * D[4]) + b2 * D[3]) - b3 * D[5];
  dphi[3 * (size_t)c + 2] = (b1 * D[6] - b2 * D[7]) + b3 * D[8];
}

// grad_lsq_qr (:1063-1152): gradient = sum_k D(:,k) (phi_k - phi_P)
__global__ void k_grad_lsq_qr(int n, int Fi, int ntot, const int *__restrict__ cf_ptr,
                              const int *__restrict__ cf, const int *__restrict__ owner,
                              const int *__restrict__ neighbour, const double *__restrict__ dmat,
                              const double *__restrict__ phi, double *__restrict__ dphi) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ntot) return;
  if (c >= n) {
    dphi[3 * (size_t)c] = 0.0;
    dphi[3 * (size_t)c + 1] = 0.0;
    dphi[3 * (size_t)c + 2] = 0.0;
    return;
  }
  const double *D = dmat + 18 * (size_t)c;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  int l = 0;
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1] && l < 6; ++k, ++l) {
    const int f = cf[k];
    double b;
    if (f < Fi) {
      const int p = owner[f], o = (c == p) ? neighbour[f] : p;
      b = phi[o] - phi[c];
    } else {
      b = phi[n + (f - Fi)] - phi[c];
    }
    s0 = s0 + D[3 * l] * b;
    s1 = s1 + D[3 * l + 1] * b;
    s2 = s2 + D[3 * l + 2] * b;
  }
  dphi[3 * (size_t)c] = s0;
  dphi[3 * (size_t)c + 1] = s1;
  dphi[3 * (size_t)c + 2] = s2;
}

LsqGeom lsq_geom(fcp_ctx *ctx) {
  return LsqGeom{ctx->xc.p, ctx->yc.p, ctx->zc.p, ctx->xf.p, ctx->yf.p, ctx->zf.p};
}

}  // namespace

// Builds and keeps the per-cell least-squares matrices of `method`
// (1 lsq, 2 lsq_dm, 3 lsq_qr).
int fcp_launch_lsq_matrix(fcp_ctx *ctx, int method, int compat) {
  const int n = ctx->n;
  const int grid = (n + BLK - 1) / BLK;
  ProfScope ps(ctx, "lsq_matrix");
  if (method == 1 || method == 2) {
    FCP_HIP(ctx, ctx->lsq_dmat.alloc(9 * (size_t)n));
    k_lsq_dmat<<<grid, BLK, 0, ctx->stream>>>(n, ctx->Fi, method == 2, ctx->cf_ptr.p,
                                              ctx->cf_face.p, ctx->owner.p, ctx->neighbour.p,
                                              lsq_geom(ctx), ctx->lsq_dmat.p);
  } else if (method == 3) {
    FCP_HIP(ctx, ctx->lsq_dmat.alloc(18 * (size_t)n));
    FCP_HIP(ctx, hipMemsetAsync(ctx->errflag.p, 0, sizeof(int), ctx->stream));
    k_lsq_qr<<<grid, BLK, 0, ctx->stream>>>(n, ctx->Fi, ctx->cf_ptr.p, ctx->cf_face.p,
                                            ctx->owner.p, ctx->neighbour.p, lsq_geom(ctx),
                                            ctx->lsq_dmat.p, ctx->errflag.p);
    int e = 0;
    FCP_HIP(ctx, hipMemcpyAsync(&e, ctx->errflag.p, sizeof(int), hipMemcpyDeviceToHost,
                                ctx->stream));
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (e) return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "lsq_qr: a cell has more than 6 faces");
  } else {
    return fcp_fail(ctx, FCP_ERR_ARG, "unknown least-squares method %d", method);
  }
  FCP_HIP(ctx, hipGetLastError());
  ctx->lsq_method = method;
  ctx->lsq_compat = compat;
  return 0;
}

int fcp_launch_grad_lsq(fcp_ctx *ctx, const double *phi, double *dphi) {
  const int n = ctx->n;
  const int ntot = n + ctx->Fb;
  const int grid = (ntot + BLK - 1) / BLK;
  ProfScope ps(ctx, "grad_lsq");
  if (ctx->lsq_method == 1 || ctx->lsq_method == 2)
    k_grad_lsq<<<grid, BLK, 0, ctx->stream>>>(n, ctx->Fi, ntot, ctx->lsq_method == 2,
                                              ctx->lsq_compat, ctx->cf_ptr.p, ctx->cf_face.p,
                                              ctx->owner.p, ctx->neighbour.p, lsq_geom(ctx),
                                              ctx->lsq_dmat.p, phi, dphi);
  else if (ctx->lsq_method == 3)
    k_grad_lsq_qr<<<grid, BLK, 0, ctx->stream>>>(n, ctx->Fi, ntot, ctx->cf_ptr.p, ctx->cf_face.p,
                                                 ctx->owner.p, ctx->neighbour.p, ctx->lsq_dmat.p,
                                                 phi, dphi);
  else
    return fcp_fail(ctx, FCP_ERR_STATE, "least-squares matrix not created");
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}
