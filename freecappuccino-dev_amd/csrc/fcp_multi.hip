// fcp_multi.hip -- batched solves of NR right-hand sides that share the
// resident matrix (the three momentum components of a segregated solver
// share one convection-diffusion operator).
//
// Vectors are interleaved: row i holds its NR values at [i*NRP, i*NRP+NR)
// (NRP = NR rounded up to 2 or 4), so one gather of a matrix column fetches
// every right-hand side with a single 8/16/32-byte load, and the operator,
// its SELL copy and the packed sweep operands are read once per batched
// product or sweep instead of once per right-hand side.
//
// Every right-hand side runs its own Krylov recurrence with its own device
// scalars and convergence flag; per-row arithmetic and the block structure
// of every reduction are those of the single-vector kernels, so each column
// of a batched solve is bit-equal to the single-vector solve of that column.
// A converged column is frozen (its vectors are no longer updated) while the
// others continue.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "fcp_device.h"

using namespace fcpk;

namespace {

template <int NR>
struct Pad {
  static constexpr int v = NR <= 1 ? 1 : (NR <= 2 ? 2 : 4);
};

constexpr int MAXNR = 4;

template <int NR>
__device__ __forceinline__ bool all_done(const SolveScalars *s) {
#pragma unroll
  for (int r = 0; r < NR; ++r)
    if (!solve_done(s + r)) return false;
  return true;
}

template <int NR>
__device__ __forceinline__ void done_mask(const SolveScalars *s, bool (&d)[NR]) {
#pragma unroll
  for (int r = 0; r < NR; ++r) d[r] = solve_done(s + r);
}

// SELL row product for NR interleaved vectors, CSR entry order per column.
template <typename T, int NR, int NRP>
__device__ __forceinline__ void msell_row(const Sell<T> &m, int i, const T *__restrict__ x,
                                          T (&acc)[NR], bool subtract) {
  const int2 at = sell_at(m.base, m.width, i);
  const int p = at.x, L = at.y;
  T av[SELL_KU];
  int cv[SELL_KU];
#pragma unroll
  for (int k = 0; k < SELL_KU; ++k) {
    cv[k] = -1;  // past the row's end (padding slots hold -1 too)
    if (k < L) {
      av[k] = m.a[p + 64 * k];
      cv[k] = m.col[p + 64 * k];
    }
  }
  // per-column diagonal: column r's own value at the diagonal entry, in the
  // entry's place of the sum
  const T *dc = m.dcol ? m.dcol + (size_t)i * NRP : nullptr;
#pragma unroll
  for (int k = 0; k < SELL_KU; ++k)
    if (cv[k] >= 0) {
      const T *xc = x + (size_t)cv[k] * NRP;
      const bool dg = dc && cv[k] == i;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const T ak = dg ? dc[r] : av[k];
        acc[r] = subtract ? acc[r] - ak * xc[r] : acc[r] + ak * xc[r];
      }
    }
  for (int k = SELL_KU; k < L; ++k) {
    const int j = m.col[p + 64 * k];
    if (j < 0) break;
    const T a = m.a[p + 64 * k];
    const T *xc = x + (size_t)j * NRP;
    const bool dg = dc && j == i;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const T ak = dg ? dc[r] : a;
      acc[r] = subtract ? acc[r] - ak * xc[r] : acc[r] + ak * xc[r];
    }
  }
}

// ---------------------------------------------------------------------------
// residual, scaling and refinement kernels
// ---------------------------------------------------------------------------
// res = b - A x per column with L1 / L2 norms and |b|_2; initialises the
// column's scalars (linear_solvers.f90:417-438, :629-638).
template <int NR, int NRP>
__global__ __launch_bounds__(BLK) void mk_residual_init(
    int n, Sell<double> m, const double *__restrict__ x, const double *__restrict__ b,
    double *__restrict__ res, double *__restrict__ reso, double *__restrict__ partials,
    int *__restrict__ counter, SolveScalars *__restrict__ s, double tol_abs) {
  double part[3 * NR];
#pragma unroll
  for (int q = 0; q < 3 * NR; ++q) part[q] = 0.0;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = b[(size_t)i * NRP + r];
    msell_row<double, NR, NRP>(m, i, x, acc, true);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      res[(size_t)i * NRP + r] = acc[r];
      if (reso) reso[(size_t)i * NRP + r] = acc[r];
      const double bi = b[(size_t)i * NRP + r];
      part[3 * r] += fabs(acc[r]);
      part[3 * r + 1] += acc[r] * acc[r];
      part[3 * r + 2] += bi * bi;
    }
  }
  double blk[3 * NR], tot[3 * NR];
  block_sum<3 * NR>(part, blk);
  if (last_block_reduce<3 * NR>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      SolveScalars &S = s[r];
      S.res0 = tot[3 * r];
      S.resl = tot[3 * r];
      S.l2 = sqrt(tot[3 * r + 1]);
      S.l2b = sqrt(tot[3 * r + 2]);
      S.s0 = FCP_S0INIT;
      S.alf = 1.0;
      S.beto = 1.0;
      S.gam = 1.0;
      S.iter = 0;
      S.status = 0;
      S.factor = 0.0;
      S.resor = tot[3 * r];
      S.done = (tot[3 * r] < tol_abs) ? 1 : 0;  // :428-432
    }
  }
}

// r32 = res / |res|_2, e32 = 0 (columns whose residual is zero get zeros).
template <int NR, int NRP>
__global__ void mk_scale_to32(int n, const double *__restrict__ res,
                              const SolveScalars *__restrict__ so, float *__restrict__ r32,
                              float *__restrict__ e32) {
  double inv[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) inv[r] = so[r].l2 > 0.0 ? 1.0 / so[r].l2 : 0.0;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      r32[(size_t)i * NRP + r] = (float)(res[(size_t)i * NRP + r] * inv[r]);
      e32[(size_t)i * NRP + r] = 0.0f;
    }
  }
}

// Inner (FP32) scalars: norms of the scaled residual; a column whose outer
// residual already meets tol_rel starts (and stays) done.  Each column's
// inner tolerance is what its own refinement step needs (0.5 tol/rel,
// clamped to [inner_tol, 0.5]), the rule of the single-column solve.
template <int NR, int NRP>
__global__ __launch_bounds__(BLK) void mk_init32(int n, const float *__restrict__ r,
                                                 float *__restrict__ reso,
                                                 double *__restrict__ partials,
                                                 int *__restrict__ counter,
                                                 SolveScalars *__restrict__ s,
                                                 const SolveScalars *__restrict__ so,
                                                 double tol_rel, double inner_tol) {
  double part[2 * NR];
#pragma unroll
  for (int q = 0; q < 2 * NR; ++q) part[q] = 0.0;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      const float v = r[(size_t)i * NRP + c];
      part[2 * c] += fabs((double)v);
      part[2 * c + 1] += (double)v * (double)v;
      if (reso) reso[(size_t)i * NRP + c] = v;
    }
  }
  double blk[2 * NR], tot[2 * NR];
  block_sum<2 * NR>(part, blk);
  if (last_block_reduce<2 * NR>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      SolveScalars &S = s[r];
      S.res0 = tot[2 * r];
      S.resl = tot[2 * r];
      S.l2 = sqrt(tot[2 * r + 1]);
      S.l2b = S.l2;
      S.s0 = FCP_S0INIT;
      S.alf = 1.0;
      S.beto = 1.0;
      S.gam = 1.0;
      S.iter = 0;
      S.status = 0;
      S.factor = 0.0;
      const double rel = so[r].l2 / (so[r].l2b > 0 ? so[r].l2b : 1.0);
      double itol = 0.5 * tol_rel / rel;
      if (itol < inner_tol) itol = inner_tol;
      if (itol > 0.5) itol = 0.5;
      S.tol = itol;
      S.done = (rel <= tol_rel || so[r].l2 == 0.0) ? 1 : 0;
    }
  }
}

// x += |res|_2 * e  per column (FP64 refinement update)
template <int NR, int NRP>
__global__ void mk_ir_update(int n, const float *__restrict__ e32,
                             const SolveScalars *__restrict__ so, double *__restrict__ x) {
  double sc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) sc[r] = so[r].l2;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      x[(size_t)i * NRP + r] = x[(size_t)i * NRP + r] + sc[r] * (double)e32[(size_t)i * NRP + r];
  }
}

template <typename T>
__global__ void mk_zero2(size_t m, T *__restrict__ a, T *__restrict__ b) {
  for (size_t i = blockIdx.x * (size_t)BLK + threadIdx.x; i < m; i += (size_t)gridDim.x * BLK) {
    a[i] = (T)0;
    if (b) b[i] = (T)0;
  }
}

// column-major (ld) <-> interleaved copies of the caller's arrays
template <int NR, int NRP>
__global__ void mk_interleave(int n, const double *__restrict__ cm, int ld,
                              double *__restrict__ il) {
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
#pragma unroll
    for (int r = 0; r < NRP; ++r)
      il[(size_t)i * NRP + r] = r < NR ? cm[(size_t)r * ld + i] : 0.0;
  }
}
template <int NR, int NRP>
__global__ void mk_deinterleave(int n, const double *__restrict__ il, double *__restrict__ cm,
                                int ld) {
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
#pragma unroll
    for (int r = 0; r < NR; ++r) cm[(size_t)r * ld + i] = il[(size_t)i * NRP + r];
  }
}

// a copy of the matrix with the diagonal of one column
__global__ void mk_set_diag(int n, const int *__restrict__ diag, const double *__restrict__ dv,
                            double *__restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[diag[i]] = dv[i];
}

// one column of an interleaved vector <-> a contiguous vector
template <typename T, int NRP>
__global__ void mk_column(int n, const T *__restrict__ src, int sstride, int soff,
                          T *__restrict__ dst, int dstride, int doff) {
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK)
    dst[(size_t)i * dstride + doff] = src[(size_t)i * sstride + soff];
}

// ---------------------------------------------------------------------------
// PCG step kernels (iccg :454-537) per column
// ---------------------------------------------------------------------------
template <typename T, int NR, int NRP>
__global__ __launch_bounds__(BLK) void mk_dot_rz(int n, const T *__restrict__ r,
                                                 const T *__restrict__ z,
                                                 double *__restrict__ partials,
                                                 int *__restrict__ counter,
                                                 SolveScalars *__restrict__ s) {
  if (all_done<NR>(s)) return;
  double part[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) part[q] = 0.0;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
#pragma unroll
    for (int q = 0; q < NR; ++q)
      part[q] += (double)r[(size_t)i * NRP + q] * (double)z[(size_t)i * NRP + q];
  }
  double blk[NR], tot[NR];
  block_sum<NR>(part, blk);
  if (last_block_reduce<NR>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      if (s[q].done) continue;
      s[q].sk = tot[q];
      s[q].bet = tot[q] / s[q].s0;
    }
  }
}

template <typename T, int NR, int NRP>
__global__ void mk_pupdate(int n, const T *__restrict__ z, T *__restrict__ p,
                           const SolveScalars *__restrict__ s) {
  if (all_done<NR>(s)) return;
  bool dn[NR];
  done_mask<NR>(s, dn);
  T bet[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) bet[q] = (T)s[q].bet;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
#pragma unroll
    for (int q = 0; q < NR; ++q)
      if (!dn[q]) p[(size_t)i * NRP + q] = z[(size_t)i * NRP + q] + bet[q] * p[(size_t)i * NRP + q];
  }
}

template <typename T, int NR, int NRP>
__global__ __launch_bounds__(BLK) void mk_spmv_pq(int n, Sell<T> m, const T *__restrict__ p,
                                                  T *__restrict__ q, double *__restrict__ partials,
                                                  int *__restrict__ counter,
                                                  SolveScalars *__restrict__ s) {
  if (all_done<NR>(s)) return;
  double part[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) part[r] = 0.0;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    T acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = (T)0;
    msell_row<T, NR, NRP>(m, i, p, acc, false);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      q[(size_t)i * NRP + r] = acc[r];
      part[r] += (double)p[(size_t)i * NRP + r] * (double)acc[r];
    }
  }
  double blk[NR], tot[NR];
  block_sum<NR>(part, blk);
  if (last_block_reduce<NR>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (s[r].done) continue;
      s[r].pkapk = tot[r];
      s[r].alf = s[r].sk / tot[r];
    }
  }
}

// x += alf p ; r -= alf q ; norms ; iteration-1 factor ; convergence per column
template <typename T, int NR, int NRP>
__global__ __launch_bounds__(BLK) void mk_xr_update(
    int n, const int *__restrict__ diag, const T *__restrict__ a, T *__restrict__ x,
    T *__restrict__ r, const T *__restrict__ p, const T *__restrict__ q,
    double *__restrict__ partials, int *__restrict__ counter, SolveScalars *__restrict__ s,
    int crit, double tol_abs, double tol_rel, int itr_max, const T *__restrict__ dcol) {
  if (all_done<NR>(s)) return;
  bool dn[NR], first[NR];
  T alf[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    dn[c] = solve_done(s + c);
    first[c] = (s[c].iter == 0);
    alf[c] = (T)s[c].alf;
  }
  double part[3 * NR];
#pragma unroll
  for (int c = 0; c < 3 * NR; ++c) part[c] = 0.0;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    const double ad0 = dcol ? 0.0 : (double)a[diag[i]];
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      const size_t o = (size_t)i * NRP + c;
      if (dn[c]) continue;
      const double ad = dcol ? (double)dcol[o] : ad0;
      const T xi = x[o] + alf[c] * p[o];
      x[o] = xi;
      const T ri = r[o] - alf[c] * q[o];
      r[o] = ri;
      part[3 * c] += fabs((double)ri);
      part[3 * c + 1] += (double)ri * (double)ri;
      if (first[c]) part[3 * c + 2] += fabs(ad * (double)xi);
    }
  }
  double blk[3 * NR], tot[3 * NR];
  block_sum<3 * NR>(part, blk);
  if (last_block_reduce<3 * NR>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      SolveScalars &S = s[c];
      if (S.done) continue;
      const double resl = tot[3 * c];
      S.resl = resl;
      S.l2 = sqrt(tot[3 * c + 1]);
      S.s0 = S.sk;
      S.iter += 1;
      if (first[c]) {
        S.factor = tot[3 * c + 2] + FCP_SMALL;
        S.resor = S.res0 / S.factor;
      }
      const double rsm = resl / (S.res0 + FCP_SMALL);
      S.rsm = rsm;
      bool conv;
      const double tr = tol_rel < 0.0 ? S.tol : tol_rel;  // < 0: per-column tolerance
      if (crit == 0)
        conv = (rsm < tr) || (resl < tol_abs);
      else
        conv = (S.l2 <= tr * S.l2b) || (resl < tol_abs);
      if (!isfinite(resl)) {
        S.status = FCP_ERR_NONFINITE;
        conv = true;
      }
      S.conv = (conv && S.status == 0) ? 1 : 0;
      if (conv || S.iter >= itr_max) S.done = 1;
    }
  }
}

// ---------------------------------------------------------------------------
// BiCGStab step kernels (bicgstab :642-773) per column
// ---------------------------------------------------------------------------
template <typename T, int NR, int NRP>
__global__ __launch_bounds__(BLK) void mk_bicg_bet(int n, const T *__restrict__ r,
                                                   const T *__restrict__ reso,
                                                   double *__restrict__ partials,
                                                   int *__restrict__ counter,
                                                   SolveScalars *__restrict__ s) {
  if (all_done<NR>(s)) return;
  double part[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) part[c] = 0.0;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
#pragma unroll
    for (int c = 0; c < NR; ++c)
      part[c] += (double)r[(size_t)i * NRP + c] * (double)reso[(size_t)i * NRP + c];
  }
  double blk[NR], tot[NR];
  block_sum<NR>(part, blk);
  if (last_block_reduce<NR>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      SolveScalars &S = s[c];
      if (S.done) continue;
      const double bet = tot[c];
      S.sk = bet;
      S.om = bet * S.gam / (S.alf * S.beto + FCP_SMALL);
      S.beto = bet;
    }
  }
}

template <typename T, int NR, int NRP>
__global__ void mk_bicg_p(int n, const T *__restrict__ r, const T *__restrict__ u,
                          T *__restrict__ p, const SolveScalars *__restrict__ s) {
  if (all_done<NR>(s)) return;
  bool dn[NR];
  done_mask<NR>(s, dn);
  T om[NR], alf[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    om[c] = (T)s[c].om;
    alf[c] = (T)s[c].alf;
  }
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      const size_t o = (size_t)i * NRP + c;
      if (!dn[c]) p[o] = r[o] + om[c] * (p[o] - alf[c] * u[o]);
    }
  }
}

template <typename T, int NR, int NRP>
__global__ __launch_bounds__(BLK) void mk_bicg_u(int n, Sell<T> m, const T *__restrict__ z,
                                                 const T *__restrict__ reso, T *__restrict__ u,
                                                 double *__restrict__ partials,
                                                 int *__restrict__ counter,
                                                 SolveScalars *__restrict__ s) {
  if (all_done<NR>(s)) return;
  bool dn[NR];
  done_mask<NR>(s, dn);
  double part[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) part[c] = 0.0;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    T acc[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) acc[c] = (T)0;
    msell_row<T, NR, NRP>(m, i, z, acc, false);
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      const size_t o = (size_t)i * NRP + c;
      if (!dn[c]) u[o] = acc[c];
      part[c] += (double)acc[c] * (double)reso[o];
    }
  }
  double blk[NR], tot[NR];
  block_sum<NR>(part, blk);
  if (last_block_reduce<NR>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      if (s[c].done) continue;
      s[c].ukreso = tot[c];
      s[c].gam = s[c].sk / tot[c];
    }
  }
}

template <typename T, int NR, int NRP>
__global__ void mk_bicg_xr1(int n, const T *__restrict__ z, const T *__restrict__ u,
                            T *__restrict__ x, T *__restrict__ r,
                            const SolveScalars *__restrict__ s) {
  if (all_done<NR>(s)) return;
  bool dn[NR];
  done_mask<NR>(s, dn);
  T gam[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) gam[c] = (T)s[c].gam;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      const size_t o = (size_t)i * NRP + c;
      if (dn[c]) continue;
      x[o] = x[o] + gam[c] * z[o];
      r[o] = r[o] - gam[c] * u[o];
    }
  }
}

template <typename T, int NR, int NRP>
__global__ __launch_bounds__(BLK) void mk_bicg_v(int n, Sell<T> m, const T *__restrict__ z,
                                                 const T *__restrict__ r, T *__restrict__ v,
                                                 double *__restrict__ partials,
                                                 int *__restrict__ counter,
                                                 SolveScalars *__restrict__ s) {
  if (all_done<NR>(s)) return;
  bool dn[NR];
  done_mask<NR>(s, dn);
  double part[2 * NR];
#pragma unroll
  for (int c = 0; c < 2 * NR; ++c) part[c] = 0.0;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    T acc[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) acc[c] = (T)0;
    msell_row<T, NR, NRP>(m, i, z, acc, false);
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      const size_t o = (size_t)i * NRP + c;
      if (!dn[c]) v[o] = acc[c];
      part[2 * c] += (double)acc[c] * (double)r[o];
      part[2 * c + 1] += (double)acc[c] * (double)acc[c];
    }
  }
  double blk[2 * NR], tot[2 * NR];
  block_sum<2 * NR>(part, blk);
  if (last_block_reduce<2 * NR>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      if (s[c].done) continue;
      s[c].vr = tot[2 * c];
      s[c].vv = tot[2 * c + 1];
      s[c].alf = tot[2 * c] / (tot[2 * c + 1] + FCP_SMALL);
    }
  }
}

// ---------------------------------------------------------------------------
// batched pipelined row-block sweep (DIC/DILU apply on NR columns)
// ---------------------------------------------------------------------------
// Same block schedule, packing and per-row arithmetic as the single-vector
// pipelined sweep; a row's NR values are kept together in LDS, published
// with one vector store and fetched with one vector load (L2-coherent
// buffer accesses, each value individually sentinel-checked).
// NRP values of T at byte offset off (4 to 32 bytes); CP = cache policy
// (CPOL_SC1 for values other workgroups publish during the kernel).
template <typename T, int NRP, int CP = CPOL_SC1>
__device__ __forceinline__ void row_load(__amdgpu_buffer_rsrc_t rs, int off, T (&v)[NRP]) {
  constexpr int BYTES = NRP * (int)sizeof(T);
  if constexpr (BYTES <= 16) {
    struct W {
      T t[NRP];
    };
    const auto u = RawVec<BYTES>::template load<CP>(rs, off);
    const W w = __builtin_bit_cast(W, u);
#pragma unroll
    for (int r = 0; r < NRP; ++r) v[r] = w.t[r];
  } else {
    struct W {
      T t[NRP / 2];
    };
    const auto u0 = RawVec<16>::template load<CP>(rs, off);
    const auto u1 = RawVec<16>::template load<CP>(rs, off + 16);
    const W w0 = __builtin_bit_cast(W, u0), w1 = __builtin_bit_cast(W, u1);
#pragma unroll
    for (int r = 0; r < NRP / 2; ++r) {
      v[r] = w0.t[r];
      v[r + NRP / 2] = w1.t[r];
    }
  }
}

template <typename T, int NRP>
__device__ __forceinline__ void row_store(__amdgpu_buffer_rsrc_t rs, int off, const T (&v)[NRP]) {
  constexpr int BYTES = NRP * (int)sizeof(T);
  if constexpr (BYTES <= 16) {
    struct W {
      T t[NRP];
    };
    W w;
#pragma unroll
    for (int r = 0; r < NRP; ++r) w.t[r] = v[r];
    RawVec<BYTES>::store(__builtin_bit_cast(typename RawVec<BYTES>::U, w), rs, off);
  } else {
    struct W {
      T t[NRP / 2];
    };
    W w0, w1;
#pragma unroll
    for (int r = 0; r < NRP / 2; ++r) {
      w0.t[r] = v[r];
      w1.t[r] = v[r + NRP / 2];
    }
    RawVec<16>::store(__builtin_bit_cast(typename RawVec<16>::U, w0), rs, off);
    RawVec<16>::store(__builtin_bit_cast(typename RawVec<16>::U, w1), rs, off + 16);
  }
}

template <typename T, int NR, int NRP>
struct MSlot {
  int row, cnt, ok, pos;
  T dv[NRP];  // the row's pivot, per column
  T xin[NRP];
  int dep[FCP_PMAXD];
  T av[FCP_PMAXD];
  T rv[FCP_PMAXD][NRP];
};

template <typename T>
struct MArgs {
  const int *prc, *pdep;
  const T *pa, *pdx, *pin;
  int n;
  // per-column pivots in processing order ([n*NRP]; null: the shared ones)
  const T *pdc = nullptr;
};

template <typename T, int NRP>
__device__ __forceinline__ void m_fetch(const MArgs<T> &P, __amdgpu_buffer_rsrc_t rin, int idx,
                                        int &row, int &cnt, T (&dvs)[NRP], T (&xin)[NRP],
                                        int *dep, T *av) {
  T dv;
  const int2 rc = reinterpret_cast<const int2 *>(P.prc)[idx];
  row = rc.x;
  cnt = rc.y & 0xFFFF;
  const int4 dp = reinterpret_cast<const int4 *>(P.pdep)[idx];
  dep[0] = dp.x;
  dep[1] = dp.y;
  dep[2] = dp.z;
  dep[3] = dp.w;
  if (sizeof(T) == 4) {
    const float4 a4 = reinterpret_cast<const float4 *>(P.pa)[idx];
    av[0] = (T)a4.x;
    av[1] = (T)a4.y;
    av[2] = (T)a4.z;
    av[3] = (T)a4.w;
    dv = (T)reinterpret_cast<const float2 *>(P.pdx)[idx].x;
  } else {
    const double2 a0 = reinterpret_cast<const double2 *>(P.pa)[2 * idx];
    const double2 a1 = reinterpret_cast<const double2 *>(P.pa)[2 * idx + 1];
    av[0] = (T)a0.x;
    av[1] = (T)a0.y;
    av[2] = (T)a1.x;
    av[3] = (T)a1.y;
    dv = (T)reinterpret_cast<const double2 *>(P.pdx)[idx].x;
  }
#pragma unroll
  for (int r = 0; r < NRP; ++r) dvs[r] = P.pdc ? P.pdc[(size_t)idx * NRP + r] : dv;
  row_load<T, NRP, 0>(rin, idx * NRP * (int)sizeof(T), xin);
}

template <typename T, int NR, int NRP>
__device__ __forceinline__ void m_load1(MSlot<T, NR, NRP> &S, const MArgs<T> &P,
                                        __amdgpu_buffer_rsrc_t rin, const int *sptr, int step,
                                        int ns) {
  const int sc = step < ns ? step : ns - 1;
  const int idx = sptr[sc] + (int)threadIdx.x;
  const bool ok = (step < ns) && (idx < sptr[sc + 1]);
  S.ok = ok ? 1 : 0;
  S.pos = ok ? idx : 0;
  m_fetch<T, NRP>(P, rin, ok ? idx : 0, S.row, S.cnt, S.dv, S.xin, S.dep, S.av);
}

// Values of dependencies in earlier blocks; lanes without one pass an
// out-of-range offset and issue no memory request.
template <typename T, int NR, int NRP>
__device__ __forceinline__ void m_load2(MSlot<T, NR, NRP> &S, __amdgpu_buffer_rsrc_t rout) {
#pragma unroll
  for (int m = 0; m < FCP_PMAXD; ++m) {
    const bool need = S.ok && m < S.cnt && S.dep[m] < 0;
    row_load<T, NRP>(rout, need ? (-S.dep[m] - 1) * NRP * (int)sizeof(T) : FCP_OOB, S.rv[m]);
  }
}

template <typename T, int NR, int NRP, int MODE>
__device__ __forceinline__ void m_row_value(int cnt, const T (&dv)[NRP], const T (&xin)[NRP],
                                            const int *dep,
                                            const T *av, const T (*rv)[NRP], const T *zl, T *out,
                                            int *err, int slp, T (&val)[NRP]) {
  using B = Bits<T>;
  T acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = (MODE == 0) ? xin[r] : xin[r] / (dv[r] + (T)FCP_SMALL);
#pragma unroll
  for (int m = 0; m < FCP_PMAXD; ++m) {
    if (m < cnt) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        T v;
        if (dep[m] >= 0)
          v = zl[dep[m] * NRP + r];
        else if (rv && B::to(rv[m][r]) != B::sent)
          v = rv[m][r];
        else
          v = wait_value<T>(out + (size_t)(-dep[m] - 1) * NRP + r, err, slp);
        acc[r] = acc[r] - av[m] * v;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < NRP; ++r) {
    T o = r < NR ? acc[r] * dv[r] : (T)0;
    if (B::to(o) == B::sent) o = B::from(B::qnan);
    val[r] = o;
  }
}

template <typename T, int NR, int NRP, int MODE>
__device__ __forceinline__ void m_compute(const MSlot<T, NR, NRP> &S, const MArgs<T> &P,
                                          __amdgpu_buffer_rsrc_t rin,
                                          __amdgpu_buffer_rsrc_t rout, const int *sptr, int step,
                                          int ns, int lo, T *zl, T *out, int *err, int slp) {
  T val[NRP];
  if (S.ok) {
    m_row_value<T, NR, NRP, MODE>(S.cnt, S.dv, S.xin, S.dep, S.av, S.rv, zl, out, err, slp, val);
#pragma unroll
    for (int r = 0; r < NRP; ++r) zl[(S.pos - lo) * NRP + r] = val[r];
    row_store<T, NRP>(rout, S.row * NRP * (int)sizeof(T), val);
  }
  if (step >= ns) return;
  const int e = sptr[step + 1];
  for (int idx = sptr[step] + (int)threadIdx.x + BLK; idx < e; idx += BLK) {
    int row, cnt, dep[FCP_PMAXD];
    T dv[NRP], xin[NRP], av[FCP_PMAXD];
    m_fetch<T, NRP>(P, rin, idx, row, cnt, dv, xin, dep, av);
    m_row_value<T, NR, NRP, MODE>(cnt, dv, xin, dep, av, nullptr, zl, out, err, slp, val);
#pragma unroll
    for (int r = 0; r < NRP; ++r) zl[(idx - lo) * NRP + r] = val[r];
    row_store<T, NRP>(rout, row * NRP * (int)sizeof(T), val);
  }
}

// Prefetch depths: each step issues 5-7 operand loads, 4-8 remote loads and
// 1-2 stores; the ring stays under 63 outstanding vector memory operations.
template <typename T, int NRP>
struct MDepth {
  static constexpr int bytes = NRP * (int)sizeof(T);
  static constexpr int PD = bytes <= 16 ? 4 : 3;
  static constexpr int PD2 = bytes <= 16 ? 2 : 1;
};

template <typename T, int NR, int NRP, int MODE>
__global__ __launch_bounds__(BLK, 1) void mk_psweep(MArgs<T> P, const int *__restrict__ step_ptr,
                                                    const int *__restrict__ blk_step,
                                                    const int *__restrict__ blk_pos, int nblocks,
                                                    int sptr_off, T *out,
                                                    int *__restrict__ ticket,
                                                    int *__restrict__ err,
                                                    const SolveScalars *__restrict__ s, int slp) {
  constexpr int PD = MDepth<T, NRP>::PD, PD2 = MDepth<T, NRP>::PD2;
  extern __shared__ unsigned char smem_raw[];
  T *zl = reinterpret_cast<T *>(smem_raw);
  int *sptr = reinterpret_cast<int *>(smem_raw + sptr_off);
  __shared__ int sb;
  if (s && all_done<NR>(s)) return;
  if (threadIdx.x == 0) sb = atomicAdd(ticket, 1);
  __syncthreads();
  const int t = sb;
  if (t >= nblocks) return;
  const int b = (MODE == 1) ? nblocks - 1 - t : t;
  const int lo = blk_pos[b];
  const int st0 = blk_step[b];
  const int ns = blk_step[b + 1] - st0;
  for (int i = threadIdx.x; i <= ns; i += BLK) sptr[i] = step_ptr[st0 + i];
  __syncthreads();
  const int vbytes = P.n * NRP * (int)sizeof(T);
  const __amdgpu_buffer_rsrc_t rin =
      __builtin_amdgcn_make_buffer_rsrc((void *)P.pin, (short)0, vbytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t rout =
      __builtin_amdgcn_make_buffer_rsrc((void *)out, (short)0, vbytes, 0x00020000);
  MSlot<T, NR, NRP> S[PD];
#pragma unroll
  for (int j = 0; j < PD; ++j) m_load1<T, NR, NRP>(S[j], P, rin, sptr, j, ns);
#pragma unroll
  for (int j = 0; j < PD2; ++j) m_load2<T, NR, NRP>(S[j], rout);
  for (int base = 0; base < ns; base += PD) {
#pragma unroll
    for (int j = 0; j < PD; ++j) {
      const int step = base + j;
      m_compute<T, NR, NRP, MODE>(S[j], P, rin, rout, sptr, step, ns, lo, zl, out, err, slp);
      __syncthreads();
      m_load2<T, NR, NRP>(S[(j + PD2) % PD], rout);
      m_load1<T, NR, NRP>(S[j], P, rin, sptr, step + PD, ns);
    }
  }
}

// packed inputs in processing order: pin[idx] = in[row] (NRP values)
template <typename T, int NR, int NRP>
__global__ void mk_gather(int n, const int *__restrict__ prc, const T *__restrict__ in,
                          T *__restrict__ pin, const SolveScalars *__restrict__ s) {
  if (s && all_done<NR>(s)) return;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int row = prc[2 * idx];
#pragma unroll
  for (int r = 0; r < NRP; ++r) pin[(size_t)idx * NRP + r] = in[(size_t)row * NRP + r];
}

template <typename T, int NR>
__global__ void mk_fill_sent(T *__restrict__ x, T *__restrict__ y, size_t m,
                             int *__restrict__ tickets, const SolveScalars *__restrict__ s) {
  if (s && all_done<NR>(s)) return;
  using B = Bits<T>;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i == 0 && tickets) {
    tickets[0] = 0;
    tickets[1] = 0;
  }
  for (size_t k = i; k < m; k += (size_t)gridDim.x * blockDim.x) {
    reinterpret_cast<typename B::U *>(x)[k] = B::sent;
    if (y) reinterpret_cast<typename B::U *>(y)[k] = B::sent;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T>
int pack_of(fcp_ctx *ctx, BlockSchedule &b);
template <>
int pack_of<double>(fcp_ctx *ctx, BlockSchedule &b) {
  return pack_block64(ctx, b);
}
template <>
int pack_of<float>(fcp_ctx *ctx, BlockSchedule &b) {
  return pack_block32(ctx, b);
}
template <typename T>
const T *pa_of(BlockSchedule &b);
template <>
const double *pa_of<double>(BlockSchedule &b) {
  return b.pa64.p;
}
template <>
const float *pa_of<float>(BlockSchedule &b) {
  return b.pa32.p;
}
template <typename T>
const T *pdx_of(BlockSchedule &b);
template <>
const double *pdx_of<double>(BlockSchedule &b) {
  return b.pdx64.p;
}
template <>
const float *pdx_of<float>(BlockSchedule &b) {
  return b.pdx32.p;
}

// Batched sweeps apply when every row has at most FCP_PMAXD dependencies
// and the block's values plus its step table fit the workgroup LDS: the
// geometric tiles when they fit, else contiguous blocks sized for NRP values
// per row.  *lo/*up/*pin receive the schedules and the input staging buffer.
template <typename T, int NRP>
int msched_for(fcp_ctx *ctx, BlockSchedule **lo, BlockSchedule **up, DBuf<char> **pin) {
  *lo = *up = nullptr;
  if (ctx->tune_sweep_kind != 0) return 0;
  // interleaved vectors are read through a buffer resource of n*NRP*sizeof(T) bytes
  if ((long long)ctx->n * NRP * (long long)sizeof(T) >= FCP_OOB) return 0;
  if (ctx->have_tiles && sched_fits(ctx, ctx->tlower, ctx->tupper, NRP * sizeof(T))) {
    *lo = &ctx->tlower;
    *up = &ctx->tupper;
    *pin = &ctx->tpin;
    return 0;
  }
  const int Rlo = fcp_block_rows(ctx, NRP * (int)sizeof(T), 0);
  const int Rup = fcp_block_rows(ctx, NRP * (int)sizeof(T), 1);
  const long long key = ((long long)Rlo << 20) | Rup;
  auto it = ctx->msched.find(key);
  if (it == ctx->msched.end()) {
    std::unique_ptr<MultiSched> m(new MultiSched());
    FCP_TRY(fcp_build_block_schedule(ctx, ctx->lower, Rlo, 0, m->lo));
    FCP_TRY(fcp_build_block_schedule(ctx, ctx->upper, Rup, 1, m->up));
    it = ctx->msched.emplace(key, std::move(m)).first;
  }
  MultiSched *m = it->second.get();
  if (!sched_fits(ctx, m->lo, m->up, NRP * sizeof(T))) return 0;
  *lo = &m->lo;
  *up = &m->up;
  *pin = &m->pin;
  return 0;
}

// z = M^{-1} r for NR interleaved columns (batched sweeps, or one column at a
// time through the single-vector sweeps when batching does not apply).
template <typename T, int NR, int NRP>
int mprecond_apply(fcp_ctx *ctx, const T *r, T *y, T *z, T *tmp3, const SolveScalars *s) {
  const int n = ctx->n;
  BlockSchedule *blo = nullptr, *bup = nullptr;
  DBuf<char> *pbuf = nullptr;
  // FP32 sweeps of the tile-local preconditioner: operands already packed with
  // its pivots, and no dependency crosses a block (no sentinel fill needed)
  const bool bj = sizeof(T) == 4 && ctx->bj_active;
  if (bj) {
    FCP_TRY(bj_pack32(ctx));
    if ((long long)n * NRP * (long long)sizeof(T) < FCP_OOB &&
        sched_fits(ctx, ctx->jlower, ctx->jupper, NRP * sizeof(T))) {
      blo = &ctx->jlower;
      bup = &ctx->jupper;
      pbuf = &ctx->jpin;
    }
  } else {
    FCP_TRY((msched_for<T, NRP>(ctx, &blo, &bup, &pbuf)));
  }
  const bool pcol = sizeof(T) == 8 && ctx->mdc_active;  // pivots of their own per column
  if (blo && (size_t)n * NRP * sizeof(T) < ((size_t)1 << 31)) {
    const size_t m = (size_t)n * NRP;
    FCP_HIP(ctx, pbuf->alloc(m * sizeof(T)));
    if (bj)
      mk_fill_sent<T, NR><<<1, 1, 0, ctx->stream>>>(y, z, 0, tickets(ctx), s);
    else
      mk_fill_sent<T, NR><<<red_grid(ctx, n), BLK, 0, ctx->stream>>>(y, z, m, tickets(ctx), s);
    for (int u = 0; u < 2; ++u) {
      BlockSchedule &b = u ? *bup : *blo;
      if (!bj) FCP_TRY(pack_of<T>(ctx, b));
      T *pin = reinterpret_cast<T *>(pbuf->p);
      const T *in = u ? y : r;
      T *out = u ? z : y;
      mk_gather<T, NR, NRP><<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(n, b.prc.p, in, pin, s);
      MArgs<T> P{b.prc.p, b.pdep.p, pa_of<T>(b), pdx_of<T>(b), pin, n};
      if constexpr (sizeof(T) == 8) {
        if (pcol) {  // the columns' pivots in this schedule's processing order
          if (ctx->mpdc_of[u] != &b) {
            FCP_HIP(ctx, ctx->mpdc[u].alloc(m));
            mk_gather<T, NR, NRP><<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(
                n, b.prc.p, ctx->mpiv.p, ctx->mpdc[u].p, nullptr);
            ctx->mpdc_of[u] = &b;
          }
          P.pdc = ctx->mpdc[u].p;
        }
      }
      const int sptr_off = (int)(((size_t)b.R * NRP * sizeof(T) + 15) & ~(size_t)15);
      const size_t lds = (size_t)sptr_off + (size_t)(b.max_steps + 2) * sizeof(int);
      ProfScope ps(ctx, u ? (sizeof(T) == 8 ? "msweep_bwd_f64" : "msweep_bwd_f32")
                          : (sizeof(T) == 8 ? "msweep_fwd_f64" : "msweep_fwd_f32"));
      if (u)
        mk_psweep<T, NR, NRP, 1><<<b.nblocks, BLK, lds, ctx->stream>>>(
            P, b.step_ptr.p, b.blk_step.p, b.blk_pos.p, b.nblocks, sptr_off, out,
            tickets(ctx) + 1, ctx->errflag.p, s, ctx->tune_sweep_sleep);
      else
        mk_psweep<T, NR, NRP, 0><<<b.nblocks, BLK, lds, ctx->stream>>>(
            P, b.step_ptr.p, b.blk_step.p, b.blk_pos.p, b.nblocks, sptr_off, out, tickets(ctx),
            ctx->errflag.p, s, ctx->tune_sweep_sleep);
      FCP_HIP(ctx, hipGetLastError());
    }
    return 0;
  }
  if (pcol)
    return fcp_fail(ctx, FCP_ERR_UNSUPPORTED,
                    "per-column diagonals need the batched sweeps (schedule does not fit)");
  // one column at a time: tmp3 holds 3 contiguous vectors of n
  T *cr = tmp3, *cy = tmp3 + n, *cz = tmp3 + 2 * (size_t)n;
  for (int c = 0; c < NR; ++c) {
    mk_column<T, NRP><<<red_grid(ctx, n), BLK, 0, ctx->stream>>>(n, r, NRP, c, cr, 1, 0);
    if (sizeof(T) == 8)
      FCP_TRY(precond_apply64(ctx, (const double *)cr, (double *)cy, (double *)cz));
    else
      FCP_TRY(precond_apply32(ctx, (const float *)cr, (float *)cy, (float *)cz));
    mk_column<T, NRP><<<red_grid(ctx, n), BLK, 0, ctx->stream>>>(n, cz, 1, 0, z, NRP, c);
  }
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

template <typename T>
int mview(fcp_ctx *ctx, Sell<T> *m);
template <>
int mview<double>(fcp_ctx *ctx, Sell<double> *m) {
  FCP_TRY(sell_view64(ctx, m));
  if (ctx->mdc_active) m->dcol = ctx->mdcol.p;  // the columns' own diagonals
  return 0;
}
template <>
int mview<float>(fcp_ctx *ctx, Sell<float> *m) {
  return sell_view32(ctx, m);
}

// Krylov vectors of one precision, NRP values per row.
template <typename T>
struct MVecs {
  T *x, *r, *reso, *p, *z, *y, *u, *v, *tmp3;
};

template <typename T>
const T *mdcol_of(fcp_ctx *ctx) {
  if constexpr (sizeof(T) == 8)
    return ctx->mdc_active ? ctx->mdcol.p : nullptr;
  else
    return nullptr;
}

template <typename T, int NR, int NRP>
int mpcg_iteration(fcp_ctx *ctx, int precond, const T *a, MVecs<T> &V, SolveScalars *s, int crit,
                   double tol_abs, double tol_rel, int itr_max) {
  const int n = ctx->n;
  const int g = red_grid(ctx, n);
  Sell<T> m;
  FCP_TRY(mview<T>(ctx, &m));
  FCP_TRY((mprecond_apply<T, NR, NRP>(ctx, V.r, V.y, V.z, V.tmp3, s)));
  mk_dot_rz<T, NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, V.r, V.z, ctx->partials.p, redctr(ctx), s);
  mk_pupdate<T, NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, V.z, V.p, s);
  mk_spmv_pq<T, NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, m, V.p, V.u, ctx->partials.p, redctr(ctx),
                                                      s);
  mk_xr_update<T, NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, ctx->diag.p, a, V.x, V.r, V.p, V.u,
                                                        ctx->partials.p, redctr(ctx), s, crit,
                                                        tol_abs, tol_rel, itr_max, mdcol_of<T>(ctx));
  (void)precond;
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

template <typename T, int NR, int NRP>
int mbicgstab_iteration(fcp_ctx *ctx, const T *a, MVecs<T> &V, SolveScalars *s, int crit,
                        double tol_abs, double tol_rel, int itr_max) {
  const int n = ctx->n;
  const int g = red_grid(ctx, n);
  Sell<T> m;
  FCP_TRY(mview<T>(ctx, &m));
  mk_bicg_bet<T, NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, V.r, V.reso, ctx->partials.p,
                                                       redctr(ctx), s);
  mk_bicg_p<T, NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, V.r, V.u, V.p, s);
  FCP_TRY((mprecond_apply<T, NR, NRP>(ctx, V.p, V.y, V.z, V.tmp3, s)));
  mk_bicg_u<T, NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, m, V.z, V.reso, V.u, ctx->partials.p,
                                                     redctr(ctx), s);
  mk_bicg_xr1<T, NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, V.z, V.u, V.x, V.r, s);
  FCP_TRY((mprecond_apply<T, NR, NRP>(ctx, V.r, V.y, V.z, V.tmp3, s)));
  mk_bicg_v<T, NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, m, V.z, V.r, V.v, ctx->partials.p,
                                                     redctr(ctx), s);
  mk_xr_update<T, NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, ctx->diag.p, a, V.x, V.r, V.z, V.v,
                                                        ctx->partials.p, redctr(ctx), s, crit,
                                                        tol_abs, tol_rel, itr_max, mdcol_of<T>(ctx));
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

// Launch iterations in batches, polling the NR device scalars in between.
template <int NR, typename F>
int miterate(fcp_ctx *ctx, SolveScalars *s_dev, SolveScalars *h, int itr_max, int check_every,
             int verbose, const char *tag, int mode, double tol, F &&one) {
  int launched = 0;
  BatchSizer bs(check_every, mode, tol);
  int next = bs.cap;
  while (launched < itr_max) {
    const int batch = (itr_max - launched) < next ? (itr_max - launched) : next;
    for (int k = 0; k < batch; ++k) FCP_TRY(one());
    launched += batch;
    FCP_TRY(poll_device(ctx, h, s_dev, NR * sizeof(SolveScalars)));
    bool all = true;
    for (int c = 0; c < NR; ++c) {
      if (verbose >= 2)
        printf("                   %s[%d] iter = %4d resl = %10.3E rsm = %10.3E\n", tag, c,
               h[c].iter, h[c].resl, h[c].rsm);
      all = all && h[c].done;
    }
    if (all) break;
    next = bs.next(h, NR);
  }
  return 0;
}

struct MultiResult {
  int iters[MAXNR], conv[MAXNR], outer;
  double res0[MAXNR], resl[MAXNR], resor[MAXNR];
};

// Batched solve; x, b interleaved FP64 device arrays (NRP per row).
template <int NR>
int msolve(fcp_ctx *ctx, int solver, double *x, const double *b, int itr_max, double tol_abs,
           double tol_rel, int prec, int crit, int check_every, int verbose, double inner_tol,
           int max_outer, MultiResult &res) {
  constexpr int NRP = Pad<NR>::v;
  const int n = ctx->n;
  const size_t m = (size_t)n * NRP;
  const int g = red_grid(ctx, n);
  const int kind = solver == 2 ? 1 : 0;
  // global pivots now for FP64 solves; a mixed solve computes them when a
  // sweep needs them (FP32 inner sweeps without the tile-local
  // preconditioner, FP64 finishing iterations)
  FCP_TRY(prepare_precond(ctx, kind, prec == FCP_PREC_MIXED, prec != FCP_PREC_MIXED));
  FCP_HIP(ctx, ctx->scal.alloc(4 + 2 * MAXNR));
  FCP_HIP(ctx, ctx->partials.alloc((size_t)MAXBLK * 16));
  SolveScalars *so = ctx->scal.p + 4;        // outer / FP64 columns
  SolveScalars *si = ctx->scal.p + 4 + MAXNR;  // inner FP32 columns
  SolveScalars h[MAXNR];
  FCP_HIP(ctx, hipMemsetAsync(ctx->counters.p, 0, CTR_INTS * sizeof(int), ctx->stream));
  FCP_HIP(ctx, hipMemsetAsync(ctx->errflag.p, 0, sizeof(int), ctx->stream));
  for (int k = 0; k < 8; ++k) FCP_HIP(ctx, ctx->mv64[k].alloc(k == 7 ? 3 * (size_t)n : m));
  MVecs<double> V64{x, ctx->mv64[0].p, ctx->mv64[1].p, ctx->mv64[2].p, ctx->mv64[3].p,
                    ctx->mv64[4].p, ctx->mv64[5].p, ctx->mv64[6].p, ctx->mv64[7].p};
  Sell<double> m64;
  FCP_TRY(mview<double>(ctx, &m64));
  for (int c = 0; c < MAXNR; ++c) {
    res.iters[c] = 0;
    res.conv[c] = 1;  // FP64 solves record their stopping test below
  }
  res.outer = 0;
  auto residual = [&](double *reso) {
    ProfScope ps(ctx, "mresidual_f64");
    mk_residual_init<NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, m64, x, b, V64.r, reso,
                                                          ctx->partials.p, redctr(ctx), so,
                                                          tol_abs);
    return hipGetLastError() == hipSuccess ? 0 : FCP_ERR_HIP;
  };
  if (prec != FCP_PREC_MIXED) {
    mk_zero2<double><<<g, BLK, 0, ctx->stream>>>(m, V64.p, V64.u);
    FCP_TRY(residual(solver == 2 ? V64.reso : nullptr));
    FCP_HIP(ctx, hipMemcpyAsync(h, so, NR * sizeof(SolveScalars), hipMemcpyDeviceToHost,
                                ctx->stream));
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < NR; ++c) res.res0[c] = h[c].res0;
    if (solver == 2) {
      auto it = [&]() {
        return mbicgstab_iteration<double, NR, NRP>(ctx, ctx->a.p, V64, so, crit, tol_abs,
                                                    tol_rel, itr_max);
      };
      FCP_TRY(miterate<NR>(ctx, so, h, itr_max, check_every, verbose, "BiCGStab",
                           crit == 0 ? BATCH_L1 : BATCH_L2, tol_rel, it));
    } else {
      auto it = [&]() {
        return mpcg_iteration<double, NR, NRP>(ctx, 0, ctx->a.p, V64, so, crit, tol_abs, tol_rel,
                                               itr_max);
      };
      FCP_TRY(miterate<NR>(ctx, so, h, itr_max, check_every, verbose, "PCG",
                           crit == 0 ? BATCH_L1 : BATCH_L2, tol_rel, it));
    }
    for (int c = 0; c < NR; ++c) {
      res.iters[c] = h[c].iter;
      res.resl[c] = h[c].resl;
      res.resor[c] = h[c].resor;
      res.conv[c] = (h[c].conv || h[c].res0 < tol_abs) ? 1 : 0;
      if (h[c].status < 0) return fcp_fail(ctx, h[c].status, "non-finite residual (column %d)", c);
    }
    res.outer = 1;
    return 0;
  }
  // FP32 inner Krylov on A e = r/|r| inside FP64 refinement, per column; the
  // inner sweeps use the tile-local preconditioner when selected (the FP64
  // refinement and the FP64 finish keep the global one)
  bool use_bj = false;
  if (ctx->inner_precond >= 1) FCP_TRY(bj_begin(ctx, kind, &use_bj));
  struct BjScope {
    fcp_ctx *c;
    ~BjScope() { c->bj_active = false; }
  } bj_scope{ctx};
  ctx->bj_active = use_bj;
  if (!use_bj) FCP_TRY(prepare_precond(ctx, kind, true, true));
  for (int k = 0; k < 9; ++k) FCP_HIP(ctx, ctx->mv32[k].alloc(k == 8 ? 3 * (size_t)n : m));
  MVecs<float> V32{ctx->mv32[0].p, ctx->mv32[1].p, ctx->mv32[2].p, ctx->mv32[3].p,
                   ctx->mv32[4].p, ctx->mv32[5].p, ctx->mv32[6].p, ctx->mv32[7].p,
                   ctx->mv32[8].p};
  FCP_TRY(residual(nullptr));
  int total = 0, outer = 0;
  double prev[MAXNR];
  for (; outer < max_outer; ++outer) {
    FCP_HIP(ctx, hipMemcpyAsync(h, so, NR * sizeof(SolveScalars), hipMemcpyDeviceToHost,
                                ctx->stream));
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    bool all = true, stall = false;
    for (int c = 0; c < NR; ++c) {
      const double rel = h[c].l2 / (h[c].l2b > 0 ? h[c].l2b : 1.0);
      if (outer == 0) res.res0[c] = h[c].res0;
      if (!std::isfinite(rel))
        return fcp_fail(ctx, FCP_ERR_NONFINITE, "non-finite residual in refinement (column %d)", c);
      if (verbose >= 2) printf("                    IR outer %d  column %d  |r|/|b| = %10.3E\n", outer, c, rel);
      if (!(rel <= tol_rel || h[c].l2 == 0.0)) {
        all = false;
        if (outer > 0 && rel > 0.5 * prev[c]) stall = true;
      }
      prev[c] = rel;
    }
    if (all || total >= itr_max) break;
    if (stall) {
      // finish every column with FP64 iterations, warm-started from x
      FCP_TRY(prepare_precond(ctx, kind, true, true));
      mk_zero2<double><<<g, BLK, 0, ctx->stream>>>(m, V64.p, V64.u);
      FCP_TRY(residual(solver == 2 ? V64.reso : nullptr));
      if (solver == 2) {
        auto it = [&]() {
          return mbicgstab_iteration<double, NR, NRP>(ctx, ctx->a.p, V64, so, 1, tol_abs, tol_rel,
                                                      itr_max - total);
        };
        FCP_TRY(miterate<NR>(ctx, so, h, itr_max - total, check_every, verbose, "BiCGStab",
                             BATCH_L2, tol_rel, it));
      } else {
        auto it = [&]() {
          return mpcg_iteration<double, NR, NRP>(ctx, 0, ctx->a.p, V64, so, 1, tol_abs, tol_rel,
                                                 itr_max - total);
        };
        FCP_TRY(miterate<NR>(ctx, so, h, itr_max - total, check_every, verbose, "PCG",
                             BATCH_L2, tol_rel, it));
      }
      int mx = 0;
      for (int c = 0; c < NR; ++c) {
        res.iters[c] += h[c].iter;
        mx = h[c].iter > mx ? h[c].iter : mx;
      }
      total += mx;
      FCP_TRY(residual(nullptr));
      continue;
    }
    if (use_bj && ctx->inner_precond == 1 && tin_usable(ctx, NRP)) {
      // inner FP32 PCG / BiCGStab in tile order (tile-local preconditioner), x updated
      if (solver == 2)
        FCP_TRY(tin_bicgstab(ctx, NR, V64.r, x, so, si, tol_rel, inner_tol, itr_max - total,
                             check_every, verbose, h));
      else
        FCP_TRY(tin_pcg(ctx, NR, V64.r, x, so, si, tol_rel, inner_tol, itr_max - total,
                        check_every, verbose, h));
      int mx = 0;
      for (int c = 0; c < NR; ++c) {
        res.iters[c] += h[c].iter;
        mx = h[c].iter > mx ? h[c].iter : mx;
      }
      total += mx;
      FCP_TRY(residual(nullptr));
      continue;
    }
    mk_scale_to32<NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, V64.r, so, V32.r, V32.x);
    mk_zero2<float><<<g, BLK, 0, ctx->stream>>>(m, V32.p, V32.u);
    mk_init32<NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, V32.r, solver == 2 ? V32.reso : nullptr,
                                                   ctx->partials.p, redctr(ctx), si, so, tol_rel,
                                                   inner_tol);
    const int budget = itr_max - total;
    const double itol = -1.0;  // per-column tolerances set by mk_init32
    if (solver == 2) {
      auto it = [&]() {
        return mbicgstab_iteration<float, NR, NRP>(ctx, ctx->a32.p, V32, si, 1, 0.0, itol, budget);
      };
      FCP_TRY(miterate<NR>(ctx, si, h, budget, check_every, verbose, "BiCGStab32",
                             BATCH_INNER, 0.0, it));
    } else {
      auto it = [&]() {
        return mpcg_iteration<float, NR, NRP>(ctx, 0, ctx->a32.p, V32, si, 1, 0.0, itol, budget);
      };
      FCP_TRY(miterate<NR>(ctx, si, h, budget, check_every, verbose, "PCG32",
                             BATCH_INNER, 0.0, it));
    }
    int mx = 0;
    for (int c = 0; c < NR; ++c) {
      res.iters[c] += h[c].iter;
      mx = h[c].iter > mx ? h[c].iter : mx;
    }
    total += mx;
    mk_ir_update<NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, V32.x, so, x);
    FCP_TRY(residual(nullptr));
  }
  res.outer = outer;
  FCP_HIP(ctx, hipMemcpyAsync(h, so, NR * sizeof(SolveScalars), hipMemcpyDeviceToHost,
                              ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int c = 0; c < NR; ++c) {
    res.resl[c] = h[c].res0;  // L1 of the final true residual
    res.resor[c] = h[c].res0;
  }
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// device entry: NR columns, column-major caller arrays (device pointers)
// ---------------------------------------------------------------------------
int fcp_solve_multi_device(fcp_ctx *ctx, int solver, int nrhs, double *fi, int ldfi,
                           const double *rhs, double *res0_out, int itr_max, double tol_abs,
                           double tol_rel, int verbose, fcp_solve_opts *opts,
                           const double *diagv) {
  FCP_TRY(fcp_ensure_workspace(ctx));
  if (nrhs < 1 || nrhs > MAXNR)
    return fcp_fail(ctx, FCP_ERR_ARG, "nrhs must be in [1, %d]", MAXNR);
  if (solver != 0 && solver != 2)
    return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "batched solves support iccg and bicgstab");
  const int n = ctx->n;
  const int prec = opts ? opts->precision : FCP_PREC_FP64;
  if (prec != FCP_PREC_FP64 && prec != FCP_PREC_MIXED)
    return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "batched solves run in FP64 or mixed precision");
  const int crit = opts ? opts->criterion : FCP_CRIT_REF_L1;
  const int ce = (opts && opts->check_every > 0) ? opts->check_every : 8;
  const double inner_tol = (opts && opts->inner_tol > 0) ? opts->inner_tol : 1e-4;
  const int max_outer = (opts && opts->max_outer > 0) ? opts->max_outer : 30;
  const int NRP = nrhs <= 1 ? 1 : (nrhs <= 2 ? 2 : 4);
  const size_t m = (size_t)n * NRP;
  FCP_HIP(ctx, ctx->mv64[8].alloc(2 * m));
  double *xi = ctx->mv64[8].p, *bi = ctx->mv64[8].p + m;
  const int g = red_grid(ctx, n);
  // per-column diagonals (velocity.f90:611-632): interleaved, and each
  // column's DIC / DILU pivots from the matrix with its diagonal
  struct DcScope {
    fcp_ctx *c;
    ~DcScope() { c->mdc_active = false; }
  } dc_scope{ctx};
  ctx->mdc_active = false;
  if (diagv) {
    if (prec != FCP_PREC_FP64)
      return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "per-column diagonals: FP64 solves only");
    FCP_TRY(fcp_ensure_workspace(ctx));
    FCP_HIP(ctx, ctx->mdcol.alloc(m));
    FCP_HIP(ctx, ctx->mpiv.alloc(m));
    FCP_HIP(ctx, ctx->mdiag_a.alloc(ctx->nnz));
    DBuf<double> piv;
    FCP_HIP(ctx, piv.alloc(n));
    FCP_HIP(ctx, hipMemsetAsync(ctx->mdcol.p, 0, m * sizeof(double), ctx->stream));
    FCP_HIP(ctx, hipMemsetAsync(ctx->mpiv.p, 0, m * sizeof(double), ctx->stream));
    const int kind = solver == 2 ? 1 : 0;
    for (int c = 0; c < nrhs; ++c) {
      const double *dvc = diagv + (size_t)c * n;
      mk_column<double, 1><<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(n, dvc, 1, 0, ctx->mdcol.p,
                                                                       NRP, c);
      FCP_HIP(ctx, hipMemcpyAsync(ctx->mdiag_a.p, ctx->a.p, sizeof(double) * ctx->nnz,
                                  hipMemcpyDeviceToDevice, ctx->stream));
      mk_set_diag<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(n, ctx->diag.p, dvc,
                                                            ctx->mdiag_a.p);
      FCP_TRY(precond_pivots64(ctx, kind, ctx->mdiag_a.p, piv.p));
      mk_column<double, 1><<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(n, piv.p, 1, 0, ctx->mpiv.p,
                                                                       NRP, c);
    }
    FCP_HIP(ctx, hipGetLastError());
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));  // piv freed on return
    ctx->mpdc_of[0] = ctx->mpdc_of[1] = nullptr;
    ctx->mdc_active = true;
  }
  hipEvent_t e0, e1;
  FCP_HIP(ctx, hipEventCreate(&e0));
  FCP_HIP(ctx, hipEventCreate(&e1));
  FCP_HIP(ctx, hipEventRecord(e0, ctx->stream));
  MultiResult R;
  int rc = 0;
  switch (nrhs) {
#define FCP_MCASE(K)                                                                            \
  case K:                                                                                      \
    mk_interleave<K, Pad<K>::v><<<g, BLK, 0, ctx->stream>>>(n, fi, ldfi, xi);                   \
    mk_interleave<K, Pad<K>::v><<<g, BLK, 0, ctx->stream>>>(n, rhs, n, bi);                     \
    rc = msolve<K>(ctx, solver, xi, bi, itr_max, tol_abs, tol_rel, prec, crit, ce, verbose,     \
                   inner_tol, max_outer, R);                                                    \
    if (rc >= 0) mk_deinterleave<K, Pad<K>::v><<<g, BLK, 0, ctx->stream>>>(n, xi, fi, ldfi);    \
    break;
    FCP_MCASE(1)
    FCP_MCASE(2)
    FCP_MCASE(3)
    FCP_MCASE(4)
#undef FCP_MCASE
  }
  if (rc < 0) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
  }
  FCP_HIP(ctx, hipGetLastError());
  FCP_HIP(ctx, hipEventRecord(e1, ctx->stream));
  // final true residual per column
  SolveScalars *sf = ctx->scal.p + 4;
  SolveScalars h[MAXNR];
  Sell<double> m64;
  FCP_TRY(mview<double>(ctx, &m64));
  switch (nrhs) {
#define FCP_MRES(K)                                                                              \
  case K:                                                                                       \
    mk_interleave<K, Pad<K>::v><<<g, BLK, 0, ctx->stream>>>(n, fi, ldfi, xi);                    \
    mk_residual_init<K, Pad<K>::v><<<g, BLK, 0, ctx->stream>>>(n, m64, xi, bi, ctx->mv64[0].p,  \
                                                               nullptr, ctx->partials.p,         \
                                                               redctr(ctx), sf, 0.0);            \
    break;
    FCP_MRES(1)
    FCP_MRES(2)
    FCP_MRES(3)
    FCP_MRES(4)
#undef FCP_MRES
  }
  FCP_HIP(ctx, hipGetLastError());
  FCP_HIP(ctx, hipMemcpyAsync(h, sf, nrhs * sizeof(SolveScalars), hipMemcpyDeviceToHost,
                              ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0.f;
  FCP_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  int worst_it = 0;
  double worst_rel = 0.0;
  bool conv = true;
  for (int c = 0; c < nrhs; ++c) {
    const double rel = h[c].l2 / (h[c].l2b > 0 ? h[c].l2b : 1.0);
    worst_rel = rel > worst_rel ? rel : worst_rel;
    worst_it = R.iters[c] > worst_it ? R.iters[c] : worst_it;
    if (res0_out) res0_out[c] = prec == FCP_PREC_MIXED ? R.res0[c] : R.resor[c];
    if (crit == FCP_CRIT_TRUE_L2 && rel > tol_rel) conv = false;
    if (!R.conv[c]) conv = false;
  }
  if (opts) {
    opts->iterations = worst_it;
    opts->outer_iterations = R.outer;
    opts->res0 = R.res0[0];
    opts->resl = R.resl[0];
    opts->resor = R.resor[0];
    opts->rel_res_l2 = worst_rel;
    opts->time_ms = ms;
  }
  return conv ? 0 : (worst_it > 0 ? worst_it : 1);
}
