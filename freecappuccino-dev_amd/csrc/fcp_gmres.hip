// fcp_gmres.hip -- restarted GMRES(m) with a true ILU(0) preconditioner
// (pmgmres_ilu, linear_solvers.f90:789-1077, with ilu_cr :1267-1385,
// lus_cr :1386-1477, ax_cr :1173-1265 and mult_givens :1479-1555) and the
// transposed product atx_cr (:1079-1171).
//
// Device layout: the Krylov basis V is column-major (column j at V + j*n),
// the ILU(0) factors sit in one nnz array in CSR order (ilu_cr's `l`), and
// the Hessenberg matrix, Givens rotations, right-hand side g and the stopping
// state live in one device-resident GmState.  The host enqueues whole restart
// cycles and reads the state back every few cycles; kernels after an inner
// exit or a satisfied stopping test return at their first instruction, so no
// per-step host synchronisation is needed.
//
// Per-row arithmetic follows the reference exactly (CSR order in products and
// sweeps, the same divisions, Gram-Schmidt updates written as the reference
// writes them), so the factors, one preconditioner application and the
// transposed product are bit-equal to the reference; inner products and
// norms are deterministic block reductions whose summation order differs from
// the sequential loop.
#include <cmath>

#include "fcp_device.h"

namespace fcpk {
namespace {

constexpr int GM_MAXMR = 32;         // largest restart length
constexpr int GM_LD = GM_MAXMR + 1;  // leading dimension of h
constexpr double GM_DELTA = 1.0e-3;  // re-orthogonalisation threshold (:903)

// Stopping rules.  GM_REF is the reference's: rho = ||r|| is taken before the
// preconditioner is applied (:970-978) and the test reads
// resl/resor <= resor*tol_rel (:1039, :1057).  GM_TRUE normalises the
// preconditioned residual by its own norm and stops on the true relative L2
// residual at a restart.
enum { GM_REF = 0, GM_TRUE = 1 };

struct GmState {
  SolveScalars gate;           // gate.done: the rest of this cycle is skipped
  double h[GM_LD * GM_MAXMR];  // h(i,k) at i + GM_LD*k, 0-based
  double c[GM_LD], s[GM_LD], g[GM_LD], y[GM_LD];
  double rho, resor, rho_tol, resl, av, htmp, l2, l2b, inner_tol;
  int itr, itr_used, k_copy, active, done, reorth, conv, status, final_check, pad;
};

__device__ __forceinline__ double &H(GmState *st, int i, int k) { return st->h[i + GM_LD * k]; }

__device__ __forceinline__ void close_solve(GmState *st, int conv, int status) {
  st->conv = conv;
  if (status) st->status = status;
  st->done = 1;
  st->active = 0;
  st->gate.done = 1;
}

// Start of a restart cycle (:166 iter_loop): zero g(2:) and h (:980-983).
// Past itr_max (GM_TRUE only: GM_REF is done by then) the cycle only forms
// the true residual of the last update and ends the solve.
__global__ void k_gm_begin(GmState *st, int mr, int itr_max) {
  if (st->done) {
    st->active = 0;
    st->gate.done = 1;
    return;
  }
  st->active = 1;
  st->gate.done = 0;
  if (st->itr >= itr_max) {
    st->final_check = 1;
    return;
  }
  st->itr += 1;
  st->reorth = 0;
  st->k_copy = 0;
  for (int i = 0; i < GM_LD * mr; ++i) st->h[i] = 0.0;
  for (int i = 0; i <= mr; ++i) st->g[i] = 0.0;
}

// r = rhs - A x (ax_cr then :960) with the L1 and squared L2 norms of r and,
// on the first cycle, ||rhs||^2.  The last block applies :962-970 (GM_REF)
// or the true-residual test (GM_TRUE).
__global__ __launch_bounds__(BLK) void k_gm_residual(int n, Sell<double> m,
                                                     const double *__restrict__ x,
                                                     const double *__restrict__ b,
                                                     double *__restrict__ r,
                                                     double *__restrict__ partials,
                                                     int *__restrict__ counter, GmState *st,
                                                     int mode, double tol_abs, double tol_rel) {
  if (!st->active) return;
  const bool first = (st->itr == 1 && !st->final_check);
  double part[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    const double ri = b[i] - sell_row<double>(m, i, x, 0.0, false);
    r[i] = ri;
    part[0] += fabs(ri);
    part[1] += ri * ri;
    if (first) part[2] += b[i] * b[i];
  }
  double blk[3], tot[3];
  block_sum<3>(part, blk);
  if (!last_block_reduce<3>(blk, partials, 0, counter, tot) || threadIdx.x != 0) return;
  if (first) {
    st->resor = tot[0];
    st->rho_tol = tot[0] * tol_rel;
    st->l2b = sqrt(tot[2]);
  }
  const double l2 = sqrt(tot[1]);
  st->l2 = l2;
  if (!isfinite(l2)) {
    close_solve(st, 0, FCP_ERR_NONFINITE);
    return;
  }
  if (mode == GM_REF) {
    st->rho = l2;
    st->g[0] = l2;
    // A zero residual would be divided by (:978); the solve ends instead.
    if (l2 == 0.0) close_solve(st, 1, 0);
  } else if (l2 == 0.0 || l2 <= tol_rel * st->l2b || tot[0] < tol_abs) {
    close_solve(st, 1, 0);
  } else if (st->final_check) {
    close_solve(st, 0, 0);
  }
}

// GM_TRUE: rho = ||M^{-1} r|| and the inner target that brings the true
// residual to tol_rel*||b|| if preconditioned and true residuals shrink alike.
__global__ __launch_bounds__(BLK) void k_gm_norm0(int n, const double *__restrict__ v,
                                                  double *__restrict__ partials,
                                                  int *__restrict__ counter, GmState *st,
                                                  double tol_rel) {
  if (solve_done(&st->gate)) return;
  double part[1] = {0.0};
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) part[0] += v[i] * v[i];
  double blk[1], tot[1];
  block_sum<1>(part, blk);
  if (!last_block_reduce<1>(blk, partials, 0, counter, tot) || threadIdx.x != 0) return;
  const double rho = sqrt(tot[0]);
  if (!isfinite(rho) || rho == 0.0) {
    close_solve(st, 0, FCP_ERR_BREAKDOWN);
    return;
  }
  st->rho = rho;
  st->g[0] = rho;
  st->inner_tol = rho * (tol_rel * st->l2b / st->l2);
}

// Level-task halves of lus_cr: forward with the unit lower factor (:1457-1461),
// backward with U divided by its pivot (:1465-1470).  Dependencies are waited
// for with the bounded NaN-sentinel protocol of the other sweeps; tasks are
// taken in level order, so a wave only waits on earlier tickets.
template <int MODE>
__global__ __launch_bounds__(BLK) void k_ilu_sweep(
    const int *__restrict__ task_ptr, const int *__restrict__ order, int ntasks,
    const int *__restrict__ ia, const int *__restrict__ diag, const int *__restrict__ ja,
    const double *__restrict__ l, const double *__restrict__ in, double *out,
    int *__restrict__ ticket, int *__restrict__ err, const SolveScalars *__restrict__ gate,
    int slp) {
  if (gate && solve_done(gate)) return;
  const int lane = threadIdx.x & 63;
  for (;;) {
    int t = 0;
    if (lane == 0) t = atomicAdd(ticket, 1);
    t = __shfl(t, 0, 64);
    if (t >= ntasks) break;
    const int idx = task_ptr[t] + lane;
    if (idx < task_ptr[t + 1]) {
      const int row = order[idx];
      const int dg = diag[row];
      double acc = in[row];
      if (MODE == 0) {
        for (int k = ia[row]; k < dg; ++k) acc = acc - l[k] * wait_value<double>(out + ja[k], err, slp);
        publish<double>(out + row, acc);
      } else {
        const int e = ia[row + 1];
        for (int k = dg + 1; k < e; ++k) acc = acc - l[k] * wait_value<double>(out + ja[k], err, slp);
        publish<double>(out + row, acc / l[dg]);
      }
    }
  }
}

__global__ void k_ilu_fill(double *__restrict__ y, double *__restrict__ z, int n,
                           int *__restrict__ tickets, const SolveScalars *__restrict__ gate) {
  if (gate && solve_done(gate)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    tickets[0] = 0;
    tickets[1] = 0;
  }
  for (int k = i; k < n; k += gridDim.x * blockDim.x) {
    reinterpret_cast<unsigned long long *>(y)[k] = FCP_SENT64;
    reinterpret_cast<unsigned long long *>(z)[k] = FCP_SENT64;
  }
}

// One level of ilu_cr (:1344-1378): every row of the level is factored by one
// thread with the reference's update order; rows of earlier levels are final.
// Row i's columns right of jrow are matched against row jrow's upper entries
// by a merge of the two ascending column lists (the iw lookup of :1338-1342).
__global__ __launch_bounds__(BLK) void k_ilu_level(const int *__restrict__ rows, int cnt,
                                                   const int *__restrict__ ia,
                                                   const int *__restrict__ ja,
                                                   const int *__restrict__ diag, double *l,
                                                   int *__restrict__ err) {
  const int t = blockIdx.x * BLK + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  const int e = ia[i + 1];
  int j = ia[i], jrow = -1;
  for (; j < e; ++j) {
    jrow = ja[j];
    if (i <= jrow) break;
    const double tl = l[j] * l[diag[jrow]];
    l[j] = tl;
    int q = j + 1;
    for (int jj = diag[jrow] + 1, ej = ia[jrow + 1]; jj < ej; ++jj) {
      const int col = ja[jj];
      while (q < e && ja[q] < col) ++q;
      if (q < e && ja[q] == col) l[q] = l[q] - tl * l[jj];
    }
  }
  if (jrow != i || l[j] == 0.0) {  // missing diagonal or zero pivot (:1361-1376)
    __hip_atomic_store(err, FCP_ERR_BREAKDOWN, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  l[j] = 1.0 / l[j];
}

// l(ua(1:n)) = 1/l(ua(1:n)) (:1382).
__global__ void k_ilu_reinvert(int n, const int *__restrict__ diag, double *__restrict__ l) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) l[diag[i]] = 1.0 / l[diag[i]];
}

__global__ void k_gm_spmv(int n, Sell<double> m, const double *__restrict__ x,
                          double *__restrict__ y, const SolveScalars *__restrict__ gate) {
  if (solve_done(gate)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = sell_row<double>(m, i, x, 0.0, false);
}

// v = v / *den unless *den is zero (:978 with den = rho, :1011-1013 with
// den = h(k+1,k)).
__global__ void k_gm_div(int n, double *__restrict__ v, const double *den,
                         const SolveScalars *__restrict__ gate) {
  if (solve_done(gate)) return;
  const double d = *den;
  if (d == 0.0) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    v[i] = v[i] / d;
}

// One fused pass of the Gram-Schmidt loop on w = v_{k+1}: for j > 0 first
// w = w - v_{j-1}*h(j-1,k) (:997), or w = w - htmp*v_{j-1} in the
// re-orthogonalisation pass (:1006); then the dot of w with v_j (j <= k:
// h(j,k) :996, htmp :1004-1005) or with itself (j = k+1: h(k+1,k) :1000,
// :1008).  The first pass's j = 0 also forms av = ||w|| (:993), and its
// j = k+1 decides whether the re-orthogonalisation pass runs (:1002).
template <bool REORTH>
__global__ __launch_bounds__(BLK) void k_gm_orth(int n, const double *__restrict__ V, double *w,
                                                 int k, int j, double *__restrict__ partials,
                                                 int *__restrict__ counter, GmState *st) {
  if (solve_done(&st->gate)) return;
  if (REORTH && !st->reorth) return;
  const double coef = j > 0 ? (REORTH ? st->htmp : st->h[(j - 1) + GM_LD * k]) : 0.0;
  const double *vm = j > 0 ? V + (size_t)(j - 1) * n : nullptr;
  const double *vj = j <= k ? V + (size_t)j * n : nullptr;
  double part[2] = {0.0, 0.0};
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    double wi = w[i];
    if (vm) {
      wi = REORTH ? wi - coef * vm[i] : wi - vm[i] * coef;
      w[i] = wi;
    }
    part[0] += vj ? wi * vj[i] : wi * wi;
    if (!REORTH && j == 0) part[1] += wi * wi;
  }
  double blk[2], tot[2];
  block_sum<2>(part, blk);
  if (!last_block_reduce<2>(blk, partials, 0, counter, tot) || threadIdx.x != 0) return;
  if (j <= k) {
    if (REORTH) {
      st->htmp = tot[0];
      H(st, j, k) = H(st, j, k) + tot[0];
    } else {
      H(st, j, k) = tot[0];
      if (j == 0) st->av = sqrt(tot[1]);
    }
  } else {
    const double hk = sqrt(tot[0]);
    H(st, k + 1, k) = hk;
    if (!REORTH) st->reorth = (st->av + GM_DELTA * hk == st->av) ? 1 : 0;
  }
}

__device__ __forceinline__ void givens(double cs, double sn, int k, double *g) {
  const double g1 = cs * g[k] - sn * g[k + 1];
  const double g2 = sn * g[k] + cs * g[k + 1];
  g[k] = g1;
  g[k + 1] = g2;
}

// Rotations of column k and the residual estimate (:1015-1041).
__global__ void k_gm_givens(GmState *st, int k, int mode, double tol_abs) {
  if (solve_done(&st->gate)) return;
  st->k_copy = k + 1;
  // y(1:k+1) = h(1:k+1,k), rotated, copied back: the same as rotating the
  // contiguous column in place
  for (int j = 0; j < k; ++j) givens(st->c[j], st->s[j], j, st->h + GM_LD * k);
  const double hkk = H(st, k, k), hk1 = H(st, k + 1, k);
  const double mu = sqrt(hkk * hkk + hk1 * hk1);
  const double cs = hkk / mu, sn = -hk1 / mu;
  st->c[k] = cs;
  st->s[k] = sn;
  H(st, k, k) = cs * hkk - sn * hk1;
  H(st, k + 1, k) = 0.0;
  givens(cs, sn, k, st->g);
  const double resl = fabs(st->g[k + 1]);
  st->resl = resl;
  st->itr_used += 1;
  if (!isfinite(resl)) {
    st->status = FCP_ERR_NONFINITE;
    st->gate.done = 1;
    st->done = 1;
    return;
  }
  if (mode == GM_REF) {
    if (resl / st->resor <= st->rho_tol || resl <= tol_abs) {
      st->conv = 1;
      st->gate.done = 1;
    }
  } else if (resl <= st->inner_tol || resl <= tol_abs) {
    st->gate.done = 1;  // inner exit; the restart's true residual decides
  }
}

// Back substitution for y (:1045-1051) and the end-of-cycle test (:1057-1059).
__global__ void k_gm_solve(GmState *st, int itr_max, int mode) {
  if (!st->active) return;
  const int m = st->k_copy - 1;
  if (m < 0) return;
  st->y[m] = st->g[m] / H(st, m, m);
  for (int i = m - 1; i >= 0; --i) {
    double s = 0.0;
    for (int q = i + 1; q <= m; ++q) s = s + H(st, i, q) * st->y[q];
    st->y[i] = (st->g[i] - s) / H(st, i, i);
  }
  if (st->conv || st->status || (mode == GM_REF && st->itr >= itr_max)) st->done = 1;
}

// x(i) = x(i) + sum_q v(i,q) y(q) in column order (:1053-1055).
__global__ void k_gm_update(int n, const double *__restrict__ V, double *__restrict__ x,
                            const GmState *st) {
  if (!st->active) return;
  const int m = st->k_copy - 1;
  if (m < 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int q = 0; q <= m; ++q) acc = acc + V[(size_t)q * n + i] * st->y[q];
  x[i] = x[i] + acc;
}

// y = A^T x (atx_cr :1162-1168): y_c sums a_ic x_i over rows i in ascending
// order, the order of the reference's row-by-row scatter, read here as a
// gather through the transposed-entry positions of row c.
__global__ void k_atx(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                      const int *__restrict__ tslot, const double *__restrict__ a,
                      const double *__restrict__ x, double *__restrict__ y,
                      int *__restrict__ err) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double acc = 0.0;
  for (int k = ia[c], e = ia[c + 1]; k < e; ++k) {
    const int t = tslot[k];
    if (t < 0) {  // the pattern is not symmetric: row c misses entries of A^T
      __hip_atomic_store(err, FCP_ERR_UNSUPPORTED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      continue;
    }
    acc = acc + a[t] * x[ja[k]];
  }
  y[c] = acc;
}

int task_grid(fcp_ctx *ctx, const Schedule &s) {
  int waves = s.ntasks;
  const int cap = ctx->tune_sweep_waves > 0 ? ctx->tune_sweep_waves
                                             : (ctx->ncu > 0 ? ctx->ncu : 256) * 2;
  if (waves > cap) waves = cap;
  const int blocks = (waves + (BLK / 64) - 1) / (BLK / 64);
  return blocks > 0 ? blocks : 1;
}

// z = (LU)^{-1} r: forward into y, backward into z (both sentinel-filled).
int ilu_apply(fcp_ctx *ctx, const double *r, double *y, double *z, const SolveScalars *gate) {
  const int n = ctx->n;
  k_ilu_fill<<<red_grid(ctx, n), BLK, 0, ctx->stream>>>(y, z, n, tickets(ctx), gate);
  {
    ProfScope ps(ctx, "ilu_fwd_f64");
    k_ilu_sweep<0><<<task_grid(ctx, ctx->lower), BLK, 0, ctx->stream>>>(
        ctx->lower.task_ptr.p, ctx->lower.order.p, ctx->lower.ntasks, ctx->ia.p, ctx->diag.p,
        ctx->ja.p, ctx->ilu.p, r, y, tickets(ctx), ctx->errflag.p, gate, ctx->tune_sweep_sleep);
  }
  {
    ProfScope ps(ctx, "ilu_bwd_f64");
    k_ilu_sweep<1><<<task_grid(ctx, ctx->upper), BLK, 0, ctx->stream>>>(
        ctx->upper.task_ptr.p, ctx->upper.order.p, ctx->upper.ntasks, ctx->ia.p, ctx->diag.p,
        ctx->ja.p, ctx->ilu.p, y, z, tickets(ctx) + 1, ctx->errflag.p, gate,
        ctx->tune_sweep_sleep);
  }
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

}  // namespace
}  // namespace fcpk

using namespace fcpk;

// ILU(0) factors of the resident matrix, recomputed when the values changed.
int fcp_ilu_factor_device(fcp_ctx *ctx) {
  if (ctx->ilu_version == ctx->a_version && ctx->ilu.p) return 0;
  const int n = ctx->n, nnz = ctx->nnz;
  const Schedule &s = ctx->lower;
  if ((int)s.lev_ptr.size() != s.nlev + 1 || s.lev_ptr.back() != n)
    return fcp_fail(ctx, FCP_ERR_STATE, "ILU(0): level schedule missing");
  FCP_HIP(ctx, ctx->ilu.alloc(nnz));
  FCP_HIP(ctx, hipMemsetAsync(ctx->errflag.p, 0, sizeof(int), ctx->stream));
  FCP_HIP(ctx, hipMemcpyAsync(ctx->ilu.p, ctx->a.p, sizeof(double) * nnz, hipMemcpyDeviceToDevice,
                              ctx->stream));
  {
    ProfScope ps(ctx, "ilu_factor");
    for (int L = 0; L < s.nlev; ++L) {
      const int lo = s.lev_ptr[L], cnt = s.lev_ptr[L + 1] - lo;
      if (cnt <= 0) continue;
      k_ilu_level<<<grid_for(cnt, BLK), BLK, 0, ctx->stream>>>(s.order.p + lo, cnt, ctx->ia.p,
                                                               ctx->ja.p, ctx->diag.p, ctx->ilu.p,
                                                               ctx->errflag.p);
    }
    k_ilu_reinvert<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(n, ctx->diag.p, ctx->ilu.p);
  }
  FCP_HIP(ctx, hipGetLastError());
  int err = 0;
  FCP_HIP(ctx, hipMemcpyAsync(&err, ctx->errflag.p, sizeof(int), hipMemcpyDeviceToHost,
                              ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (err) {
    ctx->ilu_version = ~0ull;
    return fcp_fail(ctx, FCP_ERR_BREAKDOWN, "ILU(0): zero pivot or missing diagonal");
  }
  ctx->ilu_version = ctx->a_version;
  return 0;
}

int fcp_ilu_apply_device(fcp_ctx *ctx, const double *r, double *z) {
  FCP_TRY(fcp_ensure_workspace(ctx));
  FCP_TRY(fcp_ilu_factor_device(ctx));
  FCP_HIP(ctx, hipMemsetAsync(ctx->errflag.p, 0, sizeof(int), ctx->stream));
  FCP_TRY(ilu_apply(ctx, r, ctx->yk.p, z, nullptr));
  return check_err(ctx);
}

int fcp_atx_device(fcp_ctx *ctx, const double *x, double *y) {
  FCP_TRY(fcp_ensure_workspace(ctx));
  FCP_HIP(ctx, hipMemsetAsync(ctx->errflag.p, 0, sizeof(int), ctx->stream));
  k_atx<<<grid_for(ctx->n, BLK), BLK, 0, ctx->stream>>>(ctx->n, ctx->ia.p, ctx->ja.p,
                                                        ctx->tslot.p, ctx->a.p, x, y,
                                                        ctx->errflag.p);
  FCP_HIP(ctx, hipGetLastError());
  int err = 0;
  FCP_HIP(ctx, hipMemcpyAsync(&err, ctx->errflag.p, sizeof(int), hipMemcpyDeviceToHost,
                              ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (err) return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "A^T x needs a symmetric sparsity pattern");
  return 0;
}

// x, b device pointers of length n.  Returns 0 (converged), the inner step
// count (not converged) or < 0.  opts (may be NULL) receives the iteration
// counts and the reference's residuals (resor = res0 = first L1 residual,
// resl = last estimate).
int fcp_gmres_device(fcp_ctx *ctx, double *x, const double *b, int itr_max, int mr,
                     double tol_abs, double tol_rel, int verbose, fcp_solve_opts *opts) {
  const int n = ctx->n;
  if (mr < 1 || mr > GM_MAXMR || mr >= n)
    return fcp_fail(ctx, FCP_ERR_ARG, "GMRES: restart length %d outside 1..min(%d, n-1)", mr,
                    GM_MAXMR);
  if (itr_max < 1) itr_max = 1;
  if (opts && opts->precision != FCP_PREC_FP64)
    return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "GMRES runs in FP64 only");
  const int mode = (opts && opts->criterion == FCP_CRIT_TRUE_L2) ? GM_TRUE : GM_REF;
  FCP_TRY(fcp_ensure_workspace(ctx));
  FCP_TRY(fcp_ilu_factor_device(ctx));
  FCP_HIP(ctx, ctx->gm_v.alloc((size_t)(mr + 1) * n));
  FCP_HIP(ctx, ctx->gm_state.alloc(sizeof(GmState)));
  GmState *st = reinterpret_cast<GmState *>(ctx->gm_state.p);
  const SolveScalars *gate = &st->gate;
  double *V = ctx->gm_v.p;
  FCP_HIP(ctx, hipMemsetAsync(st, 0, sizeof(GmState), ctx->stream));
  FCP_HIP(ctx, hipMemsetAsync(ctx->counters.p, 0, CTR_INTS * sizeof(int), ctx->stream));
  FCP_HIP(ctx, hipMemsetAsync(ctx->errflag.p, 0, sizeof(int), ctx->stream));
  Sell<double> m;
  FCP_TRY(sell_view64(ctx, &m));
  const int g = red_grid(ctx, n), gn = grid_for(n, BLK);
  auto cycle = [&]() -> int {
    k_gm_begin<<<1, 1, 0, ctx->stream>>>(st, mr, itr_max);
    k_gm_residual<<<g, BLK, 0, ctx->stream>>>(n, m, x, b, ctx->res.p, ctx->partials.p,
                                              redctr(ctx), st, mode, tol_abs, tol_rel);
    FCP_TRY(ilu_apply(ctx, ctx->res.p, ctx->yk.p, V, gate));
    if (mode == GM_TRUE)
      k_gm_norm0<<<g, BLK, 0, ctx->stream>>>(n, V, ctx->partials.p, redctr(ctx), st, tol_rel);
    k_gm_div<<<g, BLK, 0, ctx->stream>>>(n, V, &st->rho, gate);
    for (int k = 0; k < mr; ++k) {
      double *w = V + (size_t)(k + 1) * n;
      {
        ProfScope ps(ctx, "gmres_spmv_f64");
        k_gm_spmv<<<gn, BLK, 0, ctx->stream>>>(n, m, V + (size_t)k * n, ctx->qk.p, gate);
      }
      FCP_TRY(ilu_apply(ctx, ctx->qk.p, ctx->yk.p, w, gate));
      for (int j = 0; j <= k + 1; ++j)
        k_gm_orth<false><<<g, BLK, 0, ctx->stream>>>(n, V, w, k, j, ctx->partials.p, redctr(ctx),
                                                     st);
      for (int j = 0; j <= k + 1; ++j)
        k_gm_orth<true><<<g, BLK, 0, ctx->stream>>>(n, V, w, k, j, ctx->partials.p, redctr(ctx),
                                                    st);
      k_gm_div<<<g, BLK, 0, ctx->stream>>>(n, w, &st->h[(k + 1) + GM_LD * k], gate);
      k_gm_givens<<<1, 1, 0, ctx->stream>>>(st, k, mode, tol_abs);
    }
    k_gm_solve<<<1, 1, 0, ctx->stream>>>(st, itr_max, mode);
    k_gm_update<<<gn, BLK, 0, ctx->stream>>>(n, V, x, st);
    FCP_HIP(ctx, hipGetLastError());
    return 0;
  };
  // Host polls every few cycles; cycles enqueued after the end are no-ops.
  // GM_TRUE needs one pass beyond itr_max to test the last update.
  const int ce = (opts && opts->check_every > 0) ? opts->check_every : 8;
  const int per_poll = ce / mr > 1 ? ce / mr : 1;
  const int max_cycles = itr_max + (mode == GM_TRUE ? 1 : 0);
  GmState h;
  for (int cycles = 0;;) {
    for (int c = 0; c < per_poll && cycles < max_cycles; ++c, ++cycles) FCP_TRY(cycle());
    FCP_TRY(poll_device(ctx, &h, st, sizeof(GmState)));
    if (verbose >= 2)
      printf("  ITR = %4d  Residual = %14.6G  K = %6d  Residual = %14.6G\n", h.itr, h.rho,
             h.itr_used, h.resl);
    if (h.done || cycles >= max_cycles) break;
  }
  if (h.status < 0) return fcp_fail(ctx, h.status, "GMRES: non-finite residual or breakdown");
  if (opts) {
    opts->iterations = h.itr_used;
    opts->outer_iterations = h.itr;
    opts->res0 = h.resor;
    opts->resor = h.resor;
    opts->resl = h.resl;
  }
  if (h.conv) return 0;
  return h.itr_used > 0 ? h.itr_used : 1;
}
