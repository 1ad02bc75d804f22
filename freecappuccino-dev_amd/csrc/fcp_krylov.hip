// fcp_krylov.hip -- sparse kernels and Krylov solvers of libfcp.
//
// * SpMV: one thread per row in CSR order (same summation order as the
//   reference's row loop, linear_solvers.f90:493-498) or LANES lanes per row
//   with a shuffle reduction.
// * Reductions: two-stage -- every block writes a partial (FP64), the last
//   block to finish (atomic ticket) sums the partials in index order, so every
//   run produces the same bits.  The last block also performs the scalar part
//   of the Krylov step (alpha, beta, convergence test) on the device.
// * DIC/DILU setup and triangular sweeps: rows grouped by level of the
//   strict-lower (upper) DAG into tasks of <= 64 rows; persistent waves take
//   tasks in level order from an atomic ticket and wait for their
//   dependencies through the output array itself, pre-filled with a NaN
//   sentinel (one agent-scope store per row, no per-level launch).  A wave
//   only waits for rows of tasks with smaller tickets, which are owned by
//   waves already running, so the sweep always drains; every wait is bounded.
//   Each row is computed with the reference's operation order (:439-475),
//   so sweeps are bit-equal to the CPU loops.
// * PCG (DIC or Jacobi) and DILU-BiCGStab follow linear_solvers.f90:206-786
//   step by step; scalars stay on the device, the host only polls a status
//   word every `check_every` iterations.
// * Mixed precision: FP32 matrix/preconditioner/vectors inside an FP64
//   iterative-refinement loop (true residual, update and norms in FP64).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "fcp_device.h"

using namespace fcpk;

namespace {

template <typename T>
__global__ void k_fill_sent(T *__restrict__ x, T *__restrict__ y, int n, int *__restrict__ tickets,
                            const SolveScalars *__restrict__ s) {
  if (s && solve_done(s)) return;
  using B = Bits<T>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && tickets) {
    tickets[0] = 0;
    tickets[1] = 0;
  }
  for (int k = i; k < n; k += gridDim.x * blockDim.x) {
    reinterpret_cast<typename B::U *>(x)[k] = B::sent;
    if (y) reinterpret_cast<typename B::U *>(y)[k] = B::sent;
  }
}

// MODE 0: forward  y_i = (r_i - sum_{k<diag} a_k y_j) d_i            (:458-464)
// MODE 1: backward z_i = (y_i/(d_i+small) - sum_{k>diag} a_k z_j) d_i   (:466-475)
// MODE 2: DIC setup  d_i = 1/(a_ii - sum_{k<diag} a_k^2 d_j)          (:439-445)
// MODE 3: DILU setup d_i = 1/(a_ii - sum_{k<diag} a_k d_j a_kT)       (:613-624)
template <typename T, int MODE>
__global__ __launch_bounds__(BLK) void k_sweep(
    const int *__restrict__ task_ptr, const int *__restrict__ order, int ntasks,
    const int *__restrict__ ia, const int *__restrict__ diag, const int *__restrict__ ja,
    const T *__restrict__ a, const int *__restrict__ tslot, const T *__restrict__ d,
    const T *__restrict__ in, T *out, int *__restrict__ ticket, int *__restrict__ err,
    const SolveScalars *__restrict__ s, int slp) {
  if (s && solve_done(s)) return;
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
      if (MODE == 0) {
        T acc = in[row];
        for (int k = ia[row]; k < dg; ++k) acc = acc - a[k] * wait_value<T>(out + ja[k], err, slp);
        publish<T>(out + row, acc * d[row]);
      } else if (MODE == 1) {
        const T dr = d[row];
        T acc = in[row] / (dr + (T)FCP_SMALL);
        const int e = ia[row + 1];
        for (int k = dg + 1; k < e; ++k) acc = acc - a[k] * wait_value<T>(out + ja[k], err, slp);
        publish<T>(out + row, acc * dr);
      } else if (MODE == 2) {
        T acc = a[dg];
        for (int k = ia[row]; k < dg; ++k) {
          const T ak = a[k];
          acc = acc - ak * ak * wait_value<T>(out + ja[k], err, slp);
        }
        publish<T>(out + row, (T)1.0 / acc);
      } else {
        T acc = a[dg];
        for (int k = ia[row]; k < dg; ++k) {
          const int l = tslot[k];
          const T dj = wait_value<T>(out + ja[k], err, slp);
          acc = acc - a[k] * dj * (l >= 0 ? a[l] : (T)0);
        }
        publish<T>(out + row, (T)1.0 / acc);
      }
    }
  }
}

// Gauss-Seidel sweep (GaussSeidel, linear_solvers.f90:146-153): rows in the
// level order of the strict lower triangle.  Columns left of the diagonal
// take this sweep's values (waited for), the diagonal and the columns right
// of it take the previous sweep's, summed in ascending column order as the
// sequential loop does, so every sweep is bit-equal to it.
__global__ __launch_bounds__(BLK) void k_gs_sweep(
    const int *__restrict__ task_ptr, const int *__restrict__ order, int ntasks,
    const int *__restrict__ ia, const int *__restrict__ diag, const int *__restrict__ ja,
    const double *__restrict__ a, const double *__restrict__ rhs, const double *__restrict__ xold,
    double *xnew, double *__restrict__ res, int *__restrict__ ticket, int *__restrict__ err,
    const SolveScalars *__restrict__ s, int slp) {
  if (solve_done(s)) return;
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
      const int e = ia[row + 1];
      double acc = rhs[row];
      for (int k = ia[row]; k < dg; ++k) acc = acc - a[k] * wait_value<double>(xnew + ja[k], err, slp);
      for (int k = dg; k < e; ++k) acc = acc - a[k] * xold[ja[k]];
      res[row] = acc;
      publish<double>(xnew + row, xold[row] + acc / (a[dg] + FCP_SMALL));
    }
  }
}

// x = new sweep values; resl = |res|_1; on the first sweep res0 = resl and
// factor = sum |a_ii x_i| + small (:156-180); convergence test.
__global__ __launch_bounds__(BLK) void k_gs_update(
    int n, const int *__restrict__ diag, const double *__restrict__ a,
    const double *__restrict__ xnew, double *__restrict__ x, const double *__restrict__ res,
    double *__restrict__ partials, int *__restrict__ counter, SolveScalars *__restrict__ s,
    int crit, double tol_abs, double tol_rel, int itr_max) {
  if (solve_done(s)) return;
  const bool first = (s->iter == 0);
  double part[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    const double xi = xnew[i];
    x[i] = xi;
    const double ri = res[i];
    part[0] += fabs(ri);
    part[1] += ri * ri;
    if (first) part[2] += fabs(a[diag[i]] * xi);
  }
  double blk[3], tot[3];
  block_sum<3>(part, blk);
  if (last_block_reduce<3>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
    const double resl = tot[0];
    s->resl = resl;
    s->l2 = sqrt(tot[1]);
    s->iter += 1;
    bool conv;
    if (first) {
      s->res0 = resl;
      s->factor = tot[2] + FCP_SMALL;
      s->resor = resl / s->factor;
    }
    const double rsm = resl / (s->res0 + FCP_SMALL);
    s->rsm = rsm;
    if (first && resl < tol_abs)
      conv = true;  // :158-165
    else if (crit == 0)
      conv = (rsm < tol_rel) || (resl < tol_abs);
    else
      conv = (s->l2 <= tol_rel * s->l2b) || (resl < tol_abs);
    if (!isfinite(resl)) {
      s->status = FCP_ERR_NONFINITE;
      conv = true;
    }
    s->conv = (conv && s->status == 0) ? 1 : 0;
    if (conv || s->iter >= itr_max) s->done = 1;
  }
}

// Row-block sweep: one workgroup per block of R contiguous rows (blocks taken
// in ticket order, reversed for the backward sweep).  The block's rows are
// processed step by step (one step = one level present in the block); a
// dependency inside the block is read from LDS, written at an earlier step
// (barrier in between); a dependency in an earlier block is waited for in
// global memory like the level-task sweep.  Same per-row arithmetic as
// k_sweep, so results are bit-equal.
template <typename T, int MODE>
__global__ __launch_bounds__(BLK) void k_bsweep(
    const int *__restrict__ rows, const int *__restrict__ step_ptr,
    const int *__restrict__ blk_step, int nblocks, int R, int n, const int *__restrict__ ia,
    const int *__restrict__ diag, const int *__restrict__ ja, const T *__restrict__ a,
    const int *__restrict__ tslot, const T *__restrict__ d, const T *__restrict__ in, T *out,
    int *__restrict__ ticket, int *__restrict__ err, const SolveScalars *__restrict__ s, int slp) {
  extern __shared__ unsigned char smem_raw[];
  T *zl = reinterpret_cast<T *>(smem_raw);
  __shared__ int sb;
  if (s && solve_done(s)) return;
  if (threadIdx.x == 0) sb = atomicAdd(ticket, 1);
  __syncthreads();
  const int t = sb;
  if (t >= nblocks) return;
  const int b = (MODE == 1) ? nblocks - 1 - t : t;
  const int lo = b * R;
  const int hi = (lo + R < n) ? lo + R : n;
  for (int st = blk_step[b]; st < blk_step[b + 1]; ++st) {
    const int e = step_ptr[st + 1];
    for (int idx = step_ptr[st] + threadIdx.x; idx < e; idx += BLK) {
      const int row = rows[idx];
      const int dg = diag[row];
      T val;
      if (MODE == 0) {
        T acc = in[row];
        for (int k = ia[row]; k < dg; ++k) {
          const int j = ja[k];
          const T v = (j >= lo) ? zl[j - lo] : wait_value<T>(out + j, err, slp);
          acc = acc - a[k] * v;
        }
        val = acc * d[row];
      } else if (MODE == 1) {
        const T dr = d[row];
        T acc = in[row] / (dr + (T)FCP_SMALL);
        const int ee = ia[row + 1];
        for (int k = dg + 1; k < ee; ++k) {
          const int j = ja[k];
          const T v = (j < hi) ? zl[j - lo] : wait_value<T>(out + j, err, slp);
          acc = acc - a[k] * v;
        }
        val = acc * dr;
      } else if (MODE == 2) {
        T acc = a[dg];
        for (int k = ia[row]; k < dg; ++k) {
          const int j = ja[k];
          const T ak = a[k];
          const T v = (j >= lo) ? zl[j - lo] : wait_value<T>(out + j, err, slp);
          acc = acc - ak * ak * v;
        }
        val = (T)1.0 / acc;
      } else {
        T acc = a[dg];
        for (int k = ia[row]; k < dg; ++k) {
          const int j = ja[k];
          const int l = tslot[k];
          const T v = (j >= lo) ? zl[j - lo] : wait_value<T>(out + j, err, slp);
          acc = acc - a[k] * v * (l >= 0 ? a[l] : (T)0);
        }
        val = (T)1.0 / acc;
      }
      zl[row - lo] = val;
      publish<T>(out + row, val);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Pipelined row-block sweep on packed operands.
//
// Same block/step structure as k_bsweep, but every position's operands
// (row id, dependency count and slots, coefficients, pivot, input value) are
// stored step-major so a step's data is one contiguous load, and each thread
// fetches them PD steps ahead into a ring of registers; the values of
// dependencies in earlier blocks are fetched PD2 steps ahead (re-polled at
// use time only if still unpublished).  Steps are separated by a workgroup
// barrier; a thread handles one position per step (the schedule splits
// steps wider than the workgroup).  Prefetch loads are issued
// unconditionally (an unused slot reads position 0, or an out-of-range
// buffer offset, and is flagged invalid) so the compiler's memory-counter
// waits stay exact: a load inside a divergent branch makes every later wait
// drain all outstanding prefetches.
// ---------------------------------------------------------------------------
template <typename T>
struct PSlot {
  int row, cnt, ok, pos;
  int pubo;  // byte offset of the row in the output for an external row, else FCP_OOB
  T dv, xin;
  int dep[FCP_PMAXD];
  T av[FCP_PMAXD];
  T rv[FCP_PMAXD];
};

template <typename T>
struct PArgs {
  const int *prc, *pdep;
  const T *pa, *pdx;
  const T *in;  // sweep input, read by row
  int n;
};

// 16-byte vector loads of the packed operands (one int2, one int4, 16 or 32
// bytes of coefficients and the pivot of every position); the row's input
// value is read in the second prefetch stage, once the row id has arrived.
template <typename T>
__device__ __forceinline__ void p_fetch(const PArgs<T> &P, int idx, int &row, int &cnt, int &ext,
                                        T &dv, int *dep, T *av) {
  const int2 rc = reinterpret_cast<const int2 *>(P.prc)[idx];
  row = rc.x;
  cnt = rc.y & 0xFFFF;
  ext = rc.y >> 16;
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
    const float2 dx = reinterpret_cast<const float2 *>(P.pdx)[idx];
    dv = (T)dx.x;
  } else {
    const double2 a0 = reinterpret_cast<const double2 *>(P.pa)[2 * idx];
    const double2 a1 = reinterpret_cast<const double2 *>(P.pa)[2 * idx + 1];
    av[0] = (T)a0.x;
    av[1] = (T)a0.y;
    av[2] = (T)a1.x;
    av[3] = (T)a1.y;
    const double2 dx = reinterpret_cast<const double2 *>(P.pdx)[idx];
    dv = (T)dx.x;
  }
}

// ns >= 1; sptr holds ns+1 step offsets.
template <typename T>
__device__ __forceinline__ void p_load1(PSlot<T> &S, const PArgs<T> &P, const int *sptr, int step,
                                        int ns) {
  const int sc = step < ns ? step : ns - 1;
  const int idx = sptr[sc] + (int)threadIdx.x;
  const bool ok = (step < ns) && (idx < sptr[sc + 1]);
  S.ok = ok ? 1 : 0;
  S.pos = ok ? idx : 0;
  int ext;
  p_fetch<T>(P, ok ? idx : 0, S.row, S.cnt, ext, S.dv, S.dep, S.av);
  S.pubo = (ok && ext) ? S.row * (int)sizeof(T) : FCP_OOB;
}

// Second prefetch stage, once the row and dependency ids have arrived: the
// row's input value, and the values of dependencies in earlier blocks with
// agent-coherent buffer loads; lanes without such a dependency pass an
// out-of-range offset and issue no memory request.
template <typename T>
__device__ __forceinline__ void p_load2(PSlot<T> &S, const PArgs<T> &P,
                                        __amdgpu_buffer_rsrc_t rout) {
  using R = RawVec<sizeof(T)>;
  S.xin = P.in[S.row];
#pragma unroll
  for (int m = 0; m < FCP_PMAXD; ++m) {
    const bool need = S.ok && m < S.cnt && S.dep[m] < 0;
    const int off = need ? (-S.dep[m] - 1) * (int)sizeof(T) : FCP_OOB;
    S.rv[m] = __builtin_bit_cast(T, R::template load<CPOL_SC1>(rout, off));
  }
}

template <typename T, int MODE>
__device__ __forceinline__ T p_row_value(int cnt, T dv, T xin, const int *dep, const T *av,
                                         const T *rv, const T *zl, T *out, int *err, int slp) {
  using B = Bits<T>;
  T acc = (MODE == 0) ? xin : xin / (dv + (T)FCP_SMALL);
#pragma unroll
  for (int m = 0; m < FCP_PMAXD; ++m) {
    if (m < cnt) {
      T v;
      if (dep[m] >= 0)
        v = zl[dep[m]];
      else if (rv && B::to(rv[m]) != B::sent)
        v = rv[m];
      else
        v = wait_value<T>(out + (-dep[m] - 1), err, slp);
      acc = acc - av[m] * v;
    }
  }
  return acc * dv;
}

// One position of a step.  Fast path without branches: every dependency is
// read from LDS (slot 0 for remote ones) and the prefetched remote value is
// selected where the dependency lies in an earlier block; unused dependency
// slots are masked by a select, so the sums are formed in the same order as
// p_row_value.  Only a lane whose prefetched remote value was still
// unpublished takes p_row_value, which waits for it.  Steps never hold more
// positions than the workgroup has threads (the schedule splits them).
template <typename T, int MODE>
__device__ __forceinline__ void p_compute(const PSlot<T> &S, int lo, T *zl, int *zr, T *out,
                                          __amdgpu_buffer_rsrc_t rout, int *err, int slp) {
  using B = Bits<T>;
  if (!S.ok) return;
  T acc = (MODE == 0) ? S.xin : S.xin / (S.dv + (T)FCP_SMALL);
  unsigned miss = 0;  // bitwise flags: no short-circuit control flow
#pragma unroll
  for (int m = 0; m < FCP_PMAXD; ++m) {
    const int d = S.dep[m];
    const T lv = zl[d > 0 ? d : 0];
    const bool use = m < S.cnt;
    const bool remote = d < 0;
    miss |= (unsigned)use & (unsigned)remote & (unsigned)(B::to(S.rv[m]) == B::sent);
    const T v = remote ? S.rv[m] : lv;
    const T na = acc - S.av[m] * v;
    acc = use ? na : acc;
  }
  T val = acc * S.dv;
  if (miss) val = p_row_value<T, MODE>(S.cnt, S.dv, S.xin, S.dep, S.av, S.rv, zl, out, err, slp);
  zl[S.pos - lo] = val;
  zr[S.pos - lo] = S.row;
  // external rows are published for the other blocks now (agent-coherent
  // store; lanes of other rows pass an out-of-range offset); every row is
  // written to the output when the block ends
  using B = Bits<T>;
  typename B::U u = B::to(val);
  if (u == B::sent) u = B::qnan;
  RawVec<sizeof(T)>::store(__builtin_bit_cast(typename RawVec<sizeof(T)>::U, u), rout, S.pubo);
}

// Each step issues 8-9 vector memory loads and one store; PD steps in flight
// stay under the 63 outstanding operations a wave can count.
#ifndef FCP_PD
#define FCP_PD 6
#endif
#ifndef FCP_PD2
#define FCP_PD2 3
#endif
constexpr int PD = FCP_PD;    // steps of operand prefetch
constexpr int PD2 = FCP_PD2;  // steps of input and remote-value prefetch
static_assert(PD2 >= 1 && PD2 < PD, "remote prefetch must trail the operand prefetch");

// LDS: block values zl[R] | step table sptr[max_steps+2] | row ids zr[R].
template <typename T, int MODE>
__global__ __launch_bounds__(BLK, 1) void k_psweep(PArgs<T> P, const int *__restrict__ step_ptr,
                                                   const int *__restrict__ blk_step,
                                                   const int *__restrict__ blk_pos, int nblocks,
                                                   int sptr_off, int zr_off, T *out,
                                                   int *__restrict__ ticket, int *__restrict__ err,
                                                   const SolveScalars *__restrict__ s, int slp) {
  extern __shared__ unsigned char smem_raw[];
  T *zl = reinterpret_cast<T *>(smem_raw);
  int *sptr = reinterpret_cast<int *>(smem_raw + sptr_off);
  int *zr = reinterpret_cast<int *>(smem_raw + zr_off);
  __shared__ int sb;
  if (s && solve_done(s)) return;
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
  const __amdgpu_buffer_rsrc_t rout = raw_rsrc(out, P.n * (int)sizeof(T));
  PSlot<T> S[PD];
#pragma unroll
  for (int j = 0; j < PD; ++j) p_load1<T>(S[j], P, sptr, j, ns);
#pragma unroll
  for (int j = 0; j < PD2; ++j) p_load2<T>(S[j], P, rout);
  // Steps past ns (at most PD-1 in the last group) are empty: their slots
  // are invalid, so every group runs the same instruction stream.
  for (int base = 0; base < ns; base += PD) {
#pragma unroll
    for (int j = 0; j < PD; ++j) {
      const int step = base + j;
      p_compute<T, MODE>(S[j], lo, zl, zr, out, rout, err, slp);
      __syncthreads();
      p_load2<T>(S[(j + PD2) % PD], P, rout);
      p_load1<T>(S[j], P, sptr, step + PD, ns);
    }
  }
  // every row of the block to the output (read by later kernels)
  const int nr = blk_pos[b + 1] - lo;
  for (int i = threadIdx.x; i < nr; i += BLK) out[zr[i]] = zl[i];
}

// pa/pd of a block schedule in precision T from the resident matrix and
// pivots (repacked when either changed).
template <typename T>
__global__ void k_pack_vals(int n, const int *__restrict__ prc, const int *__restrict__ psrc,
                            const double *__restrict__ a, const double *__restrict__ d,
                            T *__restrict__ pa, T *__restrict__ pdx) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
#pragma unroll
  for (int m = 0; m < FCP_PMAXD; ++m) {
    const int k = psrc[(size_t)FCP_PMAXD * idx + m];
    pa[(size_t)FCP_PMAXD * idx + m] = k >= 0 ? (T)a[k] : (T)0;
  }
  pdx[2 * (size_t)idx] = (T)d[prc[2 * idx]];
}


// ---------------------------------------------------------------------------
// SpMV on the SELL-64 copy
// ---------------------------------------------------------------------------
__global__ void k_sell_len(const int *__restrict__ ia, int n, int nslices, int *__restrict__ len) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > nslices) return;
  if (s == nslices) {
    len[s] = 0;
    return;
  }
  int m = 0;
  for (int r = s * 64; r < s * 64 + 64 && r < n; ++r) {
    const int l = ia[r + 1] - ia[r];
    m = l > m ? l : m;
  }
  len[s] = 64 * m;
}

__global__ void k_sell_col(const int *__restrict__ ia, const int *__restrict__ ja, int n,
                           int nslices, const int *__restrict__ base, int *__restrict__ col) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nslices * 64) return;
  const int s = r >> 6, p = base[s] + (r & 63), L = (base[s + 1] - base[s]) >> 6;
  const int len = r < n ? ia[r + 1] - ia[r] : 0;
  for (int k = 0; k < L; ++k) col[p + 64 * k] = (k < len) ? ja[ia[r] + k] : -1;
}

template <typename T>
__global__ void k_sell_val(const int *__restrict__ ia, const double *__restrict__ a, int n,
                           int nslices, const int *__restrict__ base, T *__restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nslices * 64) return;
  const int s = r >> 6, p = base[s] + (r & 63), L = (base[s + 1] - base[s]) >> 6;
  const int r0 = r < n ? ia[r] : 0, len = r < n ? ia[r + 1] - r0 : 0;
  T v[SELL_KU];  // the row's entries loaded together, then stored slot by slot
#pragma unroll
  for (int k = 0; k < SELL_KU; ++k) v[k] = k < len ? (T)a[r0 + k] : (T)0;
#pragma unroll
  for (int k = 0; k < SELL_KU; ++k)
    if (k < L) out[p + 64 * k] = v[k];
  for (int k = SELL_KU; k < L; ++k) out[p + 64 * k] = (k < len) ? (T)a[r0 + k] : (T)0;
}

// ---------------------------------------------------------------------------
// SpMV
// ---------------------------------------------------------------------------
// y = A x, one thread per row, CSR order (bit-equal to the reference loop).
// The grid is a multiple of 8 workgroups; workgroups run on the eight XCDs
// round-robin (b on XCD b % 8), and XCD x takes the x-th eighth of the row
// blocks, so the x lines that neighbouring rows gather stay in one L2.
template <typename T>
__global__ void k_spmv_row(int n, Sell<T> m, const T *__restrict__ x, T *__restrict__ y) {
  const int b = (int)blockIdx.x, per = (int)gridDim.x >> 3;
  const int i = ((b & 7) * per + (b >> 3)) * (int)blockDim.x + (int)threadIdx.x;
  if (i >= n) return;
  y[i] = sell_row<T>(m, i, x, (T)0, false);
}

// The same product with R rows per thread (rows blk*R*BLK + r*BLK + tid, r <
// R, so every load stays coalesced) and, with NT, non-temporal loads of the
// matrix: its values and columns are read once and should not evict the
// gathered x lines from L2 and the Infinity Cache.  Per-row sums are those of
// sell_row (same entries, same order).
template <typename T, bool NT>
__device__ __forceinline__ T sell_ld(const T *p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

template <typename T, bool NT, int R, bool NTY = false>
__global__ __launch_bounds__(BLK) void k_spmv_rows(int n, Sell<T> m, const T *__restrict__ x,
                                                   T *__restrict__ y) {
  const int b = (int)blockIdx.x, per = (int)gridDim.x >> 3;
  const int blk = (b & 7) * per + (b >> 3);
  int row[R], p[R], L[R];
  T av[R][SELL_KU], xv[R][SELL_KU];
  int cv[R][SELL_KU];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    row[r] = (blk * R + r) * BLK + (int)threadIdx.x;
    const int i = row[r] < n ? row[r] : n - 1;
    const int2 at = sell_at(m.base, m.width, i);
    p[r] = at.x;
    L[r] = at.y;
  }
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int k = 0; k < SELL_KU; ++k) {
      cv[r][k] = -1;
      if (k < L[r]) {
        av[r][k] = sell_ld<T, NT>(m.a + p[r] + 64 * k);
        cv[r][k] = sell_ld<int, NT>(m.col + p[r] + 64 * k);
      }
    }
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int k = 0; k < SELL_KU; ++k)
      if (cv[r][k] >= 0) xv[r][k] = x[cv[r][k]];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    T acc = (T)0;
#pragma unroll
    for (int k = 0; k < SELL_KU; ++k)
      if (cv[r][k] >= 0) acc = acc + av[r][k] * xv[r][k];
    for (int k = SELL_KU; k < L[r]; ++k) {
      const int j = sell_ld<int, NT>(m.col + p[r] + 64 * k);
      if (j < 0) break;
      acc = acc + sell_ld<T, NT>(m.a + p[r] + 64 * k) * x[j];
    }
    if (row[r] < n) {
      if constexpr (NTY)
        __builtin_nontemporal_store(acc, y + row[r]);
      else
        y[row[r]] = acc;
    }
  }
}

// res = b - A x (row order as :417-422) and the L1 / squared L2 norms of res
// and the squared L2 norm of b.  The last block stores res0, l2, l2b and
// initialises the Krylov scalars.
template <typename T>
__global__ __launch_bounds__(BLK) void k_residual_init(
    int n, Sell<T> m, const T *__restrict__ x, const T *__restrict__ b, T *__restrict__ res,
    T *__restrict__ reso, double *__restrict__ partials, int *__restrict__ counter,
    SolveScalars *__restrict__ s) {
  double part[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    const T r = sell_row<T>(m, i, x, b[i], true);
    res[i] = r;
    if (reso) reso[i] = r;
    part[0] += fabs((double)r);
    part[1] += (double)r * (double)r;
    part[2] += (double)b[i] * (double)b[i];
  }
  double blk[3], tot[3];
  block_sum<3>(part, blk);
  if (last_block_reduce<3>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
    s->res0 = tot[0];
    s->resl = tot[0];
    s->l2 = sqrt(tot[1]);
    s->l2b = sqrt(tot[2]);
    s->s0 = FCP_S0INIT;
    s->alf = 1.0;
    s->beto = 1.0;
    s->gam = 1.0;
    s->iter = 0;
    s->done = 0;
    s->status = 0;
    s->factor = 0.0;
    s->resor = tot[0];
  }
}

// ---------------------------------------------------------------------------
// PCG step kernels (iccg :454-537, dpcg :282-352)
// ---------------------------------------------------------------------------
// sk = r.z ; bet = sk/s0
template <typename T>
__global__ __launch_bounds__(BLK) void k_dot_rz(int n, const T *__restrict__ r,
                                                const T *__restrict__ z,
                                                double *__restrict__ partials,
                                                int *__restrict__ counter,
                                                SolveScalars *__restrict__ s) {
  if (solve_done(s)) return;
  double part[1] = {0.0};
  const int stride = gridDim.x * BLK;
  for (int i0 = blockIdx.x * BLK + threadIdx.x; i0 < n; i0 += GS_U * stride) {
    T rv[GS_U], zv[GS_U];
#pragma unroll
    for (int u = 0; u < GS_U; ++u)
      if (i0 + u * stride < n) {
        rv[u] = r[i0 + u * stride];
        zv[u] = z[i0 + u * stride];
      }
#pragma unroll
    for (int u = 0; u < GS_U; ++u)
      if (i0 + u * stride < n) part[0] += (double)rv[u] * (double)zv[u];
  }
  double blk[1], tot[1];
  block_sum<1>(part, blk);
  if (last_block_reduce<1>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
    s->sk = tot[0];
    s->bet = tot[0] / s->s0;
  }
}

// Jacobi preconditioner z = r / a_ii fused with sk = r.z (dpcg :288-293)
template <typename T>
__global__ __launch_bounds__(BLK) void k_jacobi_dot(int n, const int *__restrict__ diag,
                                                    const T *__restrict__ a,
                                                    const T *__restrict__ r, T *__restrict__ z,
                                                    double *__restrict__ partials,
                                                    int *__restrict__ counter,
                                                    SolveScalars *__restrict__ s) {
  if (solve_done(s)) return;
  double part[1] = {0.0};
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    const T zi = r[i] / a[diag[i]];
    z[i] = zi;
    part[0] += (double)r[i] * (double)zi;
  }
  double blk[1], tot[1];
  block_sum<1>(part, blk);
  if (last_block_reduce<1>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
    s->sk = tot[0];
    s->bet = tot[0] / s->s0;
  }
}

// p = z + bet*p
template <typename T>
__global__ void k_pupdate(int n, const T *__restrict__ z, T *__restrict__ p,
                          const SolveScalars *__restrict__ s) {
  if (solve_done(s)) return;
  const T bet = (T)s->bet;
  const int stride = gridDim.x * BLK;
  for (int i0 = blockIdx.x * BLK + threadIdx.x; i0 < n; i0 += GS_U * stride) {
    T zv[GS_U], pv[GS_U];
#pragma unroll
    for (int u = 0; u < GS_U; ++u)
      if (i0 + u * stride < n) {
        zv[u] = z[i0 + u * stride];
        pv[u] = p[i0 + u * stride];
      }
#pragma unroll
    for (int u = 0; u < GS_U; ++u)
      if (i0 + u * stride < n) p[i0 + u * stride] = zv[u] + bet * pv[u];
  }
}

// q = A p ; pkapk = p.q ; alf = sk/pkapk
template <typename T>
__global__ __launch_bounds__(BLK) void k_spmv_pq(int n, Sell<T> m, const T *__restrict__ p,
                                                 T *__restrict__ q, double *__restrict__ partials,
                                                 int *__restrict__ counter,
                                                 SolveScalars *__restrict__ s) {
  if (solve_done(s)) return;
  double part[1] = {0.0};
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    const T acc = sell_row<T>(m, i, p, (T)0, false);
    q[i] = acc;
    part[0] += (double)p[i] * (double)acc;
  }
  double blk[1], tot[1];
  block_sum<1>(part, blk);
  if (last_block_reduce<1>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
    s->pkapk = tot[0];
    s->alf = s->sk / tot[0];
  }
}

// x += alf p ; r -= alf q ; resl = |r|_1 ; |r|_2 ; iteration-1 factor;
// convergence test (crit 0: reference L1, crit 1: relative L2 vs tol_rel).
template <typename T>
__global__ __launch_bounds__(BLK) void k_xr_update(
    int n, const int *__restrict__ diag, const T *__restrict__ a, T *__restrict__ x,
    T *__restrict__ r, const T *__restrict__ p, const T *__restrict__ q,
    double *__restrict__ partials, int *__restrict__ counter, SolveScalars *__restrict__ s,
    int crit, double tol_abs, double tol_rel, int itr_max) {
  if (solve_done(s)) return;
  const T alf = (T)s->alf;
  const bool first = (s->iter == 0);
  double part[3] = {0.0, 0.0, 0.0};
  const int stride = gridDim.x * BLK;
  for (int i0 = blockIdx.x * BLK + threadIdx.x; i0 < n; i0 += GS_U * stride) {
    T xv[GS_U], pv[GS_U], rv[GS_U], qv[GS_U], av[GS_U];
#pragma unroll
    for (int u = 0; u < GS_U; ++u) {
      const int i = i0 + u * stride;
      if (i < n) {
        xv[u] = x[i];
        pv[u] = p[i];
        rv[u] = r[i];
        qv[u] = q[i];
        if (first) av[u] = a[diag[i]];
      }
    }
#pragma unroll
    for (int u = 0; u < GS_U; ++u) {
      const int i = i0 + u * stride;
      if (i < n) {
        const T xi = xv[u] + alf * pv[u];
        x[i] = xi;
        const T ri = rv[u] - alf * qv[u];
        r[i] = ri;
        part[0] += fabs((double)ri);
        part[1] += (double)ri * (double)ri;
        if (first) part[2] += fabs((double)av[u] * (double)xi);
      }
    }
  }
  double blk[3], tot[3];
  block_sum<3>(part, blk);
  if (last_block_reduce<3>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
    const double resl = tot[0];
    s->resl = resl;
    s->l2 = sqrt(tot[1]);
    s->s0 = s->sk;
    s->iter += 1;
    if (first) {
      s->factor = tot[2] + FCP_SMALL;
      s->resor = s->res0 / s->factor;
    }
    const double rsm = resl / (s->res0 + FCP_SMALL);
    s->rsm = rsm;
    bool conv;
    if (crit == 0)
      conv = (rsm < tol_rel) || (resl < tol_abs);
    else
      conv = (s->l2 <= tol_rel * s->l2b) || (resl < tol_abs);
    if (!isfinite(resl)) {
      s->status = FCP_ERR_NONFINITE;
      conv = true;
    }
    s->conv = (conv && s->status == 0) ? 1 : 0;
    if (conv || s->iter >= itr_max) s->done = 1;
  }
}

// ---------------------------------------------------------------------------
// BiCGStab step kernels (bicgstab :642-773)
// ---------------------------------------------------------------------------
// bet = r.reso ; om = bet*gam/(alf*beto+small) ; beto = bet
template <typename T>
__global__ __launch_bounds__(BLK) void k_bicg_bet(int n, const T *__restrict__ r,
                                                  const T *__restrict__ reso,
                                                  double *__restrict__ partials,
                                                  int *__restrict__ counter,
                                                  SolveScalars *__restrict__ s) {
  if (solve_done(s)) return;
  double part[1] = {0.0};
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK)
    part[0] += (double)r[i] * (double)reso[i];
  double blk[1], tot[1];
  block_sum<1>(part, blk);
  if (last_block_reduce<1>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
    const double bet = tot[0];
    s->sk = bet;
    s->om = bet * s->gam / (s->alf * s->beto + FCP_SMALL);
    s->beto = bet;
  }
}

// p = r + om*(p - alf*u)
template <typename T>
__global__ void k_bicg_p(int n, const T *__restrict__ r, const T *__restrict__ u,
                         T *__restrict__ p, const SolveScalars *__restrict__ s) {
  if (solve_done(s)) return;
  const T om = (T)s->om, alf = (T)s->alf;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK)
    p[i] = r[i] + om * (p[i] - alf * u[i]);
}

// u = A z ; ukreso = u.reso ; gam = bet/ukreso
template <typename T>
__global__ __launch_bounds__(BLK) void k_bicg_u(int n, Sell<T> m, const T *__restrict__ z,
                                                const T *__restrict__ reso, T *__restrict__ u,
                                                double *__restrict__ partials,
                                                int *__restrict__ counter,
                                                SolveScalars *__restrict__ s) {
  if (solve_done(s)) return;
  double part[1] = {0.0};
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    const T acc = sell_row<T>(m, i, z, (T)0, false);
    u[i] = acc;
    part[0] += (double)acc * (double)reso[i];
  }
  double blk[1], tot[1];
  block_sum<1>(part, blk);
  if (last_block_reduce<1>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
    s->ukreso = tot[0];
    s->gam = s->sk / tot[0];
  }
}

// x += gam z ; r -= gam u
template <typename T>
__global__ void k_bicg_xr1(int n, const T *__restrict__ z, const T *__restrict__ u,
                           T *__restrict__ x, T *__restrict__ r,
                           const SolveScalars *__restrict__ s) {
  if (solve_done(s)) return;
  const T gam = (T)s->gam;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    x[i] = x[i] + gam * z[i];
    r[i] = r[i] - gam * u[i];
  }
}

// v = A z ; alf = (v.r)/(v.v + small)
template <typename T>
__global__ __launch_bounds__(BLK) void k_bicg_v(int n, Sell<T> m, const T *__restrict__ z,
                                                const T *__restrict__ r, T *__restrict__ v,
                                                double *__restrict__ partials,
                                                int *__restrict__ counter,
                                                SolveScalars *__restrict__ s) {
  if (solve_done(s)) return;
  double part[2] = {0.0, 0.0};
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    const T acc = sell_row<T>(m, i, z, (T)0, false);
    v[i] = acc;
    part[0] += (double)acc * (double)r[i];
    part[1] += (double)acc * (double)acc;
  }
  double blk[2], tot[2];
  block_sum<2>(part, blk);
  if (last_block_reduce<2>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
    s->vr = tot[0];
    s->vv = tot[1];
    s->alf = tot[0] / (tot[1] + FCP_SMALL);
  }
}

// ---------------------------------------------------------------------------
// casts and refinement updates
// ---------------------------------------------------------------------------
__global__ void k_cast_a(int nnz, const double *__restrict__ a, float *__restrict__ a32) {
  for (int i = blockIdx.x * BLK + threadIdx.x; i < nnz; i += gridDim.x * BLK) a32[i] = (float)a[i];
}

// r32 = r64 * scale, e32 = 0
__global__ void k_scale_to32(int n, const double *__restrict__ r, const SolveScalars *__restrict__ so,
                             float *__restrict__ r32, float *__restrict__ e32) {
  const double sc = 1.0 / so->l2;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    r32[i] = (float)(r[i] * sc);
    e32[i] = 0.0f;
  }
}

// x64 += l2 * e32  (FP64 update)
__global__ void k_ir_update(int n, const float *__restrict__ e32, const SolveScalars *__restrict__ so,
                            double *__restrict__ x) {
  const double sc = so->l2;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK)
    x[i] = x[i] + sc * (double)e32[i];
}

__global__ void k_zero_f32(int n, float *__restrict__ a, float *__restrict__ b) {
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    a[i] = 0.0f;
    if (b) b[i] = 0.0f;
  }
}
__global__ void k_zero_f64(int n, double *__restrict__ a, double *__restrict__ b,
                           double *__restrict__ c) {
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    a[i] = 0.0;
    if (b) b[i] = 0.0;
    if (c) c[i] = 0.0;
  }
}

// Copy of the FP32 inner-solve scalars' initial values: r32 given, x=0.
__global__ __launch_bounds__(BLK) void k_init32(int n, const float *__restrict__ r,
                                                float *__restrict__ reso,
                                                double *__restrict__ partials,
                                                int *__restrict__ counter,
                                                SolveScalars *__restrict__ s) {
  double part[2] = {0.0, 0.0};
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    part[0] += fabs((double)r[i]);
    part[1] += (double)r[i] * (double)r[i];
    if (reso) reso[i] = r[i];
  }
  double blk[2], tot[2];
  block_sum<2>(part, blk);
  if (last_block_reduce<2>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
    s->res0 = tot[0];
    s->resl = tot[0];
    s->l2 = sqrt(tot[1]);
    s->l2b = sqrt(tot[1]);
    s->s0 = FCP_S0INIT;
    s->alf = 1.0;
    s->beto = 1.0;
    s->gam = 1.0;
    s->iter = 0;
    s->done = 0;
    s->status = 0;
    s->factor = 0.0;
  }
}


int sweep_grid(fcp_ctx *ctx, const Schedule &s) {
  int waves = s.ntasks;
  int cap = ctx->tune_sweep_waves > 0 ? ctx->tune_sweep_waves : (ctx->ncu > 0 ? ctx->ncu : 256) * 2;
  if (waves > cap) waves = cap;
  int blocks = (waves + (BLK / 64) - 1) / (BLK / 64);
  return blocks > 0 ? blocks : 1;
}

}  // namespace

// ===========================================================================
// host side
// ===========================================================================
namespace fcpk {
#ifndef FCP_RED_BPC
#define FCP_RED_BPC 4
#endif
int red_grid(fcp_ctx *ctx, int n) {
  int g = grid_for(n, BLK);
  int cap = ctx->ncu > 0 ? ctx->ncu * FCP_RED_BPC : 1024;
  if (cap > MAXBLK) cap = MAXBLK;
  return g < cap ? (g > 0 ? g : 1) : cap;
}
}  // namespace fcpk

int fcp_ensure_workspace(fcp_ctx *ctx) {
  const int n = ctx->n;
  DBuf<double> *v64[] = {&ctx->d, &ctx->res, &ctx->pk, &ctx->zk, &ctx->yk, &ctx->uk, &ctx->vk,
                         &ctx->reso, &ctx->x, &ctx->rhs, &ctx->qk};
  for (auto *b : v64) FCP_HIP(ctx, b->alloc(n));
  DBuf<float> *v32[] = {&ctx->d32, &ctx->r32, &ctx->p32, &ctx->z32, &ctx->y32,   &ctx->q32,
                        &ctx->e32, &ctx->u32, &ctx->v32, &ctx->reso32};
  for (auto *b : v32) FCP_HIP(ctx, b->alloc(n));
  FCP_HIP(ctx, ctx->partials.alloc((size_t)MAXBLK * 4 * NSLOT));
  FCP_HIP(ctx, ctx->counters.alloc(CTR_INTS));
  FCP_HIP(ctx, ctx->scal.alloc(4));
  FCP_HIP(ctx, ctx->errflag.alloc(1));
  if (!ctx->h_scal) FCP_HIP(ctx, hipHostMalloc((void **)&ctx->h_scal, 4 * sizeof(SolveScalars)));
  if (!ctx->h_err) FCP_HIP(ctx, hipHostMalloc((void **)&ctx->h_err, sizeof(int)));
  return 0;
}

// Slice lengths (64 x width) of a SELL-64 copy: where padding every slice to
// the widest adds at most 2% slots, every length becomes 64 W and *width = W
// (the products then compute a row's place); else *width = 0.
int sell_uniform(fcp_ctx *ctx, int *len, int nslices, int *width) {
  *width = 0;
  if (nslices <= 0) return 0;
  std::vector<int> h(nslices);
  FCP_HIP(ctx, hipMemcpyAsync(h.data(), len, sizeof(int) * nslices, hipMemcpyDeviceToHost,
                              ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  long long tot = 0;
  int wmax = 0;
  for (int v : h) {
    tot += v;
    wmax = v > wmax ? v : wmax;
  }
  const long long uni = (long long)wmax * nslices;
  if (wmax == 0 || uni * 50 > tot * 51 || uni > (1LL << 30)) return 0;
  std::fill(h.begin(), h.end(), wmax);
  FCP_HIP(ctx, hipMemcpyAsync(len, h.data(), sizeof(int) * nslices, hipMemcpyHostToDevice,
                              ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *width = wmax / 64;
  return 0;
}

int fcp_build_sell(fcp_ctx *ctx) {
  const int n = ctx->n;
  ctx->nslices = (n + 63) / 64;
  DBuf<int> len;
  FCP_HIP(ctx, len.alloc(ctx->nslices + 1));
  FCP_HIP(ctx, ctx->sell_base.alloc(ctx->nslices + 1));
  k_sell_len<<<grid_for(ctx->nslices + 1, BLK), BLK, 0, ctx->stream>>>(ctx->ia.p, n, ctx->nslices,
                                                                       len.p);
  FCP_TRY(sell_uniform(ctx, len.p, ctx->nslices, &ctx->sell_width));
  FCP_TRY(fcp_exclusive_scan_int(ctx, len.p, ctx->sell_base.p, ctx->nslices + 1));
  int npad = 0;
  FCP_HIP(ctx, hipMemcpyAsync(&npad, ctx->sell_base.p + ctx->nslices, sizeof(int),
                              hipMemcpyDeviceToHost, ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  FCP_HIP(ctx, ctx->sell_col.alloc(npad > 0 ? npad : 1));
  k_sell_col<<<grid_for((long long)ctx->nslices * 64, BLK), BLK, 0, ctx->stream>>>(
      ctx->ia.p, ctx->ja.p, n, ctx->nslices, ctx->sell_base.p, ctx->sell_col.p);
  FCP_HIP(ctx, hipGetLastError());
  ctx->sell64_version = ctx->sell32_version = ~0ull;
  return 0;
}

namespace {
// SELL values of the resident matrix in precision T, refreshed on change.
template <typename T>
int sell_view(fcp_ctx *ctx, Sell<T> *out);
template <>
int sell_view<double>(fcp_ctx *ctx, Sell<double> *out) {
  if (ctx->sell64_version != ctx->a_version) {
    const int npad = (int)ctx->sell_col.n;
    FCP_HIP(ctx, ctx->sell_a64.alloc(npad));
    k_sell_val<double><<<grid_for((long long)ctx->nslices * 64, BLK), BLK, 0, ctx->stream>>>(
        ctx->ia.p, ctx->a.p, ctx->n, ctx->nslices, ctx->sell_base.p, ctx->sell_a64.p);
    FCP_HIP(ctx, hipGetLastError());
    ctx->sell64_version = ctx->a_version;
  }
  *out = Sell<double>{ctx->sell_base.p, ctx->sell_col.p, ctx->ia.p, ctx->sell_a64.p};
  out->width = ctx->sell_width;
  return 0;
}
template <>
int sell_view<float>(fcp_ctx *ctx, Sell<float> *out) {
  if (ctx->sell32_version != ctx->a_version) {
    const int npad = (int)ctx->sell_col.n;
    FCP_HIP(ctx, ctx->sell_a32.alloc(npad));
    k_sell_val<float><<<grid_for((long long)ctx->nslices * 64, BLK), BLK, 0, ctx->stream>>>(
        ctx->ia.p, ctx->a.p, ctx->n, ctx->nslices, ctx->sell_base.p, ctx->sell_a32.p);
    FCP_HIP(ctx, hipGetLastError());
    ctx->sell32_version = ctx->a_version;
  }
  *out = Sell<float>{ctx->sell_base.p, ctx->sell_col.p, ctx->ia.p, ctx->sell_a32.p};
  out->width = ctx->sell_width;
  return 0;
}
}  // namespace

namespace fcpk {
int check_err(fcp_ctx *ctx) { return poll_device(ctx, nullptr, nullptr, 0); }
int poll_device(fcp_ctx *ctx, void *h, const void *d, size_t bytes) {
  if (!ctx->h_err) FCP_HIP(ctx, hipHostMalloc((void **)&ctx->h_err, sizeof(int)));
  if (bytes) FCP_HIP(ctx, hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
  FCP_HIP(ctx, hipMemcpyAsync(ctx->h_err, ctx->errflag.p, sizeof(int), hipMemcpyDeviceToHost,
                              ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const int e = *ctx->h_err;
  if (e) return fcp_fail(ctx, e, "sweep dependency wait exceeded its bound");
  return 0;
}
int sell_view64(fcp_ctx *ctx, Sell<double> *out) { return sell_view<double>(ctx, out); }
int sell_view32(fcp_ctx *ctx, Sell<float> *out) { return sell_view<float>(ctx, out); }
}  // namespace fcpk

int fcp_launch_spmv64(fcp_ctx *ctx, const double *, const double *x, double *y) {
  Sell<double> m;
  FCP_TRY(sell_view<double>(ctx, &m));
  ProfScope ps(ctx, "spmv_sell_f64");
  const int n = ctx->n;
  // automatic: non-temporal matrix loads and y stores once the operator
  // (values and columns) and the two vectors exceed the 256 MB Infinity
  // Cache, so that streaming them does not evict the gathered x lines; cached
  // loads while it all fits (128^3: 31.8 against 38.6 us per SpMV, 192^3:
  // 126.1 against 142.5 us)
  int kind = ctx->tune_spmv;
  if (kind < 0) kind = (size_t)ctx->nnz * 12 + (size_t)n * 16 > ((size_t)256 << 20) ? 4 : 0;
  switch (kind) {
    case 4:
      k_spmv_rows<double, true, 1, true><<<(grid_for(n, BLK) + 7) / 8 * 8, BLK, 0, ctx->stream>>>(n, m, x, y);
      break;
    case 1:
      k_spmv_rows<double, true, 1><<<(grid_for(n, BLK) + 7) / 8 * 8, BLK, 0, ctx->stream>>>(n, m, x, y);
      break;
    case 2:
      k_spmv_rows<double, true, 2><<<(grid_for(n, 2 * BLK) + 7) / 8 * 8, BLK, 0, ctx->stream>>>(n, m, x, y);
      break;
    case 3:
      k_spmv_rows<double, false, 2><<<(grid_for(n, 2 * BLK) + 7) / 8 * 8, BLK, 0, ctx->stream>>>(n, m, x, y);
      break;
    default:
      k_spmv_row<double><<<(grid_for(n, BLK) + 7) / 8 * 8, BLK, 0, ctx->stream>>>(n, m, x, y);
  }
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

namespace {

// counters layout: [0..1] sweep tickets, [2] reduction counter

template <typename T, int MODE>
int launch_bsweep(fcp_ctx *ctx, const BlockSchedule &b, const T *a, const T *d, const T *in, T *out,
                  int *ticket, const SolveScalars *s) {
  const size_t lds = (size_t)b.R * sizeof(T);
  k_bsweep<T, MODE><<<b.nblocks, BLK, lds, ctx->stream>>>(
      b.rows.p, b.step_ptr.p, b.blk_step.p, b.nblocks, b.R, ctx->n, ctx->ia.p, ctx->diag.p,
      ctx->ja.p, a, ctx->tslot.p, d, in, out, ticket, ctx->errflag.p, s, ctx->tune_sweep_sleep);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

template <typename T>
struct PBufs;
template <>
struct PBufs<double> {
  static DBuf<double> &pa(BlockSchedule &b) { return b.pa64; }
  static DBuf<double> &pdx(BlockSchedule &b) { return b.pdx64; }
  static uint64_t &ver(BlockSchedule &b) { return b.pver64; }
};
template <>
struct PBufs<float> {
  static DBuf<float> &pa(BlockSchedule &b) { return b.pa32; }
  static DBuf<float> &pdx(BlockSchedule &b) { return b.pdx32; }
  static uint64_t &ver(BlockSchedule &b) { return b.pver32; }
};

}  // namespace

namespace fcpk {
// Blocks of one sweep pair fit the pipelined kernel: at most FCP_PMAXD
// dependencies per row, block values plus step table within the LDS limit.
bool sched_fits(fcp_ctx *ctx, const BlockSchedule &lo, const BlockSchedule &up, size_t row_bytes) {
  for (const BlockSchedule *b : {&lo, &up}) {
    if (b->nblocks == 0 || b->maxd > FCP_PMAXD) return false;
    const size_t lds = (((size_t)b->R * row_bytes + 15) & ~(size_t)15) +
                       (size_t)(b->max_steps + 2) * sizeof(int) + 256;
    if (lds > ctx->lds_cap) return false;
  }
  return true;
}
}  // namespace fcpk

namespace {

// Geometric tiles when they were built and fit, else contiguous row blocks
// sized for precision T.
template <typename T>
BlockSchedule &psched(fcp_ctx *ctx, bool upper) {
  if (ctx->have_tiles && sched_fits(ctx, ctx->tlower, ctx->tupper, sizeof(T) + sizeof(int)))
    return upper ? ctx->tupper : ctx->tlower;
  if (sizeof(T) == 4) return upper ? ctx->bupper4 : ctx->blower4;
  return upper ? ctx->bupper : ctx->blower;
}

// The pipelined sweep applies when the packed operands cover every row and
// the block plus its step table fit in LDS.
template <typename T>
bool use_psweep(fcp_ctx *ctx) {
  if (ctx->tune_sweep_kind != 0) return false;
  // remote values are read through a buffer resource of n*sizeof(T) bytes
  if ((long long)ctx->n * (long long)sizeof(T) >= FCP_OOB) return false;
  // per row: its value and its row id in LDS
  return sched_fits(ctx, psched<T>(ctx, false), psched<T>(ctx, true), sizeof(T) + sizeof(int));
}

template <typename T, int MODE>
int launch_psweep(fcp_ctx *ctx, BlockSchedule &b, const T *in, T *out, int *ticket,
                  const SolveScalars *s) {
  const int n = ctx->n;
  PArgs<T> P{b.prc.p, b.pdep.p, PBufs<T>::pa(b).p, PBufs<T>::pdx(b).p, in, n};
  const int sptr_off = (int)(((size_t)b.R * sizeof(T) + 15) & ~(size_t)15);
  const int zr_off =
      (int)(((size_t)sptr_off + (size_t)(b.max_steps + 2) * sizeof(int) + 15) & ~(size_t)15);
  const size_t lds = (size_t)zr_off + (size_t)b.R * sizeof(int);
  k_psweep<T, MODE><<<b.nblocks, BLK, lds, ctx->stream>>>(P, b.step_ptr.p, b.blk_step.p,
                                                          b.blk_pos.p, b.nblocks, sptr_off, zr_off,
                                                          out, ticket, ctx->errflag.p, s,
                                                          ctx->tune_sweep_sleep);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

// Packs matrix coefficients and pivots of a block schedule when they changed
// (keyed by the preconditioner setup epoch).
template <typename T>
int pack_block(fcp_ctx *ctx, BlockSchedule &b, const double *d, uint64_t ver) {
  const int n = ctx->n;
  if (PBufs<T>::ver(b) == ver) return 0;
  FCP_HIP(ctx, PBufs<T>::pa(b).alloc((size_t)FCP_PMAXD * n));
  FCP_HIP(ctx, PBufs<T>::pdx(b).alloc((size_t)2 * n));
  k_pack_vals<T><<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(n, b.prc.p, b.psrc.p, ctx->a.p, d,
                                                            PBufs<T>::pa(b).p,
                                                            PBufs<T>::pdx(b).p);
  FCP_HIP(ctx, hipGetLastError());
  PBufs<T>::ver(b) = ver;
  return 0;
}

template <typename T>
int pack_block(fcp_ctx *ctx, BlockSchedule &b) {
  return pack_block<T>(ctx, b, ctx->d.p, ctx->d_epoch);
}

// Packs both sweeps of the single-vector schedules.
template <typename T>
int psweep_pack(fcp_ctx *ctx) {
  for (int u = 0; u < 2; ++u) FCP_TRY(pack_block<T>(ctx, psched<T>(ctx, u)));
  return 0;
}

bool use_block_sweep(fcp_ctx *ctx, size_t elem) {
  return ctx->tune_sweep_kind == 0 && ctx->blower.nblocks > 0 &&
         (size_t)ctx->blower.R * elem + 256 <= ctx->lds_cap;
}

// Diagonal of the DIC (kind 0) or DILU (kind 1) preconditioner in FP64,
// recomputed when the matrix changed.
template <typename T>
int precond_setup(fcp_ctx *ctx, int kind, const T *a, T *d) {
  const int n = ctx->n;
  k_fill_sent<T><<<red_grid(ctx, n), BLK, 0, ctx->stream>>>(d, (T *)nullptr, n, tickets(ctx), nullptr);
  ctx->d_epoch++;
  ProfScope ps(ctx, kind ? "dilu_setup" : "dic_setup");
  if (use_block_sweep(ctx, sizeof(T))) {
    if (kind == 0)
      return launch_bsweep<T, 2>(ctx, ctx->blower, a, nullptr, nullptr, d, tickets(ctx), nullptr);
    return launch_bsweep<T, 3>(ctx, ctx->blower, a, nullptr, nullptr, d, tickets(ctx), nullptr);
  }
  if (kind == 0)
    k_sweep<T, 2><<<sweep_grid(ctx, ctx->lower), BLK, 0, ctx->stream>>>(
        ctx->lower.task_ptr.p, ctx->lower.order.p, ctx->lower.ntasks, ctx->ia.p, ctx->diag.p,
        ctx->ja.p, a, ctx->tslot.p, nullptr, nullptr, d, tickets(ctx), ctx->errflag.p, nullptr, ctx->tune_sweep_sleep);
  else
    k_sweep<T, 3><<<sweep_grid(ctx, ctx->lower), BLK, 0, ctx->stream>>>(
        ctx->lower.task_ptr.p, ctx->lower.order.p, ctx->lower.ntasks, ctx->ia.p, ctx->diag.p,
        ctx->ja.p, a, ctx->tslot.p, nullptr, nullptr, d, tickets(ctx), ctx->errflag.p, nullptr, ctx->tune_sweep_sleep);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Tile-local DIC / DILU (preconditioner of the FP32 inner solves): the same
// per-row recurrences restricted to dependencies inside a block, so every
// block of a sweep runs independently (blocks never wait on one another).
// ---------------------------------------------------------------------------

// Pivots over the block schedule jlower, one workgroup per block, steps in
// order: d_i = 1/(a_ii - sum a_ik^2 d_k) (kind 0) or
// 1/(a_ii - sum a_ik d_k a_ki) (kind 1) over the block's own k < i, summed in
// CSR order as the global setup does.
template <int KIND>
__global__ __launch_bounds__(BLK) void k_bj_setup(const int *__restrict__ prc,
                                                  const int *__restrict__ pdep,
                                                  const int *__restrict__ psrc,
                                                  const int *__restrict__ step_ptr,
                                                  const int *__restrict__ blk_step,
                                                  const int *__restrict__ blk_pos,
                                                  const int *__restrict__ diag,
                                                  const double *__restrict__ a,
                                                  const int *__restrict__ tslot,
                                                  double *__restrict__ dbj) {
  extern __shared__ double dl[];
  const int b = blockIdx.x;
  const int lo = blk_pos[b];
  for (int st = blk_step[b], se = blk_step[b + 1]; st < se; ++st) {
    const int p = step_ptr[st] + (int)threadIdx.x;
    if (p < step_ptr[st + 1]) {
      const int row = prc[2 * p];
      const int cnt = prc[2 * p + 1] & 0xFFFF;
      double acc = a[diag[row]];
      for (int m = 0; m < cnt; ++m) {
        const int k = psrc[(size_t)FCP_PMAXD * p + m];
        const double dj = dl[pdep[(size_t)FCP_PMAXD * p + m]];
        if (KIND == 0) {
          const double ak = a[k];
          acc = acc - ak * ak * dj;
        } else {
          const int l = tslot[k];
          acc = acc - a[k] * dj * (l >= 0 ? a[l] : 0.0);
        }
      }
      const double dv = 1.0 / acc;
      dl[p - lo] = dv;
      dbj[row] = dv;
    }
    __syncthreads();
  }
}

__global__ void k_reset_tickets(int *__restrict__ tickets) {
  tickets[0] = 0;
  tickets[1] = 0;
}

bool bj_usable(fcp_ctx *ctx) {
  return ctx->have_bj && ctx->tune_sweep_kind == 0 &&
         (long long)ctx->n * (long long)sizeof(float) < FCP_OOB &&
         sched_fits(ctx, ctx->jlower, ctx->jupper, sizeof(float) + sizeof(int)) &&
         (size_t)ctx->jlower.R * sizeof(double) + 256 <= ctx->lds_cap;
}

// Schedules, FP64 pivots and packed FP32 operands of the tile-local
// preconditioner of kind 0 (DIC) or 1 (DILU); false when it does not apply.
int bj_prepare(fcp_ctx *ctx, int kind, bool *ok) {
  *ok = false;
  if (!ctx->have_bj) FCP_TRY(fcp_build_bj_schedules(ctx));
  if (!bj_usable(ctx)) return 0;
  const int n = ctx->n;
  if (ctx->bj_version != ctx->a_version || ctx->bj_kind != kind) {
    FCP_HIP(ctx, ctx->dbj.alloc(n));
    const BlockSchedule &b = ctx->jlower;
    const size_t lds = (size_t)b.R * sizeof(double);
    ProfScope ps(ctx, kind ? "bj_dilu_setup" : "bj_dic_setup");
    if (kind == 0)
      k_bj_setup<0><<<b.nblocks, BLK, lds, ctx->stream>>>(
          b.prc.p, b.pdep.p, b.psrc.p, b.step_ptr.p, b.blk_step.p, b.blk_pos.p, ctx->diag.p,
          ctx->a.p, ctx->tslot.p, ctx->dbj.p);
    else
      k_bj_setup<1><<<b.nblocks, BLK, lds, ctx->stream>>>(
          b.prc.p, b.pdep.p, b.psrc.p, b.step_ptr.p, b.blk_step.p, b.blk_pos.p, ctx->diag.p,
          ctx->a.p, ctx->tslot.p, ctx->dbj.p);
    FCP_HIP(ctx, hipGetLastError());
    ctx->bj_version = ctx->a_version;
    ctx->bj_kind = kind;
    ctx->bj_epoch++;
  }
  *ok = true;
  return 0;
}

// FP32 operands of the row-gather tile-local sweeps, packed with the tile
// pivots on first use after the pivots change (the tile-order kernels read
// their own layout and never need them).
int bj_pack32_now(fcp_ctx *ctx) {
  for (BlockSchedule *b : {&ctx->jlower, &ctx->jupper})
    FCP_TRY(pack_block<float>(ctx, *b, ctx->dbj.p, ctx->bj_epoch | (1ull << 63)));
  return 0;
}

// z = M^{-1} r with the tile-local sweeps (forward into y, backward into z);
// no dependency crosses a block, so no sentinel fill is needed.
int bj_apply32(fcp_ctx *ctx, const float *r, float *y, float *z, const SolveScalars *s) {
  FCP_TRY(bj_pack32_now(ctx));
  k_reset_tickets<<<1, 1, 0, ctx->stream>>>(tickets(ctx));
  {
    ProfScope ps(ctx, "bj_sweep_fwd_f32");
    const int rc = launch_psweep<float, 0>(ctx, ctx->jlower, r, y, tickets(ctx), s);
    if (rc < 0) return rc;
  }
  ProfScope ps(ctx, "bj_sweep_bwd_f32");
  return launch_psweep<float, 1>(ctx, ctx->jupper, y, z, tickets(ctx) + 1, s);
}

// z = M^{-1} r (forward into y, backward into z), both sentinel-filled first.
template <typename T>
int precond_apply(fcp_ctx *ctx, const T *a, const T *d, const T *r, T *y, T *z,
                  const SolveScalars *s) {
  const int n = ctx->n;
  if (sizeof(T) == 4 && ctx->bj_active)
    return bj_apply32(ctx, reinterpret_cast<const float *>(r), reinterpret_cast<float *>(y),
                      reinterpret_cast<float *>(z), s);
  k_fill_sent<T><<<red_grid(ctx, n), BLK, 0, ctx->stream>>>(y, z, n, tickets(ctx), s);
  if (use_psweep<T>(ctx)) {
    FCP_TRY(psweep_pack<T>(ctx));
    {
      ProfScope ps(ctx, sizeof(T) == 8 ? "sweep_fwd_f64" : "sweep_fwd_f32");
      const int rc = launch_psweep<T, 0>(ctx, psched<T>(ctx, false), r, y, tickets(ctx), s);
      if (rc < 0) return rc;
    }
    ProfScope ps(ctx, sizeof(T) == 8 ? "sweep_bwd_f64" : "sweep_bwd_f32");
    return launch_psweep<T, 1>(ctx, psched<T>(ctx, true), y, z, tickets(ctx) + 1, s);
  }
  if (use_block_sweep(ctx, sizeof(T))) {
    {
      ProfScope ps(ctx, sizeof(T) == 8 ? "sweep_fwd_f64" : "sweep_fwd_f32");
      const int rc = launch_bsweep<T, 0>(ctx, ctx->blower, a, d, r, y, tickets(ctx), s);
      if (rc < 0) return rc;
    }
    ProfScope ps(ctx, sizeof(T) == 8 ? "sweep_bwd_f64" : "sweep_bwd_f32");
    return launch_bsweep<T, 1>(ctx, ctx->bupper, a, d, y, z, tickets(ctx) + 1, s);
  }
  {
    ProfScope ps(ctx, sizeof(T) == 8 ? "sweep_fwd_f64" : "sweep_fwd_f32");
    k_sweep<T, 0><<<sweep_grid(ctx, ctx->lower), BLK, 0, ctx->stream>>>(
        ctx->lower.task_ptr.p, ctx->lower.order.p, ctx->lower.ntasks, ctx->ia.p, ctx->diag.p,
        ctx->ja.p, a, nullptr, d, r, y, tickets(ctx), ctx->errflag.p, s, ctx->tune_sweep_sleep);
  }
  {
    ProfScope ps(ctx, sizeof(T) == 8 ? "sweep_bwd_f64" : "sweep_bwd_f32");
    k_sweep<T, 1><<<sweep_grid(ctx, ctx->upper), BLK, 0, ctx->stream>>>(
        ctx->upper.task_ptr.p, ctx->upper.order.p, ctx->upper.ntasks, ctx->ia.p, ctx->diag.p,
        ctx->ja.p, a, nullptr, d, y, z, tickets(ctx) + 1, ctx->errflag.p, s, ctx->tune_sweep_sleep);
  }
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

struct Vecs {
  void *x, *r, *p, *z, *y, *q, *u, *v, *reso;
};

// One PCG iteration (precond: 0 DIC, 1 Jacobi).
template <typename T>
int pcg_iteration(fcp_ctx *ctx, int precond, const T *a, const T *d, T *x, T *r, T *p, T *z,
                  T *y, T *q, SolveScalars *s, int crit, double tol_abs, double tol_rel,
                  int itr_max) {
  const int n = ctx->n;
  const int g = red_grid(ctx, n);
  if (precond == 0) {
    FCP_TRY(precond_apply<T>(ctx, a, d, r, y, z, s));
    ProfScope ps(ctx, "dot_rz");
    k_dot_rz<T><<<g, BLK, 0, ctx->stream>>>(n, r, z, ctx->partials.p, redctr(ctx), s);
  } else {
    ProfScope ps(ctx, "jacobi_dot");
    k_jacobi_dot<T><<<g, BLK, 0, ctx->stream>>>(n, ctx->diag.p, a, r, z, ctx->partials.p,
                                                 redctr(ctx), s);
  }
  {
    ProfScope ps(ctx, "pupdate");
    k_pupdate<T><<<g, BLK, 0, ctx->stream>>>(n, z, p, s);
  }
  {
    ProfScope ps(ctx, sizeof(T) == 8 ? "spmv_pq_f64" : "spmv_pq_f32");
    Sell<T> m;
    FCP_TRY(sell_view<T>(ctx, &m));
    k_spmv_pq<T><<<g, BLK, 0, ctx->stream>>>(n, m, p, q, ctx->partials.p, redctr(ctx), s);
  }
  {
    ProfScope ps(ctx, "xr_update");
    k_xr_update<T><<<g, BLK, 0, ctx->stream>>>(n, ctx->diag.p, a, x, r, p, q, ctx->partials.p,
                                                redctr(ctx), s, crit, tol_abs, tol_rel, itr_max);
  }
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

template <typename T>
int bicgstab_iteration(fcp_ctx *ctx, const T *a, const T *d, T *x, T *r, T *reso, T *p, T *z,
                       T *y, T *u, T *v, SolveScalars *s, int crit, double tol_abs,
                       double tol_rel, int itr_max) {
  const int n = ctx->n;
  const int g = red_grid(ctx, n);
  k_bicg_bet<T><<<g, BLK, 0, ctx->stream>>>(n, r, reso, ctx->partials.p, redctr(ctx), s);
  k_bicg_p<T><<<g, BLK, 0, ctx->stream>>>(n, r, u, p, s);
  FCP_TRY(precond_apply<T>(ctx, a, d, p, y, z, s));
  Sell<T> m;
  FCP_TRY(sell_view<T>(ctx, &m));
  k_bicg_u<T><<<g, BLK, 0, ctx->stream>>>(n, m, z, reso, u, ctx->partials.p, redctr(ctx), s);
  k_bicg_xr1<T><<<g, BLK, 0, ctx->stream>>>(n, z, u, x, r, s);
  FCP_TRY(precond_apply<T>(ctx, a, d, r, y, z, s));
  k_bicg_v<T><<<g, BLK, 0, ctx->stream>>>(n, m, z, r, v, ctx->partials.p, redctr(ctx), s);
  // x += alf z ; r -= alf v ; L1 etc. (reuse the PCG update with p=z, q=v)
  k_xr_update<T><<<g, BLK, 0, ctx->stream>>>(n, ctx->diag.p, a, x, r, z, v, ctx->partials.p,
                                              redctr(ctx), s, crit, tol_abs, tol_rel, itr_max);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

// Runs iterations until the device reports done; returns 0 or error.
template <typename F>
int iterate_until_done(fcp_ctx *ctx, SolveScalars *s_dev, int slot, int itr_max, int check_every,
                       int verbose, const char *tag, int mode, double tol, F &&one_iteration) {
  int launched = 0;
  BatchSizer bs(check_every, mode, tol);
  int next = bs.cap;
  while (launched < itr_max) {
    const int batch = (itr_max - launched) < next ? (itr_max - launched) : next;
    for (int k = 0; k < batch; ++k) FCP_TRY(one_iteration());
    launched += batch;
    FCP_TRY(poll_device(ctx, &ctx->h_scal[slot], s_dev, sizeof(SolveScalars)));
    if (verbose >= 2)
      printf("                   %s iter = %4d resl = %10.3E rsm = %10.3E\n", tag,
             ctx->h_scal[slot].iter, ctx->h_scal[slot].resl, ctx->h_scal[slot].rsm);
    if (ctx->h_scal[slot].done) break;
    next = bs.next(&ctx->h_scal[slot], 1);
  }
  return 0;
}

int solve_fp64(fcp_ctx *ctx, int solver, double *x, const double *b, int itr_max, double tol_abs,
               double tol_rel, int crit, int check_every, int verbose) {
  const int n = ctx->n;
  SolveScalars *s = ctx->scal.p;
  const double *a = ctx->a.p;
  const int g = red_grid(ctx, n);
  FCP_HIP(ctx, hipMemsetAsync(ctx->counters.p, 0, CTR_INTS * sizeof(int), ctx->stream));
  FCP_HIP(ctx, hipMemsetAsync(ctx->errflag.p, 0, sizeof(int), ctx->stream));
  k_zero_f64<<<g, BLK, 0, ctx->stream>>>(n, ctx->pk.p, ctx->uk.p, ctx->zk.p);
  {
    Sell<double> m;
    FCP_TRY(sell_view<double>(ctx, &m));
    k_residual_init<double><<<g, BLK, 0, ctx->stream>>>(n, m, x, b, ctx->res.p,
                                                         solver == 2 ? ctx->reso.p : nullptr,
                                                         ctx->partials.p, redctr(ctx), s);
  }
  FCP_HIP(ctx, hipGetLastError());
  FCP_HIP(ctx, hipMemcpyAsync(&ctx->h_scal[0], s, sizeof(SolveScalars), hipMemcpyDeviceToHost,
                              ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (solver == 3) {
    // Gauss-Seidel: its residual norms come from the sweeps themselves
    // (the early exit at :158-165 follows the first sweep).
    auto it = [&]() {
      k_fill_sent<double><<<g, BLK, 0, ctx->stream>>>(ctx->yk.p, nullptr, n, tickets(ctx), s);
      {
        ProfScope ps(ctx, "gs_sweep");
        k_gs_sweep<<<sweep_grid(ctx, ctx->lower), BLK, 0, ctx->stream>>>(
            ctx->lower.task_ptr.p, ctx->lower.order.p, ctx->lower.ntasks, ctx->ia.p, ctx->diag.p,
            ctx->ja.p, a, b, x, ctx->yk.p, ctx->res.p, tickets(ctx), ctx->errflag.p, s,
            ctx->tune_sweep_sleep);
      }
      k_gs_update<<<g, BLK, 0, ctx->stream>>>(n, ctx->diag.p, a, ctx->yk.p, x, ctx->res.p,
                                               ctx->partials.p, redctr(ctx), s, crit, tol_abs,
                                               tol_rel, itr_max);
      FCP_HIP(ctx, hipGetLastError());
      return 0;
    };
    return iterate_until_done(ctx, s, 0, itr_max, check_every, verbose, "GS",
                              crit == 0 ? BATCH_L1 : BATCH_L2, tol_rel, it);
  }
  if (ctx->h_scal[0].res0 < tol_abs) return 0;  // :428-432
  if (verbose >= 2) printf("                    res0 = %10.3E\n", ctx->h_scal[0].res0);
  if (solver == 0 || solver == 2) {
    const int kind = solver == 2 ? 1 : 0;
    if (ctx->dic_version != ctx->a_version || ctx->dic_kind != kind) {
      FCP_TRY(precond_setup<double>(ctx, kind, a, ctx->d.p));
      ctx->dic_version = ctx->a_version;
      ctx->dic_kind = kind;
    }
  }
  if (solver == 0 || solver == 1) {
    auto it = [&]() {
      return pcg_iteration<double>(ctx, solver == 0 ? 0 : 1, a, ctx->d.p, x, ctx->res.p, ctx->pk.p,
                                   ctx->zk.p, ctx->yk.p, ctx->qk.p, s, crit, tol_abs, tol_rel,
                                   itr_max);
    };
    FCP_TRY(iterate_until_done(ctx, s, 0, itr_max, check_every, verbose, "PCG",
                               crit == 0 ? BATCH_L1 : BATCH_L2, tol_rel, it));
  } else {
    auto it = [&]() {
      return bicgstab_iteration<double>(ctx, a, ctx->d.p, x, ctx->res.p, ctx->reso.p, ctx->pk.p,
                                        ctx->zk.p, ctx->yk.p, ctx->uk.p, ctx->vk.p, s, crit,
                                        tol_abs, tol_rel, itr_max);
    };
    FCP_TRY(iterate_until_done(ctx, s, 0, itr_max, check_every, verbose, "BiCGStab",
                               crit == 0 ? BATCH_L1 : BATCH_L2, tol_rel, it));
  }
  return 0;
}

}  // namespace
namespace fcpk {
int precond_pivots64(fcp_ctx *ctx, int kind, const double *a, double *d) {
  return precond_setup<double>(ctx, kind, a, d);
}
}  // namespace fcpk
namespace {

// True FP64 residual r = b - A x with norms into scalars slot.
int true_residual(fcp_ctx *ctx, const double *x, const double *b, SolveScalars *s) {
  const int n = ctx->n;
  Sell<double> m;
  FCP_TRY(sell_view<double>(ctx, &m));
  ProfScope ps(ctx, "residual_f64");
  k_residual_init<double><<<red_grid(ctx, n), BLK, 0, ctx->stream>>>(
      n, m, x, b, ctx->res.p, nullptr, ctx->partials.p, redctr(ctx), s);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

int solve_mixed(fcp_ctx *ctx, int solver, double *x, const double *b, int itr_max, double tol_abs,
                double tol_rel, int check_every, int verbose, fcp_solve_opts *opts) {
  const int n = ctx->n;
  const int g = red_grid(ctx, n);
  SolveScalars *so = ctx->scal.p;      // outer (FP64)
  SolveScalars *si = ctx->scal.p + 1;  // inner (FP32; FP64 for FCP_PREC_FP64_TILE)
  // FP64 inner iterations with the inner solves' two-level tile
  // preconditioner (PCG only), under the same refinement driver
  const bool tile64 = opts && opts->precision == FCP_PREC_FP64_TILE;
  if (tile64 && solver != 0)
    return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "FCP_PREC_FP64_TILE applies to iccg only");
  FCP_HIP(ctx, hipMemsetAsync(ctx->counters.p, 0, CTR_INTS * sizeof(int), ctx->stream));
  FCP_HIP(ctx, hipMemsetAsync(ctx->errflag.p, 0, sizeof(int), ctx->stream));
  // FP32 operator and preconditioner (diagonal computed in FP64 then rounded)
  // (the CSR-order FP32 copy of A is cast when the row-gather inner
  // iterations first need it; the tile-order kernels have their own copy)
  FCP_HIP(ctx, ctx->a32.alloc(ctx->nnz));
  const int kind = solver == 2 ? 1 : 0;
  // global DIC / DILU pivots (FP64 and their FP32 rounding), computed when a
  // sweep needs them
  auto global_pivots = [&]() -> int {
    if (solver == 1) return 0;
    if (ctx->dic_version != ctx->a_version || ctx->dic_kind != kind) {
      FCP_TRY(precond_setup<double>(ctx, kind, ctx->a.p, ctx->d.p));
      ctx->dic_version = ctx->a_version;
      ctx->dic_kind = kind;
    }
    if (ctx->d32_version != ctx->dic_version) {
      k_cast_a<<<g, BLK, 0, ctx->stream>>>(n, ctx->d.p, ctx->d32.p);
      ctx->d32_version = ctx->dic_version;
    }
    return hipGetLastError() == hipSuccess ? 0 : FCP_ERR_HIP;
  };
  // FP32 inner sweeps with the tile-local preconditioner when selected (the
  // FP64 refinement and any FP64 fallback keep the global one)
  // the coarse factorisation of the inner PCG overlaps the set-up below;
  // joined before its first use or on return
  struct CoarseJoin {
    fcp_ctx *c;
    ~CoarseJoin() { (void)tin_coarse_join(c); }
  } coarse_join{ctx};
  if (solver == 0 && (ctx->inner_precond == 1 || tile64)) FCP_TRY(tin_coarse_begin(ctx));
  bool use_bj = false;
  if (solver != 1 && (ctx->inner_precond >= 1 || tile64)) FCP_TRY(bj_prepare(ctx, kind, &use_bj));
  if (tile64 && !(use_bj && tin_usable(ctx, 1, true)))
    return fcp_fail(ctx, FCP_ERR_UNSUPPORTED,
                    "FCP_PREC_FP64_TILE: the tile preconditioner does not apply to this mesh");
  if (!use_bj) FCP_TRY(global_pivots());
  struct BjScope {
    fcp_ctx *c;
    ~BjScope() { c->bj_active = false; }
  } bj_scope{ctx};
  ctx->bj_active = use_bj;
  FCP_TRY(true_residual(ctx, x, b, so));
  const int max_outer = (opts && opts->max_outer > 0) ? opts->max_outer : 30;
  const double inner_tol = (opts && opts->inner_tol > 0) ? opts->inner_tol : 1e-4;
  int total_inner = 0, outer = 0;
  double prev = 0.0;
  bool fp64_fallback = false;
  for (; outer < max_outer; ++outer) {
    FCP_HIP(ctx, hipMemcpyAsync(&ctx->h_scal[0], so, sizeof(SolveScalars), hipMemcpyDeviceToHost,
                                ctx->stream));
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double rel = ctx->h_scal[0].l2 / (ctx->h_scal[0].l2b > 0 ? ctx->h_scal[0].l2b : 1.0);
    if (outer == 0) {
      if (opts) opts->res0 = ctx->h_scal[0].res0;
    }
    if (verbose >= 2) printf("                    IR outer %d  |r|/|b| = %10.3E\n", outer, rel);
    if (!std::isfinite(rel)) return fcp_fail(ctx, FCP_ERR_NONFINITE, "non-finite residual in refinement");
    if (rel <= tol_rel || ctx->h_scal[0].l2 == 0.0) break;
    if (total_inner >= itr_max) break;
    if (outer > 0 && rel > 0.5 * prev) fp64_fallback = true;  // stagnation
    prev = rel;
    if (fp64_fallback) {
      // finish with FP64 inner iterations, warm-started from x
      FCP_TRY(global_pivots());
      FCP_TRY(solve_fp64(ctx, solver, x, b, itr_max - total_inner, tol_abs, tol_rel, 1, check_every,
                         verbose));
      total_inner += ctx->h_scal[0].iter;
      FCP_TRY(true_residual(ctx, x, b, so));
      continue;
    }
    const int budget = itr_max - total_inner;
    if (tile64) {
      // FP64 inner PCG in tile order, reduced as far as this refinement step
      // needs (no floor: FP64 reaches it)
      FCP_TRY(tin_pcg(ctx, 1, ctx->res.p, x, so, si, tol_rel, 0.0, budget, check_every, verbose,
                      &ctx->h_scal[1], true));
      total_inner += ctx->h_scal[1].iter;
      FCP_TRY(true_residual(ctx, x, b, so));
      continue;
    }
    // inner FP32 PCG / BiCGStab in tile order (tile-local preconditioner)
    if ((solver == 0 || solver == 2) && use_bj && ctx->inner_precond == 1 && tin_usable(ctx, 1)) {
      if (solver == 0)
        FCP_TRY(tin_pcg(ctx, 1, ctx->res.p, x, so, si, tol_rel, inner_tol, budget, check_every,
                        verbose, &ctx->h_scal[1]));
      else
        FCP_TRY(tin_bicgstab(ctx, 1, ctx->res.p, x, so, si, tol_rel, inner_tol, budget,
                             check_every, verbose, &ctx->h_scal[1]));
      total_inner += ctx->h_scal[1].iter;
      FCP_TRY(true_residual(ctx, x, b, so));
      continue;
    }
    // inner FP32 solve of A e = r/|r|
    if (ctx->a32_version != ctx->a_version) {
      k_cast_a<<<red_grid(ctx, ctx->nnz), BLK, 0, ctx->stream>>>(ctx->nnz, ctx->a.p, ctx->a32.p);
      ctx->a32_version = ctx->a_version;
    }
    k_scale_to32<<<g, BLK, 0, ctx->stream>>>(n, ctx->res.p, so, ctx->r32.p, ctx->e32.p);
    k_zero_f32<<<g, BLK, 0, ctx->stream>>>(n, ctx->p32.p, solver == 2 ? ctx->u32.p : ctx->q32.p);
    k_init32<<<g, BLK, 0, ctx->stream>>>(n, ctx->r32.p, solver == 2 ? ctx->reso32.p : nullptr,
                                         ctx->partials.p, redctr(ctx), si);
    // Reduce the correction residual only as far as this refinement step
    // needs (with a 0.5 margin), never below inner_tol.
    double itol = 0.5 * tol_rel / rel;
    if (itol < inner_tol) itol = inner_tol;
    if (itol > 0.5) itol = 0.5;
    if (solver == 2) {
      auto it = [&]() {
        return bicgstab_iteration<float>(ctx, ctx->a32.p, ctx->d32.p, ctx->e32.p, ctx->r32.p,
                                         ctx->reso32.p, ctx->p32.p, ctx->z32.p, ctx->y32.p,
                                         ctx->u32.p, ctx->v32.p, si, 1, 0.0, itol, budget);
      };
      FCP_TRY(iterate_until_done(ctx, si, 1, budget, check_every, verbose, "BiCGStab32",
                                 BATCH_L2, itol, it));
    } else {
      auto it = [&]() {
        return pcg_iteration<float>(ctx, solver == 1 ? 1 : 0, ctx->a32.p, ctx->d32.p, ctx->e32.p,
                                    ctx->r32.p, ctx->p32.p, ctx->z32.p, ctx->y32.p, ctx->q32.p, si,
                                    1, 0.0, itol, budget);
      };
      FCP_TRY(iterate_until_done(ctx, si, 1, budget, check_every, verbose, "PCG32",
                                 BATCH_L2, itol, it));
    }
    total_inner += ctx->h_scal[1].iter;
    k_ir_update<<<g, BLK, 0, ctx->stream>>>(n, ctx->e32.p, so, x);
    FCP_TRY(true_residual(ctx, x, b, so));
  }
  if (opts) {
    opts->iterations = total_inner;
    opts->outer_iterations = outer;
  }
  ctx->h_scal[2].iter = total_inner;
  return 0;
}

}  // namespace

namespace fcpk {
// FP64 pivots of the DIC (kind 0) / DILU (kind 1) preconditioner of the
// resident matrix (unless pivots is false), plus the FP32 copy of A and of
// the current pivots when fp32 is set.
int prepare_precond(fcp_ctx *ctx, int kind, bool fp32, bool pivots) {
  FCP_TRY(fcp_ensure_workspace(ctx));
  if (pivots && (ctx->dic_version != ctx->a_version || ctx->dic_kind != kind)) {
    FCP_TRY(precond_setup<double>(ctx, kind, ctx->a.p, ctx->d.p));
    ctx->dic_version = ctx->a_version;
    ctx->dic_kind = kind;
  }
  if (fp32) {
    FCP_HIP(ctx, ctx->a32.alloc(ctx->nnz));
    if (ctx->a32_version != ctx->a_version) {
      k_cast_a<<<red_grid(ctx, ctx->nnz), BLK, 0, ctx->stream>>>(ctx->nnz, ctx->a.p, ctx->a32.p);
      ctx->a32_version = ctx->a_version;
    }
    if (pivots && ctx->d32_version != ctx->dic_version) {
      k_cast_a<<<red_grid(ctx, ctx->n), BLK, 0, ctx->stream>>>(ctx->n, ctx->d.p, ctx->d32.p);
      ctx->d32_version = ctx->dic_version;
    }
  }
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

// Packed coefficients and pivots of a block schedule in FP64 / FP32.
int pack_block64(fcp_ctx *ctx, BlockSchedule &b) { return pack_block<double>(ctx, b); }
int pack_block32(fcp_ctx *ctx, BlockSchedule &b) { return pack_block<float>(ctx, b); }
int bj_begin(fcp_ctx *ctx, int kind, bool *ok) { return bj_prepare(ctx, kind, ok); }
int bj_pack32(fcp_ctx *ctx) { return bj_pack32_now(ctx); }

// z = M^{-1} r with the single-vector sweeps (current pivots).
int precond_apply64(fcp_ctx *ctx, const double *r, double *y, double *z) {
  return precond_apply<double>(ctx, ctx->a.p, ctx->d.p, r, y, z, nullptr);
}
int precond_apply32(fcp_ctx *ctx, const float *r, float *y, float *z) {
  return precond_apply<float>(ctx, ctx->a32.p, ctx->d32.p, r, y, z, nullptr);
}
}  // namespace fcpk

// Tile-local DIC (kind 0) / DILU (kind 1) / DIC with the coarse correction
// (kind 2) applied in FP32 (FP64 with fp64 set) by the tile-order kernels of
// the inner solves (input rounded to FP32, output widened); d_out
// (optional) receives the FP64 pivots.
int fcp_bj32_device(fcp_ctx *ctx, int kind, const double *r, double *z, double *d_out,
                    bool fp64) {
  FCP_TRY(fcp_ensure_workspace(ctx));
  FCP_HIP(ctx, hipMemsetAsync(ctx->errflag.p, 0, sizeof(int), ctx->stream));
  FCP_HIP(ctx, hipMemsetAsync(ctx->counters.p, 0, CTR_INTS * sizeof(int), ctx->stream));
  bool ok = false;
  const bool coarse = kind == 2;  // tile-local DIC plus the coarse correction
  FCP_TRY(bj_prepare(ctx, coarse ? 0 : kind, &ok));
  const int rc = ok ? tin_apply(ctx, r, z, coarse, fp64) : FCP_ERR_UNSUPPORTED;
  if (rc == FCP_ERR_UNSUPPORTED)
    return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "tile-order preconditioner does not fit this mesh");
  FCP_TRY(rc);
  if (d_out)
    FCP_HIP(ctx, hipMemcpyAsync(d_out, ctx->dbj.p, sizeof(double) * ctx->n,
                                hipMemcpyDeviceToDevice, ctx->stream));
  FCP_TRY(check_err(ctx));
  return 0;
}

// Tile-local DIC (kind 0) / DILU (kind 1) applied in FP64, for checking the
// FP32 inner preconditioner; d_out (optional) receives its pivots.
int fcp_bj_device(fcp_ctx *ctx, int kind, const double *r, double *z, double *d_out) {
  FCP_TRY(fcp_ensure_workspace(ctx));
  FCP_HIP(ctx, hipMemsetAsync(ctx->errflag.p, 0, sizeof(int), ctx->stream));
  bool ok = false;
  FCP_TRY(bj_prepare(ctx, kind, &ok));
  if (!ok || (long long)ctx->n * 8 >= FCP_OOB ||
      !sched_fits(ctx, ctx->jlower, ctx->jupper, sizeof(double) + sizeof(int)))
    return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "tile-local preconditioner does not fit this mesh");
  for (BlockSchedule *b : {&ctx->jlower, &ctx->jupper})
    FCP_TRY(pack_block<double>(ctx, *b, ctx->dbj.p, ctx->bj_epoch | (1ull << 63)));
  k_reset_tickets<<<1, 1, 0, ctx->stream>>>(tickets(ctx));
  FCP_TRY((launch_psweep<double, 0>(ctx, ctx->jlower, r, ctx->yk.p, tickets(ctx), nullptr)));
  FCP_TRY((launch_psweep<double, 1>(ctx, ctx->jupper, ctx->yk.p, z, tickets(ctx) + 1, nullptr)));
  if (d_out)
    FCP_HIP(ctx, hipMemcpyAsync(d_out, ctx->dbj.p, sizeof(double) * ctx->n,
                                hipMemcpyDeviceToDevice, ctx->stream));
  FCP_TRY(check_err(ctx));
  return 0;
}

int fcp_precond_device(fcp_ctx *ctx, int kind, const double *r, double *z, double *d_out) {
  FCP_TRY(fcp_ensure_workspace(ctx));
  FCP_HIP(ctx, hipMemsetAsync(ctx->errflag.p, 0, sizeof(int), ctx->stream));
  if (ctx->dic_version != ctx->a_version || ctx->dic_kind != kind) {
    FCP_TRY(precond_setup<double>(ctx, kind, ctx->a.p, ctx->d.p));
    ctx->dic_version = ctx->a_version;
    ctx->dic_kind = kind;
  }
  FCP_TRY(precond_apply<double>(ctx, ctx->a.p, ctx->d.p, r, ctx->yk.p, z, nullptr));
  if (d_out)
    FCP_HIP(ctx, hipMemcpyAsync(d_out, ctx->d.p, sizeof(double) * ctx->n, hipMemcpyDeviceToDevice,
                                ctx->stream));
  FCP_TRY(check_err(ctx));
  return 0;
}

// solver: 0 iccg, 1 dpcg, 2 bicgstab, 3 gauss-seidel, 4 pmgmres (restart
// ctx->gm_restart).  x, b device pointers of length n.
int fcp_solve_device(fcp_ctx *ctx, int solver, double *x, const double *b, int itr_max,
                     double tol_abs, double tol_rel, int verbose, fcp_solve_opts *opts) {
  FCP_TRY(fcp_ensure_workspace(ctx));
  const int prec = opts ? opts->precision : FCP_PREC_FP64;
  if (prec != FCP_PREC_FP64 && prec != FCP_PREC_MIXED && prec != FCP_PREC_FP64_TILE)
    return fcp_fail(ctx, FCP_ERR_ARG, "unknown precision %d", prec);
  // FCP_PREC_MIXED and FCP_PREC_FP64_TILE run the refinement driver
  const bool refine = prec != FCP_PREC_FP64;
  if (solver == 3 && refine)
    return fcp_fail(ctx, FCP_ERR_UNSUPPORTED,
                    "gauss-seidel runs in FP64 only (no refinement variant)");
  const int crit = opts ? opts->criterion : FCP_CRIT_REF_L1;
  const int ce = (opts && opts->check_every > 0) ? opts->check_every : 8;
  hipEvent_t e0, e1;
  FCP_HIP(ctx, hipEventCreate(&e0));
  FCP_HIP(ctx, hipEventCreate(&e1));
  FCP_HIP(ctx, hipEventRecord(e0, ctx->stream));
  int rc;
  if (solver == 4) {
    // GMRES fills opts itself; the final true residual is added here.
    rc = fcp_gmres_device(ctx, x, b, itr_max, ctx->gm_restart, tol_abs, tol_rel, verbose, opts);
    if (rc >= 0) {
      FCP_HIP(ctx, hipEventRecord(e1, ctx->stream));
      SolveScalars *sf = ctx->scal.p + 2;
      FCP_TRY(true_residual(ctx, x, b, sf));
      FCP_HIP(ctx, hipMemcpyAsync(&ctx->h_scal[2], sf, sizeof(SolveScalars),
                                  hipMemcpyDeviceToHost, ctx->stream));
      FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
      float ms = 0.f;
      FCP_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
      if (opts) {
        const SolveScalars &F = ctx->h_scal[2];
        opts->rel_res_l2 = F.l2 / (F.l2b > 0 ? F.l2b : 1.0);
        opts->time_ms = ms;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
  }
  if (refine)
    rc = solve_mixed(ctx, solver, x, b, itr_max, tol_abs, tol_rel, ce, verbose, opts);
  else
    rc = solve_fp64(ctx, solver, x, b, itr_max, tol_abs, tol_rel, crit, ce, verbose);
  if (rc < 0) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
  }
  FCP_HIP(ctx, hipEventRecord(e1, ctx->stream));
  // final true residual (L2) for the report
  SolveScalars *sf = ctx->scal.p + 2;
  FCP_TRY(true_residual(ctx, x, b, sf));
  FCP_HIP(ctx, hipMemcpyAsync(&ctx->h_scal[0], ctx->scal.p, 3 * sizeof(SolveScalars),
                              hipMemcpyDeviceToHost, ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0.f;
  FCP_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const SolveScalars &S = ctx->h_scal[0];
  const SolveScalars &F = ctx->h_scal[2];
  int iters = refine ? (opts ? opts->iterations : 0) : S.iter;
  if (opts) {
    if (!refine) {
      opts->iterations = S.iter;
      opts->outer_iterations = 1;
      opts->res0 = S.res0;
      opts->resl = S.resl;
      opts->resor = S.resor;
    } else {
      opts->resl = F.res0;  // L1 of the final true residual
      opts->resor = opts->res0 / (S.factor > 0 ? S.factor : 1.0);
    }
    opts->rel_res_l2 = F.l2 / (F.l2b > 0 ? F.l2b : 1.0);
    opts->time_ms = ms;
  }
  if (!refine && S.status < 0) return fcp_fail(ctx, S.status, "non-finite residual");
  if (!refine && !S.conv && S.res0 >= tol_abs) return iters > 0 ? iters : 1;
  if (refine && F.l2 > tol_rel * F.l2b && crit == FCP_CRIT_TRUE_L2)
    return iters > 0 ? iters : 1;
  return 0;
}
