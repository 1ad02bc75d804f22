// FP32 inner PCG of the mixed-precision solves, run in the processing order
// of the tile-local preconditioner ("tile order").
//
// Position p of the tile order is the p-th row of the tile-local forward
// schedule: tiles one after another, rows of a tile by level of the strict
// lower DAG (equal levels split into steps of at most BLK rows).  Every
// neighbour pair keeps its relative order (a row's lower neighbours have
// smaller levels), so the strict lower / upper triangles of P A P^T are the
// permuted strict lower / upper triangles of A and the tile-local DIC / DILU
// is the same preconditioner, with the same per-row arithmetic, as the one
// the row-gather sweeps apply.  The backward sweep processes a tile's steps
// in reverse (a row's upper neighbours in the tile have larger levels).
//
// In tile order a tile's rows, their sweep operands and their part of every
// vector are contiguous, and a tile's preconditioner application reads only
// the tile's own residual.  One PCG iteration is two kernels:
//   tk_pre  (one workgroup per tile): x += alf p, r -= alf q with the norms
//           and stopping test of the previous iteration (iccg :506-512),
//           then z = M^-1 r with both sweeps in LDS and r.z (:458-478);
//   tk_spmv (grid-stride rows): p = z + bet p (:488) computed on the fly
//           for every operand, q = A p and p.q (:493-501).
// With the two-level preconditioner (tile-local DIC plus a coarse correction
// with one unknown per tile, no counterpart in the reference) tk_pre also
// writes each tile's residual sum and tc_zc, between the two, adds
// (A0^-1 r0)[tile] to z and completes r.z and bet.
// NR right-hand sides are interleaved (NRP values per position); every
// column's arithmetic is independent of NR, so a batched solve's columns
// equal single-column solves bit for bit.
// The PCG kernels take the value type T of the operands, vectors and sweeps:
// float for the mixed-precision inner solves, double for FP64 iterations with
// the same two-level preconditioner (FCP_PREC_FP64_TILE).
#include <algorithm>
#include <type_traits>

#include "fcp_internal.h"
#include "fcp_device.h"

using namespace fcpk;

namespace {

constexpr int TK_SLOTS = 0xFFFF;  // LDS slots a 16-bit dependency can name
#ifndef FCP_TPD
#define FCP_TPD 4
#endif

constexpr int TPD = FCP_TPD;  // sweep steps of operand prefetch
#ifndef FCP_TK_NT
#define FCP_TK_NT 256
#endif
constexpr int TK_NT = FCP_TK_NT;  // threads of a tk_pre workgroup (BLK / TK_NT positions each)

// Four values of type T in one load (the sweep coefficients of a position).
template <class T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
  __device__ static float4 make(float a, float b, float c, float d) {
    return make_float4(a, b, c, d);
  }
};
template <>
struct Vec4<double> {
  using type = double4;
  __device__ static double4 make(double a, double b, double c, double d) {
    return make_double4(a, b, c, d);
  }
};
template <class T>
using V4 = typename Vec4<T>::type;

template <int NR>
__device__ __forceinline__ bool tk_all_done(const SolveScalars *s) {
#pragma unroll
  for (int c = 0; c < NR; ++c)
    if (!solve_done(s + c)) return false;
  return true;
}

#ifndef FCP_XCD_ROWS
#define FCP_XCD_ROWS 1
#endif
// Rows of a grid-stride loop, XCD-chunked.  Workgroups are dispatched to the
// eight XCDs round-robin (workgroup b on XCD b % 8), each XCD with its own L2.
// With a grid that is a multiple of 8, XCD x takes the x-th eighth of the
// rows (in 256-row pieces), so the vector lines that neighbouring rows
// gather stay in one L2 instead of being fetched by several.
struct XRows {
  int i, end, step;
};
__device__ __forceinline__ XRows xcd_rows(int n) {
  const int G = (int)gridDim.x, b = (int)blockIdx.x, t = (int)threadIdx.x;
  if (!FCP_XCD_ROWS || (G & 7) != 0) return {b * BLK + t, n, G * BLK};
  const int chunk = ((n + 8 * BLK - 1) / (8 * BLK)) * BLK;
  const int lo = (b & 7) * chunk;
  const int hi = n < lo + chunk ? n : lo + chunk;
  return {lo + (b >> 3) * BLK + t, hi, (G >> 3) * BLK};
}

// ---------------------------------------------------------------------------
// pattern and values
// ---------------------------------------------------------------------------
__global__ void tk_iperm(int n, const int *__restrict__ perm, int *__restrict__ iperm) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) iperm[perm[p]] = p;
}

__device__ __forceinline__ int tk_block_of(const int *__restrict__ bpos, int nb, int p) {
  int lo = 0, hi = nb - 1;  // largest b with bpos[b] <= p
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bpos[mid] <= p)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int2 tk_pack16(const int (&d)[FCP_PMAXD]) {
  return make_int2((int)((unsigned)(d[0] & 0xFFFF) | ((unsigned)(d[1] & 0xFFFF) << 16)),
                   (int)((unsigned)(d[2] & 0xFFFF) | ((unsigned)(d[3] & 0xFFFF) << 16)));
}

// Dependencies in LDS slots (16 bits) of both sweeps: the forward ones from
// the tile-local forward schedule, the backward ones are the row's upper
// entries (CSR order) whose column lies in the same tile.  Unused slots name
// the zero-valued slot `dummy` (their coefficient is 0).
__global__ void tk_pattern(int n, int nb, int dummy, const int *__restrict__ perm,
                           const int *__restrict__ iperm, const int *__restrict__ bpos,
                           const int *__restrict__ prc, const int *__restrict__ pdep,
                           const int *__restrict__ ia, const int *__restrict__ diag,
                           const int *__restrict__ ja, int2 *__restrict__ dl,
                           int2 *__restrict__ du, int *__restrict__ err) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int cnt = prc[2 * p + 1] & 0xFFFF;
  int d[FCP_PMAXD];
#pragma unroll
  for (int m = 0; m < FCP_PMAXD; ++m) d[m] = m < cnt ? pdep[(size_t)FCP_PMAXD * p + m] : dummy;
  dl[p] = tk_pack16(d);
  const int b = tk_block_of(bpos, nb, p);
  const int lo = bpos[b], hi = bpos[b + 1];
  const int row = perm[p];
  int m = 0;
#pragma unroll
  for (int k = 0; k < FCP_PMAXD; ++k) d[k] = dummy;
  for (int k = diag[row] + 1; k < ia[row + 1]; ++k) {
    const int q = iperm[ja[k]];
    if (q >= lo && q < hi) {
      if (m < FCP_PMAXD) d[m] = q - lo;
      ++m;
    }
  }
  if (m > FCP_PMAXD) atomicOr(err, 1);
  du[p] = tk_pack16(d);
}

// Coefficients (FP32 roundings of the FP64 matrix) of both sweeps and the
// FP32 rounding of the tile-local pivots.
// CSR entry of every in-tile upper dependency slot of every position (-1:
// unused), in the order the backward sweep takes them (once per pattern).
__global__ void tk_upper_src(int n, int nb, const int *__restrict__ perm,
                             const int *__restrict__ iperm, const int *__restrict__ bpos,
                             const int *__restrict__ ia, const int *__restrict__ diag,
                             const int *__restrict__ ja, int *__restrict__ usrc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int b = tk_block_of(bpos, nb, p);
  const int lo = bpos[b], hi = bpos[b + 1];
  const int row = perm[p];
  int m = 0;
  for (int k = diag[row] + 1; k < ia[row + 1] && m < FCP_PMAXD; ++k) {
    const int q = iperm[ja[k]];
    if (q >= lo && q < hi) usrc[(size_t)FCP_PMAXD * p + m++] = k;
  }
  for (; m < FCP_PMAXD; ++m) usrc[(size_t)FCP_PMAXD * p + m] = -1;
}

// Distances of both sweeps' dependencies for the packed layout, 3 x 10 bits
// (0: the zero slot `dummy`); err set where a position has a fourth
// dependency or a distance beyond 1023 (the layout then does not apply).
__global__ void tk_pack_offsets(int n, int nb, int dummy, const int *__restrict__ bpos,
                                const int2 *__restrict__ dl, const int2 *__restrict__ du,
                                unsigned *__restrict__ offl, unsigned *__restrict__ offu,
                                int *__restrict__ err) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i = p - bpos[tk_block_of(bpos, nb, p)];
  bool bad = false;
#pragma unroll
  for (int up = 0; up < 2; ++up) {
    const int2 d = up ? du[p] : dl[p];
    const int sl[FCP_PMAXD] = {(int)((unsigned)d.x & 0xFFFF), (int)((unsigned)d.x >> 16),
                               (int)((unsigned)d.y & 0xFFFF), (int)((unsigned)d.y >> 16)};
    unsigned w = 0;
    if (sl[3] != dummy) bad = true;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      if (sl[m] == dummy) continue;
      const int off = up ? sl[m] - i : i - sl[m];
      if (off < 1 || off > 1023)
        bad = true;
      else
        w |= (unsigned)off << (10 * m);
    }
    (up ? offu : offl)[p] = w;
  }
  if (bad) atomicOr(err, 1);
}

template <class T>
__device__ __forceinline__ T tk_from_bits(unsigned w) {
  if constexpr (sizeof(T) == 4)
    return __uint_as_float(w);
  else
    return __longlong_as_double((long long)w);
}

// Sweep operands of the current matrix in tile order, in precision T: lower
// and upper in-tile coefficients (psrc / usrc name their CSR entries) and the
// pivots; with offl / offu (packed layout) the fourth component carries the
// dependency distances instead of a coefficient.
template <class T>
__global__ void tk_sweep_vals(int n, const int *__restrict__ perm, const int *__restrict__ psrc,
                              const int *__restrict__ usrc, const double *__restrict__ a,
                              const double *__restrict__ dbj, const unsigned *__restrict__ offl,
                              const unsigned *__restrict__ offu, V4<T> *__restrict__ al,
                              V4<T> *__restrict__ au, T *__restrict__ dv) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int4 ls = reinterpret_cast<const int4 *>(psrc)[p];
  const int4 us = reinterpret_cast<const int4 *>(usrc)[p];
  const int row = perm[p];
  auto v = [&](int k) { return k >= 0 ? (T)a[k] : (T)0; };
  al[p] = Vec4<T>::make(v(ls.x), v(ls.y), v(ls.z), offl ? tk_from_bits<T>(offl[p]) : v(ls.w));
  au[p] = Vec4<T>::make(v(us.x), v(us.y), v(us.z), offu ? tk_from_bits<T>(offu[p]) : v(us.w));
  dv[p] = (T)dbj[row];
}
// SELL-64 copy of P A P^T: slice widths, columns, values.  A row keeps its
// CSR entry order; padding entries have value 0 and the row's own column.
__global__ void tk_sell_len(int n, int nslices, const int *__restrict__ perm,
                            const int *__restrict__ ia, int *__restrict__ len) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > nslices) return;
  if (s == nslices) {
    len[s] = 0;
    return;
  }
  int w = 0;
  for (int p = s * 64; p < s * 64 + 64 && p < n; ++p) {
    const int row = perm[p];
    const int l = ia[row + 1] - ia[row];
    w = l > w ? l : w;
  }
  len[s] = 64 * w;
}

__global__ void tk_sell_col(int n, int nslices, const int *__restrict__ perm,
                            const int *__restrict__ iperm, const int *__restrict__ ia,
                            const int *__restrict__ ja, const int *__restrict__ base,
                            int *__restrict__ col) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nslices * 64) return;
  const int s = p >> 6, q = base[s] + (p & 63), L = (base[s + 1] - base[s]) >> 6;
  const int row = p < n ? perm[p] : -1;
  const int len = row >= 0 ? ia[row + 1] - ia[row] : 0;
  for (int k = 0; k < L; ++k)
    col[q + 64 * k] = k < len ? iperm[ja[ia[row] + k]] : (p < n ? p : 0);
}

// CSR start and length of the row at every tile-order position (once per
// pattern).
__global__ void tk_row_src(int n, const int *__restrict__ perm, const int *__restrict__ ia,
                           int2 *__restrict__ rs) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int row = perm[p];
  rs[p] = make_int2(ia[row], ia[row + 1] - ia[row]);
}
// Values of the SELL-64 copy in precision T, one thread per position: the
// row's CSR entries (contiguous, so each line of a is fetched once) are all
// loaded before the coalesced stores of the slice's columns.
template <class T>
__global__ void tk_sell_val(int n, int nslices, const int2 *__restrict__ rs,
                            const double *__restrict__ a, const int *__restrict__ base,
                            T *__restrict__ val) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nslices * 64) return;
  const int s = p >> 6, q = base[s] + (p & 63), L = (base[s + 1] - base[s]) >> 6;
  const int2 r = p < n ? rs[p] : make_int2(0, 0);
  T v[SELL_KU];
#pragma unroll
  for (int k = 0; k < SELL_KU; ++k) v[k] = k < r.y ? (T)a[r.x + k] : (T)0;
#pragma unroll
  for (int k = 0; k < SELL_KU; ++k)
    if (k < L) val[q + 64 * k] = v[k];
  for (int k = SELL_KU; k < L; ++k) val[q + 64 * k] = k < r.y ? (T)a[r.x + k] : (T)0;
}

// ---------------------------------------------------------------------------
// sweeps in LDS
// ---------------------------------------------------------------------------
// Operands of the PP positions (PP = BLK / NT) a thread of an NT-thread
// workgroup handles in one step: position sb.x + t + q NT for q < PP.
template <int PP, class T>
struct TSlot {
  int ok[PP], i[PP];  // position valid; its LDS slot
  int2 d[PP];         // 4 x 16-bit dependency slots
  V4<T> a[PP];
  T dv[PP];
};

// Operands for this thread of the step with position bounds sb (valid: the
// step exists).  Lanes without a position load the tile's first one: a
// branch around the loads would make the compiler drain every outstanding
// prefetch where the paths join.
template <int PP, int NT, bool DV = true, bool PACK = false, class T>
__device__ __forceinline__ void tk_load(TSlot<PP, T> &S, int2 sb, bool valid, int lo,
                                        const int2 *__restrict__ dd,
                                        const V4<T> *__restrict__ aa,
                                        const T *__restrict__ dv, int t = (int)threadIdx.x) {
#pragma unroll
  for (int q = 0; q < PP; ++q) {
    const int i = sb.x + t + q * NT;
    S.ok[q] = (valid && i < sb.y) ? 1 : 0;
    const int ii = S.ok[q] ? i : lo;
    S.i[q] = ii - lo;
    if (!PACK) S.d[q] = dd[ii];  // packed: the distances ride in the coefficients' w
    S.a[q] = aa[ii];
    if (DV) S.dv[q] = dv[ii];  // the backward sweep takes the pivot from LDS
  }
}

// The positions of a step, in place: forward (MODE 0)
// v_i = (v_i - sum_m a_m v_dep(m)) * d_i; backward (MODE 1)
// v_i = (v_i / (d_i + small) - sum_m a_m v_dep(m)) * d_i; dependencies in
// CSR order.  An unused slot subtracts 0 * 0 from the sum, which leaves every
// value (signed zeros, infinities and NaN included) unchanged, so the result
// is the one of the sum over the used slots only.  A step's positions are
// independent, so a thread's PP chains overlap.
// Packed layout (PACK, every position with at most 3 in-tile dependencies per
// sweep): coefficients in x, y, z and, in the bits of w, three 10-bit
// distances from the position to its dependency (0: unused, the zero slot
// dslot), which drops the separate 8-byte slot word per position and sweep.
// The fourth slot it lacks is always unused, so the results are those of the
// 4-slot layout bit for bit.
template <class T>
__device__ __forceinline__ unsigned tk_bits(T w) {
  if constexpr (sizeof(T) == 4)
    return __float_as_uint(w);
  else
    return (unsigned)__double_as_longlong(w);
}
template <int NR, int NRP, int MODE, int PP, bool PACK, class T>
__device__ __forceinline__ void tk_step(const TSlot<PP, T> &S, T *V, T *D, int dslot) {
  constexpr int ND = PACK ? 3 : FCP_PMAXD;
#pragma unroll
  for (int q = 0; q < PP; ++q) {
    if (!S.ok[q]) continue;
    int dep[FCP_PMAXD];
    if (PACK) {
      const unsigned w = tk_bits(S.a[q].w);
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const int off = (int)((w >> (10 * m)) & 1023u);
        dep[m] = off ? (MODE == 0 ? S.i[q] - off : S.i[q] + off) : dslot;
      }
      dep[3] = dslot;
    } else {
      const unsigned dx = (unsigned)S.d[q].x, dy = (unsigned)S.d[q].y;
      dep[0] = (int)(dx & 0xFFFF);
      dep[1] = (int)(dx >> 16);
      dep[2] = (int)(dy & 0xFFFF);
      dep[3] = (int)(dy >> 16);
    }
    const T av[FCP_PMAXD] = {S.a[q].x, S.a[q].y, S.a[q].z, PACK ? (T)0 : S.a[q].w};
    T dvq;
    if (MODE == 0) {
      dvq = S.dv[q];
      D[S.i[q]] = dvq;  // kept for the backward sweep
    } else {
      dvq = D[S.i[q]];
    }
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      T acc = V[S.i[q] * NRP + c];
      if (MODE == 1) acc = acc / (dvq + (T)FCP_SMALL);
#pragma unroll
      for (int m = 0; m < ND; ++m) acc = acc - av[m] * V[dep[m] * NRP + c];
      V[S.i[q] * NRP + c] = acc * dvq;
    }
  }
}

// Both sweeps of one tile over V (ns >= 1 steps; sb[k] = position bounds of
// step k, in LDS) by an NT-thread workgroup.  Operands are prefetched TPD
// steps ahead; the bounds of the step to prefetch are read one step early
// (always from a valid slot), so a step's critical path holds only its own
// LDS reads and write.
template <int NR, int NRP, int NT = BLK, bool PACK = false, class T>
__device__ __forceinline__ void tk_sweeps(T *V, T *D, const int2 *sb, int ns, int lo, int dslot,
                                          const int2 *__restrict__ dl,
                                          const V4<T> *__restrict__ al,
                                          const int2 *__restrict__ du,
                                          const V4<T> *__restrict__ au,
                                          const T *__restrict__ dv) {
  constexpr int PP = BLK / NT;
  TSlot<PP, T> S[TPD];
#pragma unroll
  for (int j = 0; j < TPD; ++j)
    tk_load<PP, NT, true, PACK>(S[j], sb[j < ns ? j : 0], j < ns, lo, dl, al, dv);
  int2 nxt = sb[TPD < ns ? TPD : 0];
  for (int base = 0; base < ns; base += TPD) {
#pragma unroll
    for (int j = 0; j < TPD; ++j) {
      const int nx = base + j + TPD;
      const int2 nn = sb[nx + 1 < ns ? nx + 1 : 0];
      tk_step<NR, NRP, 0, PP, PACK>(S[j], V, D, dslot);
      __syncthreads();
      tk_load<PP, NT, true, PACK>(S[j], nxt, nx < ns, lo, dl, al, dv);
      nxt = nn;
    }
  }
#pragma unroll
  for (int j = 0; j < TPD; ++j)
    tk_load<PP, NT, false, PACK>(S[j], sb[j < ns ? ns - 1 - j : 0], j < ns, lo, du, au, dv);
  nxt = sb[TPD < ns ? ns - 1 - TPD : 0];
  for (int base = 0; base < ns; base += TPD) {
#pragma unroll
    for (int j = 0; j < TPD; ++j) {
      const int nx = base + j + TPD;
      const int2 nn = sb[nx + 1 < ns ? ns - 2 - nx : 0];
      tk_step<NR, NRP, 1, PP, PACK>(S[j], V, D, dslot);
      __syncthreads();
      tk_load<PP, NT, false, PACK>(S[j], nxt, nx < ns, lo, du, au, dv);
      nxt = nn;
    }
  }
}

#ifndef FCP_TPDW
#define FCP_TPDW 8
#endif
constexpr int TPDW = FCP_TPDW;  // steps of operand prefetch of a one-wave tile sweep

// Both sweeps of one tile by one wave (steps of at most 64 positions): the
// wave's own LDS accesses are performed in program order, so no workgroup
// barrier separates the steps; a wave barrier only keeps the compiler from
// moving LDS accesses across a step.  Steps are short, so operands are
// prefetched TPDW steps ahead.
template <int NR, int NRP, bool PACK, class T>
__device__ __forceinline__ void tk_sweeps_w(T *V, T *D, const int2 *sb, int ns, int lo,
                                            int dslot, int lane,
                                            const int2 *__restrict__ dl,
                                            const V4<T> *__restrict__ al,
                                            const int2 *__restrict__ du,
                                            const V4<T> *__restrict__ au,
                                            const T *__restrict__ dv) {
  TSlot<1, T> S[TPDW];
#pragma unroll
  for (int j = 0; j < TPDW; ++j)
    tk_load<1, 64, true, PACK>(S[j], sb[j < ns ? j : 0], j < ns, lo, dl, al, dv, lane);
  int2 nxt = sb[TPDW < ns ? TPDW : 0];
  for (int base = 0; base < ns; base += TPDW) {
#pragma unroll
    for (int j = 0; j < TPDW; ++j) {
      const int nx = base + j + TPDW;
      const int2 nn = sb[nx + 1 < ns ? nx + 1 : 0];
      tk_step<NR, NRP, 0, 1, PACK>(S[j], V, D, dslot);
      __builtin_amdgcn_wave_barrier();
      tk_load<1, 64, true, PACK>(S[j], nxt, nx < ns, lo, dl, al, dv, lane);
      nxt = nn;
    }
  }
#pragma unroll
  for (int j = 0; j < TPDW; ++j)
    tk_load<1, 64, false, PACK>(S[j], sb[j < ns ? ns - 1 - j : 0], j < ns, lo, du, au, dv,
                                lane);
  nxt = sb[TPDW < ns ? ns - 1 - TPDW : 0];
  for (int base = 0; base < ns; base += TPDW) {
#pragma unroll
    for (int j = 0; j < TPDW; ++j) {
      const int nx = base + j + TPDW;
      const int2 nn = sb[nx + 1 < ns ? ns - 2 - nx : 0];
      tk_step<NR, NRP, 1, 1, PACK>(S[j], V, D, dslot);
      __builtin_amdgcn_wave_barrier();
      tk_load<1, 64, false, PACK>(S[j], nxt, nx < ns, lo, du, au, dv, lane);
      nxt = nn;
    }
  }
}

// ---------------------------------------------------------------------------
// iteration kernels
// ---------------------------------------------------------------------------
// Operands of the coarse correction when tk_pre applies it itself.
struct CoarseFuse {
  const double *G;       // ncp x ncp rows of A0^-1 (row T: tile T's coarse row)
  const int *failed;     // the coarse factor failed: no correction
  unsigned *bar;         // grid barrier: arrivals at [0], generation at [32]
  int *err;              // FCP_ERR_SWEEP when the barrier wait exceeds its bound
};

// Every workgroup of the grid arrives before any leaves.  The caller's
// agent-scope stores are complete before its arrival is counted (as in
// last_block_reduce); the wait is bounded by wall time like the sweep waits
// (FCP_ERR_SWEEP recorded and the wait abandoned), so the kernel always
// drains.  The generation is read before arriving, so a wave cannot miss the
// release of the barrier it waits for.
__device__ __forceinline__ void grid_barrier(unsigned *bar, int *err) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned gen = __hip_atomic_load(bar + 32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);  // generation read, this block's stores complete
    if (__hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
        gridDim.x - 1) {
      __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __builtin_amdgcn_s_waitcnt(0);
      __hip_atomic_fetch_add(bar + 32, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const unsigned long long t0 = wall_clock64();
      unsigned spins = 0;
      while (__hip_atomic_load(bar + 32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __builtin_amdgcn_s_sleep(2);
        if ((++spins & 63u) == 0 && wall_clock64() - t0 > FCP_WAIT_TICKS) {
          __hip_atomic_store(err, FCP_ERR_SWEEP, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
    }
  }
  __syncthreads();
}

// One workgroup per tile (tiles nb..: loop).  upd: apply the previous
// iteration's update x += alf p, r -= alf q with |r|_1, |r|_2 and its
// stopping test (relative L2 against the column's tolerance, at most
// s->imax iterations); then
// z = M^-1 r and sk = r.z, bet = sk / s0.  Columns already done are left
// untouched.
// CF (coarse fused; one workgroup per tile, every workgroup resident): after
// the sweeps every tile's residual sum is published, a grid barrier waits for
// all of them, and each workgroup adds its tile's entry of e0 = A0^-1 r0 to z
// while z is still in LDS and contributes e0[T] r0[T] to sk -- the work of
// tc_zc, with the same e0 arithmetic, without its kernel and its pass over z.
template <int NR, int NRP, int NT, bool PACK, class T, bool CF = false>
__global__ __launch_bounds__(NT) void tk_pre(
    int nb, int vslots, int dslot, const int *__restrict__ bpos, const int *__restrict__ blk_step,
    const int *__restrict__ step_ptr, const int2 *__restrict__ dl, const V4<T> *__restrict__ al,
    const int2 *__restrict__ du, const V4<T> *__restrict__ au, const T *__restrict__ dv,
    T *__restrict__ x, T *__restrict__ r, const T *__restrict__ p,
    const T *__restrict__ q, T *__restrict__ z, double *__restrict__ partials,
    int *__restrict__ counter, SolveScalars *__restrict__ s, int upd, double tol_abs,
    double *__restrict__ r0, int ncp, CoarseFuse cf = {}) {
  // [vslots] values (slot dslot = 0), the tile's pivots, then the step bounds
  extern __shared__ __align__(16) unsigned char tk_smem[];
  T *V = reinterpret_cast<T *>(tk_smem);
  T *D = V + vslots;
  int2 *sb = reinterpret_cast<int2 *>(D + ((dslot + 1) & ~1));
  if (tk_all_done<NR>(s)) return;
  bool act[NR];
  T alf[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    act[c] = !solve_done(s + c);
    alf[c] = (T)s[c].alf;
  }
  // |r|_1, |r|_2, r.z per column, then (CF) e0[T] r0[T]
  constexpr int NV = CF ? 4 * NR : 3 * NR;
  double part[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) part[k] = 0.0;
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const int lo = bpos[b], hi = bpos[b + 1];
    const int st0 = blk_step[b], ns = blk_step[b + 1] - st0;
    for (int k = threadIdx.x; k < ns; k += NT)
      sb[k] = make_int2(step_ptr[st0 + k], step_ptr[st0 + k + 1]);
    double rs[NR];  // the tile's residual sum (coarse correction)
#pragma unroll
    for (int c = 0; c < NR; ++c) rs[c] = 0.0;
    for (int i0 = lo + (int)threadIdx.x; i0 < hi; i0 += GS_U * NT) {
      T xv[GS_U][NR], pv[GS_U][NR], rv[GS_U][NR], qv[GS_U][NR];
#pragma unroll
      for (int u = 0; u < GS_U; ++u) {
        const int i = i0 + u * NT;
        if (i < hi) {
#pragma unroll
          for (int c = 0; c < NR; ++c) {
            const size_t o = (size_t)i * NRP + c;
            rv[u][c] = r[o];
            if (upd && act[c]) {
              xv[u][c] = x[o];
              pv[u][c] = p[o];
              qv[u][c] = q[o];
            }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < GS_U; ++u) {
        const int i = i0 + u * NT;
        if (i < hi) {
#pragma unroll
          for (int c = 0; c < NR; ++c) {
            const size_t o = (size_t)i * NRP + c;
            T ri = rv[u][c];
            if (upd && act[c]) {
              x[o] = xv[u][c] + alf[c] * pv[u][c];
              ri = ri - alf[c] * qv[u][c];
              r[o] = ri;
              part[3 * c] += fabs((double)ri);
              part[3 * c + 1] += (double)ri * (double)ri;
            }
            V[(i - lo) * NRP + c] = ri;
            rs[c] += (double)ri;
          }
        }
      }
    }
    if (threadIdx.x < NRP) V[dslot * NRP + threadIdx.x] = (T)0;  // the unused-dependency slot
    double ts[NR];  // the tile's residual sums (thread 0)
#pragma unroll
    for (int c = 0; c < NR; ++c) ts[c] = 0.0;
    if (r0) {  // uniform
      block_sum<NR, NT>(rs, ts);
      if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < NR; ++c) {
          if (CF)  // read by every workgroup after the grid barrier
            __hip_atomic_store(reinterpret_cast<unsigned long long *>(&r0[(size_t)c * ncp + b]),
                               (unsigned long long)__double_as_longlong(ts[c]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
          else
            r0[(size_t)c * ncp + b] = ts[c];
        }
    }
    __syncthreads();
    if (ns > 0) tk_sweeps<NR, NRP, NT, PACK>(V, D, sb, ns, lo, dslot, dl, al, du, au, dv);
    T e[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) e[c] = (T)0;
    if (CF) {
      // e0[b] = (row b of A0^-1) . r0, summed as tc_zc sums it: thread t over
      // m = t mod NT, then the waves in order
      __shared__ double SC[NR][NT / 64];
      grid_barrier(cf.bar, cf.err);
      const int t = (int)threadIdx.x, lane = t & 63, wv = t >> 6;
      const double *g = cf.G + (size_t)b * ncp;
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        double sum = 0.0;
        for (int m = t; m < ncp; m += NT)
          sum += g[m] * __hip_atomic_load(&r0[(size_t)c * ncp + m], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
        sum = wave_sum(sum);
        if (lane == 0) SC[c][wv] = sum;
      }
      __syncthreads();
      const bool off = *cf.failed != 0;
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) v += SC[c][w];
        const double E = off ? 0.0 : v;
        e[c] = (T)E;
        if (t == 0) part[3 * NR + c] = E * ts[c];
      }
    }
    // z out (plus the coarse correction); r.z with the residual this thread
    // wrote (or read) above and the tile-local z
    for (int i = lo + (int)threadIdx.x; i < hi; i += NT) {
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        const size_t o = (size_t)i * NRP + c;
        const T zi = V[(i - lo) * NRP + c];
        z[o] = (CF && act[c]) ? zi + e[c] : zi;
        part[3 * c + 2] += (double)r[o] * (double)zi;
      }
    }
    __syncthreads();  // V and sb are reused by the next tile
  }
  double blk[NV], tot[NV];
  block_sum<NV, NT>(part, blk);
  if (last_block_reduce<NV, NT>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      if (!act[c]) continue;
      SolveScalars &S = s[c];
      bool fin = false;
      if (upd) {
        const double resl = tot[3 * c];
        S.resl = resl;
        S.l2 = sqrt(tot[3 * c + 1]);
        S.s0 = S.sk;
        S.iter += 1;
        S.rsm = resl / (S.res0 + FCP_SMALL);
        bool conv = (S.l2 <= S.tol * S.l2b) || (resl < tol_abs);
        if (!isfinite(resl)) {
          S.status = FCP_ERR_NONFINITE;
          conv = true;
        }
        S.conv = (conv && S.status == 0) ? 1 : 0;
        if (conv || S.iter >= S.imax) {
          S.done = 1;
          fin = true;
        }
      }
      if (!fin) {
        S.sk = tot[3 * c + 2];
        if (CF) S.sk = S.sk + tot[3 * NR + c];  // the coarse part, as tc_zc adds it
        S.bet = S.sk / S.s0;
      }
    }
  }
}

// tk_pre with one wave per tile: four tiles per workgroup, no workgroup
// barrier inside a tile (tiles of one schedule built with steps of at most
// 64 positions).
template <int NR, int NRP, bool PACK, class T>
__global__ __launch_bounds__(BLK) void tk_pre_w(
    int nb, int vstride, int vslots, int dslot, const int *__restrict__ bpos, const int *__restrict__ blk_step,
    const int *__restrict__ step_ptr, const int2 *__restrict__ dl, const V4<T> *__restrict__ al,
    const int2 *__restrict__ du, const V4<T> *__restrict__ au, const T *__restrict__ dv,
    T *__restrict__ x, T *__restrict__ r, const T *__restrict__ p,
    const T *__restrict__ q, T *__restrict__ z, double *__restrict__ partials,
    int *__restrict__ counter, SolveScalars *__restrict__ s, int upd, double tol_abs,
    double *__restrict__ r0, int ncp) {
  // one region of vstride values per wave: [vslots] values (slot dslot = 0),
  // the tile's pivots, then the step bounds
  extern __shared__ __align__(16) unsigned char tk_smem[];
  const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
  T *V = reinterpret_cast<T *>(tk_smem) + (size_t)wv * vstride;
  T *D = V + vslots;
  int2 *sb = reinterpret_cast<int2 *>(D + ((dslot + 1) & ~1));
  if (tk_all_done<NR>(s)) return;
  bool act[NR];
  T alf[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    act[c] = !solve_done(s + c);
    alf[c] = (T)s[c].alf;
  }
  double part[3 * NR];
#pragma unroll
  for (int k = 0; k < 3 * NR; ++k) part[k] = 0.0;
  for (int b = blockIdx.x * (BLK / 64) + wv; b < nb; b += gridDim.x * (BLK / 64)) {
    const int lo = bpos[b], hi = bpos[b + 1];
    const int st0 = blk_step[b], ns = blk_step[b + 1] - st0;
    for (int k = lane; k < ns; k += 64)
      sb[k] = make_int2(step_ptr[st0 + k], step_ptr[st0 + k + 1]);
    double rs[NR];  // the tile's residual sum (coarse correction)
#pragma unroll
    for (int c = 0; c < NR; ++c) rs[c] = 0.0;
    for (int i0 = lo + lane; i0 < hi; i0 += GS_U * 64) {
      T xv[GS_U][NR], pv[GS_U][NR], rv[GS_U][NR], qv[GS_U][NR];
#pragma unroll
      for (int u = 0; u < GS_U; ++u) {
        const int i = i0 + u * 64;
        if (i < hi) {
#pragma unroll
          for (int c = 0; c < NR; ++c) {
            const size_t o = (size_t)i * NRP + c;
            rv[u][c] = r[o];
            if (upd && act[c]) {
              xv[u][c] = x[o];
              pv[u][c] = p[o];
              qv[u][c] = q[o];
            }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < GS_U; ++u) {
        const int i = i0 + u * 64;
        if (i < hi) {
#pragma unroll
          for (int c = 0; c < NR; ++c) {
            const size_t o = (size_t)i * NRP + c;
            T ri = rv[u][c];
            if (upd && act[c]) {
              x[o] = xv[u][c] + alf[c] * pv[u][c];
              ri = ri - alf[c] * qv[u][c];
              r[o] = ri;
              part[3 * c] += fabs((double)ri);
              part[3 * c + 1] += (double)ri * (double)ri;
            }
            V[(i - lo) * NRP + c] = ri;
            rs[c] += (double)ri;
          }
        }
      }
    }
    if (lane < NRP) V[dslot * NRP + lane] = (T)0;  // the unused-dependency slot
    if (r0) {  // uniform
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        const double ts = wave_sum(rs[c]);
        if (lane == 0) r0[(size_t)c * ncp + b] = ts;
      }
    }
    __builtin_amdgcn_wave_barrier();
    if (ns > 0) tk_sweeps_w<NR, NRP, PACK>(V, D, sb, ns, lo, dslot, lane, dl, al, du, au, dv);
    __builtin_amdgcn_wave_barrier();
    // z out; r.z with the residual this thread wrote (or read) above
    for (int i = lo + lane; i < hi; i += 64) {
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        const size_t o = (size_t)i * NRP + c;
        const T zi = V[(i - lo) * NRP + c];
        z[o] = zi;
        part[3 * c + 2] += (double)r[o] * (double)zi;
      }
    }
    __builtin_amdgcn_wave_barrier();  // V and sb are reused by the wave's next tile
  }
  double blk[3 * NR], tot[3 * NR];
  block_sum<3 * NR>(part, blk);
  if (last_block_reduce<3 * NR>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      if (!act[c]) continue;
      SolveScalars &S = s[c];
      bool fin = false;
      if (upd) {
        const double resl = tot[3 * c];
        S.resl = resl;
        S.l2 = sqrt(tot[3 * c + 1]);
        S.s0 = S.sk;
        S.iter += 1;
        S.rsm = resl / (S.res0 + FCP_SMALL);
        bool conv = (S.l2 <= S.tol * S.l2b) || (resl < tol_abs);
        if (!isfinite(resl)) {
          S.status = FCP_ERR_NONFINITE;
          conv = true;
        }
        S.conv = (conv && S.status == 0) ? 1 : 0;
        if (conv || S.iter >= S.imax) {
          S.done = 1;
          fin = true;
        }
      }
      if (!fin) {
        S.sk = tot[3 * c + 2];
        S.bet = S.sk / S.s0;
      }
    }
  }
}

// p_new = z + bet p_old (every operand computed where it is used), q = A p_new
// over all slice entries (padding adds 0 * p_i), pkapk = p.q, alf = sk/pkapk.
template <int NR, int NRP, class T>
__global__ __launch_bounds__(BLK) void tk_spmv(int n, const int *__restrict__ base, int W,
                                               const int *__restrict__ col,
                                               const T *__restrict__ val,
                                               const T *__restrict__ z,
                                               const T *__restrict__ pold,
                                               T *__restrict__ pnew, T *__restrict__ q,
                                               double *__restrict__ partials,
                                               int *__restrict__ counter,
                                               SolveScalars *__restrict__ s) {
  if (tk_all_done<NR>(s)) return;
  bool act[NR];
  T bet[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    act[c] = !solve_done(s + c);
    bet[c] = (T)s[c].bet;
  }
  double part[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) part[c] = 0.0;
  for (XRows X = xcd_rows(n); X.i < X.end; X.i += X.step) {
    const int i = X.i;
    const int2 at = sell_at(base, W, i);
    const int pp = at.x, L = at.y;
    T av[SELL_KU];
    int cv[SELL_KU];
#pragma unroll
    for (int k = 0; k < SELL_KU; ++k)
      if (k < L) {
        av[k] = val[pp + 64 * k];
        cv[k] = col[pp + 64 * k];
      }
    T acc[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) acc[c] = (T)0;
#pragma unroll
    for (int k = 0; k < SELL_KU; ++k)
      if (k < L) {
#pragma unroll
        for (int c = 0; c < NR; ++c) {
          const size_t o = (size_t)cv[k] * NRP + c;
          const T pj = z[o] + bet[c] * pold[o];
          acc[c] = acc[c] + av[k] * pj;
        }
      }
    for (int k = SELL_KU; k < L; ++k) {
      const T a = val[pp + 64 * k];
      const int j = col[pp + 64 * k];
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        const size_t o = (size_t)j * NRP + c;
        const T pj = z[o] + bet[c] * pold[o];
        acc[c] = acc[c] + a * pj;
      }
    }
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      if (!act[c]) continue;
      const size_t o = (size_t)i * NRP + c;
      const T pi = z[o] + bet[c] * pold[o];
      pnew[o] = pi;
      q[o] = acc[c];
      part[c] += (double)pi * (double)acc[c];
    }
  }
  double blk[NR], tot[NR];
  block_sum<NR>(part, blk);
  if (last_block_reduce<NR>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      if (!act[c]) continue;
      s[c].pkapk = tot[c];
      s[c].alf = s[c].sk / tot[c];
    }
  }
}

// r~ = P (res / |res|_2) in precision T, x~ = 0, p~ = 0 (BiCGStab: reso = r~, u = 0),
// and the inner scalars of
// every column (tolerance from the outer residual as for the row-order inner
// solves; columns whose outer residual already meets tol_rel start done).
template <int NR, int NRP, class T>
__global__ __launch_bounds__(BLK) void tk_init(int n, const int *__restrict__ perm,
                                               const double *__restrict__ res,
                                               const SolveScalars *__restrict__ so,
                                               T *__restrict__ r, T *__restrict__ x,
                                               T *__restrict__ p0,
                                               double *__restrict__ partials,
                                               int *__restrict__ counter,
                                               SolveScalars *__restrict__ s, double tol_rel,
                                               double inner_tol, int imax,
                                               T *__restrict__ reso, T *__restrict__ u0) {
  double inv[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) inv[c] = so[c].l2 > 0.0 ? 1.0 / so[c].l2 : 0.0;
  double part[2 * NR];
#pragma unroll
  for (int k = 0; k < 2 * NR; ++k) part[k] = 0.0;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    const size_t row = (size_t)perm[i];
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      const size_t o = (size_t)i * NRP + c;
      const T v = (T)(res[row * NRP + c] * inv[c]);
      r[o] = v;
      x[o] = (T)0;
      p0[o] = (T)0;
      if (reso) reso[o] = v;
      if (u0) u0[o] = (T)0;
      part[2 * c] += fabs((double)v);
      part[2 * c + 1] += (double)v * (double)v;
    }
  }
  double blk[2 * NR], tot[2 * NR];
  block_sum<2 * NR>(part, blk);
  if (last_block_reduce<2 * NR>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      SolveScalars &S = s[c];
      S.res0 = tot[2 * c];
      S.resl = tot[2 * c];
      S.l2 = sqrt(tot[2 * c + 1]);
      S.l2b = S.l2;
      S.s0 = FCP_S0INIT;
      S.alf = 1.0;
      S.beto = 1.0;
      S.gam = 1.0;
      S.iter = 0;
      S.status = 0;
      S.conv = 0;
      S.factor = 0.0;
      const double rel = so[c].l2 / (so[c].l2b > 0 ? so[c].l2b : 1.0);
      double itol = 0.5 * tol_rel / rel;
      if (itol < inner_tol) itol = inner_tol;
      if (itol > 0.5) itol = 0.5;
      S.tol = itol;
      S.done = (rel <= tol_rel || so[c].l2 == 0.0) ? 1 : 0;
      S.imax = imax;
    }
  }
}

// x += |res|_2 P^T x~ (FP64), in natural row order (x read and written
// contiguously, x~ gathered through the inverse permutation)
template <int NR, int NRP, class T>
__global__ void tk_ir_update(int n, const int *__restrict__ iperm, const T *__restrict__ e,
                             const SolveScalars *__restrict__ so, double *__restrict__ x) {
  double sc[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) sc[c] = so[c].l2;
  for (int row = blockIdx.x * BLK + threadIdx.x; row < n; row += gridDim.x * BLK) {
    const size_t i = (size_t)iperm[row];
#pragma unroll
    for (int c = 0; c < NR; ++c)
      x[(size_t)row * NRP + c] = x[(size_t)row * NRP + c] + sc[c] * (double)e[i * NRP + c];
  }
}

// ---------------------------------------------------------------------------
// BiCGStab (DILU) inner iterations in tile order (bicgstab :642-773): five
// kernels per iteration, the scalars in device memory.
//   tb_bet  (rows): x += alf z, r -= alf v with |r|_1, |r|_2 and the stopping
//           test of the previous iteration; bet = r.reso, om, beto (:647-650)
//   tb_tile<1> (tiles): p = r + om (p - alf u), then z = M^-1 p (:653-683)
//   tb_spmv<0> (rows): u = A z, ukreso = u.reso, gam = bet / ukreso (:689-700)
//   tb_tile<2> (tiles): x += gam z, r -= gam u, then z = M^-1 r (:703-732)
//   tb_spmv<1> (rows): v = A z, alf = v.r / (v.v + small) (:737-748)
// Each kernel keeps the per-row arithmetic of the row-order kernels.
// ---------------------------------------------------------------------------
template <int NR, int NRP, int MODE, bool PACK>
__global__ __launch_bounds__(BLK) void tb_tile(
    int nb, int vslots, int dslot, const int *__restrict__ bpos, const int *__restrict__ blk_step,
    const int *__restrict__ step_ptr, const int2 *__restrict__ dl, const float4 *__restrict__ al,
    const int2 *__restrict__ du, const float4 *__restrict__ au, const float *__restrict__ dv,
    float *__restrict__ x, float *__restrict__ r, float *__restrict__ p,
    const float *__restrict__ u, float *z, const SolveScalars *__restrict__ s) {
  // [vslots] values (slot dslot = 0), the tile's pivots, then the step bounds
  extern __shared__ float V[];
  float *D = V + vslots;
  int2 *sb = reinterpret_cast<int2 *>(D + ((dslot + 1) & ~1));
  if (tk_all_done<NR>(s)) return;
  bool act[NR];
  float c1[NR], c2[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    act[c] = !solve_done(s + c);
    c1[c] = MODE == 1 ? (float)s[c].om : (float)s[c].gam;
    c2[c] = (float)s[c].alf;
  }
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const int lo = bpos[b], hi = bpos[b + 1];
    const int st0 = blk_step[b], ns = blk_step[b + 1] - st0;
    for (int k = threadIdx.x; k < ns; k += BLK)
      sb[k] = make_int2(step_ptr[st0 + k], step_ptr[st0 + k + 1]);
    for (int i0 = lo + (int)threadIdx.x; i0 < hi; i0 += GS_U * BLK) {
      float av[GS_U][NR], bv[GS_U][NR], cv[GS_U][NR];
#pragma unroll
      for (int u2 = 0; u2 < GS_U; ++u2) {
        const int i = i0 + u2 * BLK;
        if (i < hi) {
#pragma unroll
          for (int c = 0; c < NR; ++c) {
            const size_t o = (size_t)i * NRP + c;
            av[u2][c] = MODE == 1 ? p[o] : r[o];
            if (act[c]) {
              bv[u2][c] = MODE == 1 ? r[o] : x[o];
              cv[u2][c] = u[o];
            }
          }
        }
      }
#pragma unroll
      for (int u2 = 0; u2 < GS_U; ++u2) {
        const int i = i0 + u2 * BLK;
        if (i < hi) {
#pragma unroll
          for (int c = 0; c < NR; ++c) {
            const size_t o = (size_t)i * NRP + c;
            float val = av[u2][c];
            if (act[c]) {
              if (MODE == 1) {  // p = r + om (p - alf u)
                val = bv[u2][c] + c1[c] * (av[u2][c] - c2[c] * cv[u2][c]);
                p[o] = val;
              } else {  // x += gam z ; r -= gam u
                x[o] = bv[u2][c] + c1[c] * z[o];
                val = av[u2][c] - c1[c] * cv[u2][c];
                r[o] = val;
              }
            }
            V[(i - lo) * NRP + c] = val;
          }
        }
      }
    }
    if (threadIdx.x < NRP) V[dslot * NRP + threadIdx.x] = 0.0f;  // the unused-dependency slot
    __syncthreads();
    if (ns > 0) tk_sweeps<NR, NRP, BLK, PACK>(V, D, sb, ns, lo, dslot, dl, al, du, au, dv);
    for (int i = lo + (int)threadIdx.x; i < hi; i += BLK) {
#pragma unroll
      for (int c = 0; c < NR; ++c) z[(size_t)i * NRP + c] = V[(i - lo) * NRP + c];
    }
    __syncthreads();  // V and sb are reused by the next tile
  }
}

// out = A z over the slice's entries (padding adds 0 * z_i); MODE 0:
// ukreso = out.w, gam = bet / ukreso (w = reso); MODE 1: vr = out.w,
// vv = out.out, alf = vr / (vv + small) (w = r).
template <int NR, int NRP, int MODE>
__global__ __launch_bounds__(BLK) void tb_spmv(int n, const int *__restrict__ base, int W,
                                               const int *__restrict__ col,
                                               const float *__restrict__ val,
                                               const float *__restrict__ z,
                                               const float *__restrict__ w, float *__restrict__ out,
                                               double *__restrict__ partials,
                                               int *__restrict__ counter,
                                               SolveScalars *__restrict__ s) {
  if (tk_all_done<NR>(s)) return;
  bool act[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) act[c] = !solve_done(s + c);
  constexpr int NV = MODE == 0 ? NR : 2 * NR;
  double part[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) part[k] = 0.0;
  for (XRows X = xcd_rows(n); X.i < X.end; X.i += X.step) {
    const int i = X.i;
    const int2 at = sell_at(base, W, i);
    const int pp = at.x, L = at.y;
    float av[SELL_KU];
    int cv[SELL_KU];
#pragma unroll
    for (int k = 0; k < SELL_KU; ++k)
      if (k < L) {
        av[k] = val[pp + 64 * k];
        cv[k] = col[pp + 64 * k];
      }
    float acc[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) acc[c] = 0.0f;
#pragma unroll
    for (int k = 0; k < SELL_KU; ++k)
      if (k < L) {
#pragma unroll
        for (int c = 0; c < NR; ++c) acc[c] = acc[c] + av[k] * z[(size_t)cv[k] * NRP + c];
      }
    for (int k = SELL_KU; k < L; ++k) {
      const float a = val[pp + 64 * k];
      const int j = col[pp + 64 * k];
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] = acc[c] + a * z[(size_t)j * NRP + c];
    }
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      if (!act[c]) continue;
      const size_t o = (size_t)i * NRP + c;
      out[o] = acc[c];
      if (MODE == 0) {
        part[c] += (double)acc[c] * (double)w[o];
      } else {
        part[2 * c] += (double)acc[c] * (double)w[o];
        part[2 * c + 1] += (double)acc[c] * (double)acc[c];
      }
    }
  }
  double blk[NV], tot[NV];
  block_sum<NV>(part, blk);
  if (last_block_reduce<NV>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      if (!act[c]) continue;
      if (MODE == 0) {
        s[c].ukreso = tot[c];
        s[c].gam = s[c].sk / tot[c];
      } else {
        s[c].vr = tot[2 * c];
        s[c].vv = tot[2 * c + 1];
        s[c].alf = tot[2 * c] / (tot[2 * c + 1] + FCP_SMALL);
      }
    }
  }
}

// upd: x += alf z, r -= alf v with |r|_1, |r|_2 and the stopping test of the
// previous iteration (relative L2 against the column's tolerance, at most
// s->imax iterations); then bet = r.reso, om = bet gam / (alf beto + small),
// beto = bet.
template <int NR, int NRP>
__global__ __launch_bounds__(BLK) void tb_bet(int n, float *__restrict__ x, float *__restrict__ r,
                                              const float *__restrict__ z,
                                              const float *__restrict__ v,
                                              const float *__restrict__ reso,
                                              double *__restrict__ partials,
                                              int *__restrict__ counter,
                                              SolveScalars *__restrict__ s, int upd,
                                              double tol_abs) {
  if (tk_all_done<NR>(s)) return;
  bool act[NR];
  float alf[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    act[c] = !solve_done(s + c);
    alf[c] = (float)s[c].alf;
  }
  double part[3 * NR];
#pragma unroll
  for (int k = 0; k < 3 * NR; ++k) part[k] = 0.0;
  const int stride = gridDim.x * BLK;
  for (int i0 = blockIdx.x * BLK + threadIdx.x; i0 < n; i0 += GS_U * stride) {
    float xv[GS_U][NR], zv[GS_U][NR], rv[GS_U][NR], vv[GS_U][NR], ov[GS_U][NR];
#pragma unroll
    for (int u = 0; u < GS_U; ++u) {
      const int i = i0 + u * stride;
      if (i < n) {
#pragma unroll
        for (int c = 0; c < NR; ++c) {
          const size_t o = (size_t)i * NRP + c;
          if (act[c]) {
            rv[u][c] = r[o];
            ov[u][c] = reso[o];
            if (upd) {
              xv[u][c] = x[o];
              zv[u][c] = z[o];
              vv[u][c] = v[o];
            }
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < GS_U; ++u) {
      const int i = i0 + u * stride;
      if (i < n) {
#pragma unroll
        for (int c = 0; c < NR; ++c) {
          if (!act[c]) continue;
          const size_t o = (size_t)i * NRP + c;
          float ri = rv[u][c];
          if (upd) {
            x[o] = xv[u][c] + alf[c] * zv[u][c];
            ri = ri - alf[c] * vv[u][c];
            r[o] = ri;
            part[3 * c] += fabs((double)ri);
            part[3 * c + 1] += (double)ri * (double)ri;
          }
          part[3 * c + 2] += (double)ri * (double)ov[u][c];
        }
      }
    }
  }
  double blk[3 * NR], tot[3 * NR];
  block_sum<3 * NR>(part, blk);
  if (last_block_reduce<3 * NR>(blk, partials, 0, counter, tot) && threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      if (!act[c]) continue;
      SolveScalars &S = s[c];
      bool fin = false;
      if (upd) {
        const double resl = tot[3 * c];
        S.resl = resl;
        S.l2 = sqrt(tot[3 * c + 1]);
        S.s0 = S.sk;
        S.iter += 1;
        S.rsm = resl / (S.res0 + FCP_SMALL);
        bool conv = (S.l2 <= S.tol * S.l2b) || (resl < tol_abs);
        if (!isfinite(resl)) {
          S.status = FCP_ERR_NONFINITE;
          conv = true;
        }
        S.conv = (conv && S.status == 0) ? 1 : 0;
        if (conv || S.iter >= S.imax) {
          S.done = 1;
          fin = true;
        }
      }
      if (!fin) {
        const double bet = tot[3 * c + 2];
        S.sk = bet;
        S.om = bet * S.gam / (S.alf * S.beto + FCP_SMALL);
        S.beto = bet;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Coarse correction of the inner PCG: z += P0 A0^-1 P0^T r with P0 the
// indicator of the tiles and A0 = P0^T A P0 (one unknown per tile, the
// additive two-level preconditioner).  A0 is factored A0 = L L^T in 64 x 64
// blocks (diagonal block factor and inverse in LDS, panel, trailing update),
// W = L^-1 is formed and A0^-1 = W^T W, so an iteration's coarse solve is one
// dense product e0 = A0^-1 r0 (r0 the tile sums of r, formed by the
// preconditioner kernel).  Every sum is taken in a fixed order
// (deterministic results).
// ---------------------------------------------------------------------------
constexpr int TCB = 64;         // block size of the coarse factorisation
constexpr int TC_MAXN = FCP_COARSE_MAXN;  // most tiles with a coarse correction
constexpr int TC_MAXDEG = 32;   // widest tile adjacency (the tile itself included)

__global__ void tc_tile_of_pos(int n, int nb, const int *__restrict__ bpos,
                               int *__restrict__ tpos) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) tpos[p] = tk_block_of(bpos, nb, p);
}

// mark[T * nc + U] = 1 where a row of tile T has an entry in tile U
__global__ void tc_mark(int n, int nc, const int *__restrict__ perm, const int *__restrict__ iperm,
                        const int *__restrict__ tpos, const int *__restrict__ ia,
                        const int *__restrict__ ja, unsigned char *__restrict__ mark) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int T = tpos[p], row = perm[p];
  for (int e = ia[row]; e < ia[row + 1]; ++e) mark[(size_t)T * nc + tpos[iperm[ja[e]]]] = 1;
}

// Slot of every CSR entry of the tiled rows in its row tile's adjacency list:
// eslot[e] = k where cadj[cptr[T] + k] is the tile of the entry's column.
__global__ void tc_eslot(int n, const int *__restrict__ perm, const int *__restrict__ iperm,
                         const int *__restrict__ tpos, const int *__restrict__ ia,
                         const int *__restrict__ ja, const int *__restrict__ cptr,
                         const int *__restrict__ cadj, unsigned char *__restrict__ eslot) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int T = tpos[p], row = perm[p], c0 = cptr[T], deg = cptr[T + 1] - c0;
  for (int e = ia[row]; e < ia[row + 1]; ++e) {
    const int U = tpos[iperm[ja[e]]];
    int k = 0;
    while (k < deg - 1 && cadj[c0 + k] != U) ++k;
    eslot[e] = (unsigned char)k;
  }
}

// A0[T][U] = sum of a_ij over rows i of tile T and columns j of tile U; one
// workgroup per tile, each thread sums its rows in order into its own LDS
// column per adjacent tile, the columns are then added in thread order.
// Dynamic LDS: maxdeg x BLK doubles.
__global__ __launch_bounds__(BLK) void tc_vals(int ncp, int maxdeg, const int *__restrict__ bpos,
                                               const int2 *__restrict__ rs,
                                               const unsigned char *__restrict__ eslot,
                                               const double *__restrict__ a,
                                               const int *__restrict__ cptr,
                                               const int *__restrict__ cadj,
                                               double *__restrict__ A) {
  extern __shared__ double SV[];
  const int T = blockIdx.x, t = threadIdx.x;
  const int c0 = cptr[T], deg = cptr[T + 1] - c0;
  const int hi = bpos[T + 1];
  for (int k = 0; k < deg; ++k) SV[(size_t)k * BLK + t] = 0.0;
  constexpr int RU = 4, KU = 8;  // rows, and entries per row, loaded together
  for (int i0 = bpos[T] + t; i0 < hi; i0 += RU * BLK) {
    int2 r[RU];
#pragma unroll
    for (int u = 0; u < RU; ++u) r[u] = i0 + u * BLK < hi ? rs[i0 + u * BLK] : make_int2(0, 0);
    double av[RU][KU];
    int sv[RU][KU];
#pragma unroll
    for (int u = 0; u < RU; ++u)
#pragma unroll
      for (int k = 0; k < KU; ++k)
        if (k < r[u].y) {
          av[u][k] = a[r[u].x + k];
          sv[u][k] = eslot[r[u].x + k];
        }
#pragma unroll
    for (int u = 0; u < RU; ++u) {  // rows in the order i0, i0 + BLK, ...
#pragma unroll
      for (int k = 0; k < KU; ++k)
        if (k < r[u].y) SV[(size_t)sv[u][k] * BLK + t] += av[u][k];
      for (int e = r[u].x + KU; e < r[u].x + r[u].y; ++e) SV[(size_t)eslot[e] * BLK + t] += a[e];
    }
  }
  __syncthreads();
  for (int k = t; k < deg; k += BLK) {
    double sum = 0.0;
    for (int u = 0; u < BLK; ++u) sum += SV[(size_t)k * BLK + u];
    A[(size_t)T * ncp + cadj[c0 + k]] = sum;
  }
}

__global__ void tc_pad(int nc, int ncp, double *__restrict__ A) {
  const int i = nc + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncp) A[(size_t)i * ncp + i] = 1.0;  // padding unknowns: identity
}

typedef double TcTile[TCB][TCB + 1];

// Operand of a block product (256 threads), stored k-major: S[t][r] =
// block[r][t] (trans) or block[t][r]; every load issued before the stores.
__device__ __forceinline__ void tc_op(TcTile &S, const double *__restrict__ M, int ncp, int bi,
                                      int bj, bool trans) {
  double v[TCB * TCB / 256];
#pragma unroll
  for (int u = 0; u < TCB * TCB / 256; ++u) {
    const int e = (int)threadIdx.x + 256 * u, r = e / TCB, c = e % TCB;
    v[u] = M[(size_t)(bi * TCB + r) * ncp + bj * TCB + c];
  }
#pragma unroll
  for (int u = 0; u < TCB * TCB / 256; ++u) {
    const int e = (int)threadIdx.x + 256 * u, r = e / TCB, c = e % TCB;
    if (trans)
      S[c][r] = v[u];
    else
      S[r][c] = v[u];
  }
}

// tc_op in two halves, so the next operand's loads can be in flight while
// the current product is formed: block (bi, bj) into registers, then into LDS.
constexpr int TC_PER = TCB * TCB / 256;
__device__ __forceinline__ void tc_fetch(double (&v)[TC_PER], const double *__restrict__ M,
                                         int ncp, int bi, int bj) {
#pragma unroll
  for (int u = 0; u < TC_PER; ++u) {
    const int e = (int)threadIdx.x + 256 * u, r = e / TCB, c = e % TCB;
    v[u] = M[(size_t)(bi * TCB + r) * ncp + bj * TCB + c];
  }
}
__device__ __forceinline__ void tc_put(TcTile &S, const double (&v)[TC_PER], bool trans) {
#pragma unroll
  for (int u = 0; u < TC_PER; ++u) {
    const int e = (int)threadIdx.x + 256 * u, r = e / TCB, c = e % TCB;
    if (trans)
      S[c][r] = v[u];
    else
      S[r][c] = v[u];
  }
}

// acc[i][j] += sum_t A[t][ty + 16 i] B[t][tx + 16 j] (256 threads, thread
// (ty, tx) = (tid / 16, tid % 16) holding a 4 x 4 output set; t ascending).
__device__ __forceinline__ void tc_mm(const TcTile &A, const TcTile &B, double (&acc)[4][4]) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  for (int t = 0; t < TCB; ++t) {
    double av[4], bv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) av[i] = A[t][ty + 16 * i];
#pragma unroll
    for (int j = 0; j < 4; ++j) bv[j] = B[t][tx + 16 * j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] += av[i] * bv[j];
  }
}

__device__ __forceinline__ void tc_zero(double (&acc)[4][4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
}

// block (bi, bj) of M = sign * acc
__device__ __forceinline__ void tc_store(double *__restrict__ M, int ncp, int bi, int bj,
                                         const double (&acc)[4][4], double sign) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      M[(size_t)(bi * TCB + ty + 16 * i) * ncp + bj * TCB + tx + 16 * j] = sign * acc[i][j];
}

// Diagonal block k: Cholesky, written back with its upper triangle zeroed,
// and the inverse of the factor written to W's block (k, k).  The factor is
// right-looking on the unscaled columns (thread (ty, tx) holds rows
// ty + 16 i, columns tx + 16 j in registers): step j subtracts
// v[r][j] v[c][j] / d_j from the trailing block, with column j and its pivot
// d_j = v[j][j] published by the column's owners at the end of step j - 1
// into one of two LDS buffers, so a step costs one barrier; at the end
// L[r][j] = v[r][j] / sqrt(d_j), L[j][j] = sqrt(d_j).  The inverse is formed
// row by row, X[r][c] = -(sum over c <= m < r of L[r][m] X[m][c]) / L[r][r],
// by the four lanes of column c (one wave: no workgroup barrier).
__global__ __launch_bounds__(256) void tc_potrf(int ncp, int k, double *__restrict__ A,
                                                double *__restrict__ W, int *__restrict__ err) {
  __shared__ TcTile B;
  __shared__ TcTile X;
  __shared__ double col[2][TCB], piv[TCB], d0s[TCB];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double v[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int q = 0; q < 4; ++q)
      v[i][q] = A[(size_t)(k * TCB + ty + 16 * i) * ncp + k * TCB + tx + 16 * q];
  if (tx == ty) {
#pragma unroll
    for (int i = 0; i < 4; ++i) d0s[ty + 16 * i] = v[i][i];  // original diagonal
  }
  if (tx == 0) {  // column 0 (unscaled) for step 0
#pragma unroll
    for (int i = 0; i < 4; ++i) col[0][ty + 16 * i] = v[i][0];
  }
  for (int j = 0; j < TCB; ++j) {
    __syncthreads();
    const double *cj = col[j & 1];
    const double d = cj[j];
    // uniform: a pivot lost to cancellation marks a singular coarse operator
    if (!(d > 1e-10 * d0s[j])) {
      if (threadIdx.x == 0) atomicOr(err, 2);
      return;
    }
    if (threadIdx.x == 0) piv[j] = d;
    const double rd = 1.0 / d;
    double cr[4], cc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) cr[i] = cj[ty + 16 * i];
#pragma unroll
    for (int q = 0; q < 4; ++q) cc[q] = cj[tx + 16 * q];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = ty + 16 * i, c = tx + 16 * q;
        if (c > j && c <= r) v[i][q] = v[i][q] - (cr[i] * cc[q]) * rd;
      }
    const int jn = j + 1;
    if (jn < TCB && tx == (jn & 15)) {  // owners of column j + 1 publish it
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        double cv = v[i][0];
#pragma unroll
        for (int q = 1; q < 4; ++q) cv = q == (jn >> 4) ? v[i][q] : cv;
        col[jn & 1][ty + 16 * i] = cv;  // read only for rows >= j + 1
      }
    }
  }
  __syncthreads();
  // scale the columns: L[r][c] = v[r][c] / sqrt(d_c) below the diagonal
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = ty + 16 * i, c = tx + 16 * q;
      const double l = sqrt(piv[c]);
      B[r][c] = c < r ? v[i][q] / l : (c == r ? l : 0.0);
      X[r][c] = 0.0;
    }
  __syncthreads();
  const int c = threadIdx.x >> 2, part = threadIdx.x & 3;
  for (int r = 0; r < TCB; ++r) {
    double sum = 0.0;
#pragma unroll
    for (int u = 0; u < TCB / 4; ++u) {  // every load issued, the terms masked
      const int m = 4 * u + part;
      const double w = (m >= c && m < r) ? 1.0 : 0.0;
      sum += (B[r][m] * X[m][c]) * w;  // finite operands: a masked term adds 0
    }
    sum += __shfl_xor(sum, 1);
    sum += __shfl_xor(sum, 2);
    if (part == 0) {
      if (c < r) X[r][c] = -sum / B[r][r];
      if (c == r) X[r][c] = 1.0 / B[r][r];
    }
    // column c's lanes are one wave and LDS operations of a wave complete in
    // order: the next row reads what this one wrote without a barrier
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  for (int e = threadIdx.x; e < TCB * TCB; e += blockDim.x) {
    const int r = e / TCB, cc = e % TCB;
    const size_t o = (size_t)(k * TCB + r) * ncp + k * TCB + cc;
    A[o] = B[r][cc];
    W[o] = X[r][cc];
  }
}

// Panel below diagonal block k: L_ik = A_ik L_kk^-T = A_ik W_kk^T (block row
// i = k+1+blockIdx.x).
__global__ __launch_bounds__(256) void tc_panel(int ncp, int k, double *__restrict__ A,
                                                const double *__restrict__ W) {
  __shared__ TcTile P;
  __shared__ TcTile Q;
  const int i = k + 1 + blockIdx.x;
  tc_op(P, A, ncp, i, k, true);
  tc_op(Q, W, ncp, k, k, true);
  __syncthreads();
  double acc[4][4];
  tc_zero(acc);
  tc_mm(P, Q, acc);
  tc_store(A, ncp, i, k, acc, 1.0);
}

// Trailing update after block column k: A_ij -= L_ik L_jk^T for k < j <= i
// (blockIdx.x enumerates the pairs row by row).
__global__ __launch_bounds__(256) void tc_update(int ncp, int k, double *__restrict__ A) {
  __shared__ TcTile P;
  __shared__ TcTile Q;
  int e0 = blockIdx.x, ii = 0;
  while (e0 > ii) {
    e0 -= ii + 1;
    ++ii;
  }
  const int i = k + 1 + ii, j = k + 1 + e0;
  tc_op(P, A, ncp, i, k, true);
  tc_op(Q, A, ncp, j, k, true);
  __syncthreads();
  double acc[4][4];
  tc_zero(acc);
  tc_mm(P, Q, acc);
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const size_t o = (size_t)(i * TCB + ty + 16 * a) * ncp + j * TCB + tx + 16 * b;
      A[o] = A[o] - acc[a][b];
    }
}

// Block row i of W = L^-1: W_ij = -W_ii sum_{m=j..i-1} L_im W_mj (j = blockIdx.x).
__global__ __launch_bounds__(256) void tc_inv(int ncp, int i, const double *__restrict__ A,
                                              double *__restrict__ W) {
  __shared__ TcTile P;
  __shared__ TcTile Q;
  const int j = blockIdx.x;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[4][4];
  tc_zero(acc);
  double pv[TC_PER], qv[TC_PER];
  tc_fetch(pv, A, ncp, i, j);
  tc_fetch(qv, W, ncp, j, j);
  for (int m = j; m < i; ++m) {
    tc_put(P, pv, true);
    tc_put(Q, qv, false);
    __syncthreads();
    if (m + 1 < i) {  // the next operands load during this product
      tc_fetch(pv, A, ncp, i, m + 1);
      tc_fetch(qv, W, ncp, m + 1, j);
    } else {
      tc_fetch(pv, W, ncp, i, i);
    }
    tc_mm(P, Q, acc);
    __syncthreads();
  }
  // S (acc) to LDS k-major, then -W_ii S
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) Q[ty + 16 * a][tx + 16 * b] = acc[a][b];
  tc_put(P, pv, true);
  __syncthreads();
  tc_zero(acc);
  tc_mm(P, Q, acc);
  tc_store(W, ncp, i, j, acc, -1.0);
}

// A0^-1 = W^T W (block (bi, bj) = sum over block rows m >= max(bi, bj) of
// W_m,bi^T W_m,bj), written over the factor.
__global__ __launch_bounds__(256) void tc_gram(int ncp, const double *__restrict__ W,
                                               double *__restrict__ G) {
  __shared__ TcTile P;
  __shared__ TcTile Q;
  const int bi = blockIdx.x, bj = blockIdx.y, nbk = ncp / TCB;
  double acc[4][4];
  tc_zero(acc);
  const int m0 = bi > bj ? bi : bj;
  double pv[TC_PER], qv[TC_PER];
  tc_fetch(pv, W, ncp, m0, bi);
  tc_fetch(qv, W, ncp, m0, bj);
  for (int m = m0; m < nbk; ++m) {
    tc_put(P, pv, false);
    tc_put(Q, qv, false);
    __syncthreads();
    if (m + 1 < nbk) {  // the next operands load during this product
      tc_fetch(pv, W, ncp, m + 1, bi);
      tc_fetch(qv, W, ncp, m + 1, bj);
    }
    tc_mm(P, Q, acc);
    __syncthreads();
  }
  tc_store(G, ncp, bi, bj, acc, 1.0);
}

// z += e0[T] on the positions of tile T, e0 = A0^-1 r0 the coarse correction
// (r0 the tile sums of r, ncp per column), then sk += e0 . r0 (the coarse
// part of r.z) and bet = sk / s0.  Workgroup w updates part w % parts of
// tile w / parts; every workgroup of a tile forms its entry of e0 with the
// same summation order (thread t sums the m = t mod BLK, then the waves in
// order).
template <int NR, int NRP, class VT>
__global__ __launch_bounds__(BLK) void tc_zc(int parts, int ncp, const int *__restrict__ bpos,
                                             const int *__restrict__ failed,
                                             const double *__restrict__ G,
                                             const double *__restrict__ r0,
                                             VT *__restrict__ z, double *__restrict__ partials,
                                             int *__restrict__ counter,
                                             SolveScalars *__restrict__ s) {
  __shared__ double S[NR][BLK / 64];
  if (tk_all_done<NR>(s)) return;
  bool act[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) act[c] = !solve_done(s + c);
  const int T = blockIdx.x / parts, part = blockIdx.x % parts;
  const int t = (int)threadIdx.x, lane = t & 63, wv = t >> 6;
  const int lo = bpos[T], len = bpos[T + 1] - lo;
  const int chunk = (len + parts - 1) / parts;
  const int i0 = lo + part * chunk, i1 = min(lo + len, i0 + chunk);
  // the first positions' z loaded while the coarse row is summed
  VT zv[GS_U][NR];
#pragma unroll
  for (int u = 0; u < GS_U; ++u) {
    const int i = i0 + t + u * BLK;
    if (i < i1)
#pragma unroll
      for (int c = 0; c < NR; ++c) zv[u][c] = z[(size_t)i * NRP + c];
  }
  const double *g = G + (size_t)T * ncp;
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    double sum = 0.0;
    for (int m = t; m < ncp; m += BLK) sum += g[m] * r0[(size_t)c * ncp + m];
    sum = wave_sum(sum);
    if (lane == 0) S[c][wv] = sum;
  }
  __syncthreads();
  const bool off = *failed != 0;  // no factor: tile-local only
  double E[NR];
  VT e[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < BLK / 64; ++w) v += S[c][w];
    E[c] = off ? 0.0 : v;
    e[c] = (VT)E[c];
  }
#pragma unroll
  for (int u = 0; u < GS_U; ++u) {
    const int i = i0 + t + u * BLK;
    if (i < i1)
#pragma unroll
      for (int c = 0; c < NR; ++c)
        if (act[c]) z[(size_t)i * NRP + c] = zv[u][c] + e[c];
  }
  for (int i = i0 + t + GS_U * BLK; i < i1; i += BLK) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      if (!act[c]) continue;
      const size_t o = (size_t)i * NRP + c;
      z[o] = z[o] + e[c];
    }
  }
  double pt[NR], tot[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) pt[c] = part == 0 ? E[c] * r0[(size_t)c * ncp + T] : 0.0;
  if (last_block_reduce<NR>(pt, partials, 0, counter, tot) && threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      if (!act[c]) continue;
      s[c].sk = s[c].sk + tot[c];
      s[c].bet = s[c].sk / s[c].s0;
    }
  }
}

template <class T>
__global__ void tk_gather_in(int n, const int *__restrict__ perm, const double *__restrict__ r,
                             T *__restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (T)r[perm[i]];
}
template <class T>
__global__ void tk_scatter_out(int n, const int *__restrict__ perm, const T *__restrict__ v,
                               double *__restrict__ z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[perm[i]] = (double)v[i];
}

// LDS of one wave's tile region, in values of ts bytes (values, pivots, then
// the step bounds).
int tk_wave_stride(const fcp_ctx *ctx, int nrp, int ts = sizeof(float)) {
  const BlockSchedule &b = ctx->jlower;
  return (((b.R + 1) * nrp + 1) & ~1) + ((b.R + 1) & ~1) +
         (b.max_steps + 1) * (int)(sizeof(int2) / ts);
}

size_t tk_lds(const fcp_ctx *ctx, int nrp, int ts = sizeof(float)) {
  const BlockSchedule &b = ctx->jlower;
  if (ctx->bj_wave) return (size_t)(BLK / 64) * tk_wave_stride(ctx, nrp, ts) * ts;
  // values (R slots and the zero slot), pivots, then 8-byte aligned step bounds
  return (size_t)((((b.R + 1) * nrp + 1) & ~1) + ((b.R + 1) & ~1)) * ts +
         (size_t)(b.max_steps + 1) * sizeof(int2);
}

#ifndef FCP_TK_SPMV_BLOCKS
#define FCP_TK_SPMV_BLOCKS 1024
#endif
// Workgroups of the SpMV (grid-stride beyond that).
int tk_spmv_grid(const fcp_ctx *ctx) {
  const int g = grid_for(ctx->n, BLK);
  const int cap = FCP_TK_SPMV_BLOCKS < MAXBLK ? FCP_TK_SPMV_BLOCKS : MAXBLK;
  return g < cap ? (g > 0 ? g : 1) : cap;
}

int tk_grid(const fcp_ctx *ctx) {
  const int nb = ctx->jlower.nblocks;
  return nb < MAXBLK ? nb : MAXBLK;
}

// Pattern (once per tile-local schedule) and values in FP32, or in FP64 with
// fp64 set (per pivot epoch).
int tk_prepare(fcp_ctx *ctx, bool fp64 = false) {
  const int n = ctx->n;
  const BlockSchedule &b = ctx->jlower;
  const int g = grid_for(n, BLK);
  if (!ctx->tin_have) {
    FCP_HIP(ctx, ctx->tin_iperm.alloc(n));
    FCP_HIP(ctx, ctx->tin_dl.alloc(2 * (size_t)n));
    FCP_HIP(ctx, ctx->tin_du.alloc(2 * (size_t)n));
    tk_iperm<<<g, BLK, 0, ctx->stream>>>(n, b.rows.p, ctx->tin_iperm.p);
    FCP_HIP(ctx, hipMemsetAsync(ctx->errflag.p, 0, sizeof(int), ctx->stream));
    tk_pattern<<<g, BLK, 0, ctx->stream>>>(
        n, b.nblocks, b.R, b.rows.p, ctx->tin_iperm.p, b.blk_pos.p, b.prc.p, b.pdep.p, ctx->ia.p,
        ctx->diag.p, ctx->ja.p, reinterpret_cast<int2 *>(ctx->tin_dl.p),
        reinterpret_cast<int2 *>(ctx->tin_du.p), ctx->errflag.p);
    const int nsl = (n + 63) / 64;
    DBuf<int> len;
    FCP_HIP(ctx, len.alloc(nsl + 1));
    FCP_HIP(ctx, ctx->tin_base.alloc(nsl + 1));
    tk_sell_len<<<grid_for(nsl + 1, BLK), BLK, 0, ctx->stream>>>(n, nsl, b.rows.p, ctx->ia.p,
                                                                 len.p);
    FCP_TRY(sell_uniform(ctx, len.p, nsl, &ctx->tin_width));
    FCP_TRY(fcp_exclusive_scan_int(ctx, len.p, ctx->tin_base.p, nsl + 1));
    int npad = 0;
    FCP_HIP(ctx, hipMemcpyAsync(&npad, ctx->tin_base.p + nsl, sizeof(int), hipMemcpyDeviceToHost,
                                ctx->stream));
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int err = 0;
    FCP_HIP(ctx, hipMemcpy(&err, ctx->errflag.p, sizeof(int), hipMemcpyDeviceToHost));
    if (err) return FCP_ERR_UNSUPPORTED;  // more upper dependencies than slots
    FCP_HIP(ctx, ctx->tin_rs.alloc(2 * (size_t)n));
    tk_row_src<<<g, BLK, 0, ctx->stream>>>(n, b.rows.p, ctx->ia.p,
                                          reinterpret_cast<int2 *>(ctx->tin_rs.p));
    FCP_HIP(ctx, ctx->tin_usrc.alloc((size_t)FCP_PMAXD * n));
    tk_upper_src<<<g, BLK, 0, ctx->stream>>>(n, b.nblocks, b.rows.p, ctx->tin_iperm.p,
                                            b.blk_pos.p, ctx->ia.p, ctx->diag.p, ctx->ja.p,
                                            ctx->tin_usrc.p);
    FCP_HIP(ctx, ctx->tin_col.alloc(npad));
    FCP_HIP(ctx, ctx->tin_val.alloc(npad));
    tk_sell_col<<<grid_for(nsl * 64, BLK), BLK, 0, ctx->stream>>>(
        n, nsl, b.rows.p, ctx->tin_iperm.p, ctx->ia.p, ctx->ja.p, ctx->tin_base.p,
        ctx->tin_col.p);
    FCP_HIP(ctx, hipGetLastError());
    ctx->tin_pack = false;
    if (ctx->tune_tile_pack) {
      FCP_HIP(ctx, ctx->tin_offl.alloc(n));
      FCP_HIP(ctx, ctx->tin_offu.alloc(n));
      FCP_HIP(ctx, hipMemsetAsync(ctx->errflag.p, 0, sizeof(int), ctx->stream));
      tk_pack_offsets<<<g, BLK, 0, ctx->stream>>>(
          n, b.nblocks, b.R, b.blk_pos.p, reinterpret_cast<const int2 *>(ctx->tin_dl.p),
          reinterpret_cast<const int2 *>(ctx->tin_du.p), ctx->tin_offl.p, ctx->tin_offu.p,
          ctx->errflag.p);
      FCP_HIP(ctx, hipGetLastError());
      FCP_HIP(ctx, hipMemcpyAsync(&err, ctx->errflag.p, sizeof(int), hipMemcpyDeviceToHost,
                                  ctx->stream));
      FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
      ctx->tin_pack = err == 0;
    }
    ctx->tin_have = true;
    ctx->tin_epoch = ~0ull;
    ctx->tin_epoch64 = ~0ull;
  }
  auto values = [&](auto *al, auto *au, auto *dv, auto *val, uint64_t &epoch) -> int {
    using T = std::remove_pointer_t<decltype(dv)>;
    if (epoch == ctx->bj_epoch) return 0;
    tk_sweep_vals<T><<<g, BLK, 0, ctx->stream>>>(
        n, b.rows.p, b.psrc.p, ctx->tin_usrc.p, ctx->a.p, ctx->dbj.p,
        ctx->tin_pack ? ctx->tin_offl.p : nullptr, ctx->tin_pack ? ctx->tin_offu.p : nullptr,
        reinterpret_cast<V4<T> *>(al), reinterpret_cast<V4<T> *>(au), dv);
    const int nsl = (n + 63) / 64;
    tk_sell_val<T><<<grid_for((size_t)nsl * 64, BLK), BLK, 0, ctx->stream>>>(
        n, nsl, reinterpret_cast<const int2 *>(ctx->tin_rs.p), ctx->a.p, ctx->tin_base.p, val);
    FCP_HIP(ctx, hipGetLastError());
    epoch = ctx->bj_epoch;
    return 0;
  };
  if (fp64) {
    FCP_HIP(ctx, ctx->tin_al64.alloc(4 * (size_t)n));
    FCP_HIP(ctx, ctx->tin_au64.alloc(4 * (size_t)n));
    FCP_HIP(ctx, ctx->tin_dv64.alloc(n));
    FCP_HIP(ctx, ctx->tin_val64.alloc(ctx->tin_col.n));
    return values(ctx->tin_al64.p, ctx->tin_au64.p, ctx->tin_dv64.p, ctx->tin_val64.p,
                  ctx->tin_epoch64);
  }
  FCP_HIP(ctx, ctx->tin_al.alloc(4 * (size_t)n));
  FCP_HIP(ctx, ctx->tin_au.alloc(4 * (size_t)n));
  FCP_HIP(ctx, ctx->tin_dv.alloc(n));
  return values(ctx->tin_al.p, ctx->tin_au.p, ctx->tin_dv.p, ctx->tin_val.p, ctx->tin_epoch);
}

// Vectors, operands and coarse buffers of a tile-order PCG in precision T.
template <class T>
struct TkBufs {
  T *x, *r, *z, *q, *p[2];
  double *r0 = nullptr;  // tile sums of r (null: no coarse correction)
  int ncp = 0;
  bool fuse = false;     // tk_pre applies the coarse correction itself
  const int2 *dl = nullptr, *du = nullptr;
  const V4<T> *al = nullptr, *au = nullptr;
  const T *dv = nullptr, *val = nullptr;
};

// Vectors (nr columns) and packed operands of the current values in
// precision T (tk_prepare done for T).
int tk_bufs(fcp_ctx *ctx, size_t m, TkBufs<float> &B) {
  for (int k = 0; k < 6; ++k) FCP_HIP(ctx, ctx->tin_vec[k].alloc(m));
  B.x = ctx->tin_vec[0].p;
  B.r = ctx->tin_vec[1].p;
  B.z = ctx->tin_vec[2].p;
  B.q = ctx->tin_vec[3].p;
  B.p[0] = ctx->tin_vec[4].p;
  B.p[1] = ctx->tin_vec[5].p;
  B.al = reinterpret_cast<const float4 *>(ctx->tin_al.p);
  B.au = reinterpret_cast<const float4 *>(ctx->tin_au.p);
  B.dv = ctx->tin_dv.p;
  B.val = ctx->tin_val.p;
  B.dl = reinterpret_cast<const int2 *>(ctx->tin_dl.p);
  B.du = reinterpret_cast<const int2 *>(ctx->tin_du.p);
  return 0;
}
int tk_bufs(fcp_ctx *ctx, size_t m, TkBufs<double> &B) {
  for (int k = 0; k < 6; ++k) FCP_HIP(ctx, ctx->tin_vec64[k].alloc(m));
  B.x = ctx->tin_vec64[0].p;
  B.r = ctx->tin_vec64[1].p;
  B.z = ctx->tin_vec64[2].p;
  B.q = ctx->tin_vec64[3].p;
  B.p[0] = ctx->tin_vec64[4].p;
  B.p[1] = ctx->tin_vec64[5].p;
  B.al = reinterpret_cast<const double4 *>(ctx->tin_al64.p);
  B.au = reinterpret_cast<const double4 *>(ctx->tin_au64.p);
  B.dv = ctx->tin_dv64.p;
  B.val = ctx->tin_val64.p;
  B.dl = reinterpret_cast<const int2 *>(ctx->tin_dl.p);
  B.du = reinterpret_cast<const int2 *>(ctx->tin_du.p);
  return 0;
}

// A0 of the current matrix, its factor and A0^-1 (over the factor), on
// stream st; the factor's failure goes to tc_flag on the device.
int tc_factor(fcp_ctx *ctx, hipStream_t st) {
  const BlockSchedule &b = ctx->jlower;
  const int nc = ctx->tc_nc, ncp = ctx->tc_ncp, nbk = ncp / TCB;
  const size_t nn = (size_t)ncp * ncp;
  FCP_HIP(ctx, ctx->tc_L.alloc(nn));
  FCP_HIP(ctx, ctx->tc_W.alloc(nn));
  FCP_HIP(ctx, ctx->tc_flag.alloc(1));
  FCP_HIP(ctx, hipMemsetAsync(ctx->tc_L.p, 0, nn * sizeof(double), st));
  FCP_HIP(ctx, hipMemsetAsync(ctx->tc_W.p, 0, nn * sizeof(double), st));
  FCP_HIP(ctx, hipMemsetAsync(ctx->tc_flag.p, 0, sizeof(int), st));
  const size_t lds = (size_t)ctx->tc_maxdeg * BLK * sizeof(double);
  tc_vals<<<nc, BLK, lds, st>>>(ncp, ctx->tc_maxdeg, b.blk_pos.p,
                                reinterpret_cast<const int2 *>(ctx->tin_rs.p), ctx->tc_eslot.p,
                                ctx->a.p, ctx->tc_cptr.p, ctx->tc_cadj.p, ctx->tc_L.p);
  if (ncp > nc) tc_pad<<<1, TCB, 0, st>>>(nc, ncp, ctx->tc_L.p);
  for (int k = 0; k < nbk; ++k) {
    tc_potrf<<<1, 256, 0, st>>>(ncp, k, ctx->tc_L.p, ctx->tc_W.p, ctx->tc_flag.p);
    const int m = nbk - k - 1;
    if (m == 0) break;
    tc_panel<<<m, 256, 0, st>>>(ncp, k, ctx->tc_L.p, ctx->tc_W.p);
    tc_update<<<m * (m + 1) / 2, 256, 0, st>>>(ncp, k, ctx->tc_L.p);
  }
  for (int i = 1; i < nbk; ++i) tc_inv<<<i, 256, 0, st>>>(ncp, i, ctx->tc_L.p, ctx->tc_W.p);
  tc_gram<<<dim3(nbk, nbk), 256, 0, st>>>(ncp, ctx->tc_W.p, ctx->tc_L.p);
  FCP_HIP(ctx, hipGetLastError());
  ctx->tc_epoch = ctx->a_version;
  return 0;
}

// Coarse operator of the current tile schedule and matrix: the tile
// adjacency once per schedule, A0, its factor and inverse once per matrix.
// Leaves tc_nc = 0 (no coarse correction) for fewer than two or more than
// TC_MAXN tiles or a tile adjacency wider than TC_MAXDEG.  A coarse operator
// that is not positive definite sets tc_flag on the device (no host wait);
// the iterations then add no correction.
int tc_join(fcp_ctx *ctx);
int tc_stream_init(fcp_ctx *ctx);

int tc_prepare(fcp_ctx *ctx) {
  const BlockSchedule &b = ctx->jlower;
  const int n = ctx->n, nc = b.nblocks;
  FCP_TRY(tc_join(ctx));
  if (!ctx->tc_have) {
    ctx->tc_nc = 0;
    ctx->tc_have = true;
    if (nc < 2 || nc > TC_MAXN) return 0;
    FCP_HIP(ctx, ctx->tc_tpos.alloc(n));
    tc_tile_of_pos<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(n, nc, b.blk_pos.p,
                                                             ctx->tc_tpos.p);
    DBuf<unsigned char> mark;
    FCP_HIP(ctx, mark.alloc((size_t)nc * nc));
    FCP_HIP(ctx, hipMemsetAsync(mark.p, 0, (size_t)nc * nc, ctx->stream));
    tc_mark<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(n, nc, b.rows.p, ctx->tin_iperm.p,
                                                      ctx->tc_tpos.p, ctx->ia.p, ctx->ja.p,
                                                      mark.p);
    FCP_HIP(ctx, hipGetLastError());
    std::vector<unsigned char> hm((size_t)nc * nc);
    FCP_HIP(ctx, hipMemcpyAsync(hm.data(), mark.p, hm.size(), hipMemcpyDeviceToHost,
                                ctx->stream));
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<int> cptr(nc + 1, 0), cadj;
    int maxdeg = 0;
    for (int T = 0; T < nc; ++T) {
      for (int U = 0; U < nc; ++U)
        if (hm[(size_t)T * nc + U]) cadj.push_back(U);
      cptr[T + 1] = (int)cadj.size();
      maxdeg = std::max(maxdeg, cptr[T + 1] - cptr[T]);
    }
    if (maxdeg > TC_MAXDEG) return 0;
    FCP_HIP(ctx, ctx->tc_cptr.alloc(nc + 1));
    FCP_HIP(ctx, ctx->tc_cadj.alloc(cadj.size()));
    FCP_HIP(ctx, hipMemcpy(ctx->tc_cptr.p, cptr.data(), sizeof(int) * (nc + 1),
                           hipMemcpyHostToDevice));
    FCP_HIP(ctx, hipMemcpy(ctx->tc_cadj.p, cadj.data(), sizeof(int) * cadj.size(),
                           hipMemcpyHostToDevice));
    FCP_HIP(ctx, ctx->tc_eslot.alloc(ctx->nnz));
    tc_eslot<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(n, b.rows.p, ctx->tin_iperm.p,
                                                       ctx->tc_tpos.p, ctx->ia.p, ctx->ja.p,
                                                       ctx->tc_cptr.p, ctx->tc_cadj.p,
                                                       ctx->tc_eslot.p);
    FCP_HIP(ctx, hipGetLastError());
    ctx->tc_nc = nc;
    ctx->tc_ncp = (nc + TCB - 1) / TCB * TCB;
    ctx->tc_maxdeg = maxdeg;
    ctx->tc_epoch = ~0ull;
    FCP_TRY(tc_stream_init(ctx));
  }
  if (ctx->tc_nc == 0 || ctx->tc_epoch == ctx->a_version) return 0;
  ProfScope ps(ctx, "coarse_setup");
  return tc_factor(ctx, ctx->stream);
}

// The side stream of the coarse factorisation and its events, created (and
// its queue brought up by a first operation) with the adjacency, so that the
// first overlapped factorisation does not pay for them.
int tc_stream_init(fcp_ctx *ctx) {
  if (ctx->tc_stream) return 0;
  FCP_HIP(ctx, hipStreamCreateWithFlags(&ctx->tc_stream, hipStreamNonBlocking));
  FCP_HIP(ctx, hipEventCreateWithFlags(&ctx->tc_ready, hipEventDisableTiming));
  FCP_HIP(ctx, hipEventCreateWithFlags(&ctx->tc_done, hipEventDisableTiming));
  FCP_HIP(ctx, ctx->tc_flag.alloc(1));
  FCP_HIP(ctx, hipMemsetAsync(ctx->tc_flag.p, 0, sizeof(int), ctx->tc_stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->tc_stream));
  return 0;
}

// Waits (on the library stream) for a coarse factorisation started on the
// side stream.
int tc_join(fcp_ctx *ctx) {
  if (!ctx->tc_pending) return 0;
  ctx->tc_pending = false;
  FCP_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->tc_done, 0));
  return 0;
}

// Starts the coarse factorisation of the current matrix on a side stream
// (default priority), after the work already queued on the library stream,
// so that it overlaps the inner solve's own set-up; tc_bufs joins it.  Needs
// the adjacency of the current tile schedule (built by an earlier solve);
// otherwise tc_prepare does the whole set-up in order.
int tc_begin(fcp_ctx *ctx) {
  if (!ctx->inner_coarse || !ctx->have_bj || !ctx->tc_have || ctx->tc_nc == 0 ||
      ctx->tc_pending || ctx->tc_epoch == ctx->a_version)
    return 0;
  FCP_TRY(tc_stream_init(ctx));
  FCP_HIP(ctx, hipEventRecord(ctx->tc_ready, ctx->stream));
  FCP_HIP(ctx, hipStreamWaitEvent(ctx->tc_stream, ctx->tc_ready, 0));
  FCP_TRY(tc_factor(ctx, ctx->tc_stream));
  FCP_HIP(ctx, hipEventRecord(ctx->tc_done, ctx->tc_stream));
  ctx->tc_pending = true;
  return 0;
}

// Coarse buffers of an inner solve of nr columns (r0 cleared: entries past
// the last tile stay zero), or none where the correction does not apply.
template <class T>
int tc_bufs(fcp_ctx *ctx, int nr, TkBufs<T> &B) {
  if (!ctx->inner_coarse) return 0;
  FCP_TRY(tc_prepare(ctx));
  if (ctx->tc_nc == 0) return 0;
  const size_t m = (size_t)nr * ctx->tc_ncp;
  FCP_HIP(ctx, ctx->tc_r0.alloc(m));
  FCP_HIP(ctx, hipMemsetAsync(ctx->tc_r0.p, 0, m * sizeof(double), ctx->stream));
  B.r0 = ctx->tc_r0.p;
  B.ncp = ctx->tc_ncp;
  return 0;
}

template <int NR, int NRP, class T>
int tk_launch_pre(fcp_ctx *ctx, const TkBufs<T> &B, const T *p, SolveScalars *si, int upd) {
  const BlockSchedule &b = ctx->jlower;
  constexpr int TS = sizeof(T);
  ProfScope ps(ctx, TS == 8 ? "tin_pre64" : "tin_pre");
  auto go = [&](auto pack) {
    constexpr bool PK = decltype(pack)::value;
    if (B.fuse) {
      const CoarseFuse cf{ctx->tc_L.p, ctx->tc_flag.p, ctx->tin_bar.p, ctx->errflag.p};
      tk_pre<NR, NRP, TK_NT, PK, T, true><<<b.nblocks, TK_NT, tk_lds(ctx, NRP, TS), ctx->stream>>>(
          b.nblocks, ((b.R + 1) * NRP + 1) & ~1, b.R, b.blk_pos.p, b.blk_step.p, b.step_ptr.p,
          B.dl, B.al, B.du, B.au, B.dv, B.x, B.r, p, B.q, B.z, ctx->partials.p, redctr(ctx), si,
          upd, 0.0, B.r0, B.ncp, cf);
    } else if (ctx->bj_wave) {
      int g = (b.nblocks + BLK / 64 - 1) / (BLK / 64);
      g = g < MAXBLK ? g : MAXBLK;
      tk_pre_w<NR, NRP, PK, T><<<g, BLK, tk_lds(ctx, NRP, TS), ctx->stream>>>(
          b.nblocks, tk_wave_stride(ctx, NRP, TS), ((b.R + 1) * NRP + 1) & ~1, b.R, b.blk_pos.p,
          b.blk_step.p, b.step_ptr.p, B.dl, B.al, B.du, B.au, B.dv, B.x, B.r, p, B.q, B.z,
          ctx->partials.p, redctr(ctx), si, upd, 0.0, B.r0, B.ncp);
    } else {
      tk_pre<NR, NRP, TK_NT, PK, T><<<tk_grid(ctx), TK_NT, tk_lds(ctx, NRP, TS), ctx->stream>>>(
          b.nblocks, ((b.R + 1) * NRP + 1) & ~1, b.R, b.blk_pos.p, b.blk_step.p, b.step_ptr.p,
          B.dl, B.al, B.du, B.au, B.dv, B.x, B.r, p, B.q, B.z, ctx->partials.p, redctr(ctx), si,
          upd, 0.0, B.r0, B.ncp);
    }
    return hipGetLastError() == hipSuccess ? 0 : FCP_ERR_HIP;
  };
  return ctx->tin_pack ? go(std::true_type()) : go(std::false_type());
}

// Inner iteration k: p_new = z + bet p_old, q = A p_new (tk_spmv), then the
// update, stopping test and next preconditioner application (tk_pre).
// With the coarse correction, tc_zc first adds it to z and completes r.z and
// bet.
// Workgroups per tile of tc_zc: about 1024 in all (at 512 tiles, 2 per tile
// took 11.5 us per launch, 4 per tile 12.4 us, 8 per tile 16.6 us).
int tc_parts(const fcp_ctx *ctx) {
  const int nb = ctx->jlower.nblocks, p = 1024 / nb;
  return p < 1 ? 1 : (p > 8 ? 8 : p);
}

template <int NR, int NRP, class T>
int tc_launch_zc(fcp_ctx *ctx, const TkBufs<T> &B, SolveScalars *si) {
  ProfScope ps(ctx, "tin_coarse");
  const int parts = tc_parts(ctx);
  tc_zc<NR, NRP, T><<<ctx->jlower.nblocks * parts, BLK, 0, ctx->stream>>>(
      parts, B.ncp, ctx->jlower.blk_pos.p, ctx->tc_flag.p, ctx->tc_L.p, B.r0, B.z,
      ctx->partials.p, redctr(ctx), si);
  return hipGetLastError() == hipSuccess ? 0 : FCP_ERR_HIP;
}

// Whether tk_pre can apply the coarse correction itself: one workgroup per
// tile (not the one-wave tiles), every workgroup resident at once (the grid
// barrier), the reduction's values within the partials' slots.
template <int NR, int NRP, class T>
bool tk_can_fuse(fcp_ctx *ctx) {
  const BlockSchedule &b = ctx->jlower;
  if (!ctx->tune_coarse_fuse || ctx->bj_wave || 4 * NR > 4 * NSLOT || b.nblocks > MAXBLK)
    return false;
  const size_t lds = tk_lds(ctx, NRP, sizeof(T));
  int per_cu = 0;
  const void *k = ctx->tin_pack
                      ? reinterpret_cast<const void *>(tk_pre<NR, NRP, TK_NT, true, T, true>)
                      : reinterpret_cast<const void *>(tk_pre<NR, NRP, TK_NT, false, T, true>);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, TK_NT, lds) != hipSuccess)
    return false;
  return (long long)per_cu * ctx->ncu >= b.nblocks;
}

template <int NR, int NRP, class T>
int tk_iteration(fcp_ctx *ctx, const TkBufs<T> &B, SolveScalars *si, int k) {
  if (B.r0 && !B.fuse) FCP_TRY((tc_launch_zc<NR, NRP, T>(ctx, B, si)));
  T *pold = B.p[k & 1], *pnew = B.p[(k + 1) & 1];
  {
    ProfScope ps(ctx, sizeof(T) == 8 ? "tin_spmv64" : "tin_spmv");
    tk_spmv<NR, NRP, T><<<tk_spmv_grid(ctx), BLK, 0, ctx->stream>>>(
        ctx->n, ctx->tin_base.p, ctx->tin_width, ctx->tin_col.p, B.val, B.z, pold, pnew, B.q,
        ctx->partials.p,
        redctr(ctx), si);
  }
  return tk_launch_pre<NR, NRP, T>(ctx, B, pnew, si, 1);
}

// Executable graph of two inner iterations (k even, then odd) for the current
// buffers, captured once and reused while every pointer it holds is unchanged.
template <int NR, int NRP, class T>
int tk_graph(fcp_ctx *ctx, const TkBufs<T> &B, SolveScalars *si, hipGraphExec_t *out) {
  const BlockSchedule &b = ctx->jlower;
  const std::vector<const void *> key = {
      (const void *)(size_t)NR, (const void *)sizeof(T), (const void *)(size_t)ctx->tin_pack,
      (const void *)(size_t)B.fuse, ctx->tin_bar.p, ctx->errflag.p,
      B.x, B.r, B.z, B.q, B.p[0], B.p[1], si,
      ctx->partials.p, ctx->counters.p, B.dl, B.du, B.al, B.au, B.dv,
      ctx->tin_base.p, (const void *)(size_t)ctx->tin_width, ctx->tin_col.p, B.val, b.blk_pos.p,
      b.blk_step.p, b.step_ptr.p,
      (const void *)(size_t)b.nblocks, (const void *)(size_t)b.R, (const void *)(size_t)b.max_steps,
      (const void *)(size_t)ctx->n, B.r0, (const void *)(size_t)B.ncp, ctx->tc_L.p};
  if (ctx->tin_graph && ctx->tin_graph_key == key) {
    *out = ctx->tin_graph;
    return 0;
  }
  if (ctx->tin_graph) {
    FCP_HIP(ctx, hipGraphExecDestroy(ctx->tin_graph));
    ctx->tin_graph = nullptr;
  }
  hipGraph_t g = nullptr;
  FCP_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = tk_iteration<NR, NRP, T>(ctx, B, si, 0);
  if (rc == 0) rc = tk_iteration<NR, NRP, T>(ctx, B, si, 1);
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
  FCP_TRY(rc);
  FCP_HIP(ctx, ec);
  const hipError_t ei = hipGraphInstantiate(&ctx->tin_graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  FCP_HIP(ctx, ei);
  ctx->tin_graph_key = key;
  *out = ctx->tin_graph;
  return 0;
}

// The pinned slots and events of the pipelined polls (once per context).
int poll_slots(fcp_ctx *ctx) {
  if (!ctx->h_poll) FCP_HIP(ctx, hipHostMalloc((void **)&ctx->h_poll, 8 * sizeof(SolveScalars)));
  if (!ctx->h_perr) FCP_HIP(ctx, hipHostMalloc((void **)&ctx->h_perr, 2 * sizeof(int)));
  for (auto &e : ctx->poll_ev)
    if (!e) FCP_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return 0;
}

template <int NR, int NRP, class T>
int tk_pcg(fcp_ctx *ctx, const double *res, double *x, const SolveScalars *so, SolveScalars *si,
           double tol_rel, double inner_tol, int budget, int check_every, int verbose,
           SolveScalars *h) {
  const int n = ctx->n;
  TkBufs<T> B;
  FCP_TRY(tk_bufs(ctx, (size_t)n * NRP, B));
  FCP_TRY(tc_bufs(ctx, NR, B));
  if (B.r0 && tk_can_fuse<NR, NRP, T>(ctx)) {
    FCP_HIP(ctx, ctx->tin_bar.alloc(64));
    FCP_HIP(ctx, hipMemsetAsync(ctx->tin_bar.p, 0, 64 * sizeof(unsigned), ctx->stream));
    B.fuse = true;
  }
  const int g = red_grid(ctx, n);
  tk_init<NR, NRP, T><<<g, BLK, 0, ctx->stream>>>(
      n, ctx->jlower.rows.p, res, so, B.r, B.x, B.p[0], ctx->partials.p, redctr(ctx), si,
      tol_rel, inner_tol, budget, (T *)nullptr, (T *)nullptr);
  FCP_TRY((tk_launch_pre<NR, NRP, T>(ctx, B, B.p[0], si, 0)));
  BatchSizer bs(check_every, BATCH_INNER, 0.0);
  hipGraphExec_t ge = nullptr;
  if (ctx->tune_graph && !ctx->prof) FCP_TRY((tk_graph<NR, NRP, T>(ctx, B, si, &ge)));
  int launched = 0, k = 0;
  auto run = [&](int batch) -> int {
    int j = 0;
    if (ge && (k & 1) == 0)  // the graph starts from an even iteration
      for (; j + 2 <= batch; j += 2, k += 2) FCP_HIP(ctx, hipGraphLaunch(ge, ctx->stream));
    for (; j < batch; ++j, ++k) FCP_TRY((tk_iteration<NR, NRP, T>(ctx, B, si, k)));
    launched += batch;
    return 0;
  };
  // Pipelined polls: the scalars after every batch are copied to a pinned
  // slot, and the next batch is queued before the host waits for the
  // previous slot, so the device does not idle while the host reads it.
  // Batches queued after the device-side stopping test exit at their first
  // instruction and leave the scalars unchanged.  A batch is queued ahead
  // only while the convergence rate seen so far predicts that the work
  // already queued does not reach the tolerance.
  FCP_TRY(poll_slots(ctx));
  auto post = [&](int slot) -> int {
    FCP_HIP(ctx, hipMemcpyAsync(ctx->h_poll + 4 * slot, si, NR * sizeof(SolveScalars),
                                hipMemcpyDeviceToHost, ctx->stream));
    FCP_HIP(ctx, hipMemcpyAsync(ctx->h_perr + slot, ctx->errflag.p, sizeof(int),
                                hipMemcpyDeviceToHost, ctx->stream));
    FCP_HIP(ctx, hipEventRecord(ctx->poll_ev[slot], ctx->stream));
    return 0;
  };
  const int cap = bs.cap;
  FCP_TRY(run(budget < cap ? budget : cap));
  FCP_TRY(post(0));
  int slot = 0;     // the slot of the oldest poll not yet read
  int ahead = cap;  // iterations predicted to be needed after those queued
  const SolveScalars *hs = nullptr;
  for (;;) {
    int spec = 0;
    if (ahead > 0 && launched < budget) {
      spec = ahead < cap ? ahead : cap;
      if (spec > budget - launched) spec = budget - launched;
      FCP_TRY(run(spec));
      FCP_TRY(post(1 - slot));
    }
    FCP_HIP(ctx, hipEventSynchronize(ctx->poll_ev[slot]));
    if (ctx->h_perr[slot])
      return fcp_fail(ctx, ctx->h_perr[slot], "sweep dependency wait exceeded its bound");
    hs = ctx->h_poll + 4 * slot;
    bool all = true;
    for (int c = 0; c < NR; ++c) {
      if (verbose >= 2)
        printf("                   PCG%d[%d] iter = %4d resl = %10.3E rsm = %10.3E\n",
               (int)(8 * sizeof(T)), c, hs[c].iter, hs[c].resl, hs[c].rsm);
      all = all && hs[c].done;
    }
    if (all) break;
    bs.next(hs, NR);
    const int need = bs.last_need;  // iterations predicted past this poll
    if (spec > 0) {
      ahead = need - spec;  // the batch queued since is part of them
      slot = 1 - slot;
      continue;
    }
    if (launched >= budget) break;  // not converged within the budget
    int b = need < cap ? need : cap;
    if (b > budget - launched) b = budget - launched;
    FCP_TRY(run(b));
    FCP_TRY(post(slot));
    ahead = need - b;
  }
  // the scalars are final once the stopping test has fired (or the budget
  // is spent); wait for any batch still queued before they are returned
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int c = 0; c < NR; ++c) h[c] = hs[c];
  tk_ir_update<NR, NRP, T><<<g, BLK, 0, ctx->stream>>>(n, ctx->tin_iperm.p, B.x, so, x);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}


template <int NR, int NRP>
int tb_bicgstab(fcp_ctx *ctx, const double *res, double *x, const SolveScalars *so,
                SolveScalars *si, double tol_rel, double inner_tol, int budget, int check_every,
                int verbose, SolveScalars *h) {
  const int n = ctx->n;
  const size_t m = (size_t)n * NRP;
  for (int k = 0; k < 7; ++k) FCP_HIP(ctx, ctx->tin_vec[k].alloc(m));
  float *X = ctx->tin_vec[0].p, *R = ctx->tin_vec[1].p, *Z = ctx->tin_vec[2].p,
        *P = ctx->tin_vec[3].p, *U = ctx->tin_vec[4].p, *Vv = ctx->tin_vec[5].p,
        *Reso = ctx->tin_vec[6].p;
  const BlockSchedule &b = ctx->jlower;
  const int g = red_grid(ctx, n);
  const size_t lds = tk_lds(ctx, NRP);
  const int vslots = ((b.R + 1) * NRP + 1) & ~1;
  const int2 *dl = reinterpret_cast<const int2 *>(ctx->tin_dl.p);
  const int2 *du = reinterpret_cast<const int2 *>(ctx->tin_du.p);
  const float4 *al = reinterpret_cast<const float4 *>(ctx->tin_al.p);
  const float4 *au = reinterpret_cast<const float4 *>(ctx->tin_au.p);
  tk_init<NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, b.rows.p, res, so, R, X, P, ctx->partials.p,
                                               redctr(ctx), si, tol_rel, inner_tol, budget, Reso,
                                               U);
  auto bet = [&](int upd) {
    ProfScope ps(ctx, "tinb_bet");
    tb_bet<NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, X, R, Z, Vv, Reso, ctx->partials.p,
                                                redctr(ctx), si, upd, 0.0);
  };
  auto tile = [&](auto mode) {
    constexpr int MODE = decltype(mode)::value;
    ProfScope ps(ctx, MODE == 1 ? "tinb_pz" : "tinb_xz");
    auto go = [&](auto pack) {
      tb_tile<NR, NRP, MODE, decltype(pack)::value><<<tk_grid(ctx), BLK, lds, ctx->stream>>>(
          b.nblocks, vslots, b.R, b.blk_pos.p, b.blk_step.p, b.step_ptr.p, dl, al, du, au,
          ctx->tin_dv.p, X, R, P, U, Z, si);
    };
    if (ctx->tin_pack)
      go(std::true_type());
    else
      go(std::false_type());
  };
  auto spmv = [&](auto mode) {
    constexpr int MODE = decltype(mode)::value;
    ProfScope ps(ctx, MODE == 0 ? "tinb_u" : "tinb_v");
    tb_spmv<NR, NRP, MODE><<<tk_spmv_grid(ctx), BLK, 0, ctx->stream>>>(
        n, ctx->tin_base.p, ctx->tin_width, ctx->tin_col.p, ctx->tin_val.p, Z, MODE == 0 ? Reso : R,
        MODE == 0 ? U : Vv, ctx->partials.p, redctr(ctx), si);
  };
  bet(0);
  BatchSizer bs(check_every, BATCH_INNER, 0.0);
  int next = bs.cap;
  int launched = 0;
  while (launched < budget) {
    const int batch = (budget - launched) < next ? (budget - launched) : next;
    for (int j = 0; j < batch; ++j) {
      tile(std::integral_constant<int, 1>());
      spmv(std::integral_constant<int, 0>());
      tile(std::integral_constant<int, 2>());
      spmv(std::integral_constant<int, 1>());
      bet(1);
    }
    FCP_HIP(ctx, hipGetLastError());
    launched += batch;
    FCP_TRY(poll_device(ctx, h, si, NR * sizeof(SolveScalars)));
    bool all = true;
    for (int c = 0; c < NR; ++c) {
      if (verbose >= 2)
        printf("                   BiCGStab32[%d] iter = %4d resl = %10.3E rsm = %10.3E\n", c,
               h[c].iter, h[c].resl, h[c].rsm);
      all = all && h[c].done;
    }
    if (all) break;
    next = bs.next(h, NR);
  }
  tk_ir_update<NR, NRP><<<g, BLK, 0, ctx->stream>>>(n, ctx->tin_iperm.p, X, so, x);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

}  // namespace

namespace fcpk {
// The tile-order inner PCG applies: tile-local preconditioner built, at most
// FCP_PMAXD dependencies per row in either sweep, slots addressable in 16
// bits, one tile's values and step table in LDS.
bool tin_usable(fcp_ctx *ctx, int nrp, bool fp64) {
  const BlockSchedule &b = ctx->jlower;
  const int ts = fp64 ? (int)sizeof(double) : (int)sizeof(float);
  return ctx->have_bj && b.nblocks > 0 && b.maxd <= FCP_PMAXD && ctx->jupper.maxd <= FCP_PMAXD &&
         b.R < TK_SLOTS && tk_lds(ctx, nrp, ts) + 256 <= ctx->lds_cap;
}

// Inner PCG (FP32, or FP64 with fp64 set) of nr interleaved columns
// (NRP = 1, 2, 4) on the FP64 residual res (scaled by so[c].l2), adding the
// correction to x; the tile-local pivots must be current (bj_begin).  h
// receives the final inner scalars.  FCP_ERR_UNSUPPORTED: the mesh does not
// fit (nothing launched).
int tin_pcg(fcp_ctx *ctx, int nr, const double *res, double *x, const SolveScalars *so,
            SolveScalars *si, double tol_rel, double inner_tol, int budget, int check_every,
            int verbose, SolveScalars *h, bool fp64) {
  const int nrp = nr <= 1 ? 1 : (nr <= 2 ? 2 : 4);
  if (!tin_usable(ctx, nrp, fp64)) return FCP_ERR_UNSUPPORTED;
  // the coarse factorisation overlaps the tile operands' set-up and the first
  // preconditioner application; joined before its first use or on return
  FCP_TRY(tc_begin(ctx));
  struct Join {
    fcp_ctx *c;
    ~Join() { (void)tc_join(c); }
  } join{ctx};
  const int rc = tk_prepare(ctx, fp64);
  if (rc == FCP_ERR_UNSUPPORTED) {
    ctx->tin_have = false;
    return rc;
  }
  FCP_TRY(rc);
  if (fp64) {
    if (nr != 1) return FCP_ERR_ARG;  // FP64 tile-order iterations: one column
    return tk_pcg<1, 1, double>(ctx, res, x, so, si, tol_rel, inner_tol, budget, check_every,
                                verbose, h);
  }
  switch (nr) {
    case 1:
      return tk_pcg<1, 1, float>(ctx, res, x, so, si, tol_rel, inner_tol, budget, check_every,
                                 verbose, h);
    case 2:
      return tk_pcg<2, 2, float>(ctx, res, x, so, si, tol_rel, inner_tol, budget, check_every,
                                 verbose, h);
    case 3:
      return tk_pcg<3, 4, float>(ctx, res, x, so, si, tol_rel, inner_tol, budget, check_every,
                                 verbose, h);
    case 4:
      return tk_pcg<4, 4, float>(ctx, res, x, so, si, tol_rel, inner_tol, budget, check_every,
                                 verbose, h);
    default:
      return FCP_ERR_ARG;
  }
}

// Starts the coarse factorisation of the inner PCG for the current matrix on
// the side stream (nothing when it is current, does not apply or its tile
// adjacency is not built yet); tin_coarse_join makes the library stream wait
// for it.
int tin_coarse_begin(fcp_ctx *ctx) { return tc_begin(ctx); }
int tin_coarse_join(fcp_ctx *ctx) { return tc_join(ctx); }

// Inner FP32 DILU-BiCGStab in tile order; arguments as tin_pcg.
int tin_bicgstab(fcp_ctx *ctx, int nr, const double *res, double *x, const SolveScalars *so,
                 SolveScalars *si, double tol_rel, double inner_tol, int budget, int check_every,
                 int verbose, SolveScalars *h) {
  const int nrp = nr <= 1 ? 1 : (nr <= 2 ? 2 : 4);
  if (!tin_usable(ctx, nrp)) return FCP_ERR_UNSUPPORTED;
  const int rc = tk_prepare(ctx);
  if (rc == FCP_ERR_UNSUPPORTED) {
    ctx->tin_have = false;
    return rc;
  }
  FCP_TRY(rc);
  switch (nr) {
    case 1:
      return tb_bicgstab<1, 1>(ctx, res, x, so, si, tol_rel, inner_tol, budget, check_every,
                               verbose, h);
    case 2:
      return tb_bicgstab<2, 2>(ctx, res, x, so, si, tol_rel, inner_tol, budget, check_every,
                               verbose, h);
    case 3:
      return tb_bicgstab<3, 4>(ctx, res, x, so, si, tol_rel, inner_tol, budget, check_every,
                               verbose, h);
    case 4:
      return tb_bicgstab<4, 4>(ctx, res, x, so, si, tol_rel, inner_tol, budget, check_every,
                               verbose, h);
    default:
      return FCP_ERR_ARG;
  }
}

// z = M^-1 r with the tile-local preconditioner in FP32 (FP64 with fp64 set)
// exactly as the inner solves apply it (one tile-order application, pivots
// current); coarse: plus the coarse correction (FCP_ERR_UNSUPPORTED where it
// does not apply).
int tin_apply(fcp_ctx *ctx, const double *r, double *z, bool coarse, bool fp64) {
  if (!tin_usable(ctx, 1, fp64)) return FCP_ERR_UNSUPPORTED;
  FCP_TRY(tk_prepare(ctx, fp64));
  auto run = [&](auto tag) -> int {
    using T = decltype(tag);
    const int n = ctx->n;
    TkBufs<T> B;
    FCP_TRY(tk_bufs(ctx, n, B));
    if (coarse) {
      const int on = ctx->inner_coarse;
      ctx->inner_coarse = 1;
      const int rc = tc_bufs(ctx, 1, B);
      ctx->inner_coarse = on;
      FCP_TRY(rc);
      if (!B.r0) return FCP_ERR_UNSUPPORTED;
      int failed = 0;
      FCP_HIP(ctx, hipMemcpyAsync(&failed, ctx->tc_flag.p, sizeof(int), hipMemcpyDeviceToHost,
                                  ctx->stream));
      FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
      if (failed) return FCP_ERR_UNSUPPORTED;
    }
    // scalars in the spare slot: not done, unit tolerance, no update
    SolveScalars *si = ctx->scal.p + 3;
    FCP_HIP(ctx, hipMemsetAsync(si, 0, sizeof(SolveScalars), ctx->stream));
    tk_gather_in<T><<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(n, ctx->jlower.rows.p, r, B.r);
    FCP_TRY((tk_launch_pre<1, 1, T>(ctx, B, B.p[0], si, 0)));
    if (coarse) FCP_TRY((tc_launch_zc<1, 1, T>(ctx, B, si)));
    tk_scatter_out<T><<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(n, ctx->jlower.rows.p, B.z, z);
    FCP_HIP(ctx, hipGetLastError());
    return 0;
  };
  return fp64 ? run(0.0) : run(0.0f);
}
}  // namespace fcpk

// Matrix writers join a pending coarse factorisation before they change the
// matrix (the side stream reads it); the Laplacian assembly starts the next
// one (for a context whose inner PCG built the coarse adjacency), so that it
// overlaps the gradient and source-term work before the solve.
int fcp_coarse_join(fcp_ctx *ctx) { return tc_join(ctx); }
// ... and before buffers it reads are freed or reallocated
int fcp_coarse_wait(fcp_ctx *ctx) {
  FCP_TRY(tc_join(ctx));
  if (ctx->tc_stream) FCP_HIP(ctx, hipStreamSynchronize(ctx->tc_stream));
  return 0;
}
int fcp_coarse_begin(fcp_ctx *ctx) { return tc_begin(ctx); }
