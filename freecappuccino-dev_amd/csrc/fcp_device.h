// fcp_device.h -- device helpers shared by the sparse/Krylov translation
// units of libfcp: deterministic block reductions, the NaN-sentinel
// publish/wait protocol of the sync-free sweeps and the SELL-64 row product.
#pragma once
#include <hip/hip_runtime.h>

#include "fcp_internal.h"

namespace fcpk {

constexpr int BLK = 256;
constexpr int MAXBLK = 4096;
// Arrival counters of the block reductions: RED_GROUPS sub-counters and one
// top counter, each on its own 128-byte line, after the two sweep tickets.
constexpr int RED_GROUPS = 32;
constexpr int CTR_INTS = 32 + (RED_GROUPS + 1) * 32;  // blocks of a reduction kernel (partials per slot)

enum Slot { SLOT_A = 0, SLOT_B = 1, SLOT_C = 2, NSLOT = 4 };

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum of up to 3 values into lane 0 of warp 0 (deterministic order).
template <int NV, int NT = BLK>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (&out)[NV]) {
  __shared__ double sh[NV][NT / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = wave_sum(v[i]);
    if (lane == 0) sh[i][w] = s;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
    if (threadIdx.x == 0)
      for (int k = 0; k < NT / 64; ++k) s += sh[i][k];
    out[i] = s;
  }
}

// Publish NV block partials; returns true in the single last-arriving block,
// whose thread 0 then holds the totals (summed in block order) in tot[].
// The partials are written with agent-coherent stores and read back with
// agent-coherent loads; waiting for the stores to complete before counting
// the arrival orders them without a device fence (which would write back
// every dirty L2 line in every block).  Arrivals are counted in two levels
// (block b at sub-counter b % RED_GROUPS, the last of each group at the top
// counter), so no single address takes more than ~gridDim / RED_GROUPS
// serialized atomics; every counter is reset by its last arrival.  counter
// points to (RED_GROUPS + 1) lines of 32 ints: the top counter, then the
// sub-counters.
template <int NV, int NT = BLK>
__device__ bool last_block_reduce(double (&part)[NV], double *__restrict__ partials, int slot,
                                  int *__restrict__ counter, double (&tot)[NV]) {
  __shared__ int is_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i)
      __hip_atomic_store(
          reinterpret_cast<unsigned long long *>(&partials[(size_t)(slot + i) * MAXBLK + blockIdx.x]),
          (unsigned long long)__double_as_longlong(part[i]), __ATOMIC_RELAXED,
          __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);  // every memory counter at zero: the stores are complete
    const int nb = (int)gridDim.x;
    const int g = (int)blockIdx.x % RED_GROUPS;
    const int members = (nb - g + RED_GROUPS - 1) / RED_GROUPS;
    const int groups = nb < RED_GROUPS ? nb : RED_GROUPS;
    int *sub = counter + 32 * (1 + g);
    int last = 0;
    if (__hip_atomic_fetch_add(sub, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == members - 1) {
      // every member's stores completed before its arrival was counted
      __hip_atomic_store(sub, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      last = __hip_atomic_fetch_add(counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
             groups - 1;
    }
    is_last = last;
  }
  __syncthreads();
  if (!is_last) return false;
  // all threads of the last block reduce the partials in fixed order
  double loc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += NT)
      s += __hip_atomic_load(&partials[(size_t)(slot + i) * MAXBLK + b], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    loc[i] = s;
  }
  block_sum<NV, NT>(loc, tot);
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

__device__ __forceinline__ bool solve_done(const SolveScalars *s) {
  return __hip_atomic_load(&s->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// ---------------------------------------------------------------------------
// sentinel wait/publish
// ---------------------------------------------------------------------------
template <typename T>
struct Bits;
template <>
struct Bits<double> {
  using U = unsigned long long;
  static constexpr U sent = FCP_SENT64;
  static __device__ __forceinline__ U to(double v) { return __double_as_longlong(v); }
  static __device__ __forceinline__ double from(U u) { return __longlong_as_double(u); }
  static constexpr U qnan = 0x7FF8000000000000ull;
};
template <>
struct Bits<float> {
  using U = unsigned int;
  static constexpr U sent = FCP_SENT32;
  static __device__ __forceinline__ U to(float v) { return __float_as_uint(v); }
  static __device__ __forceinline__ float from(U u) { return __uint_as_float(u); }
  static constexpr U qnan = 0x7FC00000u;
};

template <typename T>
__device__ __forceinline__ T wait_value(T *p, int *err, int slp) {
  using B = Bits<T>;
  typename B::U *q = reinterpret_cast<typename B::U *>(p);
  typename B::U v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (v != B::sent) return B::from(v);
  // Bounded by wall time (100 MHz constant clock): a dependency that is not
  // published within FCP_WAIT_TICKS records FCP_ERR_SWEEP and the wave
  // continues with NaN, so the kernel always drains.
  const unsigned long long t0 = wall_clock64();
  unsigned spins = 0;
  while (v == B::sent) {
    // s_sleep n idles the wave for about 64*n clocks between polls
    if (slp >= 32)
      __builtin_amdgcn_s_sleep(32);
    else if (slp >= 16)
      __builtin_amdgcn_s_sleep(16);
    else if (slp >= 8)
      __builtin_amdgcn_s_sleep(8);
    else if (slp >= 4)
      __builtin_amdgcn_s_sleep(4);
    else if (slp >= 2)
      __builtin_amdgcn_s_sleep(2);
    else
      __builtin_amdgcn_s_sleep(1);
    v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((++spins & 63u) == 0 && wall_clock64() - t0 > FCP_WAIT_TICKS) {
      __hip_atomic_store(err, FCP_ERR_SWEEP, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      v = B::qnan;
    }
  }
  return B::from(v);
}

template <typename T>
__device__ __forceinline__ void publish(T *p, T v) {
  using B = Bits<T>;
  typename B::U u = B::to(v);
  if (u == B::sent) u = B::qnan;
  __hip_atomic_store(reinterpret_cast<typename B::U *>(p), u, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// raw buffer accesses: 32-bit byte offsets from one SGPR resource; an offset
// at or beyond the resource's size reads 0 and issues no memory request, so
// lanes that need nothing can take part in a load without diverging
// ---------------------------------------------------------------------------
constexpr int CPOL_SC1 = 16;        // cache-policy bit of agent-coherent accesses
constexpr int FCP_OOB = 0x7FFFFF00;  // an offset beyond every resource (bytes < FCP_OOB)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void *p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, bytes, 0x00020000);
}

template <int BYTES>
struct RawVec;
template <>
struct RawVec<4> {
  typedef unsigned int U;
  template <int CP>
  static __device__ __forceinline__ U load(__amdgpu_buffer_rsrc_t rs, int off) {
    return __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, CP);
  }
  static __device__ __forceinline__ void store(U v, __amdgpu_buffer_rsrc_t rs, int off) {
    __builtin_amdgcn_raw_buffer_store_b32(v, rs, off, 0, CPOL_SC1);
  }
};
template <>
struct RawVec<8> {
  typedef unsigned int U __attribute__((ext_vector_type(2)));
  template <int CP>
  static __device__ __forceinline__ U load(__amdgpu_buffer_rsrc_t rs, int off) {
    return __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, CP);
  }
  static __device__ __forceinline__ void store(U v, __amdgpu_buffer_rsrc_t rs, int off) {
    __builtin_amdgcn_raw_buffer_store_b64(v, rs, off, 0, CPOL_SC1);
  }
};
template <>
struct RawVec<16> {
  typedef unsigned int U __attribute__((ext_vector_type(4)));
  template <int CP>
  static __device__ __forceinline__ U load(__amdgpu_buffer_rsrc_t rs, int off) {
    return __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, CP);
  }
  static __device__ __forceinline__ void store(U v, __amdgpu_buffer_rsrc_t rs, int off) {
    __builtin_amdgcn_raw_buffer_store_b128(v, rs, off, 0, CPOL_SC1);
  }
};

// ---------------------------------------------------------------------------
// SELL-64 copy of the operator
// ---------------------------------------------------------------------------
// Slice s holds rows 64s..64s+63; entry k of its row r sits at
// base[s] + 64k + (r & 63), for k below the slice width (base[s+1]-base[s])/64.
// A row's entries fill k = 0.. in CSR order; the slots after them hold the
// column -1 and the value 0, so a product finds the row's end from the
// columns it loads anyway.  Where padding every slice to the widest costs
// little, all slices have one width W (base[s] = 64 W s) and the products
// compute a row's place instead of loading it.
template <typename T>
struct Sell {
  const int *base, *col, *ia;
  const T *a;
  // batched products with a diagonal of their own per column: [n*NRP]
  // interleaved, used in place of the diagonal entry (null: the matrix's)
  const T *dcol = nullptr;
  int width = 0;  // the uniform slice width, 0: widths differ (read base)
};

// Entry 0 of row i of a SELL-64 array (.x) and its slice width (.y).
__device__ __forceinline__ int2 sell_at(const int *__restrict__ base, int W, int i) {
  const int sl = i >> 6;
  if (W > 0) return make_int2(sl * 64 * W + (i & 63), W);
  const int b0 = base[sl];
  return make_int2(b0 + (i & 63), (base[sl + 1] - b0) >> 6);
}

// Entries of a row whose loads are issued together (rows longer than this
// finish in a plain loop).
constexpr int SELL_KU = 8;

// Grid-stride trips taken GS_U at a time by the streaming kernels: the loads
// of GS_U trips are issued before their first use, the trips themselves stay
// in the plain loop's order (i, i + stride, ...), so per-thread sums do not
// change.
constexpr int GS_U = 4;

// sum_k a_ik x_k of row i in CSR entry order (the reference's row loop,
// linear_solvers.f90:493-498); loads of a wave's 64 rows are contiguous.
// The first SELL_KU entries of the slice (padding: a = 0, col = row) are
// loaded before any product, so a row costs two memory latencies.
template <typename T>
__device__ __forceinline__ T sell_row(const Sell<T> &m, int i, const T *__restrict__ x, T init,
                                      bool subtract) {
  const int2 at = sell_at(m.base, m.width, i);
  const int p = at.x, L = at.y;  // slice width: k < L exists for all its rows
  T av[SELL_KU], xv[SELL_KU];
  int cv[SELL_KU];
#pragma unroll
  for (int k = 0; k < SELL_KU; ++k) {
    cv[k] = -1;
    if (k < L) {
      av[k] = m.a[p + 64 * k];
      cv[k] = m.col[p + 64 * k];
    }
  }
#pragma unroll
  for (int k = 0; k < SELL_KU; ++k)
    if (cv[k] >= 0) xv[k] = x[cv[k]];
  T acc = init;
#pragma unroll
  for (int k = 0; k < SELL_KU; ++k)
    if (cv[k] >= 0) acc = subtract ? acc - av[k] * xv[k] : acc + av[k] * xv[k];
  for (int k = SELL_KU; k < L; ++k) {
    const int j = m.col[p + 64 * k];
    if (j < 0) break;
    acc = subtract ? acc - m.a[p + 64 * k] * x[j] : acc + m.a[p + 64 * k] * x[j];
  }
  return acc;
}

// ---------------------------------------------------------------------------
// host helpers (fcp_krylov.hip)
// ---------------------------------------------------------------------------
int red_grid(fcp_ctx *ctx, int n);
int sell_view64(fcp_ctx *ctx, Sell<double> *out);
int sell_view32(fcp_ctx *ctx, Sell<float> *out);
inline int *tickets(fcp_ctx *ctx) { return ctx->counters.p; }
inline int *redctr(fcp_ctx *ctx) { return ctx->counters.p + 32; }
int check_err(fcp_ctx *ctx);
// Copies bytes of device scalars d to host memory h and the sweep error flag
// to its pinned mirror, waits once for both, and fails on a recorded error.
int poll_device(fcp_ctx *ctx, void *h, const void *d, size_t bytes);
int prepare_precond(fcp_ctx *ctx, int kind, bool fp32, bool pivots = true);
bool sched_fits(fcp_ctx *ctx, const BlockSchedule &lo, const BlockSchedule &up, size_t row_bytes);
int pack_block64(fcp_ctx *ctx, BlockSchedule &b);
int pack_block32(fcp_ctx *ctx, BlockSchedule &b);
// Pivots of the tile-local preconditioner (schedules
// jlower/jupper, pivots dbj); *ok false when it does not fit this mesh.
int bj_begin(fcp_ctx *ctx, int kind, bool *ok);
// Packs the row-gather tile sweeps' FP32 operands if the tile pivots changed.
int bj_pack32(fcp_ctx *ctx);
// Inner PCG in tile order (fcp_inner.hip): FP32, or FP64 with fp64 set
// (values per position of sizeof(double) in LDS).
bool tin_usable(fcp_ctx *ctx, int nrp, bool fp64 = false);
int tin_pcg(fcp_ctx *ctx, int nr, const double *res, double *x, const SolveScalars *so,
            SolveScalars *si, double tol_rel, double inner_tol, int budget, int check_every,
            int verbose, SolveScalars *h, bool fp64 = false);
int tin_bicgstab(fcp_ctx *ctx, int nr, const double *res, double *x, const SolveScalars *so,
                 SolveScalars *si, double tol_rel, double inner_tol, int budget, int check_every,
                 int verbose, SolveScalars *h);
int tin_apply(fcp_ctx *ctx, const double *r, double *z, bool coarse, bool fp64 = false);
int tin_coarse_begin(fcp_ctx *ctx);
int tin_coarse_join(fcp_ctx *ctx);
int precond_apply64(fcp_ctx *ctx, const double *r, double *y, double *z);
// DIC (kind 0) / DILU (kind 1) pivots of the matrix values a (the resident
// pattern) into d, as the solves compute them.
int precond_pivots64(fcp_ctx *ctx, int kind, const double *a, double *d);
int precond_apply32(fcp_ctx *ctx, const float *r, float *y, float *z);

}  // namespace fcpk
