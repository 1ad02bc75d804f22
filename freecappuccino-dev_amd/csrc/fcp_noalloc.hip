// fcp_noalloc.hip -- the Jacobi-preconditioned CG of the explicit-array
// entry point fcp_dpcg (dpcg, linear_solvers.f90:206-359) for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).
//
// The caller passes device arrays only (the matrix in the reference's 1-based
// CSR form, the solution and the right-hand side); every work vector, the
// block-reduction partials, the arrival counters and the solve scalars live
// in one file-scope device array, `scratch`, of fixed size: systems of up to
// FCP_NOALLOC_MAX_ROWS rows.  Host code never uses the array's name as a
// device address: its address on each device is taken once, with that device
// current, by hipGetSymbolAddress and kept per device.
//
// Every call on a device shares that device's array and leaves its contents
// behind; calls that overlap on one device (two host threads, or two streams
// without an ordering between them) race on it.  Nothing here calls hipMalloc,
// hipFree or an allocating library routine.
#include <mutex>

#include "fcp_device.h"

using namespace fcpk;

namespace {

constexpr int NA_MAXN = FCP_NOALLOC_MAX_ROWS;
constexpr int NA_GRID = 1024;  // blocks of every kernel (<= MAXBLK partials per slot)
constexpr int NA_MAXDEV = 64;

// byte layout of the scratch array: r, z (z = r / diag, then q = A p), p,
// two partial slots, the scalars, the arrival counters, the input-check flag
constexpr size_t NA_VEC = (size_t)NA_MAXN * sizeof(double);
constexpr size_t OFF_R = 0;
constexpr size_t OFF_Z = OFF_R + NA_VEC;
constexpr size_t OFF_P = OFF_Z + NA_VEC;
constexpr size_t OFF_PART = OFF_P + NA_VEC;
constexpr size_t OFF_SCAL = OFF_PART + (size_t)2 * MAXBLK * sizeof(double);
constexpr size_t OFF_CTR = OFF_SCAL + 256;
constexpr size_t OFF_FLAG = OFF_CTR + (size_t)(RED_GROUPS + 1) * 32 * sizeof(int);
constexpr size_t NA_BYTES = OFF_FLAG + 256;
static_assert(sizeof(SolveScalars) <= 256, "scalars fit their slot");
static_assert(NA_GRID <= MAXBLK, "partials per slot");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

struct NaBufs {
  double *r, *z, *p, *part;
  SolveScalars *s;
  int *ctr, *flag;
};

NaBufs carve(unsigned char *base) {
  NaBufs B;
  B.r = reinterpret_cast<double *>(base + OFF_R);
  B.z = reinterpret_cast<double *>(base + OFF_Z);
  B.p = reinterpret_cast<double *>(base + OFF_P);
  B.part = reinterpret_cast<double *>(base + OFF_PART);
  B.s = reinterpret_cast<SolveScalars *>(base + OFF_SCAL);
  B.ctr = reinterpret_cast<int *>(base + OFF_CTR);
  B.flag = reinterpret_cast<int *>(base + OFF_FLAG);
  return B;
}

// Scalars, arrival counters and the check flag start from zero on every call
// (a fill kernel, so that the sequence stays capturable).
__global__ void k_na_clear(NaBufs B) {
  const int nctr = (RED_GROUPS + 1) * 32;
  for (int i = threadIdx.x; i < nctr; i += blockDim.x) B.ctr[i] = 0;
  int *s = reinterpret_cast<int *>(B.s);
  for (int i = threadIdx.x; i < (int)(sizeof(SolveScalars) / sizeof(int)); i += blockDim.x) s[i] = 0;
  if (threadIdx.x == 0) *B.flag = 0;
}

// The 1-based CSR arrays describe n rows and nnz entries: ia starts at 1,
// rises and ends at nnz + 1, every column lies in 1..n, every diag(i) points
// into row i at column i.  Reads stay inside [0, nnz) whatever ia holds.
__global__ void k_na_check(int n, int nnz, const int *__restrict__ ia, const int *__restrict__ ja,
                           const int *__restrict__ diag, NaBufs B) {
  int bad = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) bad |= (ia[0] != 1) | (ia[n] != nnz + 1);
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    const int lo = ia[i] - 1, hi = ia[i + 1] - 1;
    if (lo < 0 || hi < lo || hi > nnz) {
      bad = 1;
      continue;
    }
    for (int k = lo; k < hi; ++k) {
      const int j = ja[k];
      bad |= (j < 1) | (j > n);
    }
    const int d = diag[i] - 1;
    if (d < lo || d >= hi)
      bad = 1;
    else
      bad |= ja[d] != i + 1;
  }
  if (bad) atomicOr(B.flag, 1);
}

// r = b - A x in the reference's row order, p = 0, |r|_1 (:253-260).
__global__ __launch_bounds__(BLK) void k_na_init(int n, const int *__restrict__ ia,
                                                 const int *__restrict__ ja,
                                                 const double *__restrict__ a,
                                                 const double *__restrict__ x,
                                                 const double *__restrict__ b, NaBufs B,
                                                 double tol_abs) {
  double part[1] = {0.0};
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    double r = b[i];
    for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k) r = r - a[k] * x[ja[k] - 1];
    B.r[i] = r;
    B.p[i] = 0.0;
    part[0] += fabs(r);
  }
  double blk[1], tot[1];
  block_sum<1>(part, blk);
  if (last_block_reduce<1>(blk, B.part, 0, B.ctr, tot) && threadIdx.x == 0) {
    SolveScalars &S = *B.s;
    S.res0 = tot[0];
    S.resl = tot[0];
    S.s0 = FCP_S0INIT;
    if (tot[0] < tol_abs) {  // :264-268
      S.resor = tot[0];
      S.conv = 1;
      S.done = 1;
    }
    if (!isfinite(tot[0])) {
      S.status = FCP_ERR_NONFINITE;
      S.done = 1;
    }
  }
}

// z = r / a_ii (no +small, :288-290), r.z, beta = r.z / s0.
__global__ __launch_bounds__(BLK) void k_na_z(int n, const int *__restrict__ diag,
                                              const double *__restrict__ a, NaBufs B) {
  if (solve_done(B.s)) return;
  double part[1] = {0.0};
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    const double z = B.r[i] / a[diag[i] - 1];
    B.z[i] = z;
    part[0] += B.r[i] * z;
  }
  double blk[1], tot[1];
  block_sum<1>(part, blk);
  if (last_block_reduce<1>(blk, B.part, 0, B.ctr, tot) && threadIdx.x == 0) {
    SolveScalars &S = *B.s;
    S.sk = tot[0];
    S.bet = S.sk / S.s0;
  }
}

// p = z + beta p
__global__ __launch_bounds__(BLK) void k_na_p(int n, NaBufs B) {
  if (solve_done(B.s)) return;
  const double bet = B.s->bet;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK)
    B.p[i] = B.z[i] + bet * B.p[i];
}

// q = A p (into z), p.q, alpha = r.z / p.q
__global__ __launch_bounds__(BLK) void k_na_q(int n, const int *__restrict__ ia,
                                              const int *__restrict__ ja,
                                              const double *__restrict__ a, NaBufs B) {
  if (solve_done(B.s)) return;
  double part[1] = {0.0};
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    double q = 0.0;
    for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k) q = q + a[k] * B.p[ja[k] - 1];
    B.z[i] = q;
    part[0] += B.p[i] * q;
  }
  double blk[1], tot[1];
  block_sum<1>(part, blk);
  if (last_block_reduce<1>(blk, B.part, 0, B.ctr, tot) && threadIdx.x == 0) {
    SolveScalars &S = *B.s;
    S.pkapk = tot[0];
    S.alf = S.sk / S.pkapk;
  }
}

// x += alpha p, r -= alpha q, |r|_1; after the first iteration the
// normalisation factor sum |a_ii x_i| + small and resor = res0 / factor;
// the reference's L1 stopping test (:318-340).
__global__ __launch_bounds__(BLK) void k_na_xr(int n, const int *__restrict__ diag,
                                               const double *__restrict__ a, double *__restrict__ x,
                                               NaBufs B, double tol_abs, double tol_rel,
                                               int itr_max) {
  if (solve_done(B.s)) return;
  const double alf = B.s->alf;
  const bool first = B.s->iter == 0;
  double part[2] = {0.0, 0.0};
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    const double xi = x[i] + alf * B.p[i];
    x[i] = xi;
    const double r = B.r[i] - alf * B.z[i];
    B.r[i] = r;
    part[0] += fabs(r);
    if (first) part[1] += fabs(a[diag[i] - 1] * xi);
  }
  double blk[2], tot[2];
  block_sum<2>(part, blk);
  if (last_block_reduce<2>(blk, B.part, 0, B.ctr, tot) && threadIdx.x == 0) {
    SolveScalars &S = *B.s;
    const double resl = tot[0];
    S.resl = resl;
    S.s0 = S.sk;
    S.iter += 1;
    if (first) {
      S.factor = tot[1] + FCP_SMALL;
      S.resor = S.res0 / S.factor;
    }
    S.rsm = resl / (S.res0 + FCP_SMALL);
    bool conv = (S.rsm < tol_rel) || (resl < tol_abs);
    if (!isfinite(resl)) {
      S.status = FCP_ERR_NONFINITE;
      conv = true;
    }
    S.conv = (conv && S.status == 0) ? 1 : 0;
    if (conv || S.iter >= itr_max) S.done = 1;
  }
}

std::mutex g_na_mu;
unsigned char *g_na_ptr[NA_MAXDEV];
thread_local char g_na_err[256];

int na_fail(int code, const char *msg) {
  snprintf(g_na_err, sizeof g_na_err, "%s", msg);
  return code;
}

// Device address of `scratch` on the current device, taken once per device.
int na_scratch(int dev, unsigned char **out) {
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return na_fail(FCP_ERR_HIP, "hipGetSymbolAddress of the scratch array failed");
    }
    g_na_ptr[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_na_ptr[dev];
  return 0;
}

// p is device memory on device `dev` and holds at least `bytes` bytes from p on.
bool na_device_span(const void *p, size_t bytes, int dev) {
  if (!p) return false;
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  if (attr.type != hipMemoryTypeDevice || attr.device != dev) return false;
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  const char *b = static_cast<const char *>(base);
  const char *q = static_cast<const char *>(p);
  return q >= b && (size_t)(q - b) + bytes <= size;
}

#define NA_HIP(x)                                                          \
  do {                                                                     \
    const hipError_t e_ = (x);                                             \
    if (e_ != hipSuccess) return na_fail(FCP_ERR_HIP, hipGetErrorString(e_)); \
  } while (0)

}  // namespace

namespace fcpk {
// Message of the last failed call of the entry point below on this thread
// (read through fcp_last_error(NULL)).
const char *noalloc_message() { return g_na_err; }
}  // namespace fcpk

extern "C" {

int fcp_dpcg_noalloc(int n, int nnz, const int *ia, const int *ja, const double *a,
                     const int *diag, double *fi, const double *rhs, double *resor, int itr_max,
                     double tol_abs, double tol_rel, void *stream) {
  g_na_err[0] = '\0';
  if (n <= 0 || nnz <= 0 || itr_max < 0)
    return na_fail(FCP_ERR_ARG, "fcp_dpcg_noalloc: n, nnz must be positive, itr_max >= 0");
  if (n > NA_MAXN)
    return na_fail(FCP_ERR_ARG,
                   "fcp_dpcg_noalloc: n exceeds FCP_NOALLOC_MAX_ROWS (the fixed scratch size)");
  int dev = 0;
  NA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= NA_MAXDEV) return na_fail(FCP_ERR_ARG, "fcp_dpcg_noalloc: device id");
  if (!na_device_span(ia, sizeof(int) * ((size_t)n + 1), dev) ||
      !na_device_span(ja, sizeof(int) * (size_t)nnz, dev) ||
      !na_device_span(a, sizeof(double) * (size_t)nnz, dev) ||
      !na_device_span(diag, sizeof(int) * (size_t)n, dev) ||
      !na_device_span(fi, sizeof(double) * (size_t)n, dev) ||
      !na_device_span(rhs, sizeof(double) * (size_t)n, dev))
    return na_fail(FCP_ERR_ARG,
                   "fcp_dpcg_noalloc: ia, ja, a, diag, fi, rhs must be device arrays of the "
                   "current device, at least n+1, nnz, nnz, n, n, n long");
  unsigned char *base = nullptr;
  const int rc0 = na_scratch(dev, &base);
  if (rc0) return rc0;
  const NaBufs B = carve(base);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int g = (n + BLK - 1) / BLK < NA_GRID ? (n + BLK - 1) / BLK : NA_GRID;
  k_na_clear<<<1, BLK, 0, st>>>(B);
  k_na_check<<<g, BLK, 0, st>>>(n, nnz, ia, ja, diag, B);
  NA_HIP(hipGetLastError());
  int flag = 0;
  NA_HIP(hipMemcpyAsync(&flag, B.flag, sizeof(int), hipMemcpyDeviceToHost, st));
  NA_HIP(hipStreamSynchronize(st));
  if (flag)
    return na_fail(FCP_ERR_ARG, "fcp_dpcg_noalloc: ia/ja/diag are not a 1-based CSR matrix of "
                                "n rows and nnz entries with diag(i) at column i");
  k_na_init<<<g, BLK, 0, st>>>(n, ia, ja, a, fi, rhs, B, tol_abs);
  NA_HIP(hipGetLastError());
  SolveScalars h;
  BatchSizer bs(8, BATCH_L1, tol_rel);
  int launched = 0;
  for (;;) {
    NA_HIP(hipMemcpyAsync(&h, B.s, sizeof h, hipMemcpyDeviceToHost, st));
    NA_HIP(hipStreamSynchronize(st));
    if (h.done || launched >= itr_max) break;
    const int next = bs.next(&h, 1);
    const int batch = (itr_max - launched) < next ? (itr_max - launched) : next;
    for (int k = 0; k < batch; ++k) {
      k_na_z<<<g, BLK, 0, st>>>(n, diag, a, B);
      k_na_p<<<g, BLK, 0, st>>>(n, B);
      k_na_q<<<g, BLK, 0, st>>>(n, ia, ja, a, B);
      k_na_xr<<<g, BLK, 0, st>>>(n, diag, a, fi, B, tol_abs, tol_rel, itr_max);
    }
    NA_HIP(hipGetLastError());
    launched += batch;
  }
  if (resor) *resor = h.resor;
  if (h.status < 0) return na_fail(h.status, "fcp_dpcg_noalloc: non-finite residual");
  if (h.conv || h.res0 < tol_abs) return 0;
  return h.iter > 0 ? h.iter : 1;
}

}  // extern "C"
