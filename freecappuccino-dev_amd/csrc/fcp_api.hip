// fcp_api.hip -- extern "C" entry points of libfcp (see include/fcp.h).
#include <hip/hip_runtime.h>

#include <cctype>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <memory>
#include <map>
#include <mutex>
#include <tuple>

#include "fcp_internal.h"

// ---------------------------------------------------------------------------
// errors, pointer kinds, staging
// ---------------------------------------------------------------------------
int fcp_fail(fcp_ctx *ctx, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

bool fcp_is_device_ptr(const void *p) {
  if (!p) return false;
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

namespace {

// An input/output array that is either used in place (device pointer) or
// staged through a device buffer (host pointer).
template <typename T>
struct Arg {
  T *dev = nullptr;
  T *host = nullptr;
  size_t n = 0;
  DBuf<T> own;
};

template <typename T>
int arg_in(fcp_ctx *ctx, const T *p, size_t n, Arg<T> &a) {
  a.n = n;
  if (!p) return 0;
  if (fcp_is_device_ptr(p)) {
    a.dev = const_cast<T *>(p);
    return 0;
  }
  a.host = const_cast<T *>(p);
  FCP_HIP(ctx, a.own.alloc(n > 0 ? n : 1));
  a.dev = a.own.p;
  if (n) FCP_HIP(ctx, hipMemcpyAsync(a.dev, p, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return 0;
}

template <typename T>
int arg_out(fcp_ctx *ctx, T *p, size_t n, Arg<T> &a) {
  a.n = n;
  if (!p) return 0;
  if (fcp_is_device_ptr(p)) {
    a.dev = p;
    return 0;
  }
  a.host = p;
  FCP_HIP(ctx, a.own.alloc(n > 0 ? n : 1));
  a.dev = a.own.p;
  return 0;
}

template <typename T>
int arg_back(fcp_ctx *ctx, Arg<T> &a) {
  if (a.host && a.n)
    FCP_HIP(ctx, hipMemcpyAsync(a.host, a.dev, a.n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
  return 0;
}

__global__ void k_sub1(const int *__restrict__ in, int *__restrict__ out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] - 1;
}
__global__ void k_add1(const int *__restrict__ in, int *__restrict__ out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] + 1;
}
__global__ void k_node_sub1(const int *__restrict__ in, const int *__restrict__ nn,
                            int *__restrict__ out, int F) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  for (int k = 0; k < FCP_NOMAX; ++k) {
    const int v = in[(size_t)f * FCP_NOMAX + k];
    out[(size_t)f * FCP_NOMAX + k] = (k < nn[f]) ? v - 1 : 0;
  }
}
__global__ void k_bpatch(const int *__restrict__ start, const int *__restrict__ nf, int np, int Fi,
                         int *__restrict__ bpatch) {
  int ib = blockIdx.x;
  if (ib >= np) return;
  for (int i = threadIdx.x; i < nf[ib]; i += blockDim.x) bpatch[start[ib] - Fi + i] = ib;
}

// Bounds checks of caller-supplied connectivity (1-based), run on the device
// before any kernel indexes with it; violations only set a flag.
__global__ void k_check_mesh(const int *__restrict__ nn, const int *__restrict__ node,
                             const int *__restrict__ owner, const int *__restrict__ neighbour,
                             int F, int Fi, int nNodes, int n, int *__restrict__ bad) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const int m = nn[f];
  bool ok = (m >= 3 && m <= FCP_NOMAX);
  if (ok)
    for (int k = 0; k < m; ++k) {
      const int v = node[(size_t)f * FCP_NOMAX + k];
      ok = ok && v >= 1 && v <= nNodes;
    }
  ok = ok && owner[f] >= 1 && owner[f] <= n;
  if (f < Fi) ok = ok && neighbour[f] >= 1 && neighbour[f] <= n && neighbour[f] != owner[f];
  if (!ok) atomicOr(bad, 1);
}

bool csr_ok_host(int n, int nnz, const int *ia, const int *ja, const int *diag) {
  if (ia[0] != 1 || ia[n] != nnz + 1) return false;
  for (int i = 0; i < n; ++i) {
    if (ia[i + 1] < ia[i]) return false;
    const int d = diag[i];
    if (d < ia[i] || d >= ia[i + 1]) return false;
    if (ja[d - 1] != i + 1) return false;
    // columns strictly ascending inside a row: the strict lower part is
    // [ia(i), diag(i)) as the reference's sweeps and ilu_cr assume
    for (int k = ia[i]; k + 1 < ia[i + 1]; ++k)
      if (ja[k - 1] >= ja[k]) return false;
  }
  for (int k = 0; k < nnz; ++k)
    if (ja[k] < 1 || ja[k] > n) return false;
  return true;
}

int out_index_plus1(fcp_ctx *ctx, const int *dev0, int *dst, size_t n) {
  if (!dst || n == 0) return 0;
  Arg<int> o;
  FCP_TRY(arg_out(ctx, dst, n, o));
  k_add1<<<grid_for((long long)n, 256), 256, 0, ctx->stream>>>(dev0, o.dev, n);
  FCP_HIP(ctx, hipGetLastError());
  FCP_TRY(arg_back(ctx, o));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int out_copy(fcp_ctx *ctx, const double *dev, double *dst, size_t n) {
  if (!dst || n == 0) return 0;
  hipMemcpyKind kind = fcp_is_device_ptr(dst) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  FCP_HIP(ctx, hipMemcpyAsync(dst, dev, n * sizeof(double), kind, ctx->stream));
  return 0;
}

int need(fcp_ctx *ctx, bool mesh, bool geom, bool csr) {
  if (!ctx) return FCP_ERR_ARG;
  if (mesh && !ctx->have_mesh) return fcp_fail(ctx, FCP_ERR_STATE, "no mesh uploaded");
  if (geom && !ctx->have_geom) return fcp_fail(ctx, FCP_ERR_STATE, "geometry not computed");
  if (csr && !ctx->have_csr) return fcp_fail(ctx, FCP_ERR_STATE, "CSR pattern not built");
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// profiling
// ---------------------------------------------------------------------------
void fcp_prof_begin(fcp_ctx *ctx, const char *, hipEvent_t *ev) {
  *ev = nullptr;
  if (!ctx->prof) return;
  hipEvent_t e;
  if (!ctx->ev_pool.empty()) {
    e = ctx->ev_pool.back();
    ctx->ev_pool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    return;
  }
  (void)hipEventRecord(e, ctx->stream);
  *ev = e;
}

static void prof_flush(fcp_ctx *ctx) {
  if (ctx->prof_pending.empty()) return;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &p : ctx->prof_pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      auto &e = ctx->prof_tab[p.name];
      e.launches++;
      e.total_ms += ms;
    }
    ctx->ev_pool.push_back(p.a);
    ctx->ev_pool.push_back(p.b);
  }
  ctx->prof_pending.clear();
}

void fcp_prof_end(fcp_ctx *ctx, const char *name, hipEvent_t ev) {
  if (!ctx->prof || !ev) return;
  hipEvent_t e;
  if (!ctx->ev_pool.empty()) {
    e = ctx->ev_pool.back();
    ctx->ev_pool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    return;
  }
  (void)hipEventRecord(e, ctx->stream);
  ctx->prof_pending.push_back({name, ev, e});
  if (ctx->prof_pending.size() > 4096) prof_flush(ctx);
}

extern "C" {

int fcp_profile(fcp_ctx *ctx, int enable) {
  if (!ctx) return FCP_ERR_ARG;
  prof_flush(ctx);
  ctx->prof = enable != 0;
  if (enable) ctx->prof_tab.clear();
  return 0;
}

int fcp_profile_report(fcp_ctx *ctx, char *buf, int len) {
  if (!ctx || !buf || len <= 0) return FCP_ERR_ARG;
  prof_flush(ctx);
  std::string s;
  char line[256];
  for (auto &kv : ctx->prof_tab) {
    snprintf(line, sizeof line, "%s %d %.6f\n", kv.first.c_str(), kv.second.launches,
             kv.second.total_ms);
    s += line;
  }
  snprintf(buf, (size_t)len, "%s", s.c_str());
  return (int)s.size();
}

int fcp_tune_set(fcp_ctx *ctx, int param, int value) {
  if (!ctx) return FCP_ERR_ARG;
  switch (param) {
    case FCP_TUNE_SWEEP_WAVES:
      if (value < 0 || value > (1 << 20)) return fcp_fail(ctx, FCP_ERR_ARG, "sweep waves out of range");
      ctx->tune_sweep_waves = value;
      return 0;
    case FCP_TUNE_SWEEP_SLEEP:
      if (value < 0 || value > 127) return fcp_fail(ctx, FCP_ERR_ARG, "sleep out of range");
      ctx->tune_sweep_sleep = value;
      return 0;
    case FCP_TUNE_SWEEP_KIND:
      if (value != 0 && value != 1) return fcp_fail(ctx, FCP_ERR_ARG, "sweep kind is 0 or 1");
      ctx->tune_sweep_kind = value;
      return 0;
    case FCP_TUNE_SWEEP_TILE:
      if (value != 0 && value != 1 && (value < 64 || value > 32768))
        return fcp_fail(ctx, FCP_ERR_ARG, "tile rows must be 0, 1 or in [64, 32768]");
      if (value != ctx->tune_sweep_tile) {
        ctx->tune_sweep_tile = value;
        if (ctx->have_csr) return fcp_build_tile_schedules(ctx);
      }
      return 0;
    case FCP_TUNE_SPMV:
      if (value < -1 || value > 4) return fcp_fail(ctx, FCP_ERR_ARG, "SpMV variant is -1..4");
      ctx->tune_spmv = value;
      return 0;
    case FCP_TUNE_TILE_PACK:
      if (value != 0 && value != 1) return fcp_fail(ctx, FCP_ERR_ARG, "tile packing is 0 or 1");
      if (value != ctx->tune_tile_pack) {
        FCP_TRY(fcp_coarse_wait(ctx));
        ctx->tune_tile_pack = value;
        ctx->tin_have = false;  // operands rebuilt in the chosen layout
      }
      return 0;
    case FCP_TUNE_ASM_BLOCKS:
      if (value != 0 && value != 1) return fcp_fail(ctx, FCP_ERR_ARG, "assembly blocks is 0 or 1");
      ctx->tune_asm_blk = value;
      return 0;
    case FCP_TUNE_COARSE_FUSE:
      if (value != 0 && value != 1) return fcp_fail(ctx, FCP_ERR_ARG, "coarse fusion is 0 or 1");
      ctx->tune_coarse_fuse = value;
      return 0;
    case FCP_TUNE_GRAPH:
      if (value != 0 && value != 1) return fcp_fail(ctx, FCP_ERR_ARG, "graph replay is 0 or 1");
      ctx->tune_graph = value;
      return 0;
    case FCP_TUNE_INNER_TILE:
      if (value != 0 && (value < 64 || value > 32768))
        return fcp_fail(ctx, FCP_ERR_ARG, "inner tile rows must be 0 or in [64, 32768]");
      if (value != ctx->tune_inner_tile) {
        ctx->tune_inner_tile = value;
        ctx->have_bj = false;
        ctx->tin_have = false;
        ctx->tc_have = false;
      }
      return 0;
    case FCP_TUNE_INNER_COARSE:
      if (value != 0 && value != 1) return fcp_fail(ctx, FCP_ERR_ARG, "coarse correction is 0 or 1");
      ctx->inner_coarse = value;
      return 0;
    case FCP_TUNE_INNER_PRECOND:
      if (value < 0 || value > 2)
        return fcp_fail(ctx, FCP_ERR_ARG, "inner preconditioner must be 0, 1 or 2");
      ctx->inner_precond = value;
      return 0;
    case FCP_TUNE_SWEEP_ROWS:
      if (value != 0 && (value < 64 || value > 32768 || (value & 63)))
        return fcp_fail(ctx, FCP_ERR_ARG, "sweep rows must be 0 or a multiple of 64 in [64, 32768]");
      if (value != ctx->tune_sweep_rows) {
        ctx->tune_sweep_rows = value;
        if (ctx->have_csr) return fcp_build_block_schedules(ctx);
      }
      return 0;
  }
  return fcp_fail(ctx, FCP_ERR_ARG, "unknown tuning parameter %d", param);
}

void *fcp_stream(fcp_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
int fcp_init(int device, fcp_ctx **out) {
  if (!out) return FCP_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return FCP_ERR_HIP;
  if (device < 0 || device >= ndev) return FCP_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return FCP_ERR_HIP;
  fcp_ctx *ctx = new fcp_ctx();
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    ctx->ncu = prop.multiProcessorCount;
    if (prop.sharedMemPerBlock > 0) ctx->lds_cap = prop.sharedMemPerBlock;
  }
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return FCP_ERR_HIP;
  }
  *out = ctx;
  return 0;
}

void fcp_finalize(fcp_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  prof_flush(ctx);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->tc_stream) {
    (void)hipStreamSynchronize(ctx->tc_stream);
    (void)hipStreamDestroy(ctx->tc_stream);
    (void)hipEventDestroy(ctx->tc_ready);
    (void)hipEventDestroy(ctx->tc_done);
  }
  for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
  if (ctx->h_scal) (void)hipHostFree(ctx->h_scal);
  if (ctx->h_err) (void)hipHostFree(ctx->h_err);
  if (ctx->h_poll) (void)hipHostFree(ctx->h_poll);
  if (ctx->h_perr) (void)hipHostFree(ctx->h_perr);
  for (auto e : ctx->poll_ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->tin_graph) (void)hipGraphExecDestroy(ctx->tin_graph);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *fcp_last_error(const fcp_ctx *ctx) {
  if (ctx) return ctx->err.c_str();
  const char *m = fcpk::noalloc_message();
  return m[0] ? m : "null context";
}

void fcp_solve_opts_default(fcp_solve_opts *o) {
  if (!o) return;
  memset(o, 0, sizeof *o);
  o->precision = FCP_PREC_FP64;
  o->criterion = FCP_CRIT_REF_L1;
  o->inner_tol = 1e-4;
  o->max_outer = 30;
  o->check_every = 8;
}

// ---------------------------------------------------------------------------
// mesh / geometry / CSR
// ---------------------------------------------------------------------------
int fcp_mesh_upload(fcp_ctx *ctx, int nNodes, int nCells, int nInnerFaces, int nBoundaryFaces,
                    int nBoundaries, const double *xyz, const int *nnodes, const int *node,
                    const int *owner, const int *neighbour, const int *nfaces,
                    const int *startFace, const int *bctype) {
  if (!ctx) return FCP_ERR_ARG;
  if (nNodes <= 0 || nCells <= 0 || nInnerFaces < 0 || nBoundaryFaces < 0 || nBoundaries < 0 ||
      !xyz || !nnodes || !node || !owner || (nInnerFaces > 0 && !neighbour) ||
      (nBoundaries > 0 && (!nfaces || !startFace || !bctype)))
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_mesh_upload: bad sizes or null arrays");
  (void)hipSetDevice(ctx->device);
  const int F = nInnerFaces + nBoundaryFaces;
  ctx->n = nCells;
  ctx->F = F;
  ctx->Fi = nInnerFaces;
  ctx->Fb = nBoundaryFaces;
  ctx->nNodes = nNodes;
  ctx->nPatches = nBoundaries;
  ctx->h_nfaces.assign(nfaces, nfaces + nBoundaries);
  ctx->h_start.assign(startFace, startFace + nBoundaries);
  ctx->h_bctype.assign(bctype, bctype + nBoundaries);
  ctx->tr_kind.release();  // the transport step's boundary table belongs to the old mesh
  ctx->tr_wf.clear();
  ctx->h_twin.clear();  // periodic twins come with fcp_mesh_periodic
  ctx->nper = 0;
  long long tot = 0;
  for (int ib = 0; ib < nBoundaries; ++ib) {
    if (startFace[ib] < nInnerFaces || startFace[ib] + nfaces[ib] > F || nfaces[ib] < 0)
      return fcp_fail(ctx, FCP_ERR_ARG, "patch %d: faces outside the boundary range", ib);
    if (ib > 0 && startFace[ib] < startFace[ib - 1] + nfaces[ib - 1])
      return fcp_fail(ctx, FCP_ERR_ARG, "patches must be ordered by startFace");
    tot += nfaces[ib];
  }
  if (tot != nBoundaryFaces) return fcp_fail(ctx, FCP_ERR_ARG, "patch faces do not sum to nBoundaryFaces");
  ctx->have_mesh = ctx->have_geom = ctx->have_csr = false;
  FCP_HIP(ctx, ctx->points.alloc((size_t)3 * nNodes));
  FCP_HIP(ctx, ctx->nnodes.alloc(F));
  FCP_HIP(ctx, ctx->node.alloc((size_t)FCP_NOMAX * F));
  FCP_HIP(ctx, ctx->owner.alloc(F));
  FCP_HIP(ctx, ctx->neighbour.alloc(nInnerFaces > 0 ? nInnerFaces : 1));
  FCP_HIP(ctx, ctx->bpatch.alloc(nBoundaryFaces > 0 ? nBoundaryFaces : 1));
  FCP_HIP(ctx, ctx->bctype.alloc(nBoundaries > 0 ? nBoundaries : 1));
  {
    Arg<double> p;
    FCP_TRY(arg_in(ctx, xyz, (size_t)3 * nNodes, p));
    FCP_HIP(ctx, hipMemcpyAsync(ctx->points.p, p.dev, sizeof(double) * 3 * nNodes,
                                hipMemcpyDeviceToDevice, ctx->stream));
    Arg<int> nn, nd, ow, nb;
    FCP_TRY(arg_in(ctx, nnodes, F, nn));
    FCP_TRY(arg_in(ctx, node, (size_t)FCP_NOMAX * F, nd));
    FCP_TRY(arg_in(ctx, owner, F, ow));
    FCP_HIP(ctx, hipMemcpyAsync(ctx->nnodes.p, nn.dev, sizeof(int) * F, hipMemcpyDeviceToDevice,
                                ctx->stream));
    if (nInnerFaces > 0) FCP_TRY(arg_in(ctx, neighbour, nInnerFaces, nb));
    {
      DBuf<int> bad;
      FCP_HIP(ctx, bad.alloc(1));
      FCP_HIP(ctx, hipMemsetAsync(bad.p, 0, sizeof(int), ctx->stream));
      k_check_mesh<<<grid_for(F, 256), 256, 0, ctx->stream>>>(nn.dev, nd.dev, ow.dev, nb.dev, F,
                                                              nInnerFaces, nNodes, nCells, bad.p);
      int hbad = 0;
      FCP_HIP(ctx, hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
      if (hbad)
        return fcp_fail(ctx, FCP_ERR_ARG,
                        "fcp_mesh_upload: face vertex/owner/neighbour index out of range");
    }
    k_node_sub1<<<grid_for(F, 256), 256, 0, ctx->stream>>>(nd.dev, nn.dev, ctx->node.p, F);
    k_sub1<<<grid_for(F, 256), 256, 0, ctx->stream>>>(ow.dev, ctx->owner.p, F);
    if (nInnerFaces > 0) {
      k_sub1<<<grid_for(nInnerFaces, 256), 256, 0, ctx->stream>>>(nb.dev, ctx->neighbour.p,
                                                                   nInnerFaces);
    }
    if (nBoundaries > 0) {
      DBuf<int> st, nf;
      FCP_HIP(ctx, st.alloc(nBoundaries));
      FCP_HIP(ctx, nf.alloc(nBoundaries));
      FCP_HIP(ctx, hipMemcpyAsync(st.p, startFace, sizeof(int) * nBoundaries, hipMemcpyHostToDevice,
                                  ctx->stream));
      FCP_HIP(ctx, hipMemcpyAsync(nf.p, nfaces, sizeof(int) * nBoundaries, hipMemcpyHostToDevice,
                                  ctx->stream));
      FCP_HIP(ctx, hipMemcpyAsync(ctx->bctype.p, bctype, sizeof(int) * nBoundaries,
                                  hipMemcpyHostToDevice, ctx->stream));
      k_bpatch<<<nBoundaries, 256, 0, ctx->stream>>>(st.p, nf.p, nBoundaries, nInnerFaces,
                                                     ctx->bpatch.p);
      FCP_HIP(ctx, hipGetLastError());
      FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    FCP_HIP(ctx, hipGetLastError());
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  FCP_TRY(fcp_build_cell_faces(ctx));
  ctx->have_mesh = true;
  return 0;
}

int fcp_mesh_periodic(fcp_ctx *ctx, int nPeriodic, const int *startFaceTwin) {
  FCP_TRY(need(ctx, false, false, false));
  if (!ctx->have_mesh) return fcp_fail(ctx, FCP_ERR_STATE, "fcp_mesh_periodic: no mesh");
  std::vector<int> per;
  for (int ib = 0; ib < ctx->nPatches; ++ib)
    if (ctx->h_bctype[ib] == FCP_BC_PERIODIC) per.push_back(ib);
  if (nPeriodic != (int)per.size() || (nPeriodic > 0 && !startFaceTwin))
    return fcp_fail(ctx, FCP_ERR_ARG,
                    "fcp_mesh_periodic: %d twins given for %d periodic patches", nPeriodic,
                    (int)per.size());
  std::vector<int> used(ctx->nPatches, 0);
  for (int k = 0; k < nPeriodic; ++k) {
    const int ib = per[k];
    int tb = -1;
    for (int jb = 0; jb < ctx->nPatches; ++jb)
      if (ctx->h_start[jb] == startFaceTwin[k]) tb = jb;
    // the twin is a whole patch of the same size, not periodic itself and
    // the twin of no other patch (then the reference's per-patch mappings
    // touch disjoint faces)
    if (tb < 0 || tb == ib || ctx->h_nfaces[tb] != ctx->h_nfaces[ib] ||
        ctx->h_bctype[tb] == FCP_BC_PERIODIC || used[tb]++)
      return fcp_fail(ctx, FCP_ERR_ARG, "fcp_mesh_periodic: bad twin start face %d for patch %d",
                      startFaceTwin[k], ib);
  }
  (void)hipSetDevice(ctx->device);
  FCP_TRY(fcp_coarse_wait(ctx));
  ctx->h_twin.assign(startFaceTwin, startFaceTwin + nPeriodic);
  FCP_HIP(ctx, ctx->owner0.alloc(ctx->F));
  FCP_HIP(ctx, hipMemcpyAsync(ctx->owner0.p, ctx->owner.p, sizeof(int) * ctx->F,
                              hipMemcpyDeviceToDevice, ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->have_geom = ctx->have_csr = false;  // face mapping runs with the geometry
  ctx->tr_kind.release();
  ctx->tr_wf.clear();
  return 0;
}

int fcp_geometry(fcp_ctx *ctx, double *xc, double *yc, double *zc, double *vol, double *arx,
                 double *ary, double *arz, double *xf, double *yf, double *zf, double *facint,
                 double *Df, double *dnb, double *srdb) {
  FCP_TRY(need(ctx, true, false, false));
  (void)hipSetDevice(ctx->device);
  FCP_TRY(fcp_compute_geometry(ctx));
  const size_t n = ctx->n, F = ctx->F, Fi = ctx->Fi, Fb = ctx->Fb;
  FCP_TRY(out_copy(ctx, ctx->xc.p, xc, n));
  FCP_TRY(out_copy(ctx, ctx->yc.p, yc, n));
  FCP_TRY(out_copy(ctx, ctx->zc.p, zc, n));
  FCP_TRY(out_copy(ctx, ctx->vol.p, vol, n));
  FCP_TRY(out_copy(ctx, ctx->arx.p, arx, F));
  FCP_TRY(out_copy(ctx, ctx->ary.p, ary, F));
  FCP_TRY(out_copy(ctx, ctx->arz.p, arz, F));
  FCP_TRY(out_copy(ctx, ctx->xf.p, xf, F));
  FCP_TRY(out_copy(ctx, ctx->yf.p, yf, F));
  FCP_TRY(out_copy(ctx, ctx->zf.p, zf, F));
  FCP_TRY(out_copy(ctx, ctx->facint.p, facint, Fi));
  FCP_TRY(out_copy(ctx, ctx->Df.p, Df, Fi));
  FCP_TRY(out_copy(ctx, ctx->dnb.p, dnb, Fb));
  FCP_TRY(out_copy(ctx, ctx->srdb.p, srdb, Fb));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_csr_build(fcp_ctx *ctx, int *nnz, int *ia, int *ja, int *diag, int *slot_pn, int *slot_np) {
  FCP_TRY(need(ctx, true, false, false));
  (void)hipSetDevice(ctx->device);
  FCP_TRY(fcp_build_csr(ctx));
  if (nnz) *nnz = ctx->nnz;
  FCP_TRY(out_index_plus1(ctx, ctx->ia.p, ia, (size_t)ctx->n + 1));
  FCP_TRY(out_index_plus1(ctx, ctx->ja.p, ja, ctx->nnz));
  FCP_TRY(out_index_plus1(ctx, ctx->diag.p, diag, ctx->n));
  FCP_TRY(out_index_plus1(ctx, ctx->slot_pn.p, slot_pn, (size_t)ctx->Fi + ctx->nper));
  FCP_TRY(out_index_plus1(ctx, ctx->slot_np.p, slot_np, (size_t)ctx->Fi + ctx->nper));
  return 0;
}

int fcp_levels(fcp_ctx *ctx, int upper, int *lev, int *nlev) {
  FCP_TRY(need(ctx, true, false, true));
  const Schedule &s = upper ? ctx->upper : ctx->lower;
  if (nlev) *nlev = s.nlev;
  if (lev) {
    hipMemcpyKind kind = fcp_is_device_ptr(lev) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    FCP_HIP(ctx, hipMemcpyAsync(lev, s.lev.p, sizeof(int) * ctx->n, kind, ctx->stream));
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return 0;
}

// ---------------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------------
int fcp_laplacian(fcp_ctx *ctx, const double *mu, const double *phi, double *a, double *su) {
  FCP_TRY(need(ctx, true, true, true));
  if (!mu || !phi || !su) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_laplacian: null argument");
  (void)hipSetDevice(ctx->device);
  Arg<double> m, p, s;
  FCP_TRY(arg_in(ctx, mu, ctx->n, m));
  FCP_TRY(arg_in(ctx, phi, (size_t)ctx->n + ctx->Fb, p));
  FCP_TRY(arg_in(ctx, (const double *)su, ctx->n, s));
  FCP_TRY(fcp_launch_laplacian(ctx, m.dev, nullptr, nullptr, p.dev, s.dev));
  FCP_TRY(arg_back(ctx, s));
  FCP_TRY(out_copy(ctx, ctx->a.p, a, ctx->nnz));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_laplacian_face(fcp_ctx *ctx, const double *mu_f, const unsigned char *bc_kind,
                       const double *phi, double *a, double *su) {
  FCP_TRY(need(ctx, true, true, true));
  if (!mu_f || !phi || !su) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_laplacian_face: null argument");
  (void)hipSetDevice(ctx->device);
  Arg<double> m, p, s;
  Arg<unsigned char> bk;
  FCP_TRY(arg_in(ctx, mu_f, ctx->F, m));
  FCP_TRY(arg_in(ctx, phi, (size_t)ctx->n + ctx->Fb, p));
  FCP_TRY(arg_in(ctx, (const double *)su, ctx->n, s));
  if (bc_kind) FCP_TRY(arg_in(ctx, bc_kind, ctx->Fb, bk));
  FCP_TRY(fcp_launch_laplacian(ctx, nullptr, m.dev, bk.dev, p.dev, s.dev));
  FCP_TRY(arg_back(ctx, s));
  FCP_TRY(out_copy(ctx, ctx->a.p, a, ctx->nnz));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_nonorth_correction(fcp_ctx *ctx, const double *gamma_f, const double *grad, double *su) {
  FCP_TRY(need(ctx, true, true, false));
  if (!gamma_f || !grad || !su) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_nonorth_correction: null argument");
  (void)hipSetDevice(ctx->device);
  Arg<double> g, gr, s;
  FCP_TRY(arg_in(ctx, gamma_f, ctx->F, g));
  FCP_TRY(arg_in(ctx, grad, (size_t)3 * (ctx->n + ctx->Fb), gr));
  FCP_TRY(arg_in(ctx, (const double *)su, ctx->n, s));
  FCP_TRY(fcp_launch_nonorth(ctx, g.dev, gr.dev, s.dev));
  FCP_TRY(arg_back(ctx, s));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_convdiff(fcp_ctx *ctx, const double *flmass, const double *dcoef, double *phi,
                 const double *grad, double gam, int scheme, const unsigned char *bc_kind,
                 double urf, double *a, double *su, double *sp) {
  FCP_TRY(need(ctx, true, true, true));
  if (!flmass || !dcoef || !phi || !grad || !bc_kind || !su || !sp || (scheme != 0 && scheme != 1) ||
      !(urf > 0.0))
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_convdiff: bad argument");
  (void)hipSetDevice(ctx->device);
  const size_t ntot = (size_t)ctx->n + ctx->Fb;
  Arg<double> fm, dc, ph, gr, s, p;
  Arg<unsigned char> bk;
  FCP_TRY(arg_in(ctx, flmass, ctx->F, fm));
  FCP_TRY(arg_in(ctx, dcoef, ctx->F, dc));
  FCP_TRY(arg_in(ctx, (const double *)phi, ntot, ph));
  FCP_TRY(arg_in(ctx, grad, 3 * ntot, gr));
  FCP_TRY(arg_in(ctx, bc_kind, ctx->Fb, bk));
  FCP_TRY(arg_in(ctx, (const double *)su, ctx->n, s));
  FCP_TRY(arg_in(ctx, (const double *)sp, ctx->n, p));
  FCP_TRY(fcp_launch_convdiff(ctx, fm.dev, dc.dev, ph.dev, gr.dev, gam, scheme, bk.dev, urf, s.dev,
                              p.dev));
  FCP_TRY(arg_back(ctx, ph));
  FCP_TRY(arg_back(ctx, s));
  FCP_TRY(arg_back(ctx, p));
  FCP_TRY(out_copy(ctx, ctx->a.p, a, ctx->nnz));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_convdiff_cell(fcp_ctx *ctx, const double *flmass, const double *diffcoef, double *phi,
                      const double *grad, double gam, int scheme, const unsigned char *bc_kind,
                      double urf, double *a, double *su, double *sp) {
  FCP_TRY(need(ctx, true, true, true));
  if (!diffcoef) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_convdiff_cell: bad argument");
  (void)hipSetDevice(ctx->device);
  Arg<double> dc;
  FCP_TRY(arg_in(ctx, diffcoef, ctx->n, dc));
  FCP_HIP(ctx, ctx->cd_gamf.alloc(ctx->F > 0 ? ctx->F : 1));
  FCP_TRY(fcp_launch_face_gamma_cells(ctx, dc.dev, ctx->cd_gamf.p));
  return fcp_convdiff(ctx, flmass, ctx->cd_gamf.p, phi, grad, gam, scheme, bc_kind, urf, a, su,
                      sp);
}

static int pcorr_check(fcp_ctx *ctx, const char *who) {
  for (int ib = 0; ib < ctx->nPatches; ++ib)
    if (ctx->h_bctype[ib] == FCP_BC_PRESSURE)
      return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "%s: pressure patches are not supported", who);
  return 0;
}

int fcp_pressure_assemble(fcp_ctx *ctx, double *u, double *v, double *w, const double *p,
                          const double *dpdxi, const double *den, const double *apu,
                          const double *apv, const double *apw, double *flmass, int const_mflux,
                          double flomas, double *a, double *su) {
  FCP_TRY(need(ctx, true, true, true));
  if (!u || !v || !w || !p || !dpdxi || !den || !apu || !apv || !apw || !flmass || !su)
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_pressure_assemble: null argument");
  FCP_TRY(pcorr_check(ctx, "fcp_pressure_assemble"));
  (void)hipSetDevice(ctx->device);
  const size_t n = ctx->n, nt = n + ctx->Fb;
  Arg<double> U, V, W, P, D, R, AU, AV, AW, FM, S;
  FCP_TRY(arg_in(ctx, (const double *)u, nt, U));
  FCP_TRY(arg_in(ctx, (const double *)v, nt, V));
  FCP_TRY(arg_in(ctx, (const double *)w, nt, W));
  FCP_TRY(arg_in(ctx, p, nt, P));
  FCP_TRY(arg_in(ctx, dpdxi, 3 * nt, D));
  FCP_TRY(arg_in(ctx, den, nt, R));
  FCP_TRY(arg_in(ctx, apu, n, AU));
  FCP_TRY(arg_in(ctx, apv, n, AV));
  FCP_TRY(arg_in(ctx, apw, n, AW));
  FCP_TRY(arg_in(ctx, (const double *)flmass, ctx->F, FM));
  FCP_TRY(arg_out(ctx, su, n, S));
  FCP_TRY(fcp_launch_pcorr_assemble(ctx, U.dev, V.dev, W.dev, P.dev, D.dev, R.dev, AU.dev, AV.dev,
                                    AW.dev, FM.dev, const_mflux, flomas, S.dev));
  if (!const_mflux) {
    FCP_TRY(arg_back(ctx, U));
    FCP_TRY(arg_back(ctx, V));
    FCP_TRY(arg_back(ctx, W));
  }
  FCP_TRY(arg_back(ctx, FM));
  FCP_TRY(arg_back(ctx, S));
  FCP_TRY(out_copy(ctx, ctx->a.p, a, ctx->nnz));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_pressure_correct(fcp_ctx *ctx, double *pp, double *flmass, double *u, double *v,
                         double *w, double *p, double *dpdxi, const double *apu,
                         const double *apv, const double *apw, double urfp, int pref_cell,
                         double *su, double *sv, double *sw) {
  FCP_TRY(need(ctx, true, true, true));
  if (!pp || !flmass || !u || !v || !w || !p || !dpdxi || !apu || !apv || !apw)
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_pressure_correct: null argument");
  if (pref_cell < 1 || pref_cell > ctx->n)
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_pressure_correct: pRefCell %d outside 1..%d",
                    pref_cell, ctx->n);
  FCP_TRY(pcorr_check(ctx, "fcp_pressure_correct"));
  (void)hipSetDevice(ctx->device);
  const size_t n = ctx->n, nt = n + ctx->Fb;
  Arg<double> PP, FM, U, V, W, P, D, AU, AV, AW, SU, SV, SW;
  FCP_TRY(arg_in(ctx, (const double *)pp, nt, PP));
  FCP_TRY(arg_in(ctx, (const double *)flmass, ctx->F, FM));
  FCP_TRY(arg_in(ctx, (const double *)u, nt, U));
  FCP_TRY(arg_in(ctx, (const double *)v, nt, V));
  FCP_TRY(arg_in(ctx, (const double *)w, nt, W));
  FCP_TRY(arg_in(ctx, (const double *)p, nt, P));
  FCP_TRY(arg_in(ctx, (const double *)dpdxi, 3 * nt, D));
  FCP_TRY(arg_in(ctx, apu, n, AU));
  FCP_TRY(arg_in(ctx, apv, n, AV));
  FCP_TRY(arg_in(ctx, apw, n, AW));
  // su, sv, sw: the caller's arrays or scratch of the context
  FCP_HIP(ctx, ctx->pc_s.alloc(3 * n > 0 ? 3 * n : 1));
  double *s3[3] = {su, sv, sw};
  Arg<double> SA[3];
  for (int k = 0; k < 3; ++k) {
    if (s3[k]) {
      FCP_TRY(arg_out(ctx, s3[k], n, SA[k]));
    } else {
      SA[k].dev = ctx->pc_s.p + k * n;
    }
  }
  FCP_TRY(fcp_launch_pcorr_correct(ctx, PP.dev, FM.dev, U.dev, V.dev, W.dev, P.dev, D.dev,
                                   AU.dev, AV.dev, AW.dev, urfp, pref_cell - 1, SA[0].dev,
                                   SA[1].dev, SA[2].dev));
  for (Arg<double> *x : {&PP, &FM, &U, &V, &W, &P, &D}) FCP_TRY(arg_back(ctx, *x));
  for (int k = 0; k < 3; ++k)
    if (s3[k]) FCP_TRY(arg_back(ctx, SA[k]));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_momentum_assemble(fcp_ctx *ctx, double *u, double *v, double *w, const double *uo,
                          const double *vo, const double *wo, const double *uoo,
                          const double *voo, const double *woo, double *p, const double *den,
                          const double *vis, const double *visw, const double *flmass,
                          double viscos, double dt, int time_scheme, int conv_scheme, double gds,
                          const double *urf, double *dpdxi, double *grad, double *rhs,
                          double *diagv, double *ap, double *spv, double *src, double *a) {
  FCP_TRY(need(ctx, true, true, true));
  if (!u || !v || !w || !p || !den || !vis || !flmass || !urf || !rhs || !diagv ||
      (time_scheme > 0 && (!uo || !vo || !wo)) || (time_scheme == 2 && (!uoo || !voo || !woo)))
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_momentum_assemble: null argument");
  if (time_scheme < 0 || time_scheme > 3 || (conv_scheme != 0 && conv_scheme != 1) ||
      (time_scheme > 0 && !(dt > 0.0)) || !(urf[0] > 0.0) || !(urf[1] > 0.0) || !(urf[2] > 0.0))
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_momentum_assemble: bad argument");
  (void)hipSetDevice(ctx->device);
  const size_t n = ctx->n, nt = n + ctx->Fb;
  // wall index of every boundary face, in patch order (the reference's iWall)
  std::vector<int> wall(ctx->Fb > 0 ? ctx->Fb : 1, -1);
  int nwal = 0;
  for (int ib = 0; ib < ctx->nPatches; ++ib)
    if (ctx->h_bctype[ib] == FCP_BC_WALL)
      for (int i = 0; i < ctx->h_nfaces[ib]; ++i) wall[ctx->h_start[ib] - ctx->Fi + i] = nwal++;
  if (nwal > 0 && !visw) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_momentum_assemble: visw needed");
  FCP_HIP(ctx, ctx->mom_wall.alloc(wall.size()));
  FCP_HIP(ctx, hipMemcpyAsync(ctx->mom_wall.p, wall.data(), sizeof(int) * wall.size(),
                              hipMemcpyHostToDevice, ctx->stream));
  FCP_HIP(ctx, ctx->mom_grad.alloc(9 * nt));
  FCP_HIP(ctx, ctx->mom_s3.alloc(3 * n > 0 ? 3 * n : 1));
  FCP_HIP(ctx, ctx->mom_dp.alloc(3 * nt));
  for (auto &b : ctx->mom_w3) FCP_HIP(ctx, b.alloc(3 * n > 0 ? 3 * n : 1));
  Arg<double> U, V, W, UO, VO, WO, UOO, VOO, WOO, P, R, VI, VW, FM, D, GR, RH, DG, AP, SPV, SRC;
  FCP_TRY(arg_in(ctx, (const double *)u, nt, U));
  FCP_TRY(arg_in(ctx, (const double *)v, nt, V));
  FCP_TRY(arg_in(ctx, (const double *)w, nt, W));
  FCP_TRY(arg_in(ctx, uo, uo ? n : 0, UO));
  FCP_TRY(arg_in(ctx, vo, vo ? n : 0, VO));
  FCP_TRY(arg_in(ctx, wo, wo ? n : 0, WO));
  FCP_TRY(arg_in(ctx, uoo, uoo ? n : 0, UOO));
  FCP_TRY(arg_in(ctx, voo, voo ? n : 0, VOO));
  FCP_TRY(arg_in(ctx, woo, woo ? n : 0, WOO));
  FCP_TRY(arg_in(ctx, (const double *)p, nt, P));
  FCP_TRY(arg_in(ctx, den, nt, R));
  FCP_TRY(arg_in(ctx, vis, nt, VI));
  FCP_TRY(arg_in(ctx, visw, visw ? (size_t)(nwal > 0 ? nwal : 1) : 0, VW));
  FCP_TRY(arg_in(ctx, flmass, ctx->F, FM));
  // optional outputs go to scratch of the context when not asked for
  auto out_or = [&](double *user, size_t cnt, Arg<double> &A, double *scratch) -> int {
    if (user) return arg_out(ctx, user, cnt, A);
    A.dev = scratch;
    return 0;
  };
  if (dpdxi) {
    FCP_TRY(arg_in(ctx, (const double *)dpdxi, 3 * nt, D));
  } else {
    D.dev = ctx->mom_dp.p;
  }
  FCP_TRY(out_or(grad, 9 * nt, GR, ctx->mom_grad.p));
  FCP_TRY(arg_out(ctx, rhs, 3 * n, RH));
  FCP_TRY(arg_out(ctx, diagv, 3 * n, DG));
  FCP_TRY(out_or(ap, 3 * n, AP, ctx->mom_w3[0].p));
  FCP_TRY(out_or(spv, 3 * n, SPV, ctx->mom_w3[1].p));
  FCP_TRY(out_or(src, 3 * n, SRC, ctx->mom_w3[2].p));
  FCP_TRY(fcp_launch_momentum(ctx, U.dev, V.dev, W.dev, UO.dev, VO.dev, WO.dev, UOO.dev, VOO.dev,
                              WOO.dev, P.dev, R.dev, VI.dev, VW.dev ? VW.dev : ctx->mom_s3.p,
                              ctx->mom_wall.p, FM.dev, viscos, dt, time_scheme, conv_scheme, gds,
                              urf, D.dev, GR.dev, ctx->mom_s3.p, RH.dev, DG.dev, AP.dev, SPV.dev,
                              SRC.dev));
  for (Arg<double> *x : {&U, &V, &W, &P, &D, &GR, &RH, &DG, &AP, &SPV, &SRC})
    FCP_TRY(arg_back(ctx, *x));
  FCP_TRY(out_copy(ctx, ctx->a.p, a, ctx->nnz));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_ddt(fcp_ctx *ctx, int scheme, const double *den, double dt, const double *phi_o,
            const double *phi_oo, double *su, double *sp) {
  FCP_TRY(need(ctx, true, true, false));
  if (!den || !phi_o || !su || !sp || (scheme == 2 && !phi_oo) || (scheme != 1 && scheme != 2) ||
      !(dt > 0.0))
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_ddt: bad argument");
  (void)hipSetDevice(ctx->device);
  Arg<double> d, fo, foo, s, p;
  FCP_TRY(arg_in(ctx, den, ctx->n, d));
  FCP_TRY(arg_in(ctx, phi_o, ctx->n, fo));
  if (phi_oo) FCP_TRY(arg_in(ctx, phi_oo, ctx->n, foo));
  FCP_TRY(arg_in(ctx, (const double *)su, ctx->n, s));
  FCP_TRY(arg_in(ctx, (const double *)sp, ctx->n, p));
  FCP_TRY(fcp_launch_ddt(ctx, scheme, d.dev, dt, fo.dev, foo.dev, s.dev, p.dev));
  FCP_TRY(arg_back(ctx, s));
  FCP_TRY(arg_back(ctx, p));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_grad_gauss(fcp_ctx *ctx, const double *u, double *dudxi) {
  FCP_TRY(need(ctx, true, true, false));
  if (!u || !dudxi) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_grad_gauss: null argument");
  (void)hipSetDevice(ctx->device);
  const size_t ntot = (size_t)ctx->n + ctx->Fb;
  Arg<double> in, out;
  FCP_TRY(arg_in(ctx, u, ntot, in));
  FCP_TRY(arg_out(ctx, dudxi, 3 * ntot, out));
  FCP_TRY(fcp_launch_grad_gauss(ctx, in.dev, out.dev));
  FCP_TRY(arg_back(ctx, out));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

namespace {
__global__ void k_plus_one(int n, int *__restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] += 1;
}
}  // namespace

int fcp_rcm(fcp_ctx *ctx, int pp_iters, int *perm) {
  FCP_TRY(need(ctx, true, false, true));
  if (!perm) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_rcm: null argument");
  if (pp_iters < 0 || pp_iters > 64) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_rcm: pp_iters");
  (void)hipSetDevice(ctx->device);
  Arg<int> out;
  FCP_TRY(arg_out(ctx, perm, (size_t)ctx->n, out));
  FCP_TRY(fcp_rcm_device(ctx, pp_iters, out.dev));
  k_plus_one<<<(ctx->n + 255) / 256, 256, 0, ctx->stream>>>(ctx->n, out.dev);
  FCP_HIP(ctx, hipGetLastError());
  FCP_TRY(arg_back(ctx, out));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_matrix_transpose(fcp_ctx *ctx) {
  FCP_TRY(need(ctx, true, false, true));
  (void)hipSetDevice(ctx->device);
  FCP_TRY(fcp_launch_transpose(ctx));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_convdiff_backward(fcp_ctx *ctx, const double *flmass, const double *dcoef,
                          const double *phi_old, const double *grad_old, double gam, int scheme,
                          const unsigned char *bc_kind, const double *lam, const double *phi_new,
                          double *dcoef_bar, double *phi_old_bar, double *grad_old_bar) {
  FCP_TRY(need(ctx, true, true, false));
  if (!flmass || !dcoef || !phi_old || !grad_old || !bc_kind || !lam || !phi_new || !dcoef_bar ||
      !phi_old_bar || !grad_old_bar)
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_convdiff_backward: null argument");
  if (scheme != 0 && scheme != 1) return fcp_fail(ctx, FCP_ERR_ARG, "scheme must be 0 or 1");
  (void)hipSetDevice(ctx->device);
  const size_t n = ctx->n, F = (size_t)ctx->Fi + ctx->Fb, nt = n + ctx->Fb;
  Arg<double> fm, dc, po, go, la, pn, dcb, pob, gob;
  Arg<unsigned char> bk;
  FCP_TRY(arg_in(ctx, flmass, F, fm));
  FCP_TRY(arg_in(ctx, dcoef, F, dc));
  FCP_TRY(arg_in(ctx, phi_old, nt, po));
  FCP_TRY(arg_in(ctx, grad_old, 3 * nt, go));
  FCP_TRY(arg_in(ctx, bc_kind, (size_t)ctx->Fb, bk));
  FCP_TRY(arg_in(ctx, lam, n, la));
  FCP_TRY(arg_in(ctx, phi_new, n, pn));
  FCP_TRY(arg_in(ctx, (const double *)dcoef_bar, F, dcb));
  FCP_TRY(arg_in(ctx, (const double *)phi_old_bar, nt, pob));
  FCP_TRY(arg_in(ctx, (const double *)grad_old_bar, 3 * nt, gob));
  FCP_TRY(fcp_launch_convdiff_bwd(ctx, fm.dev, dc.dev, po.dev, go.dev, gam, scheme, bk.dev, la.dev,
                                  pn.dev, dcb.dev, pob.dev, gob.dev));
  FCP_TRY(arg_back(ctx, dcb));
  FCP_TRY(arg_back(ctx, pob));
  FCP_TRY(arg_back(ctx, gob));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_laplacian_face_backward(fcp_ctx *ctx, const double *mu_f, const unsigned char *bc_kind,
                                const double *phi_bc, const double *gamma_f, const double *grad,
                                const double *lam, const double *phi, double *mu_f_bar,
                                double *phi_bc_bar, double *gamma_bar, double *grad_bar) {
  FCP_TRY(need(ctx, true, true, false));
  if (!mu_f || !phi_bc || !lam || !phi || !mu_f_bar || !phi_bc_bar)
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_laplacian_face_backward: null argument");
  if (gamma_f && (!grad || !gamma_bar || !grad_bar))
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_laplacian_face_backward: correction outputs missing");
  (void)hipSetDevice(ctx->device);
  const size_t n = ctx->n, F = (size_t)ctx->Fi + ctx->Fb, nt = n + ctx->Fb;
  Arg<double> mf, pb, gf, gr, la, ph, mfb, pbb, gfb, grb;
  Arg<unsigned char> bk;
  FCP_TRY(arg_in(ctx, mu_f, F, mf));
  FCP_TRY(arg_in(ctx, bc_kind, (size_t)ctx->Fb, bk));
  FCP_TRY(arg_in(ctx, phi_bc, nt, pb));
  FCP_TRY(arg_in(ctx, gamma_f, F, gf));
  FCP_TRY(arg_in(ctx, grad, 3 * nt, gr));
  FCP_TRY(arg_in(ctx, lam, n, la));
  FCP_TRY(arg_in(ctx, phi, n, ph));
  FCP_TRY(arg_in(ctx, (const double *)mu_f_bar, F, mfb));
  FCP_TRY(arg_in(ctx, (const double *)phi_bc_bar, nt, pbb));
  FCP_TRY(arg_in(ctx, (const double *)gamma_bar, F, gfb));
  FCP_TRY(arg_in(ctx, (const double *)grad_bar, 3 * nt, grb));
  FCP_TRY(fcp_launch_laplacian_bwd(ctx, mf.dev, bk.dev, pb.dev, gf.dev, gr.dev, la.dev, ph.dev,
                                   mfb.dev, pbb.dev, gfb.dev, grb.dev));
  FCP_TRY(arg_back(ctx, mfb));
  FCP_TRY(arg_back(ctx, pbb));
  FCP_TRY(arg_back(ctx, gfb));
  FCP_TRY(arg_back(ctx, grb));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_grad_gauss_backward(fcp_ctx *ctx, const double *grad_bar, double *phi_bar) {
  FCP_TRY(need(ctx, true, true, false));
  if (!grad_bar || !phi_bar) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_grad_gauss_backward: null");
  (void)hipSetDevice(ctx->device);
  const size_t nt = (size_t)ctx->n + ctx->Fb;
  Arg<double> gb, pb;
  FCP_TRY(arg_in(ctx, grad_bar, 3 * nt, gb));
  FCP_TRY(arg_in(ctx, (const double *)phi_bar, nt, pb));
  FCP_TRY(fcp_launch_grad_gauss_bwd(ctx, gb.dev, pb.dev));
  FCP_TRY(arg_back(ctx, pb));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_create_lsq_grad_matrix(fcp_ctx *ctx, int method, int compat, double *dmat) {
  FCP_TRY(need(ctx, true, true, false));
  if (method < 1 || method > 3)
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_create_lsq_grad_matrix: method %d", method);
  (void)hipSetDevice(ctx->device);
  FCP_TRY(fcp_ensure_workspace(ctx));
  FCP_TRY(fcp_launch_lsq_matrix(ctx, method, compat ? 1 : 0));
  if (dmat) {
    const size_t cnt = (method == 3 ? 18 : 9) * (size_t)ctx->n;
    FCP_HIP(ctx, hipMemcpyAsync(dmat, ctx->lsq_dmat.p, cnt * sizeof(double), hipMemcpyDefault,
                                ctx->stream));
  }
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_grad(fcp_ctx *ctx, int method, const double *phi, double *dphidxi) {
  FCP_TRY(need(ctx, true, true, false));
  if (!phi || !dphidxi) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_grad: null argument");
  if (method == FCP_GRAD_GAUSS) return fcp_grad_gauss(ctx, phi, dphidxi);
  if (method == FCP_GRAD_GAUSS_FVX || method == FCP_GRAD_GAUSS_CORRECTED) {
    (void)hipSetDevice(ctx->device);
    const size_t ntot = (size_t)ctx->n + ctx->Fb;
    FCP_HIP(ctx, ctx->fvx_old.alloc(3 * ntot));
    Arg<double> in, io;
    FCP_TRY(arg_in(ctx, phi, ntot, in));
    const double *dfo = nullptr;
    if (method == FCP_GRAD_GAUSS_CORRECTED) {
      FCP_TRY(arg_in(ctx, (const double *)dphidxi, 3 * ntot, io));
      FCP_HIP(ctx, hipMemcpyAsync(ctx->fvx_old.p, io.dev, sizeof(double) * 3 * ctx->n,
                                  hipMemcpyDeviceToDevice, ctx->stream));
      dfo = ctx->fvx_old.p;
    } else {
      FCP_TRY(arg_out(ctx, dphidxi, 3 * ntot, io));
    }
    FCP_TRY(fcp_launch_grad_fvx(ctx, in.dev, dfo, ctx->fvx_old.p, io.dev));
    FCP_TRY(arg_back(ctx, io));
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
  }
  if (method != ctx->lsq_method)
    return fcp_fail(ctx, FCP_ERR_STATE, "fcp_grad: least-squares matrices of method %d not created",
                    method);
  (void)hipSetDevice(ctx->device);
  const size_t ntot = (size_t)ctx->n + ctx->Fb;
  Arg<double> in, out;
  FCP_TRY(arg_in(ctx, phi, ntot, in));
  FCP_TRY(arg_out(ctx, dphidxi, 3 * ntot, out));
  FCP_TRY(fcp_launch_grad_lsq(ctx, in.dev, out.dev));
  FCP_TRY(arg_back(ctx, out));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_update_boundary(fcp_ctx *ctx, double *phi) {
  FCP_TRY(need(ctx, true, false, false));
  if (!phi) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_update_boundary: null argument");
  (void)hipSetDevice(ctx->device);
  Arg<double> p;
  FCP_TRY(arg_in(ctx, (const double *)phi, (size_t)ctx->n + ctx->Fb, p));
  FCP_TRY(fcp_launch_update_boundary(ctx, p.dev));
  FCP_TRY(arg_back(ctx, p));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// linear algebra
// ---------------------------------------------------------------------------
int fcp_matrix_set(fcp_ctx *ctx, const double *a) {
  FCP_TRY(need(ctx, true, false, true));
  if (!a) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_matrix_set: null argument");
  FCP_TRY(fcp_coarse_join(ctx));
  hipMemcpyKind kind = fcp_is_device_ptr(a) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  FCP_HIP(ctx, hipMemcpyAsync(ctx->a.p, a, sizeof(double) * ctx->nnz, kind, ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->a_version++;
  return 0;
}

int fcp_spmv(fcp_ctx *ctx, const double *x, double *y) {
  FCP_TRY(need(ctx, true, false, true));
  if (!x || !y) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_spmv: null argument");
  Arg<double> xi, yo;
  FCP_TRY(arg_in(ctx, x, ctx->n, xi));
  FCP_TRY(arg_out(ctx, y, ctx->n, yo));
  FCP_TRY(fcp_launch_spmv64(ctx, ctx->a.p, xi.dev, yo.dev));
  FCP_TRY(arg_back(ctx, yo));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_precond_apply(fcp_ctx *ctx, int kind, const double *r, double *z, double *d) {
  FCP_TRY(need(ctx, true, false, true));
  if (!r || !z || kind < 0 || kind > 9)
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_precond_apply: bad argument");
  if (kind == 2) d = nullptr;  // ILU(0) has no pivot vector to return
  (void)hipSetDevice(ctx->device);
  Arg<double> ri, zo, dout;
  FCP_TRY(arg_in(ctx, r, ctx->n, ri));
  FCP_TRY(arg_out(ctx, z, ctx->n, zo));
  if (d) FCP_TRY(arg_out(ctx, d, ctx->n, dout));
  if (kind == 2)
    FCP_TRY(fcp_ilu_apply_device(ctx, ri.dev, zo.dev));
  else if (kind >= 8)  // FP64 tile-order kernel: DIC (8), plus the coarse correction (9)
    FCP_TRY(fcp_bj32_device(ctx, kind == 8 ? 0 : 2, ri.dev, zo.dev, dout.dev, true));
  else if (kind >= 5)
    FCP_TRY(fcp_bj32_device(ctx, kind - 5, ri.dev, zo.dev, dout.dev));
  else if (kind >= 3)
    FCP_TRY(fcp_bj_device(ctx, kind - 3, ri.dev, zo.dev, dout.dev));
  else
    FCP_TRY(fcp_precond_device(ctx, kind, ri.dev, zo.dev, dout.dev));
  FCP_TRY(arg_back(ctx, zo));
  FCP_TRY(arg_back(ctx, dout));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

static int solver_code(const char *s, int len) {
  std::string name(s, s + (len > 0 ? len : strlen(s)));
  while (!name.empty() && name.back() == ' ') name.pop_back();
  if (name == "iccg") return 0;
  if (name == "dpcg") return 1;
  if (name == "bicgstab") return 2;
  if (name == "gauss-seidel") return 3;
  if (name == "pmgmres") return 4;
  return -1;
}

int fcp_spmv_transpose(fcp_ctx *ctx, const double *x, double *y) {
  FCP_TRY(need(ctx, true, false, true));
  if (!x || !y) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_spmv_transpose: null argument");
  (void)hipSetDevice(ctx->device);
  Arg<double> xi, yo;
  FCP_TRY(arg_in(ctx, x, ctx->n, xi));
  FCP_TRY(arg_out(ctx, y, ctx->n, yo));
  FCP_TRY(fcp_atx_device(ctx, xi.dev, yo.dev));
  FCP_TRY(arg_back(ctx, yo));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_sweep_tiles(fcp_ctx *ctx, int *tile) {
  FCP_TRY(need(ctx, true, false, true));
  if (!tile) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_sweep_tiles: null argument");
  (void)hipSetDevice(ctx->device);
  if (!ctx->have_bj) FCP_TRY(fcp_build_bj_schedules(ctx));
  if (fcp_is_device_ptr(tile)) {
    FCP_HIP(ctx, hipMemcpy(tile, ctx->bj_tile.data(), sizeof(int) * ctx->n, hipMemcpyHostToDevice));
  } else {
    std::copy(ctx->bj_tile.begin(), ctx->bj_tile.end(), tile);
  }
  return 0;
}

int fcp_ilu_factor(fcp_ctx *ctx, double *l) {
  FCP_TRY(need(ctx, true, false, true));
  if (!l) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_ilu_factor: null argument");
  (void)hipSetDevice(ctx->device);
  FCP_TRY(fcp_ensure_workspace(ctx));
  FCP_TRY(fcp_ilu_factor_device(ctx));
  FCP_TRY(out_copy(ctx, ctx->ilu.p, l, ctx->nnz));
  return 0;
}

int fcp_csrsolve(fcp_ctx *ctx, const char *solver, int solver_len, double *fi, const double *rhs,
                 double *res0, int itr_max, double tol_abs, double tol_rel, const char *chvar,
                 int chvar_len, int verbose, fcp_solve_opts *opts) {
  FCP_TRY(need(ctx, true, false, true));
  if (!solver || !fi || !rhs) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_csrsolve: null argument");
  const int code = solver_code(solver, solver_len);
  if (code < 0)
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_csrsolve: unknown solver '%.*s'",
                    solver_len > 0 ? solver_len : (int)strlen(solver), solver);
  (void)hipSetDevice(ctx->device);
  Arg<double> x, b;
  FCP_TRY(arg_in(ctx, (const double *)fi, ctx->n, x));
  FCP_TRY(arg_in(ctx, rhs, ctx->n, b));
  fcp_solve_opts local;
  fcp_solve_opts_default(&local);
  fcp_solve_opts *o = opts ? opts : &local;
  const int rc = fcp_solve_device(ctx, code, x.dev, b.dev, itr_max, tol_abs, tol_rel, verbose, o);
  if (rc < 0) return rc;
  FCP_TRY(arg_back(ctx, x));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (res0) *res0 = o->resor;
  if (verbose >= 1) {
    static const char *names[] = {"PCG(IC0)", "PCG(Jacobi)", "BiCGStab(ILU(0))", "Gauss-Seidel",
                                  "PMGMRES_ILU(4)"};
    std::string var = chvar ? std::string(chvar, chvar + (chvar_len > 0 ? chvar_len : strlen(chvar)))
                            : std::string("x");
    while (!var.empty() && var.back() == ' ') var.pop_back();
    const double factor = (o->res0 > 0 && o->resor > 0) ? o->res0 / o->resor : 1.0;
    printf("  %s:  Solving for %s, Initial residual = %10.3E, Final residual = %10.3E, No Iterations %d\n",
           names[code], var.c_str(), o->resor, o->resl / factor, o->iterations);
  }
  return rc;
}

// ---------------------------------------------------------------------------
// LIS replacement (LIS_linear_solvers.f90:38-200): the LIS option syntax
// mapped onto the library's solvers; the residual report of lis_spsolve.
// ---------------------------------------------------------------------------
namespace {

bool lis_word(const std::string &v, const char *name, int code) {
  return v == name || v == std::to_string(code);
}

int lis_parse(fcp_ctx *ctx, const std::string &text, fcp_ctx::LisSolver &L) {
  std::vector<std::string> tok;
  size_t i = 0;
  while (i < text.size()) {
    while (i < text.size() && isspace((unsigned char)text[i])) ++i;
    size_t j = i;
    while (j < text.size() && !isspace((unsigned char)text[j])) ++j;
    if (j > i) tok.push_back(text.substr(i, j - i));
    i = j;
  }
  std::string solver = "cg", precond = "none";  // LIS defaults
  L = fcp_ctx::LisSolver();
  for (size_t k = 0; k < tok.size(); ++k) {
    const std::string &o = tok[k];
    if (o == "-print" || o == "-scale" || o == "-omp_num_threads" || o == "-storage" ||
        o == "-storage_block" || o == "-adds" || o == "-adds_iter") {
      // output / threading / storage choices of LIS: no effect on the result here
      if (o != "-print" || (k + 1 < tok.size() && tok[k + 1][0] != '-')) ++k;
      continue;
    }
    if (k + 1 >= tok.size())
      return fcp_fail(ctx, FCP_ERR_ARG, "LIS options: %s needs a value", o.c_str());
    const std::string &v = tok[++k];
    char *end = nullptr;
    if (o == "-i") {
      solver = v;
    } else if (o == "-p") {
      precond = v;
    } else if (o == "-tol") {
      L.tol = strtod(v.c_str(), &end);
      if (!end || *end || !(L.tol > 0.0)) return fcp_fail(ctx, FCP_ERR_ARG, "LIS options: bad -tol");
    } else if (o == "-maxiter") {
      L.maxiter = (int)strtol(v.c_str(), &end, 10);
      if (!end || *end || L.maxiter < 1) return fcp_fail(ctx, FCP_ERR_ARG, "LIS options: bad -maxiter");
    } else if (o == "-restart") {
      L.restart = (int)strtol(v.c_str(), &end, 10);
      if (!end || *end || L.restart < 1) return fcp_fail(ctx, FCP_ERR_ARG, "LIS options: bad -restart");
    } else if (o == "-initx_zeros") {
      if (v == "true" || v == "1") L.initx_zeros = true;
      else if (v == "false" || v == "0") L.initx_zeros = false;
      else return fcp_fail(ctx, FCP_ERR_ARG, "LIS options: bad -initx_zeros");
    } else if (o == "-f") {  // working precision: double only
      if (v != "double" && v != "0")
        return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "LIS options: -f %s (double only)", v.c_str());
    } else if (o == "-conv_cond") {
      if (v != "0" && v != "1") return fcp_fail(ctx, FCP_ERR_ARG, "LIS options: -conv_cond 0 or 1");
      L.conv_cond = v[0] - '0';
    } else {
      return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "LIS options: %s is not supported", o.c_str());
    }
  }
  // (LIS solver, LIS preconditioner) -> the library's solver of the same
  // Krylov method and preconditioner family
  const bool ilu = lis_word(precond, "ilu", 2), jac = lis_word(precond, "jacobi", 1),
             none = lis_word(precond, "none", 0);
  if (lis_word(solver, "cg", 1) && ilu) L.code = 0;               // IC(0)-type PCG (iccg)
  else if (lis_word(solver, "cg", 1) && jac) L.code = 1;          // Jacobi PCG (dpcg)
  else if (lis_word(solver, "bicgstab", 4) && ilu) L.code = 2;    // BiCGStab, DILU
  else if (lis_word(solver, "gs", 11) && none) L.code = 3;        // Gauss-Seidel
  else if (lis_word(solver, "gmres", 9) && ilu) L.code = 4;       // GMRES(m), ILU(0)
  else
    return fcp_fail(ctx, FCP_ERR_UNSUPPORTED,
                    "LIS options: -i %s -p %s has no counterpart (cg+ilu, cg+jacobi, "
                    "bicgstab+ilu, gs+none, gmres+ilu)",
                    solver.c_str(), precond.c_str());
  L.options = text;
  L.have = true;
  return 0;
}

}  // namespace

int fcp_lis_create_solver(fcp_ctx *ctx, const char *options, int len) {
  if (!ctx) return FCP_ERR_ARG;
  if (!options) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_lis_create_solver: null options");
  std::string text(options, options + (len > 0 ? len : strlen(options)));
  while (!text.empty() && text.back() == ' ') text.pop_back();
  fcp_ctx::LisSolver L;
  FCP_TRY(lis_parse(ctx, text, L));
  ctx->lis = L;
  return 0;
}

int fcp_lis_destroy_solver(fcp_ctx *ctx) {
  if (!ctx) return FCP_ERR_ARG;
  ctx->lis = fcp_ctx::LisSolver();
  return 0;
}

int fcp_lis_spsolve(fcp_ctx *ctx, double *fi, const double *rhs, const char *chvar,
                    int chvar_len, double *res0_out, double *resid_out, int *iter_out) {
  FCP_TRY(need(ctx, true, false, true));
  if (!ctx->lis.have) return fcp_fail(ctx, FCP_ERR_STATE, "fcp_lis_spsolve: no solver created");
  if (!fi || !rhs) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_lis_spsolve: null argument");
  (void)hipSetDevice(ctx->device);
  const fcp_ctx::LisSolver &L = ctx->lis;
  const int n = ctx->n;
  Arg<double> x, b;
  FCP_TRY(arg_in(ctx, (const double *)fi, n, x));
  FCP_TRY(arg_in(ctx, rhs, n, b));
  // initial residual of the field passed in, L1 as lis_spsolve sums it
  // (:135-143), and its L2 norm with ||b||_2 for the convergence condition
  DBuf<double> ax;
  FCP_HIP(ctx, ax.alloc(n));
  FCP_TRY(fcp_launch_spmv64(ctx, ctx->a.p, x.dev, ax.p));
  std::vector<double> hax(n), hb(n), hx(n);
  FCP_HIP(ctx, hipMemcpyAsync(hax.data(), ax.p, sizeof(double) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
  FCP_HIP(ctx, hipMemcpyAsync(hb.data(), b.dev, sizeof(double) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double res0 = 0.0, r2 = 0.0, b2 = 0.0;
  for (int i = 0; i < n; ++i) {
    const double r = hb[i] - hax[i];
    res0 = res0 + fabs(r);
    r2 += r * r;
    b2 += hb[i] * hb[i];
  }
  r2 = sqrt(r2);
  b2 = sqrt(b2);
  // LIS starts from zero unless -initx_zeros false; convergence condition 0
  // is ||r||_2 <= tol ||r_0||_2, 1 is ||r||_2 <= tol ||b||_2
  if (L.initx_zeros) FCP_HIP(ctx, hipMemsetAsync(x.dev, 0, sizeof(double) * n, ctx->stream));
  const double r0n = L.initx_zeros ? b2 : r2;
  const double rel = (L.conv_cond == 0 && !L.initx_zeros && b2 > 0.0) ? L.tol * r2 / b2 : L.tol;
  fcp_solve_opts o;
  fcp_solve_opts_default(&o);
  o.criterion = FCP_CRIT_TRUE_L2;
  int itmax = L.maxiter;
  if (L.code == 4) itmax = (L.maxiter + L.restart - 1) / L.restart;  // GMRES counts restart cycles
  const int keep = ctx->gm_restart;
  ctx->gm_restart = L.restart;
  const int rc = fcp_solve_device(ctx, L.code, x.dev, b.dev, itmax, 0.0, rel, 0, &o);
  ctx->gm_restart = keep;
  if (rc < 0) return rc;
  FCP_TRY(arg_back(ctx, x));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const double denom = L.conv_cond == 0 ? r0n : b2;
  const double resid = denom > 0.0 ? o.rel_res_l2 * b2 / denom : 0.0;
  const int iters = o.iterations;
  std::string var = chvar ? std::string(chvar, chvar + (chvar_len > 0 ? chvar_len : strlen(chvar)))
                          : std::string("x");
  while (!var.empty() && var.back() == ' ') var.pop_back();
  printf("  LIS[%s]:  Solving for %s, Initial residual = %10.3E, Final residual = %10.3E, "
         "No Iterations %d\n",
         L.options.c_str(), var.c_str(), res0, res0 * resid, iters);
  if (res0_out) *res0_out = res0;
  if (resid_out) *resid_out = resid;
  if (iter_out) *iter_out = iters;
  return rc;
}

int fcp_transport(fcp_ctx *ctx, double *phi, const double *phi_o, const double *phi_oo,
                  const double *den, const double *vis, const double *flmass,
                  const double *wall_frac, double viscos, double pranl, double sigt, double urf,
                  double gds, double dt, int time_scheme, int conv_scheme, int grad_method,
                  const char *solver, int solver_len, int itr_max, double tol_abs, double tol_rel,
                  int clip, double *grad, double *a, double *su, double *sp, double *resor,
                  double *fimin, double *fimax, fcp_solve_opts *opts) {
  FCP_TRY(need(ctx, true, true, true));
  const bool transient = time_scheme != 0;
  if (!phi || !vis || !flmass || !solver || time_scheme < 0 || time_scheme > 3 ||
      (conv_scheme != 0 && conv_scheme != 1) || !(urf > 0.0) || !(pranl != 0.0) ||
      !(sigt != 0.0) || (transient && (!den || !phi_o || !(dt > 0.0))) ||
      (time_scheme == 2 && !phi_oo) || (ctx->nPatches > 0 && !wall_frac))
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_transport: bad argument");
  if (grad_method != FCP_GRAD_GAUSS && grad_method != ctx->lsq_method)
    return fcp_fail(ctx, FCP_ERR_STATE,
                    "fcp_transport: least-squares matrices of method %d not created", grad_method);
  const int code = solver_code(solver, solver_len);
  if (code < 0) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_transport: unknown solver");
  (void)hipSetDevice(ctx->device);
  const int n = ctx->n, Fb = ctx->Fb, np = ctx->nPatches;
  const size_t ntot = (size_t)n + Fb;
  // boundary treatment per face (temperature.f90:168-320), rebuilt when the
  // wall fractions change
  std::vector<double> wf(np);
  if (np > 0) {
    if (fcp_is_device_ptr(wall_frac)) {
      FCP_HIP(ctx, hipMemcpy(wf.data(), wall_frac, sizeof(double) * np, hipMemcpyDeviceToHost));
    } else {
      std::copy(wall_frac, wall_frac + np, wf.begin());
    }
  }
  if (!ctx->tr_kind.p || ctx->tr_wf != wf) {
    std::vector<unsigned char> kind(Fb > 0 ? Fb : 1, 3);
    std::vector<int> srd(Fb > 0 ? Fb : 1);
    std::vector<int> walls;
    for (int ib = 0; ib < np; ++ib)
      if (ctx->h_bctype[ib] == FCP_BC_WALL)
        for (int i = 0; i < ctx->h_nfaces[ib]; ++i) walls.push_back(ctx->h_start[ib] - ctx->Fi + i);
    size_t iwall = 0;
    for (int ib = 0; ib < np; ++ib) {
      const int t = ctx->h_bctype[ib], s0 = ctx->h_start[ib] - ctx->Fi, nf = ctx->h_nfaces[ib];
      for (int i = 0; i < nf; ++i) srd[s0 + i] = s0 + i;
      unsigned char k = 3;
      if (t == FCP_BC_PERIODIC) k = 5;
      else if (t == FCP_BC_INLET || t == FCP_BC_OUTLET || t == FCP_BC_PRESSURE) k = 2;
      else if ((t == FCP_BC_WALL && wf[ib] == 1.0) || t == FCP_BC_WALL_QFLUX) k = t == FCP_BC_WALL ? 0 : 4;
      else if (t == FCP_BC_WALL && wf[ib] == 0.0) k = 1;
      if (k == 0 || k == 4) {
        // srdw(iWall): the counter runs over these faces, srdw over the wall patches' faces
        if (iwall + nf > walls.size())
          return fcp_fail(ctx, FCP_ERR_ARG, "fcp_transport: more prescribed-wall faces than wall faces");
        for (int i = 0; i < nf; ++i) srd[s0 + i] = walls[iwall + i];
        iwall += nf;
      }
      for (int i = 0; i < nf; ++i) kind[s0 + i] = k;
    }
    FCP_HIP(ctx, ctx->tr_kind.alloc(kind.size()));
    FCP_HIP(ctx, ctx->tr_srd.alloc(srd.size()));
    FCP_HIP(ctx, hipMemcpyAsync(ctx->tr_kind.p, kind.data(), kind.size(), hipMemcpyHostToDevice,
                                ctx->stream));
    FCP_HIP(ctx, hipMemcpyAsync(ctx->tr_srd.p, srd.data(), srd.size() * sizeof(int),
                                hipMemcpyHostToDevice, ctx->stream));
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->tr_wf = wf;
  }
  FCP_HIP(ctx, ctx->tr_grad.alloc(3 * ntot));
  FCP_HIP(ctx, ctx->tr_dcoef.alloc(ctx->F));
  FCP_HIP(ctx, ctx->tr_su.alloc(n));
  FCP_HIP(ctx, ctx->tr_sp.alloc(n));
  FCP_HIP(ctx, ctx->tr_mm.alloc(2 * 1024));
  Arg<double> ph, fo, foo, dn, vs, fm;
  FCP_TRY(arg_in(ctx, (const double *)phi, ntot, ph));
  FCP_TRY(arg_in(ctx, vis, n, vs));
  FCP_TRY(arg_in(ctx, flmass, ctx->F, fm));
  if (transient) {
    FCP_TRY(arg_in(ctx, den, n, dn));
    FCP_TRY(arg_in(ctx, phi_o, n, fo));
    if (time_scheme == 2) FCP_TRY(arg_in(ctx, phi_oo, n, foo));
  }
  // grad(T, dTdxi) (:72)
  if (grad_method == FCP_GRAD_GAUSS)
    FCP_TRY(fcp_launch_grad_gauss(ctx, ph.dev, ctx->tr_grad.p));
  else
    FCP_TRY(fcp_launch_grad_lsq(ctx, ph.dev, ctx->tr_grad.p));
  FCP_TRY(fcp_launch_transport_dcoef(ctx, vs.dev, ctx->tr_kind.p, viscos, pranl, sigt,
                                     ctx->tr_dcoef.p));
  FCP_HIP(ctx, hipMemsetAsync(ctx->tr_su.p, 0, sizeof(double) * n, ctx->stream));
  FCP_HIP(ctx, hipMemsetAsync(ctx->tr_sp.p, 0, sizeof(double) * n, ctx->stream));
  if (transient)  // :82-102; Crank-Nicolson shares BDF1's volume terms
    FCP_TRY(fcp_launch_ddt(ctx, time_scheme == 2 ? 2 : 1, dn.dev, dt, fo.dev, foo.dev,
                           ctx->tr_su.p, ctx->tr_sp.p));
  FCP_TRY(fcp_launch_convdiff(ctx, fm.dev, ctx->tr_dcoef.p, ph.dev, ctx->tr_grad.p, gds,
                              conv_scheme, ctx->tr_kind.p, urf, ctx->tr_su.p, ctx->tr_sp.p,
                              ctx->tr_srd.p, time_scheme == 3, dn.dev, dt, fo.dev));
  FCP_TRY(out_copy(ctx, ctx->a.p, a, ctx->nnz));
  FCP_TRY(out_copy(ctx, ctx->tr_su.p, su, n));
  FCP_TRY(out_copy(ctx, ctx->tr_sp.p, sp, n));
  // csrsolve(lSolverT, T, su, resor(7), ...) (:369) from the current cell values
  fcp_solve_opts local;
  fcp_solve_opts_default(&local);
  fcp_solve_opts *o = opts ? opts : &local;
  const int rc = fcp_solve_device(ctx, code, ph.dev, ctx->tr_su.p, itr_max, tol_abs, tol_rel, 0, o);
  if (rc < 0) return rc;
  FCP_TRY(fcp_launch_update_boundary(ctx, ph.dev));  // :373
  FCP_TRY(fcp_launch_minmax_clip(ctx, ph.dev, clip, ctx->tr_mm.p));  // :377-381
  double mm[2];
  FCP_HIP(ctx, hipMemcpyAsync(mm, ctx->tr_mm.p, sizeof mm, hipMemcpyDeviceToHost, ctx->stream));
  FCP_TRY(arg_back(ctx, ph));
  FCP_TRY(out_copy(ctx, ctx->tr_grad.p, grad, 3 * ntot));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (resor) *resor = o->resor;
  if (fimin) *fimin = mm[0];
  if (fimax) *fimax = mm[1];
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// autotuner of the tile size of the FP32 inner solves' preconditioner
// ---------------------------------------------------------------------------
namespace {
// 64-bit finaliser of MurmurHash3 (public domain): mixes one pattern entry.
__device__ __forceinline__ unsigned long long mix64(unsigned long long v) {
  v ^= v >> 33;
  v *= 0xff51afd7ed558ccdULL;
  v ^= v >> 33;
  v *= 0xc4ceb9fe1a85ec53ULL;
  v ^= v >> 33;
  return v;
}

// Per-block sums (mod 2^64) of the mixed (position, value) pairs of ia and
// ja: an order-independent signature of the CSR pattern.
__global__ void k_pattern_sig(const int *__restrict__ ia, const int *__restrict__ ja, int n,
                              int nnz, unsigned long long *__restrict__ part) {
  __shared__ unsigned long long sh[256];
  unsigned long long s = 0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += gridDim.x * blockDim.x)
    s += mix64(((unsigned long long)(unsigned)ja[k] << 32) | (unsigned)k);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k <= n; k += gridDim.x * blockDim.x)
    s += mix64((((unsigned long long)(unsigned)ia[k] << 32) | (unsigned)k) ^ 0x9e3779b97f4a7c15ULL);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_fill64(double *__restrict__ x, int n, double v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = v;
}

// (cells, entries, pattern signature, solver) -> chosen tile rows
std::mutex g_tune_mu;
std::map<std::tuple<int, int, unsigned long long, int>, int> g_tune_cache;

void set_inner_tile(fcp_ctx *ctx, int t) {
  if (t == ctx->tune_inner_tile) return;
  ctx->tune_inner_tile = t;
  ctx->have_bj = false;
  ctx->tin_have = false;
  ctx->tc_have = false;
}
}  // namespace

extern "C" {

int fcp_autotune(fcp_ctx *ctx, const char *solver, int solver_len, int *tile_out, int *cached) {
  FCP_TRY(need(ctx, true, false, true));
  if (!solver) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_autotune: null solver");
  const int code = solver_code(solver, solver_len);
  if (code != 0 && code != 2)
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_autotune: solver must be iccg or bicgstab");
  (void)hipSetDevice(ctx->device);
  const int n = ctx->n;
  constexpr int G = 256;
  DBuf<unsigned long long> part;
  FCP_HIP(ctx, part.alloc(G));
  k_pattern_sig<<<G, 256, 0, ctx->stream>>>(ctx->ia.p, ctx->ja.p, n, ctx->nnz, part.p);
  std::vector<unsigned long long> hp(G);
  FCP_HIP(ctx, hipMemcpyAsync(hp.data(), part.p, sizeof(unsigned long long) * G,
                              hipMemcpyDeviceToHost, ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  unsigned long long sig = 0;
  for (unsigned long long v : hp) sig += v;
  const auto key = std::make_tuple(n, ctx->nnz, sig, code);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      set_inner_tile(ctx, it->second);
      if (tile_out) *tile_out = it->second;
      if (cached) *cached = 1;
      return 0;
    }
  }
  // trials: the mixed solve of A x = A 1 to a relative L2 residual of 1e-6,
  // once to build the tile schedules and once timed, per candidate
  DBuf<double> ones, b, x;
  FCP_HIP(ctx, ones.alloc(n));
  FCP_HIP(ctx, b.alloc(n));
  FCP_HIP(ctx, x.alloc(n));
  k_fill64<<<grid_for(n, 256) < 1024 ? grid_for(n, 256) : 1024, 256, 0, ctx->stream>>>(ones.p, n, 1.0);
  FCP_TRY(fcp_launch_spmv64(ctx, ctx->a.p, ones.p, b.p));
  const int prev = ctx->tune_inner_tile;
  int best = prev;
  double best_ms = 0.0;
  bool found = false;
  for (int t : {1024, 2048, 4096, 8192}) {
    if (n < 4 * t) continue;
    set_inner_tile(ctx, t);
    for (int rep = 0; rep < 2; ++rep) {
      k_fill64<<<grid_for(n, 256) < 1024 ? grid_for(n, 256) : 1024, 256, 0, ctx->stream>>>(
          x.p, n, 0.0);
      fcp_solve_opts o;
      fcp_solve_opts_default(&o);
      o.precision = FCP_PREC_MIXED;
      o.criterion = FCP_CRIT_TRUE_L2;
      o.check_every = 16;
      const int rc = fcp_solve_device(ctx, code, x.p, b.p, 20000, 0.0, 1e-6, 0, &o);
      if (rc < 0) {
        set_inner_tile(ctx, prev);
        return rc;
      }
      if (rep == 1 && rc == 0 && (!found || o.time_ms < best_ms)) {
        found = true;
        best = t;
        best_ms = o.time_ms;
      }
    }
  }
  set_inner_tile(ctx, found ? best : prev);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = ctx->tune_inner_tile;
  }
  if (tile_out) *tile_out = ctx->tune_inner_tile;
  if (cached) *cached = 0;
  return 0;
}

int fcp_csrsolve_multi(fcp_ctx *ctx, const char *solver, int solver_len, int nrhs, double *fi,
                       int ldfi, const double *rhs, double *res0, int itr_max, double tol_abs,
                       double tol_rel, int verbose, fcp_solve_opts *opts) {
  FCP_TRY(need(ctx, true, false, true));
  if (!solver || !fi || !rhs) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_csrsolve_multi: null argument");
  if (nrhs < 1 || nrhs > 4) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_csrsolve_multi: nrhs %d", nrhs);
  if (ldfi < ctx->n) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_csrsolve_multi: ldfi < numCells");
  const int code = solver_code(solver, solver_len);
  if (code != 0 && code != 2)
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_csrsolve_multi: solver must be iccg or bicgstab");
  (void)hipSetDevice(ctx->device);
  Arg<double> x, b;
  FCP_TRY(arg_in(ctx, (const double *)fi, (size_t)ldfi * nrhs, x));
  FCP_TRY(arg_in(ctx, rhs, (size_t)ctx->n * nrhs, b));
  fcp_solve_opts local;
  fcp_solve_opts_default(&local);
  fcp_solve_opts *o = opts ? opts : &local;
  const int rc = fcp_solve_multi_device(ctx, code, nrhs, x.dev, ldfi, b.dev, res0, itr_max,
                                        tol_abs, tol_rel, verbose, o);
  if (rc < 0) return rc;
  FCP_TRY(arg_back(ctx, x));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (verbose >= 1)
    printf("  %s x%d:  Initial residual = %10.3E, worst |b-Ax|/|b| = %10.3E, No Iterations %d\n",
           code == 0 ? "PCG(IC0)" : "BiCGStab(ILU(0))", nrhs, o->res0, o->rel_res_l2, o->iterations);
  return rc;
}

namespace {
__global__ void k_set_diag(int n, const int *__restrict__ diag, const double *__restrict__ dv,
                           double *__restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[diag[i]] = dv[i];
}
}  // namespace

int fcp_csrsolve_multi_diag(fcp_ctx *ctx, const char *solver, int solver_len, int nrhs,
                            double *fi, int ldfi, const double *rhs, const double *diagv,
                            double *res0, int itr_max, double tol_abs, double tol_rel,
                            int verbose, fcp_solve_opts *opts) {
  FCP_TRY(need(ctx, true, false, true));
  if (!solver || !fi || !rhs || !diagv)
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_csrsolve_multi_diag: null argument");
  if (nrhs < 1 || nrhs > 4)
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_csrsolve_multi_diag: nrhs %d", nrhs);
  if (ldfi < ctx->n) return fcp_fail(ctx, FCP_ERR_ARG, "fcp_csrsolve_multi_diag: ldfi < numCells");
  const int code = solver_code(solver, solver_len);
  if (code != 0 && code != 2)
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_csrsolve_multi_diag: solver must be iccg or bicgstab");
  (void)hipSetDevice(ctx->device);
  Arg<double> x, b, d;
  FCP_TRY(arg_in(ctx, (const double *)fi, (size_t)ldfi * nrhs, x));
  FCP_TRY(arg_in(ctx, rhs, (size_t)ctx->n * nrhs, b));
  FCP_TRY(arg_in(ctx, diagv, (size_t)ctx->n * nrhs, d));
  fcp_solve_opts local;
  fcp_solve_opts_default(&local);
  fcp_solve_opts *o = opts ? opts : &local;
  int rc = fcp_solve_multi_device(ctx, code, nrhs, x.dev, ldfi, b.dev, res0, itr_max, tol_abs,
                                  tol_rel, verbose, o, d.dev);
  if (rc == FCP_ERR_UNSUPPORTED && o->precision == FCP_PREC_FP64) {
    // the batched sweeps do not apply to this pattern (e.g. their schedule
    // does not fit): one column at a time on the matrix with that column's
    // diagonal, as the reference's separate csrsolve calls run; the resident
    // matrix keeps the last column's diagonal
    ctx->mdc_active = false;
    int worst = 0;
    double rel = 0.0;
    int its = 0;
    for (int c = 0; c < nrhs; ++c) {
      k_set_diag<<<grid_for(ctx->n, 256), 256, 0, ctx->stream>>>(
          ctx->n, ctx->diag.p, d.dev + (size_t)c * ctx->n, ctx->a.p);
      FCP_HIP(ctx, hipGetLastError());
      ctx->a_version++;
      const int rc1 = fcp_solve_device(ctx, code, x.dev + (size_t)c * ldfi,
                                       b.dev + (size_t)c * ctx->n, itr_max, tol_abs, tol_rel,
                                       verbose, o);
      if (rc1 < 0) return rc1;
      if (res0) res0[c] = o->res0;
      worst = rc1 > worst ? rc1 : worst;
      rel = o->rel_res_l2 > rel ? o->rel_res_l2 : rel;
      its = o->iterations > its ? o->iterations : its;
    }
    o->rel_res_l2 = rel;
    o->iterations = its;
    rc = worst;
  }
  if (rc < 0) return rc;
  FCP_TRY(arg_back(ctx, x));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (verbose >= 1)
    printf("  %s x%d (own diagonals):  Initial residual = %10.3E, worst |b-Ax|/|b| = %10.3E, "
           "No Iterations %d\n",
           code == 0 ? "PCG(IC0)" : "BiCGStab(ILU(0))", nrhs, o->res0, o->rel_res_l2,
           o->iterations);
  return rc;
}

// Explicit-array entry points on a library-owned context (device 0).
static std::mutex g_explicit_mu;
static fcp_ctx *g_explicit_ctx = nullptr;

static int explicit_solve(int code, int n, int nnz, const int *ia, const int *ja, const double *a,
                          const int *diag, double *fi, const double *rhs, double *resor,
                          int itr_max, double tol_abs, double tol_rel, int verbose, int mr = 4) {
  std::lock_guard<std::mutex> lk(g_explicit_mu);
  if (n <= 0 || nnz <= 0 || !ia || !ja || !a || !diag || !fi || !rhs) return FCP_ERR_ARG;
  if (fcp_is_device_ptr(ia) || fcp_is_device_ptr(ja) || fcp_is_device_ptr(diag)) return FCP_ERR_ARG;
  if (!csr_ok_host(n, nnz, ia, ja, diag)) return FCP_ERR_ARG;
  if (!g_explicit_ctx) {
    int rc = fcp_init(0, &g_explicit_ctx);
    if (rc < 0) return rc;
  }
  fcp_ctx *ctx = g_explicit_ctx;
  (void)hipSetDevice(ctx->device);
  ctx->have_mesh = true;
  ctx->have_csr = false;
  ctx->n = n;
  ctx->nnz = nnz;
  ctx->Fi = 0;
  ctx->Fb = 0;
  FCP_HIP(ctx, ctx->ia.alloc(n + 1));
  FCP_HIP(ctx, ctx->ja.alloc(nnz));
  FCP_HIP(ctx, ctx->diag.alloc(n));
  FCP_HIP(ctx, ctx->tslot.alloc(nnz));
  FCP_HIP(ctx, ctx->row_of.alloc(nnz));
  FCP_HIP(ctx, ctx->a.alloc(nnz));
  {
    Arg<int> i1, j1, d1;
    FCP_TRY(arg_in(ctx, ia, (size_t)n + 1, i1));
    FCP_TRY(arg_in(ctx, ja, nnz, j1));
    FCP_TRY(arg_in(ctx, diag, n, d1));
    k_sub1<<<grid_for(n + 1, 256), 256, 0, ctx->stream>>>(i1.dev, ctx->ia.p, (size_t)n + 1);
    k_sub1<<<grid_for(nnz, 256), 256, 0, ctx->stream>>>(j1.dev, ctx->ja.p, nnz);
    k_sub1<<<grid_for(n, 256), 256, 0, ctx->stream>>>(d1.dev, ctx->diag.p, n);
    FCP_HIP(ctx, hipGetLastError());
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  // transposed-entry positions for DILU; a missing one marks the pattern
  // non-symmetric (the sweeps then publish every row)
  {
    std::vector<int> hia(n + 1), hja(nnz), ht(nnz);
    for (int i = 0; i <= n; ++i) hia[i] = ia[i] - 1;
    for (int k = 0; k < nnz; ++k) hja[k] = ja[k] - 1;
    bool sym = true;
    for (int i = 0; i < n; ++i)
      for (int k = hia[i]; k < hia[i + 1]; ++k) {
        const int j = hja[k];
        int l = -1;
        for (int m = hia[j]; m < hia[j + 1]; ++m)
          if (hja[m] == i) {
            l = m;
            break;
          }
        ht[k] = l;
        if (l < 0) sym = false;
      }
    ctx->sym_pattern = sym;
    FCP_HIP(ctx, hipMemcpyAsync(ctx->tslot.p, ht.data(), sizeof(int) * nnz, hipMemcpyHostToDevice,
                                ctx->stream));
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  FCP_TRY(fcp_build_schedules(ctx));
  FCP_TRY(fcp_build_sell(ctx));
  ctx->have_csr = true;
  ctx->asm_state = 0;
  FCP_TRY(fcp_matrix_set(ctx, a));
  fcp_solve_opts o;
  fcp_solve_opts_default(&o);
  Arg<double> x, b;
  FCP_TRY(arg_in(ctx, (const double *)fi, n, x));
  FCP_TRY(arg_in(ctx, rhs, n, b));
  ctx->gm_restart = mr;
  const int rc = fcp_solve_device(ctx, code, x.dev, b.dev, itr_max, tol_abs, tol_rel, verbose, &o);
  ctx->gm_restart = 4;
  if (rc < 0) return rc;
  FCP_TRY(arg_back(ctx, x));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (resor) *resor = o.resor;
  return rc;
}

int fcp_iccg(int n, int nnz, const int *ia, const int *ja, const double *a, const int *diag,
             double *fi, const double *rhs, double *resor, int itr_max, double tol_abs,
             double tol_rel, int verbose) {
  return explicit_solve(0, n, nnz, ia, ja, a, diag, fi, rhs, resor, itr_max, tol_abs, tol_rel, verbose);
}
int fcp_dpcg(int n, int nnz, const int *ia, const int *ja, const double *a, const int *diag,
             double *fi, const double *rhs, double *resor, int itr_max, double tol_abs,
             double tol_rel, int verbose) {
  return explicit_solve(1, n, nnz, ia, ja, a, diag, fi, rhs, resor, itr_max, tol_abs, tol_rel, verbose);
}
int fcp_bicgstab(int n, int nnz, const int *ia, const int *ja, const double *a, const int *diag,
                 double *fi, const double *rhs, double *resor, int itr_max, double tol_abs,
                 double tol_rel, int verbose) {
  return explicit_solve(2, n, nnz, ia, ja, a, diag, fi, rhs, resor, itr_max, tol_abs, tol_rel, verbose);
}
int fcp_gauss_seidel(int n, int nnz, const int *ia, const int *ja, const double *a,
                     const int *diag, double *fi, const double *rhs, double *resor, int itr_max,
                     double tol_abs, double tol_rel, int verbose) {
  return explicit_solve(3, n, nnz, ia, ja, a, diag, fi, rhs, resor, itr_max, tol_abs, tol_rel, verbose);
}
int fcp_pmgmres(int n, int nnz, const int *ia, const int *ja, const double *a, const int *diag,
                double *fi, const double *rhs, double *resor, int itr_max, int mr, double tol_abs,
                double tol_rel, int verbose) {
  return explicit_solve(4, n, nnz, ia, ja, a, diag, fi, rhs, resor, itr_max, tol_abs, tol_rel,
                        verbose, mr);
}

}  // extern "C"
