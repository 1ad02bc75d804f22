// fcp_mesh.hip -- mesh connectivity, FP64 geometry, CSR pattern, face->slot
// maps and level schedules, all built on the GPU.
//
// Determinism: the reference scatters per face into owner/neighbour cells in
// ascending face order (geometry.f90:416-530).  Here every cell gathers its
// faces in ascending face id from the cell->face adjacency, performing the
// same floating-point operations in the same order, so results are bit-equal
// without atomics on floating-point data.
#include <algorithm>
#include <cmath>
#include <tuple>
#include <vector>

#include "fcp_internal.h"

namespace {

constexpr int BLK = 256;

__global__ void k_cf_count(const int *__restrict__ owner, const int *__restrict__ neighbour,
                           int F, int Fi, int *__restrict__ cnt) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  atomicAdd(&cnt[owner[f]], 1);
  if (f < Fi) atomicAdd(&cnt[neighbour[f]], 1);
}

__global__ void k_cf_fill(const int *__restrict__ owner, const int *__restrict__ neighbour,
                          int F, int Fi, const int *__restrict__ ptr, int *__restrict__ fill,
                          int *__restrict__ cf) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  int c = owner[f];
  cf[ptr[c] + atomicAdd(&fill[c], 1)] = f;
  if (f < Fi) {
    c = neighbour[f];
    cf[ptr[c] + atomicAdd(&fill[c], 1)] = f;
  }
}

// Sort each cell's face list ascending (insertion sort, short lists).
__global__ void k_cf_sort(const int *__restrict__ ptr, int *__restrict__ cf, int n) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  int b = ptr[c], e = ptr[c + 1];
  for (int k = b + 1; k < e; ++k) {
    int v = cf[k];
    int l = k - 1;
    while (l >= b && cf[l] > v) {
      cf[l + 1] = cf[l];
      --l;
    }
    cf[l + 1] = v;
  }
}

// Face area vectors and centroids by fan triangulation from the first vertex
// (geometry.f90:416-516, triangle_area_vector :1721-1749).
__global__ void k_face_geom(const double *__restrict__ P, const int *__restrict__ nnodes,
                            const int *__restrict__ node, int F, double *__restrict__ arx,
                            double *__restrict__ ary, double *__restrict__ arz,
                            double *__restrict__ xf, double *__restrict__ yf,
                            double *__restrict__ zf) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const int *nd = node + (size_t)f * FCP_NOMAX;
  const double third = 1.0 / 3.0;
  const double x1 = P[3 * nd[0]], y1 = P[3 * nd[0] + 1], z1 = P[3 * nd[0] + 2];
  double sx = 0.0, sy = 0.0, sz = 0.0, fx = 0.0, fy = 0.0, fz = 0.0, areasum = 0.0;
  const int nt = nnodes[f] - 2;
  for (int i = 0; i < nt; ++i) {
    const double *pb = P + 3 * nd[i + 1];
    const double *pc = P + 3 * nd[i + 2];
    const double px = pb[0] - x1, py = pb[1] - y1, pz = pb[2] - z1;
    const double qx = pc[0] - x1, qy = pc[1] - y1, qz = pc[2] - z1;
    const double nx = 0.5 * (py * qz - pz * qy);
    const double ny = 0.5 * (pz * qx - px * qz);
    const double nz = 0.5 * (px * qy - py * qx);
    sx += nx;
    sy += ny;
    sz += nz;
    const double cx = third * ((pc[0] + pb[0]) + x1);
    const double cy = third * ((pc[1] + pb[1]) + y1);
    const double cz = third * ((pc[2] + pb[2]) + z1);
    const double are = sqrt((nx * nx + ny * ny) + nz * nz);
    fx += are * cx;
    fy += are * cy;
    fz += are * cz;
    areasum += are;
  }
  arx[f] = sx;
  ary[f] = sy;
  arz[f] = sz;
  xf[f] = fx / areasum;
  yf[f] = fy / areasum;
  zf[f] = fz / areasum;
}

// Cell volume and centroid (Wang 1999 as in geometry.f90:467-530): gather
// over the cell's faces in ascending face id, triangles in fan order.
__global__ void k_cell_geom(const double *__restrict__ P, const int *__restrict__ nnodes,
                            const int *__restrict__ node, const int *__restrict__ owner,
                            const int *__restrict__ cf_ptr, const int *__restrict__ cf, int n,
                            double *__restrict__ xc, double *__restrict__ yc,
                            double *__restrict__ zc, double *__restrict__ vol) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double third = 1.0 / 3.0;
  double v = 0.0, sx = 0.0, sy = 0.0, sz = 0.0, den = 0.0;
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int f = cf[k];
    const bool own = (owner[f] == c);
    const int *nd = node + (size_t)f * FCP_NOMAX;
    const double x1 = P[3 * nd[0]], y1 = P[3 * nd[0] + 1], z1 = P[3 * nd[0] + 2];
    const int nt = nnodes[f] - 2;
    for (int i = 0; i < nt; ++i) {
      const double *pb = P + 3 * nd[i + 1];
      const double *pc = P + 3 * nd[i + 2];
      const double px = pb[0] - x1, py = pb[1] - y1, pz = pb[2] - z1;
      const double qx = pc[0] - x1, qy = pc[1] - y1, qz = pc[2] - z1;
      const double nx = 0.5 * (py * qz - pz * qy);
      const double ny = 0.5 * (pz * qx - px * qz);
      const double nz = 0.5 * (px * qy - py * qx);
      const double cx = third * ((pc[0] + pb[0]) + x1);
      const double cy = third * ((pc[1] + pb[1]) + y1);
      const double cz = third * ((pc[2] + pb[2]) + z1);
      double riSi = (cx * nx + cy * ny) + cz * nz;
      if (!own) riSi = -riSi;
      v += third * riSi;
      sx += 0.75 * riSi * cx;
      sy += 0.75 * riSi * cy;
      sz += 0.75 * riSi * cz;
      den += riSi;
    }
  }
  vol[c] = v;
  xc[c] = sx / den;
  yc[c] = sy / den;
  zc[c] = sz / den;
}

// Interpolation factor (variant 2, geometry.f90:581-606), Df (:648-664) and
// boundary normal distance / |S|/dn (:701-754).
__global__ void k_face_geom2(const int *__restrict__ owner, const int *__restrict__ neighbour,
                             int F, int Fi, const double *__restrict__ xc,
                             const double *__restrict__ yc, const double *__restrict__ zc,
                             const double *__restrict__ arx, const double *__restrict__ ary,
                             const double *__restrict__ arz, const double *__restrict__ xf,
                             const double *__restrict__ yf, const double *__restrict__ zf,
                             double *__restrict__ facint, double *__restrict__ Df,
                             double *__restrict__ dnb, double *__restrict__ srdb) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const int p = owner[f];
  if (f < Fi) {
    const int q = neighbour[f];
    double dx = xf[f] - xc[p], dy = yf[f] - yc[p], dz = zf[f] - zc[p];
    const double djp = sqrt((dx * dx + dy * dy) + dz * dz);
    dx = xf[f] - xc[q];
    dy = yf[f] - yc[q];
    dz = zf[f] - zc[q];
    const double djn = sqrt((dx * dx + dy * dy) + dz * dz);
    facint[f] = djp / (djp + djn);
    const double xpn = xc[q] - xc[p], ypn = yc[q] - yc[p], zpn = zc[q] - zc[p];
    const double are = (arx[f] * arx[f] + ary[f] * ary[f]) + arz[f] * arz[f];
    Df[f] = are / ((arx[f] * xpn + ary[f] * ypn) + arz[f] * zpn);
  } else {
    const double are = sqrt((arx[f] * arx[f] + ary[f] * ary[f]) + arz[f] * arz[f]);
    const double nxf = arx[f] / are, nyf = ary[f] / are, nzf = arz[f] / are;
    const double dn = ((xf[f] - xc[p]) * nxf + (yf[f] - yc[p]) * nyf) + (zf[f] - zc[p]) * nzf;
    dnb[f - Fi] = dn;
    srdb[f - Fi] = are / dn;
  }
}

// ---- periodic face mapping (geometry.f90:1848-1997) -------------------------
// Face i of a periodic patch (faces s0..s0+nf-1) takes the twin face (t0..)
// whose owner centre is nearest to its centre: L1 distance summed x, y, z,
// first minimum below the reference's start value 1e+30 (a default-real
// literal).  The search reads the twin owners before any is moved.
__global__ void k_face_map_search(int nf, int s0, int t0, const int *__restrict__ owner,
                                  const double *__restrict__ xf, const double *__restrict__ yf,
                                  const double *__restrict__ zf, const double *__restrict__ xc,
                                  const double *__restrict__ yc, const double *__restrict__ zc,
                                  int *__restrict__ tmp, int *__restrict__ itmp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  const double fx = xf[s0 + i], fy = yf[s0 + i], fz = zf[s0 + i];
  double mindist = (double)1e30f;
  int bt = owner[t0 + i], bi = i;
  for (int it = 0; it < nf; ++it) {
    const int ijn = owner[t0 + it];
    const double dist = (fabs(fx - xc[ijn]) + fabs(fy - yc[ijn])) + fabs(fz - zc[ijn]);
    if (dist < mindist) {
      mindist = dist;
      bt = ijn;
      bi = it;
    }
  }
  tmp[i] = bt;
  itmp[i] = bi;
}

// The twin faces' centres and area vectors, copied aside (6 x nf)...
__global__ void k_face_map_save(int nf, int t0, const double *__restrict__ xf,
                                const double *__restrict__ yf, const double *__restrict__ zf,
                                const double *__restrict__ arx, const double *__restrict__ ary,
                                const double *__restrict__ arz, double *__restrict__ save) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  save[i] = xf[t0 + i];
  save[nf + i] = yf[t0 + i];
  save[2 * nf + i] = zf[t0 + i];
  save[3 * nf + i] = arx[t0 + i];
  save[4 * nf + i] = ary[t0 + i];
  save[5 * nf + i] = arz[t0 + i];
}
// ...then twin face i takes the owner and the geometry of twin face itmp[i].
__global__ void k_face_map_apply(int nf, int t0, const int *__restrict__ tmp,
                                 const int *__restrict__ itmp, const double *__restrict__ save,
                                 int *__restrict__ owner, double *__restrict__ xf,
                                 double *__restrict__ yf, double *__restrict__ zf,
                                 double *__restrict__ arx, double *__restrict__ ary,
                                 double *__restrict__ arz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  const int j = itmp[i];
  owner[t0 + i] = tmp[i];
  xf[t0 + i] = save[j];
  yf[t0 + i] = save[nf + j];
  zf[t0 + i] = save[2 * nf + j];
  arx[t0 + i] = save[3 * nf + j];
  ary[t0 + i] = save[4 * nf + j];
  arz[t0 + i] = save[5 * nf + j];
}

// ---- CSR ------------------------------------------------------------------
// Entries of row c: the diagonal, the other cell of each inner face and the
// partner of each periodic pair it belongs to (tw: pairs whose twin face it
// owns; its periodic faces are found through kof).
__global__ void k_row_count(const int *__restrict__ cf_ptr, const int *__restrict__ cf, int n,
                            int Fi, PerP per, int *__restrict__ cnt) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > n) return;
  if (c == n) {
    cnt[n] = 0;
    return;
  }
  int m = 1;
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int f = cf[k];
    m += (f < Fi) || (per.np > 0 && per.kof[f - Fi] >= 0);
  }
  if (per.np > 0) m += per.tw_ptr[c + 1] - per.tw_ptr[c];
  cnt[c] = m;
}

// Row c holds {c} and the other cell of each inner face and periodic pair,
// sorted ascending (the unique result of the reference's lexical COO sort,
// utils.f90:443); a repeated column (a periodic pair doubling an entry)
// raises err.
__global__ void k_row_fill(const int *__restrict__ cf_ptr, const int *__restrict__ cf,
                           const int *__restrict__ owner, const int *__restrict__ neighbour,
                           int n, int Fi, PerP per, const int *__restrict__ ia,
                           int *__restrict__ ja, int *__restrict__ diag, int *__restrict__ err) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int b = ia[c];
  int e = b;
  ja[e++] = c;
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int f = cf[k];
    if (f >= Fi) {
      const int kp = per.np > 0 ? per.kof[f - Fi] : -1;
      if (kp >= 0) ja[e++] = owner[per.ft[kp]];
      continue;
    }
    ja[e++] = (owner[f] == c) ? neighbour[f] : owner[f];
  }
  if (per.np > 0)
    for (int k = per.tw_ptr[c]; k < per.tw_ptr[c + 1]; ++k) ja[e++] = owner[per.fp[per.tw_k[k]]];
  for (int k = b + 1; k < e; ++k) {
    int v = ja[k];
    int l = k - 1;
    while (l >= b && ja[l] > v) {
      ja[l + 1] = ja[l];
      --l;
    }
    ja[l + 1] = v;
  }
  for (int k = b + 1; k < e; ++k)
    if (ja[k] == ja[k - 1]) atomicOr(err, 1);
  int dpos = -1;
  for (int k = b; k < e; ++k)
    if (ja[k] == c) {
      dpos = k;
      break;
    }
  diag[c] = dpos;
}

__device__ __forceinline__ int first_in_row(const int *ja, int b, int e, int col) {
  // lower_bound: first position with ja >= col; rows are sorted ascending
  while (b < e) {
    int m = (b + e) >> 1;
    if (ja[m] < col)
      b = m + 1;
    else
      e = m;
  }
  return b;
}

// csr_to_k (utils.f90:96-146): first match of (row,col).
__global__ void k_slots(const int *__restrict__ owner, const int *__restrict__ neighbour, int Fi,
                        const int *__restrict__ ia, const int *__restrict__ ja,
                        int *__restrict__ slot_pn, int *__restrict__ slot_np) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= Fi) return;
  const int p = owner[f], q = neighbour[f];
  int k = first_in_row(ja, ia[p], ia[p + 1], q);
  slot_pn[f] = (k < ia[p + 1] && ja[k] == q) ? k : -1;
  k = first_in_row(ja, ia[q], ia[q + 1], p);
  slot_np[f] = (k < ia[q + 1] && ja[k] == p) ? k : -1;
}

// The same for periodic pair k: (owner of its face, owner of its twin face)
// (sparse_matrix.f90:262-293), stored after the inner faces' slots.
__global__ void k_slots_periodic(PerP per, const int *__restrict__ owner, int Fi,
                                 const int *__restrict__ ia, const int *__restrict__ ja,
                                 int *__restrict__ slot_pn, int *__restrict__ slot_np) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= per.np) return;
  const int p = owner[per.fp[k]], q = owner[per.ft[k]];
  int j = first_in_row(ja, ia[p], ia[p + 1], q);
  slot_pn[Fi + k] = (j < ia[p + 1] && ja[j] == q) ? j : -1;
  j = first_in_row(ja, ia[q], ia[q + 1], p);
  slot_np[Fi + k] = (j < ia[q + 1] && ja[j] == p) ? j : -1;
}

__global__ void k_row_of_tslot(const int *__restrict__ ia, const int *__restrict__ ja, int n,
                               int *__restrict__ row_of, int *__restrict__ tslot) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = ia[i]; k < ia[i + 1]; ++k) {
    row_of[k] = i;
    const int j = ja[k];
    int l = first_in_row(ja, ia[j], ia[j + 1], i);
    tslot[k] = (l < ia[j + 1] && ja[l] == i) ? l : -1;
  }
}

// ---- level sets -------------------------------------------------------------
// lev[i] = 0 without strict-lower (upper) entries, else 1 + max over them.
// In-place monotone relaxation: every value stays a lower bound of the true
// level, so the fixed point is the exact level and independent of timing.
__global__ void k_levels(const int *__restrict__ ia, const int *__restrict__ diag,
                         const int *__restrict__ ja, int n, int upper, int *__restrict__ lev,
                         int *__restrict__ changed) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = upper ? diag[i] + 1 : ia[i];
  const int e = upper ? ia[i + 1] : diag[i];
  int l = 0;
  for (int k = b; k < e; ++k) {
    const int v = lev[ja[k]] + 1;
    l = v > l ? v : l;
  }
  if (l != lev[i]) {
    lev[i] = l;
    *changed = 1;
  }
}

__global__ void k_iota(int *__restrict__ v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

__global__ void k_hist(const int *__restrict__ lev, int n, int *__restrict__ cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&cnt[lev[i]], 1);
}

__global__ void k_max(const int *__restrict__ v, int n, int *__restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicMax(out, v[i]);
}

}  // namespace

int fcp_build_cell_faces(fcp_ctx *ctx) {
  const int n = ctx->n, F = ctx->F, Fi = ctx->Fi;
  DBuf<int> cnt, fill;
  FCP_HIP(ctx, cnt.alloc(n + 1));
  FCP_HIP(ctx, fill.alloc(n));
  FCP_HIP(ctx, ctx->cf_ptr.alloc(n + 1));
  FCP_HIP(ctx, ctx->cf_face.alloc((size_t)2 * Fi + (F - Fi)));
  FCP_HIP(ctx, hipMemsetAsync(cnt.p, 0, sizeof(int) * (n + 1), ctx->stream));
  FCP_HIP(ctx, hipMemsetAsync(fill.p, 0, sizeof(int) * n, ctx->stream));
  k_cf_count<<<grid_for(F, BLK), BLK, 0, ctx->stream>>>(ctx->owner.p, ctx->neighbour.p, F, Fi, cnt.p);
  FCP_TRY(fcp_exclusive_scan_int(ctx, cnt.p, ctx->cf_ptr.p, n + 1));
  k_cf_fill<<<grid_for(F, BLK), BLK, 0, ctx->stream>>>(ctx->owner.p, ctx->neighbour.p, F, Fi,
                                                         ctx->cf_ptr.p, fill.p, ctx->cf_face.p);
  k_cf_sort<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(ctx->cf_ptr.p, ctx->cf_face.p, n);
  FCP_HIP(ctx, hipGetLastError());
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

PerP fcp_periodic_view(const fcp_ctx *ctx) {
  if (ctx->nper == 0) return PerP{0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  return PerP{ctx->nper,        ctx->per_kof.p, ctx->per_twb.p,   ctx->per_fp.p,
              ctx->per_ft.p,    ctx->per_ii.p,  ctx->per_twptr.p, ctx->per_twk.p};
}

// The periodic pairs after face mapping (host tables uploaded once per mesh).
static int build_periodic(fcp_ctx *ctx) {
  const int n = ctx->n, F = ctx->F, Fi = ctx->Fi, Fb = ctx->Fb;
  std::vector<int> own(F);
  FCP_HIP(ctx, hipMemcpy(own.data(), ctx->owner.p, sizeof(int) * F, hipMemcpyDeviceToHost));
  std::vector<int> kof(Fb > 0 ? Fb : 1, -1), twb(Fb > 0 ? Fb : 1, -1), fp, ft, ii, cnt(n + 1, 0);
  std::vector<int> patch_of(Fb > 0 ? Fb : 1, -1);
  for (int ib = 0; ib < ctx->nPatches; ++ib)
    for (int i = 0; i < ctx->h_nfaces[ib]; ++i) patch_of[ctx->h_start[ib] - Fi + i] = ib;
  int iper = 0;
  for (int ib = 0; ib < ctx->nPatches; ++ib) {
    if (ctx->h_bctype[ib] != FCP_BC_PERIODIC) continue;
    const int s0 = ctx->h_start[ib], t0 = ctx->h_twin[iper++], nf = ctx->h_nfaces[ib];
    if (nf > Fi)  // the reference reads facint / Df of inner face i for periodic face i
      return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "periodic patch with more faces than inner faces");
    const int tb = patch_of[t0 - Fi], tt = ctx->h_bctype[tb];
    // the twin's own zero-gradient copy (updateBoundary.f90:31-44) wins when
    // its patch comes after this one
    const bool own_last = tb > ib && (tt == FCP_BC_OUTLET || tt == FCP_BC_SYMMETRY ||
                                      tt == FCP_BC_PRESSURE || tt == FCP_BC_EMPTY);
    for (int i = 0; i < nf; ++i) {
      const int k = (int)fp.size();
      if (own[s0 + i] == own[t0 + i])
        return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "periodic pair inside one cell");
      kof[s0 - Fi + i] = k;
      if (!own_last) twb[t0 - Fi + i] = k;
      fp.push_back(s0 + i);
      ft.push_back(t0 + i);
      ii.push_back(i);
      cnt[own[t0 + i] + 1]++;
    }
  }
  const int np = (int)fp.size();
  std::vector<int> twptr(n + 1, 0), twk(np > 0 ? np : 1);
  for (int c = 0; c < n; ++c) twptr[c + 1] = twptr[c] + cnt[c + 1];
  std::vector<int> fill(twptr.begin(), twptr.end() - 1);
  for (int k = 0; k < np; ++k) twk[fill[own[ft[k]]]++] = k;  // ascending k = ascending face
  ctx->nper = np;
  auto up = [&](DBuf<int> &d, const std::vector<int> &h) -> int {
    FCP_HIP(ctx, d.alloc(h.size()));
    FCP_HIP(ctx, hipMemcpy(d.p, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    return 0;
  };
  FCP_TRY(up(ctx->per_kof, kof));
  FCP_TRY(up(ctx->per_twb, twb));
  FCP_TRY(up(ctx->per_fp, fp));
  FCP_TRY(up(ctx->per_ft, ft));
  FCP_TRY(up(ctx->per_ii, ii));
  FCP_TRY(up(ctx->per_twptr, twptr));
  FCP_TRY(up(ctx->per_twk, twk));
  return 0;
}

// face_mapping after the geometry (read_mesh, geometry.f90:110-112): patch
// by patch in order, as the reference maps them, then the cell -> face
// lists of the mapped owners.
static int map_periodic_faces(fcp_ctx *ctx) {
  int iper = 0;
  for (int ib = 0; ib < ctx->nPatches; ++ib) {
    if (ctx->h_bctype[ib] != FCP_BC_PERIODIC) continue;
    const int s0 = ctx->h_start[ib], t0 = ctx->h_twin[iper++], nf = ctx->h_nfaces[ib];
    if (nf == 0) continue;
    DBuf<int> tmp, itmp;
    DBuf<double> save;
    FCP_HIP(ctx, tmp.alloc(nf));
    FCP_HIP(ctx, itmp.alloc(nf));
    FCP_HIP(ctx, save.alloc((size_t)6 * nf));
    k_face_map_search<<<grid_for(nf, BLK), BLK, 0, ctx->stream>>>(
        nf, s0, t0, ctx->owner.p, ctx->xf.p, ctx->yf.p, ctx->zf.p, ctx->xc.p, ctx->yc.p,
        ctx->zc.p, tmp.p, itmp.p);
    k_face_map_save<<<grid_for(nf, BLK), BLK, 0, ctx->stream>>>(
        nf, t0, ctx->xf.p, ctx->yf.p, ctx->zf.p, ctx->arx.p, ctx->ary.p, ctx->arz.p, save.p);
    k_face_map_apply<<<grid_for(nf, BLK), BLK, 0, ctx->stream>>>(
        nf, t0, tmp.p, itmp.p, save.p, ctx->owner.p, ctx->xf.p, ctx->yf.p, ctx->zf.p, ctx->arx.p,
        ctx->ary.p, ctx->arz.p);
    FCP_HIP(ctx, hipGetLastError());
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));  // tmp, itmp, save freed on return
  }
  FCP_TRY(fcp_build_cell_faces(ctx));
  return build_periodic(ctx);
}

int fcp_compute_geometry(fcp_ctx *ctx) {
  const int n = ctx->n, F = ctx->F, Fi = ctx->Fi, Fb = ctx->Fb;
  ctx->nper = 0;
  if (!ctx->h_twin.empty()) {  // geometry from the owners as uploaded
    FCP_HIP(ctx, hipMemcpyAsync(ctx->owner.p, ctx->owner0.p, sizeof(int) * F,
                                hipMemcpyDeviceToDevice, ctx->stream));
    FCP_TRY(fcp_build_cell_faces(ctx));
  }
  FCP_HIP(ctx, ctx->xc.alloc(n));
  FCP_HIP(ctx, ctx->yc.alloc(n));
  FCP_HIP(ctx, ctx->zc.alloc(n));
  FCP_HIP(ctx, ctx->vol.alloc(n));
  FCP_HIP(ctx, ctx->arx.alloc(F));
  FCP_HIP(ctx, ctx->ary.alloc(F));
  FCP_HIP(ctx, ctx->arz.alloc(F));
  FCP_HIP(ctx, ctx->xf.alloc(F));
  FCP_HIP(ctx, ctx->yf.alloc(F));
  FCP_HIP(ctx, ctx->zf.alloc(F));
  FCP_HIP(ctx, ctx->facint.alloc(Fi > 0 ? Fi : 1));
  FCP_HIP(ctx, ctx->Df.alloc(Fi > 0 ? Fi : 1));
  FCP_HIP(ctx, ctx->dnb.alloc(Fb > 0 ? Fb : 1));
  FCP_HIP(ctx, ctx->srdb.alloc(Fb > 0 ? Fb : 1));
  {
    ProfScope ps(ctx, "face_geom");
    k_face_geom<<<grid_for(F, BLK), BLK, 0, ctx->stream>>>(ctx->points.p, ctx->nnodes.p, ctx->node.p,
                                                             F, ctx->arx.p, ctx->ary.p, ctx->arz.p,
                                                             ctx->xf.p, ctx->yf.p, ctx->zf.p);
  }
  {
    ProfScope ps(ctx, "cell_geom");
    k_cell_geom<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(
        ctx->points.p, ctx->nnodes.p, ctx->node.p, ctx->owner.p, ctx->cf_ptr.p, ctx->cf_face.p, n,
        ctx->xc.p, ctx->yc.p, ctx->zc.p, ctx->vol.p);
  }
  {
    ProfScope ps(ctx, "face_geom2");
    k_face_geom2<<<grid_for(F, BLK), BLK, 0, ctx->stream>>>(
        ctx->owner.p, ctx->neighbour.p, F, Fi, ctx->xc.p, ctx->yc.p, ctx->zc.p, ctx->arx.p,
        ctx->ary.p, ctx->arz.p, ctx->xf.p, ctx->yf.p, ctx->zf.p, ctx->facint.p, ctx->Df.p,
        ctx->dnb.p, ctx->srdb.p);
  }
  FCP_HIP(ctx, hipGetLastError());
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (!ctx->h_twin.empty()) FCP_TRY(map_periodic_faces(ctx));
  ctx->have_geom = true;
  return 0;
}

int fcp_build_csr(fcp_ctx *ctx) {
  FCP_TRY(fcp_coarse_wait(ctx));
  const int n = ctx->n, Fi = ctx->Fi;
  // (owner, neighbour) and (neighbour, owner) for every inner face and
  // periodic pair
  ctx->sym_pattern = true;
  int nperp = 0;
  for (int ib = 0; ib < ctx->nPatches; ++ib) nperp += ctx->h_bctype[ib] == FCP_BC_PERIODIC;
  if ((int)ctx->h_twin.size() != nperp)
    return fcp_fail(ctx, FCP_ERR_STATE, "periodic patches without their twins (fcp_mesh_periodic)");
  if (nperp > 0 && !ctx->have_geom)
    return fcp_fail(ctx, FCP_ERR_STATE, "periodic patches: compute the geometry (face mapping) first");
  const PerP per = fcp_periodic_view(ctx);
  const int np = per.np;
  const long long nnz = 2LL * (Fi + np) + n;  // sparse_matrix.f90:110
  if (nnz > 0x7fffffffLL) return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "nnz exceeds int32");
  ctx->nnz = (int)nnz;
  DBuf<int> cnt;
  FCP_HIP(ctx, cnt.alloc(n + 1));
  FCP_HIP(ctx, ctx->ia.alloc(n + 1));
  FCP_HIP(ctx, ctx->ja.alloc(nnz));
  FCP_HIP(ctx, ctx->diag.alloc(n));
  FCP_HIP(ctx, ctx->slot_pn.alloc(Fi + np > 0 ? Fi + np : 1));
  FCP_HIP(ctx, ctx->slot_np.alloc(Fi + np > 0 ? Fi + np : 1));
  FCP_HIP(ctx, ctx->row_of.alloc(nnz));
  FCP_HIP(ctx, ctx->tslot.alloc(nnz));
  DBuf<int> dup;
  FCP_HIP(ctx, dup.alloc(1));
  FCP_HIP(ctx, hipMemsetAsync(dup.p, 0, sizeof(int), ctx->stream));
  k_row_count<<<grid_for(n + 1, BLK), BLK, 0, ctx->stream>>>(ctx->cf_ptr.p, ctx->cf_face.p, n, Fi,
                                                               per, cnt.p);
  FCP_TRY(fcp_exclusive_scan_int(ctx, cnt.p, ctx->ia.p, n + 1));
  k_row_fill<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(ctx->cf_ptr.p, ctx->cf_face.p, ctx->owner.p,
                                                          ctx->neighbour.p, n, Fi, per, ctx->ia.p,
                                                          ctx->ja.p, ctx->diag.p, dup.p);
  if (Fi > 0)
    k_slots<<<grid_for(Fi, BLK), BLK, 0, ctx->stream>>>(ctx->owner.p, ctx->neighbour.p, Fi, ctx->ia.p,
                                                          ctx->ja.p, ctx->slot_pn.p, ctx->slot_np.p);
  if (np > 0)
    k_slots_periodic<<<grid_for(np, BLK), BLK, 0, ctx->stream>>>(
        per, ctx->owner.p, Fi, ctx->ia.p, ctx->ja.p, ctx->slot_pn.p, ctx->slot_np.p);
  k_row_of_tslot<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(ctx->ia.p, ctx->ja.p, n, ctx->row_of.p,
                                                              ctx->tslot.p);
  FCP_HIP(ctx, hipGetLastError());
  int last = 0;
  FCP_HIP(ctx, hipMemcpyAsync(&last, ctx->ia.p + n, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (last != ctx->nnz) return fcp_fail(ctx, FCP_ERR_STATE, "CSR size mismatch %d vs %d", last, ctx->nnz);
  int hdup = 0;
  FCP_HIP(ctx, hipMemcpy(&hdup, dup.p, sizeof(int), hipMemcpyDeviceToHost));
  if (hdup)
    return fcp_fail(ctx, FCP_ERR_UNSUPPORTED,
                    "a periodic pair repeats a matrix entry (cells already coupled)");
  FCP_HIP(ctx, ctx->a.alloc(nnz));
  FCP_HIP(ctx, hipMemsetAsync(ctx->a.p, 0, sizeof(double) * nnz, ctx->stream));
  ctx->a_version++;
  FCP_TRY(fcp_build_schedules(ctx));
  FCP_TRY(fcp_build_sell(ctx));
  ctx->have_csr = true;
  ctx->asm_state = 0;
  return 0;
}

static int build_one_schedule(fcp_ctx *ctx, int upper, Schedule &s) {
  const int n = ctx->n;
  FCP_HIP(ctx, s.lev.alloc(n));
  FCP_HIP(ctx, s.order.alloc(n));
  DBuf<int> flag, iota, keys_out;
  FCP_HIP(ctx, flag.alloc(4));
  FCP_HIP(ctx, iota.alloc(n));
  FCP_HIP(ctx, keys_out.alloc(n));
  FCP_HIP(ctx, hipMemsetAsync(s.lev.p, 0, sizeof(int) * n, ctx->stream));
  int *h_changed = nullptr;
  FCP_HIP(ctx, hipHostMalloc(&h_changed, sizeof(int)));
  // Relax in batches of passes; each pass can only raise values, so a
  // quiescent batch has reached the fixed point.
  for (int pass = 0; pass <= n; pass += 8) {
    FCP_HIP(ctx, hipMemsetAsync(flag.p, 0, sizeof(int), ctx->stream));
    for (int r = 0; r < 8; ++r)
      k_levels<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(ctx->ia.p, ctx->diag.p, ctx->ja.p, n, upper,
                                                            s.lev.p, flag.p);
    FCP_HIP(ctx, hipMemcpyAsync(h_changed, flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!*h_changed) break;
  }
  (void)hipHostFree(h_changed);
  // number of levels
  FCP_HIP(ctx, hipMemsetAsync(flag.p, 0, sizeof(int), ctx->stream));
  k_max<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(s.lev.p, n, flag.p);
  int maxlev = 0;
  FCP_HIP(ctx, hipMemcpyAsync(&maxlev, flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  s.nlev = maxlev + 1;
  // rows sorted by level (stable: ascending row id inside a level)
  k_iota<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(iota.p, n);
  FCP_TRY(fcp_sort_pairs_int(ctx, s.lev.p, keys_out.p, iota.p, s.order.p, n, maxlev));
  DBuf<int> cnt;
  FCP_HIP(ctx, cnt.alloc(s.nlev));
  FCP_HIP(ctx, hipMemsetAsync(cnt.p, 0, sizeof(int) * s.nlev, ctx->stream));
  k_hist<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(s.lev.p, n, cnt.p);
  std::vector<int> h_cnt(s.nlev);
  FCP_HIP(ctx, hipMemcpyAsync(h_cnt.data(), cnt.p, sizeof(int) * s.nlev, hipMemcpyDeviceToHost,
                              ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<int> tp;
  tp.reserve(n / 64 + s.nlev + 1);
  s.lev_ptr.assign(1, 0);
  int off = 0;
  for (int L = 0; L < s.nlev; ++L) {
    for (int k = 0; k < h_cnt[L]; k += 64) tp.push_back(off + k);
    off += h_cnt[L];
    s.lev_ptr.push_back(off);
  }
  tp.push_back(off);
  s.ntasks = (int)tp.size() - 1;
  FCP_HIP(ctx, s.task_ptr.alloc(tp.size()));
  FCP_HIP(ctx, hipMemcpyAsync(s.task_ptr.p, tp.data(), sizeof(int) * tp.size(), hipMemcpyHostToDevice,
                              ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

namespace {
__global__ void k_block_key(const int *__restrict__ order, int n, int R,
                            const int *__restrict__ rank, int *__restrict__ key) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) key[i] = rank ? rank[order[i]] : order[i] / R;
}

__global__ void k_pos_of(const int *__restrict__ rows, int n, int *__restrict__ pos) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pos[rows[i]] = i;
}

// Static part of the packed operands: dependency list of every position in
// processing order (strict-lower entries for the forward sweep, strict-upper
// for the backward sweep), in CSR order.  A dependency inside the row's
// block is stored as its position offset in the block (its LDS slot).
// (row, count | external << 16) per position: external rows have a consumer
// in another block (their values are published during the sweep).  With a
// symmetric pattern the consumers of a forward-sweep row are its strict-upper
// neighbours, those of a backward-sweep row its strict-lower ones; without
// one (sym = 0) every row counts as external.
__global__ void k_pack_static(const int *__restrict__ rows, int n, int upper, int sym, int local,
                              const int *__restrict__ blk_of, const int *__restrict__ blk_pos,
                              const int *__restrict__ pos_of, const int *__restrict__ ia,
                              const int *__restrict__ diag, const int *__restrict__ ja,
                              int *__restrict__ prc, int *__restrict__ pdep,
                              int *__restrict__ psrc, int *__restrict__ maxd) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int row = rows[idx];
  const int blk = blk_of[row];
  const int lo = blk_pos[blk];
  const int b = upper ? diag[row] + 1 : ia[row];
  const int e = upper ? ia[row + 1] : diag[row];
  int c = e - b;
  if (local) {  // only the dependencies inside the block count
    c = 0;
    for (int k = b; k < e; ++k) c += blk_of[ja[k]] == blk;
  }
  int ext = local ? 0 : 1;
  if (sym && !local) {
    ext = 0;
    const int cb = upper ? ia[row] : diag[row] + 1;
    const int ce = upper ? diag[row] : ia[row + 1];
    for (int k = cb; k < ce; ++k)
      if (blk_of[ja[k]] != blk) ext = 1;
  }
  prc[2 * idx] = row;
  prc[2 * idx + 1] = c | (ext << 16);
  atomicMax(maxd, c);
  // the m-th counted dependency, in CSR order
  int k = b;
  for (int m = 0; m < FCP_PMAXD; ++m) {
    int dep = 0, src = -1;
    if (m < c) {
      if (local)
        while (blk_of[ja[k]] != blk) ++k;
      const int j = ja[k];
      dep = (blk_of[j] == blk) ? (pos_of[j] - lo) : -(j + 1);
      src = k;
      ++k;
    }
    pdep[(size_t)FCP_PMAXD * idx + m] = dep;
    psrc[(size_t)FCP_PMAXD * idx + m] = src;
  }
}
}  // namespace

// Row-block schedule from a level schedule: a stable sort of the level-major
// row order by block id gives every block's rows in level order; steps are
// the runs of equal level inside a block.  Blocks are contiguous runs of R
// rows, or -- with tile_rank -- the given tiles in rank order.
int fcp_build_block_schedule(fcp_ctx *ctx, const Schedule &s, int R, int upper, BlockSchedule &b,
                             const int *tile_rank, int ntiles, bool local_only, int step_max) {
  const int n = ctx->n;
  b.tiled = tile_rank != nullptr;
  b.nblocks = tile_rank ? ntiles : (n + R - 1) / R;
  DBuf<int> key, key_out, rank_dev;
  FCP_HIP(ctx, key.alloc(n));
  FCP_HIP(ctx, key_out.alloc(n));
  FCP_HIP(ctx, b.rows.alloc(n));
  if (tile_rank) {
    FCP_HIP(ctx, rank_dev.alloc(n));
    FCP_HIP(ctx, hipMemcpyAsync(rank_dev.p, tile_rank, sizeof(int) * n, hipMemcpyHostToDevice,
                                ctx->stream));
  }
  k_block_key<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(s.order.p, n, R, rank_dev.p, key.p);
  FCP_TRY(fcp_sort_pairs_int(ctx, key.p, key_out.p, s.order.p, b.rows.p, n, b.nblocks));
  std::vector<int> rows(n), lev(n), blk_of(n);
  FCP_HIP(ctx, hipMemcpyAsync(rows.data(), b.rows.p, sizeof(int) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
  FCP_HIP(ctx, hipMemcpyAsync(lev.data(), s.lev.p, sizeof(int) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int r = 0; r < n; ++r) blk_of[r] = tile_rank ? tile_rank[r] : r / R;
  // step table, block starts and the largest block
  std::vector<int> sp, bs(b.nblocks + 1, 0), bp(b.nblocks + 1, n);
  sp.reserve(b.nblocks * 64 + 1);
  int prev_blk = -1, prev_lev = -1;
  for (int i = 0; i < n; ++i) {
    const int r = rows[i], blk = blk_of[r];
    if (blk != prev_blk) {
      for (int q = prev_blk + 1; q <= blk; ++q) {
        bs[q] = (int)sp.size();
        bp[q] = i;
      }
      prev_blk = blk;
      prev_lev = -1;
    }
    // a new step at every level change, and after step_max rows of one
    // level: rows of a level are independent, and a step the sweep kernels
    // take with one row per thread is pipelined (wider steps are not)
    if (lev[r] != prev_lev || i - sp.back() >= step_max) {
      sp.push_back(i);
      prev_lev = lev[r];
    }
  }
  for (int q = prev_blk + 1; q <= b.nblocks; ++q) {
    bs[q] = (int)sp.size();
    bp[q] = n;
  }
  b.nsteps = (int)sp.size();
  sp.push_back(n);
  // the pipelined sweeps take one position per thread and step
  for (size_t k = 0; k + 1 < sp.size(); ++k)
    if (sp[k + 1] - sp[k] > step_max)
      return fcp_fail(ctx, FCP_ERR_STATE, "block schedule: step of %d rows", sp[k + 1] - sp[k]);
  b.max_steps = 0;
  b.R = 0;
  for (int q = 0; q < b.nblocks; ++q) {
    if (bs[q + 1] - bs[q] > b.max_steps) b.max_steps = bs[q + 1] - bs[q];
    if (bp[q + 1] - bp[q] > b.R) b.R = bp[q + 1] - bp[q];
  }
  FCP_HIP(ctx, b.step_ptr.alloc(sp.size()));
  FCP_HIP(ctx, b.blk_step.alloc(bs.size()));
  FCP_HIP(ctx, b.blk_pos.alloc(bp.size()));
  FCP_HIP(ctx, hipMemcpyAsync(b.step_ptr.p, sp.data(), sizeof(int) * sp.size(), hipMemcpyHostToDevice,
                              ctx->stream));
  FCP_HIP(ctx, hipMemcpyAsync(b.blk_step.p, bs.data(), sizeof(int) * bs.size(), hipMemcpyHostToDevice,
                              ctx->stream));
  FCP_HIP(ctx, hipMemcpyAsync(b.blk_pos.p, bp.data(), sizeof(int) * bp.size(), hipMemcpyHostToDevice,
                              ctx->stream));
  // packed static operands
  FCP_HIP(ctx, b.prc.alloc((size_t)2 * n));
  FCP_HIP(ctx, b.pdep.alloc((size_t)FCP_PMAXD * n));
  FCP_HIP(ctx, b.psrc.alloc((size_t)FCP_PMAXD * n));
  {
    DBuf<int> mx, blk_dev, pos;
    FCP_HIP(ctx, mx.alloc(1));
    FCP_HIP(ctx, blk_dev.alloc(n));
    FCP_HIP(ctx, pos.alloc(n));
    FCP_HIP(ctx, hipMemcpyAsync(blk_dev.p, blk_of.data(), sizeof(int) * n, hipMemcpyHostToDevice,
                                ctx->stream));
    FCP_HIP(ctx, hipMemsetAsync(mx.p, 0, sizeof(int), ctx->stream));
    k_pos_of<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(b.rows.p, n, pos.p);
    k_pack_static<<<grid_for(n, BLK), BLK, 0, ctx->stream>>>(b.rows.p, n, upper,
                                                             ctx->sym_pattern ? 1 : 0,
                                                             local_only ? 1 : 0, blk_dev.p,
                                                             b.blk_pos.p, pos.p, ctx->ia.p,
                                                             ctx->diag.p, ctx->ja.p, b.prc.p,
                                                             b.pdep.p, b.psrc.p, mx.p);
    FCP_HIP(ctx, hipGetLastError());
    FCP_HIP(ctx, hipMemcpyAsync(&b.maxd, mx.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  b.pver64 = b.pver32 = ~0ull;
  return 0;
}

// Geometric tiles: cells binned by centroid on a d0 x d1 x d2 grid of the
// bounding box with about `target` cells per tile, tiles ranked by
// (tx+ty+tz, tz, ty, tx).  Usable only when every strict-lower dependency
// lies in a tile of lower or equal rank and every strict-upper one in a tile
// of higher or equal rank (then the forward sweep's tickets, and the
// reversed backward tickets, only ever wait on earlier tickets).  Returns
// the number of tiles, 0 when the check fails.
namespace {
struct TileInput {
  std::vector<double> x, y, z;
  std::vector<int> ia, ja, dg;
};

int tile_ranks(const TileInput &in, int n, int target, std::vector<int> &rank,
               bool check_acyclic = true) {
  double lo[3] = {in.x[0], in.y[0], in.z[0]}, hi[3] = {in.x[0], in.y[0], in.z[0]};
  for (int i = 0; i < n; ++i) {
    const double v[3] = {in.x[i], in.y[i], in.z[i]};
    for (int a = 0; a < 3; ++a) {
      lo[a] = v[a] < lo[a] ? v[a] : lo[a];
      hi[a] = v[a] > hi[a] ? v[a] : hi[a];
    }
  }
  // grid: cells per tile ~ target, tile edges proportional to the box
  const double ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  const double vol = ext[0] * ext[1] * ext[2];
  const double ntile = (double)n / target;
  int d[3];
  for (int a = 0; a < 3; ++a) {
    const double k = vol > 0 ? ext[a] * std::cbrt(ntile / vol) : std::cbrt(ntile);
    d[a] = k < 1.0 ? 1 : (int)std::lround(k);
  }
  const int nt = d[0] * d[1] * d[2];
  if (nt < 2) return 0;
  std::vector<int> tid(n);
  for (int i = 0; i < n; ++i) {
    const double v[3] = {in.x[i], in.y[i], in.z[i]};
    int t[3];
    for (int a = 0; a < 3; ++a) {
      const double u = ext[a] > 0 ? (v[a] - lo[a]) / ext[a] : 0.0;
      int q = (int)(u * d[a]);
      t[a] = q < 0 ? 0 : (q >= d[a] ? d[a] - 1 : q);
    }
    tid[i] = t[0] + d[0] * (t[1] + d[1] * t[2]);
  }
  // rank tiles by (tx+ty+tz, tz, ty, tx)
  std::vector<int> tiles(nt), trank(nt);
  for (int t = 0; t < nt; ++t) tiles[t] = t;
  auto key = [&](int t) {
    const int tx = t % d[0], ty = (t / d[0]) % d[1], tz = t / (d[0] * d[1]);
    return std::make_tuple(tx + ty + tz, tz, ty, tx);
  };
  std::sort(tiles.begin(), tiles.end(), [&](int a, int b) { return key(a) < key(b); });
  for (int r = 0; r < nt; ++r) trank[tiles[r]] = r;
  rank.assign(n, 0);
  for (int i = 0; i < n; ++i) rank[i] = trank[tid[i]];
  // acyclic only if every lower dependency has a rank <= its row's, every
  // upper dependency a rank >= it
  for (int i = 0; i < n && check_acyclic; ++i) {
    for (int k = in.ia[i]; k < in.dg[i]; ++k)
      if (rank[in.ja[k]] > rank[i]) return 0;
    for (int k = in.dg[i] + 1; k < in.ia[i + 1]; ++k)
      if (rank[in.ja[k]] < rank[i]) return 0;
  }
  // compact away empty tiles (keep rank order)
  std::vector<int> used(nt, 0), remap(nt, -1);
  for (int i = 0; i < n; ++i) used[rank[i]] = 1;
  int m = 0;
  for (int r = 0; r < nt; ++r)
    if (used[r]) remap[r] = m++;
  for (int i = 0; i < n; ++i) rank[i] = remap[rank[i]];
  return m;
}
}  // namespace

static int tile_input(fcp_ctx *ctx, TileInput &in) {
  const int n = ctx->n;
  in.x.resize(n);
  in.y.resize(n);
  in.z.resize(n);
  in.ia.resize(n + 1);
  in.ja.resize(ctx->nnz);
  in.dg.resize(n);
  FCP_HIP(ctx, hipMemcpyAsync(in.x.data(), ctx->xc.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  FCP_HIP(ctx, hipMemcpyAsync(in.y.data(), ctx->yc.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  FCP_HIP(ctx, hipMemcpyAsync(in.z.data(), ctx->zc.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  FCP_HIP(ctx, hipMemcpyAsync(in.ia.data(), ctx->ia.p, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, ctx->stream));
  FCP_HIP(ctx, hipMemcpyAsync(in.ja.data(), ctx->ja.p, sizeof(int) * ctx->nnz, hipMemcpyDeviceToHost, ctx->stream));
  FCP_HIP(ctx, hipMemcpyAsync(in.dg.data(), ctx->diag.p, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
  FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int fcp_build_tile_schedules(fcp_ctx *ctx) {
  FCP_TRY(fcp_coarse_wait(ctx));  // a pending coarse factorisation reads the schedules
  ctx->have_tiles = false;
  ctx->have_bj = false;
  ctx->tc_have = false;
  ctx->msched.clear();
  if (!ctx->have_geom || ctx->tune_sweep_tile == 1) return 0;
  const int n = ctx->n;
  const int target = ctx->tune_sweep_tile >= 64 ? ctx->tune_sweep_tile : 4096;
  if (n < 4 * target) return 0;
  TileInput in;
  FCP_TRY(tile_input(ctx, in));
  std::vector<int> rank;
  const int m = tile_ranks(in, n, target, rank);
  if (m == 0) return 0;
  FCP_TRY(fcp_build_block_schedule(ctx, ctx->lower, 0, 0, ctx->tlower, rank.data(), m));
  FCP_TRY(fcp_build_block_schedule(ctx, ctx->upper, 0, 1, ctx->tupper, rank.data(), m));
  ctx->have_tiles = true;
  return 0;
}

// Blocks of the tile-local preconditioner: geometric tiles of about 4096
// cells (more above 2048 such tiles) when the mesh has geometry (any
// numbering: blocks never wait on one another), else runs of contiguous rows
// of the same size.
int fcp_build_bj_schedules(fcp_ctx *ctx) {
  FCP_TRY(fcp_coarse_wait(ctx));  // a pending coarse factorisation reads the schedules
  ctx->have_bj = false;
  ctx->tin_have = false;
  ctx->tc_have = false;
  const int n = ctx->n;
  int target = ctx->tune_inner_tile >= 64 ? ctx->tune_inner_tile : 4096;
  // by default, tiles large enough that their count -- rounded to the tile
  // grid, so with a factor 2 of margin -- stays within the coarse
  // correction's dense limit (192^3 gets 8192-cell tiles, 256^3 16384)
  if (ctx->tune_inner_tile < 64)
    while ((long long)n > (long long)target * (FCP_COARSE_MAXN / 2) && target < 65536)
      target *= 2;
  std::vector<int> rank;
  int m = 0;
  if (ctx->have_geom && ctx->tune_sweep_tile != 1 && n >= 4 * target) {
    TileInput in;
    FCP_TRY(tile_input(ctx, in));
    m = tile_ranks(in, n, target, rank, false);
  }
  if (m == 0) {
    rank.resize(n);
    for (int i = 0; i < n; ++i) rank[i] = i / target;
    m = (n + target - 1) / target;
  }
  // tiles of up to 2048 cells are swept by one wave each: steps of at most
  // 64 rows
  const bool wave = target <= 2048;
  const int step_max = wave ? 64 : 256;
  FCP_TRY(fcp_build_block_schedule(ctx, ctx->lower, 0, 0, ctx->jlower, rank.data(), m, true,
                                   step_max));
  FCP_TRY(fcp_build_block_schedule(ctx, ctx->upper, 0, 1, ctx->jupper, rank.data(), m, true,
                                   step_max));
  ctx->bj_wave = wave;
  ctx->bj_tile.swap(rank);
  ctx->bj_version = ~0ull;
  ctx->have_bj = true;
  return 0;
}

int fcp_build_block_schedules(fcp_ctx *ctx) {
  FCP_TRY(fcp_coarse_wait(ctx));  // a pending coarse factorisation reads the schedules
  ctx->msched.clear();
  FCP_TRY(fcp_build_block_schedule(ctx, ctx->lower, fcp_block_rows(ctx, 8, 0), 0, ctx->blower));
  FCP_TRY(fcp_build_block_schedule(ctx, ctx->upper, fcp_block_rows(ctx, 8, 1), 1, ctx->bupper));
  FCP_TRY(fcp_build_block_schedule(ctx, ctx->lower, fcp_block_rows(ctx, 4, 0), 0, ctx->blower4));
  FCP_TRY(fcp_build_block_schedule(ctx, ctx->upper, fcp_block_rows(ctx, 4, 1), 1, ctx->bupper4));
  return 0;
}

// Estimated critical path, in pipelined steps, of one sweep over contiguous
// blocks of R rows: blocks run concurrently and a block's step for a level
// follows its steps for the levels before, so the path is the sum over
// levels of the widest block's sub-steps at that level (ceil(rows/BLK)).
static long long sweep_path_steps(const std::vector<int> &lev, int nlev, int n, int R,
                                  std::vector<int> &cnt, std::vector<int> &widest) {
  std::fill(widest.begin(), widest.end(), 0);
  for (int b0 = 0; b0 < n; b0 += R) {
    const int b1 = n < b0 + R ? n : b0 + R;
    for (int i = b0; i < b1; ++i) ++cnt[lev[i]];
    for (int i = b0; i < b1; ++i) {
      const int L = lev[i];
      if (cnt[L] == 0) continue;
      const int sub = (cnt[L] + BLK - 1) / BLK;
      widest[L] = sub > widest[L] ? sub : widest[L];
      cnt[L] = 0;
    }
  }
  long long steps = 0;
  for (int L = 0; L < nlev; ++L) steps += widest[L];
  return steps;
}

// Rows per contiguous block of one triangle: the largest candidate (Rmax
// halved down to 256 rows) whose estimated path is within 10% of the best.
// Numberings with wide level fronts (RCM) need small blocks, lexicographic
// ones do not.
static int choose_block_rows(const std::vector<int> &lev, int nlev, int n, int Rmax) {
  std::vector<int> cand;
  for (int R = Rmax; R >= 256; R = ((R / 2) / 64) * 64) cand.push_back(R);
  if (cand.empty()) return Rmax;
  std::vector<int> cnt(nlev > 0 ? nlev : 1, 0), widest(nlev > 0 ? nlev : 1, 0);
  std::vector<long long> cost;
  long long best = -1;
  for (int R : cand) {
    cost.push_back(sweep_path_steps(lev, nlev, n, R, cnt, widest));
    if (best < 0 || cost.back() < best) best = cost.back();
  }
  for (size_t k = 0; k < cand.size(); ++k)
    if (cost[k] * 10 <= best * 11) return cand[k];
  return cand.back();
}

static int set_block_rows_caps(fcp_ctx *ctx) {
  const int n = ctx->n;
  std::vector<int> lev(n);
  const long long rmax = ((long long)ctx->lds_cap - 4096) / 4 / 64 * 64;
  const int Rmax = (int)(rmax > 65536 ? 65536 : (rmax < 64 ? 64 : rmax));
  for (int u = 0; u < 2; ++u) {
    const Schedule &s = u ? ctx->upper : ctx->lower;
    FCP_HIP(ctx, hipMemcpyAsync(lev.data(), s.lev.p, sizeof(int) * n, hipMemcpyDeviceToHost,
                                ctx->stream));
    FCP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->block_rows_cap[u] = choose_block_rows(lev, s.nlev, n, Rmax);
  }
  return 0;
}

int fcp_build_schedules(fcp_ctx *ctx) {
  FCP_TRY(fcp_coarse_wait(ctx));  // a pending coarse factorisation reads the schedules
  ctx->have_bj = false;
  ctx->tc_have = false;
  FCP_TRY(build_one_schedule(ctx, 0, ctx->lower));
  FCP_TRY(build_one_schedule(ctx, 1, ctx->upper));
  FCP_TRY(set_block_rows_caps(ctx));
  FCP_TRY(fcp_build_block_schedules(ctx));
  FCP_TRY(fcp_build_tile_schedules(ctx));
  return 0;
}
