// fcp_rcm.hip -- reverse Cuthill-McKee renumbering of the matrix graph.
//
// The reference has no renumbering.  The ordering rule (bit-identical to the
// CPU specification in the oracle): each connected component starts from its
// unvisited vertex of minimum (degree, id), moved by at most pp_iters
// George-Liu steps to the minimum-(degree, id) vertex of the last BFS level
// while that increases the eccentricity; Cuthill-McKee then processes a
// FIFO queue, every dequeued vertex appending its unvisited neighbours in
// (degree, id) order; the order is finally reversed.
//
// Level-synchronous form: all vertices of one BFS level are dequeued before
// any of the next, so the next level in FIFO order is the set of unvisited
// neighbours of the level, each attached to its first (lowest-position)
// parent, sorted by (parent position, degree, id) -- one 64-bit radix sort
// per level.  Degree = off-diagonal count of the CSR row.
#include <hip/hip_runtime.h>

#include <climits>
#include <vector>

#include "fcp_internal.h"

namespace {

constexpr int BLK = 256;
typedef unsigned long long u64;

__global__ void k_rcm_degree(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                             int *__restrict__ deg, int *__restrict__ maxdeg) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  int d = 0;
  for (int k = ia[i]; k < ia[i + 1]; ++k) d += (ja[k] != i);
  deg[i] = d;
  atomicMax(maxdeg, d);
}

__global__ void k_rcm_init(int n, int *__restrict__ a, int va, int *__restrict__ b, int vb) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  a[i] = va;
  if (b) b[i] = vb;
}

// min over unvisited vertices (list == nullptr) or over list[0..m) of
// deg*n + id
__global__ void k_rcm_minkey(int m, const int *__restrict__ list, const char *__restrict__ done,
                             const int *__restrict__ deg, int n, u64 *__restrict__ out) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  u64 best = ~0ull;
  if (i < m) {
    const int v = list ? list[i] : i;
    if (list || !done[v]) best = (u64)deg[v] * (u64)n + (u64)v;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const u64 t = __shfl_xor(best, o, 64);
    best = t < best ? t : best;
  }
  if ((threadIdx.x & 63) == 0 && best != ~0ull) atomicMin(out, best);
}

// one BFS level over vertices not yet numbered (distances only)
__global__ void k_rcm_bfs(int fsize, const int *__restrict__ front, const int *__restrict__ ia,
                          const int *__restrict__ ja, const char *__restrict__ done,
                          int *__restrict__ dist, int lev, int *__restrict__ next,
                          int *__restrict__ next_cnt) {
  const int q = blockIdx.x * BLK + threadIdx.x;
  if (q >= fsize) return;
  const int v = front[q];
  for (int k = ia[v]; k < ia[v + 1]; ++k) {
    const int u = ja[k];
    if (u == v || done[u]) continue;
    if (atomicCAS(&dist[u], -1, lev + 1) == -1) next[atomicAdd(next_cnt, 1)] = u;
  }
}

// Cuthill-McKee level: attach every unvisited neighbour of the level (in
// queue order q) to its first parent.
__global__ void k_rcm_claim(int fsize, const int *__restrict__ front, const int *__restrict__ ia,
                            const int *__restrict__ ja, const char *__restrict__ done,
                            int *__restrict__ parent, int *__restrict__ claimed,
                            int *__restrict__ cand, int *__restrict__ cand_cnt) {
  const int q = blockIdx.x * BLK + threadIdx.x;
  if (q >= fsize) return;
  const int v = front[q];
  for (int k = ia[v]; k < ia[v + 1]; ++k) {
    const int u = ja[k];
    if (u == v || done[u]) continue;
    atomicMin(&parent[u], q);
    if (atomicExch(&claimed[u], 1) == 0) cand[atomicAdd(cand_cnt, 1)] = u;
  }
}

__global__ void k_rcm_keys(int m, const int *__restrict__ cand, const int *__restrict__ parent,
                           const int *__restrict__ deg, u64 dmul, u64 n,
                           u64 *__restrict__ keys) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= m) return;
  const int u = cand[i];
  keys[i] = ((u64)parent[u] * dmul + (u64)deg[u]) * n + (u64)u;
}

__global__ void k_rcm_emit(int m, const u64 *__restrict__ keys, u64 n, int pos,
                           int *__restrict__ order, char *__restrict__ done,
                           int *__restrict__ front) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= m) return;
  const int u = (int)(keys[i] % n);
  order[pos + i] = u;
  done[u] = 1;
  front[i] = u;
}

__global__ void k_rcm_root(int r, int pos, int *__restrict__ order, char *__restrict__ done,
                           int *__restrict__ front) {
  order[pos] = r;
  done[r] = 1;
  front[0] = r;
}

__global__ void k_rcm_reverse(int n, const int *__restrict__ order, int *__restrict__ new_of_old) {
  const int k = blockIdx.x * BLK + threadIdx.x;
  if (k < n) new_of_old[order[k]] = n - 1 - k;
}

int grid(int m) { return m > 0 ? (m + BLK - 1) / BLK : 1; }

}  // namespace

int fcp_rcm_device(fcp_ctx *ctx, int pp_iters, int *new_of_old) {
  const int n = ctx->n;
  hipStream_t st = ctx->stream;
  DBuf<int> deg, dist, parent, claimed, fa, fb, order, cnt;
  DBuf<char> done;
  DBuf<u64> keys, keys2, mk;
  FCP_HIP(ctx, deg.alloc(n));
  FCP_HIP(ctx, dist.alloc(n));
  FCP_HIP(ctx, parent.alloc(n));
  FCP_HIP(ctx, claimed.alloc(n));
  FCP_HIP(ctx, fa.alloc(n));
  FCP_HIP(ctx, fb.alloc(n));
  FCP_HIP(ctx, order.alloc(n));
  FCP_HIP(ctx, cnt.alloc(2));
  FCP_HIP(ctx, done.alloc(n));
  FCP_HIP(ctx, keys.alloc(n));
  FCP_HIP(ctx, keys2.alloc(n));
  FCP_HIP(ctx, mk.alloc(1));
  FCP_HIP(ctx, hipMemsetAsync(cnt.p, 0, 2 * sizeof(int), st));
  k_rcm_degree<<<grid(n), BLK, 0, st>>>(n, ctx->ia.p, ctx->ja.p, deg.p, cnt.p);
  int maxdeg = 0;
  FCP_HIP(ctx, hipMemcpyAsync(&maxdeg, cnt.p, sizeof(int), hipMemcpyDeviceToHost, st));
  FCP_HIP(ctx, hipMemsetAsync(done.p, 0, n, st));
  k_rcm_init<<<grid(n), BLK, 0, st>>>(n, parent.p, INT_MAX, claimed.p, 0);
  FCP_HIP(ctx, hipStreamSynchronize(st));
  const u64 dmul = (u64)maxdeg + 1, nn = (u64)n;
  // key = (parent * dmul + deg) * n + id must fit 64 bits
  const long double span = (long double)n * (long double)dmul * (long double)n;
  if (span >= 18446744073709551615.0L)
    return fcp_fail(ctx, FCP_ERR_UNSUPPORTED, "rcm: graph too large for 64-bit sort keys");
  int end_bit = 1;
  while (end_bit < 64 && ((long double)((u64)1 << end_bit)) < span) ++end_bit;

  auto min_key = [&](int m, const int *list, u64 *out) -> int {
    FCP_HIP(ctx, hipMemsetAsync(mk.p, 0xFF, sizeof(u64), st));
    k_rcm_minkey<<<grid(m), BLK, 0, st>>>(m, list, done.p, deg.p, n, mk.p);
    FCP_HIP(ctx, hipMemcpyAsync(out, mk.p, sizeof(u64), hipMemcpyDeviceToHost, st));
    FCP_HIP(ctx, hipStreamSynchronize(st));
    return 0;
  };
  // BFS from r over unnumbered vertices: eccentricity and min key of the last level
  auto bfs = [&](int r, int *ecc, int *last_min) -> int {
    k_rcm_init<<<grid(n), BLK, 0, st>>>(n, dist.p, -1, nullptr, 0);
    int zero = 0;
    FCP_HIP(ctx, hipMemcpyAsync(dist.p + r, &zero, sizeof(int), hipMemcpyHostToDevice, st));
    FCP_HIP(ctx, hipMemcpyAsync(fa.p, &r, sizeof(int), hipMemcpyHostToDevice, st));
    int *cur = fa.p, *nxt = fb.p;
    int fsize = 1, lev = 0;
    for (;;) {
      FCP_HIP(ctx, hipMemsetAsync(cnt.p, 0, sizeof(int), st));
      k_rcm_bfs<<<grid(fsize), BLK, 0, st>>>(fsize, cur, ctx->ia.p, ctx->ja.p, done.p, dist.p, lev,
                                            nxt, cnt.p);
      int m = 0;
      FCP_HIP(ctx, hipMemcpyAsync(&m, cnt.p, sizeof(int), hipMemcpyDeviceToHost, st));
      FCP_HIP(ctx, hipStreamSynchronize(st));
      if (m == 0) break;
      std::swap(cur, nxt);
      fsize = m;
      ++lev;
    }
    u64 k = 0;
    FCP_TRY(min_key(fsize, cur, &k));
    *ecc = lev;
    *last_min = (int)(k % nn);
    return 0;
  };

  int pos = 0;
  while (pos < n) {
    u64 k0 = 0;
    FCP_TRY(min_key(n, nullptr, &k0));
    int r = (int)(k0 % nn), ecc = 0, last = 0;
    FCP_TRY(bfs(r, &ecc, &last));
    for (int it = 0; it < pp_iters; ++it) {
      int ecc2 = 0, last2 = 0;
      const int x = last;
      FCP_TRY(bfs(x, &ecc2, &last2));
      if (ecc2 > ecc) {
        r = x;
        ecc = ecc2;
        last = last2;
      } else {
        break;
      }
    }
    k_rcm_root<<<1, 1, 0, st>>>(r, pos, order.p, done.p, fa.p);
    ++pos;
    int *cur = fa.p, *nxt = fb.p;
    int fsize = 1;
    for (;;) {
      FCP_HIP(ctx, hipMemsetAsync(cnt.p, 0, sizeof(int), st));
      k_rcm_claim<<<grid(fsize), BLK, 0, st>>>(fsize, cur, ctx->ia.p, ctx->ja.p, done.p, parent.p,
                                              claimed.p, nxt, cnt.p);
      int m = 0;
      FCP_HIP(ctx, hipMemcpyAsync(&m, cnt.p, sizeof(int), hipMemcpyDeviceToHost, st));
      FCP_HIP(ctx, hipStreamSynchronize(st));
      if (m == 0) break;
      k_rcm_keys<<<grid(m), BLK, 0, st>>>(m, nxt, parent.p, deg.p, dmul, nn, keys.p);
      FCP_TRY(fcp_sort_keys_u64(ctx, keys.p, keys2.p, m, end_bit));
      k_rcm_emit<<<grid(m), BLK, 0, st>>>(m, keys2.p, nn, pos, order.p, done.p, cur);
      FCP_HIP(ctx, hipGetLastError());
      pos += m;
      fsize = m;  // the next level is now in `cur`, in queue order
    }
  }
  k_rcm_reverse<<<grid(n), BLK, 0, st>>>(n, order.p, new_of_old);
  FCP_HIP(ctx, hipGetLastError());
  FCP_HIP(ctx, hipStreamSynchronize(st));
  return 0;
}
