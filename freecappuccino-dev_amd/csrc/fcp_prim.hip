// fcp_prim.hip -- setup-time device primitives (integer scan, stable key sort)
// on rocPRIM.  Used only when a mesh/pattern is uploaded, never per step.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "fcp_internal.h"

int fcp_exclusive_scan_int(fcp_ctx *ctx, const int *in, int *out, int n) {
  size_t bytes = 0;
  FCP_HIP(ctx, rocprim::exclusive_scan(nullptr, bytes, in, out, 0, (size_t)n,
                                       rocprim::plus<int>(), ctx->stream));
  FCP_HIP(ctx, ctx->stage[3].alloc(bytes + 16));
  FCP_HIP(ctx, rocprim::exclusive_scan(ctx->stage[3].p, bytes, in, out, 0, (size_t)n,
                                       rocprim::plus<int>(), ctx->stream));
  return 0;
}

int fcp_sort_pairs_int(fcp_ctx *ctx, const int *keys_in, int *keys_out, const int *vals_in,
                       int *vals_out, int n, int max_key) {
  unsigned end_bit = 1;
  while (end_bit < 31 && (1u << end_bit) <= (unsigned)max_key) ++end_bit;
  size_t bytes = 0;
  FCP_HIP(ctx, rocprim::radix_sort_pairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out,
                                         (size_t)n, 0u, end_bit, ctx->stream));
  FCP_HIP(ctx, ctx->stage[3].alloc(bytes + 16));
  FCP_HIP(ctx, rocprim::radix_sort_pairs(ctx->stage[3].p, bytes, keys_in, keys_out, vals_in,
                                         vals_out, (size_t)n, 0u, end_bit, ctx->stream));
  return 0;
}

int fcp_sort_keys_u64(fcp_ctx *ctx, const unsigned long long *keys_in,
                      unsigned long long *keys_out, int n, int end_bit) {
  size_t bytes = 0;
  FCP_HIP(ctx, rocprim::radix_sort_keys(nullptr, bytes, keys_in, keys_out, (size_t)n, 0u,
                                        (unsigned)end_bit, ctx->stream));
  FCP_HIP(ctx, ctx->stage[3].alloc(bytes + 16));
  FCP_HIP(ctx, rocprim::radix_sort_keys(ctx->stage[3].p, bytes, keys_in, keys_out, (size_t)n, 0u,
                                        (unsigned)end_bit, ctx->stream));
  return 0;
}
