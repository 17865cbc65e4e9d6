// The device's streaming rate, for rooflines (no reference counterpart: the
// reference launches no kernels).  y = x over n f64 with 16-byte nontemporal loads
// and stores, kVec vectors per lane issued before the first store, so each lane
// keeps kVec loads in flight; one pass, no grid-stride loop.  Half the traffic is
// reads and half writes, the mix of the Model C step (φ, u read; φ, u written).
#include "lgd_internal.h"

namespace {

constexpr int kThreads = 256;
typedef double dvec2 __attribute__((ext_vector_type(2)));

template <int kVec>
__global__ __launch_bounds__(kThreads) void stream_copy_kernel(const dvec2 *__restrict__ x,
                                                               dvec2 *__restrict__ y, int64_t n2) {
  const int64_t base = (int64_t)blockIdx.x * (kThreads * kVec) + threadIdx.x;
  dvec2 v[kVec];
#pragma unroll
  for (int k = 0; k < kVec; ++k) {
    const int64_t i = base + (int64_t)k * kThreads;
    if (i < n2) v[k] = __builtin_nontemporal_load(x + i);
  }
#pragma unroll
  for (int k = 0; k < kVec; ++k) {
    const int64_t i = base + (int64_t)k * kThreads;
    if (i < n2) __builtin_nontemporal_store(v[k], y + i);
  }
}

template <int kVec>
int launch(lgd_ctx *ctx, const double *x, double *y, int64_t n2) {
  const int64_t per = (int64_t)kThreads * kVec;
  stream_copy_kernel<kVec><<<(unsigned)((n2 + per - 1) / per), kThreads, 0, ctx->stream>>>(
      (const dvec2 *)x, (dvec2 *)y, n2);
  return lgd::launch_check(ctx, "stream copy");
}

}  // namespace

using namespace lgd;

extern "C" int lgd_stream_copy(lgd_ctx *ctx, const lgd_buf *x, lgd_buf *y, int32_t vec) {
  if (!ctx || !x || !y) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  if (x->len != y->len || (x->len & 1))
    return fail(ctx, LGD_ERR_SIZE_MISMATCH, "stream copy: equal, even lengths");
  const int64_t n2 = (int64_t)(x->len / 2);
  if (n2 == 0) return LGD_OK;
  switch (vec) {
    case 1: return launch<1>(ctx, x->data, y->data, n2);
    case 2: return launch<2>(ctx, x->data, y->data, n2);
    case 4: return launch<4>(ctx, x->data, y->data, n2);
    case 8: return launch<8>(ctx, x->data, y->data, n2);
    default: return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "stream copy: vec is 1, 2, 4 or 8");
  }
}
