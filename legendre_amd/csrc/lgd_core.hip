// Context, device storage (the Allocator/StorageBackend backend) and the
// Cartesian block table.
#include <cmath>
#include <cstdio>
#include <new>

#include "lgd_internal.h"

namespace lgd {

int fail(lgd_ctx *ctx, int status, const std::string &msg) {
  if (ctx) ctx->last_error = msg;
  return status;
}

int hip_fail(lgd_ctx *ctx, hipError_t e, const char *what) {
  std::string msg = std::string(what) + ": " + hipGetErrorString(e);
  return fail(ctx, e == hipErrorOutOfMemory ? LGD_ERR_OUT_OF_MEMORY : LGD_ERR_HIP, msg);
}

int check_field(lgd_ctx *ctx, const lgd_grid *grid, const lgd_buf *buf, uint32_t ghost,
                const char *name) {
  if (!buf) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, std::string(name) + " is null");
  int64_t want = grid->cart.nblocks * slab_len(grid->cart, ghost);
  if ((int64_t)buf->len != want) {
    char m[256];
    snprintf(m, sizeof m, "%s holds %zu f64 but the grid needs %lld (ghost %u)", name, buf->len,
             (long long)want, ghost);
    return fail(ctx, LGD_ERR_SIZE_MISMATCH, m);
  }
  return LGD_OK;
}

// fill_ghosts_generic's guard (src/geometry/cartesian.rs:456-462).
int check_ghost(lgd_ctx *ctx, const lgd_grid *grid, uint32_t ghost) {
  for (int d = 0; d < grid->cart.dim; ++d) {
    if ((int64_t)ghost > grid->cart.n[d]) {
      char m[200];
      snprintf(m, sizeof m,
               "ghost width %u exceeds block extent %lld in dimension %d; blocks must be at "
               "least as thick as the widest stencil",
               ghost, (long long)grid->cart.n[d], d);
      return fail(ctx, LGD_ERR_GHOST_TOO_WIDE, m);
    }
  }
  return LGD_OK;
}

}  // namespace lgd

lgd_kr_consts lgd_make_kr_consts(const lgd_modelc_params &p, const lgd_cart &c) {
  lgd_kr_consts k;
  k.a_bar = std::fma(3.0, -p.eps4, 1.0);
  k.eps_prime = 3.0 * p.eps4 / k.a_bar;
  k.a12 = 4.0 * k.a_bar * k.eps_prime;
  k.fv_tol = p.fv_tol;
  k.model_tol = p.anisotropy_tol;
  k.inv_h2 = 1.0 / (c.h[0] * c.h[1]);
  k.inv_hx2 = 1.0 / (c.h[0] * c.h[0]);
  k.inv_hy2 = 1.0 / (c.h[1] * c.h[1]);
  k.lambda = p.lambda;
  k.d_thermal = p.d_thermal;
  return k;
}

using namespace lgd;

extern "C" {

uint32_t lgd_abi_version(void) { return LGD_ABI_VERSION; }

int lgd_ctx_create(int32_t device, void *stream, lgd_ctx **out) {
  if (!out) return LGD_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  lgd_ctx *ctx = new (std::nothrow) lgd_ctx();
  if (!ctx) return LGD_ERR_OUT_OF_MEMORY;
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    delete ctx;
    return LGD_ERR_HIP;
  }
  if (stream) {
    ctx->stream = (hipStream_t)stream;
    ctx->owns_stream = false;
  } else {
    // Default priority only: streams of mixed priorities are not used anywhere.
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete ctx;
      return LGD_ERR_HIP;
    }
    ctx->owns_stream = true;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess)
    ctx->num_cus = cus;
  // The AMR hierarchy's device arrays come from the device's stream-ordered memory
  // pool: keep freed blocks in the pool for the next regrid instead of returning
  // them to the driver at every synchronisation.
  hipMemPool_t pool = nullptr;
  if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
    uint64_t keep = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  *out = ctx;
  return LGD_OK;
}

int lgd_ctx_destroy(lgd_ctx *ctx) {
  if (!ctx) return LGD_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto &a : ctx->arenas) (void)hipFree(a.first);
  if (ctx->pinned_done) (void)hipEventDestroy(ctx->pinned_done);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  if (ctx->owns_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return LGD_OK;
}

int lgd_ctx_sync(lgd_ctx *ctx) {
  if (!ctx) return LGD_ERR_INVALID_ARGUMENT;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
  return LGD_OK;
}

void *lgd_ctx_stream(lgd_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

const char *lgd_ctx_last_error(const lgd_ctx *ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

int lgd_ctx_set_tuning(lgd_ctx *ctx, const char *key, int64_t value) {
  if (!ctx || !key) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  std::string k(key);
  if (value < 0) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "tuning values are non-negative");
  ++ctx->tuning_epoch;
  if (k == "modelc_kernel") {
    if (value > 3) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "modelc_kernel is 0, 1, 2 or 3");
    ctx->modelc_kernel = value;
  } else if (k == "modelc_strips") {
    ctx->modelc_strips = value;
  } else if (k == "modelc_prefetch") {
    if (value > 2) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "modelc_prefetch is 0, 1 or 2");
    ctx->modelc_prefetch = value;
  } else if (k == "amr_per_block") {
    if (value > 1) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "amr_per_block is 0 or 1");
    ctx->amr_per_block = value;
  } else if (k == "amr_split_fill") {
    if (value > 1) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "amr_split_fill is 0 or 1");
    ctx->amr_split_fill = value;
  } else if (k == "amr_tile") {
    if (value > 6) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "amr_tile is 0..6");
    ctx->amr_tile = value;
  } else if (k == "amr_wg") {
    if (value != 0 && value != 64 && value != 128 && value != 256 && value != 512 && value != 1024)
      return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "amr_wg is 0, 64, 128, 256, 512 or 1024");
    ctx->amr_wg = value;
  } else if (k == "amr_kernel") {
    if (value > 2) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "amr_kernel is 0, 1 or 2");
    ctx->amr_kernel = value;
  } else if (k == "amr_rows") {
    ctx->amr_rows = value;
  } else if (k == "adjoint_kernel") {
    if (value > 2) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "adjoint_kernel is 0, 1 or 2");
    ctx->adjoint_kernel = value;
  } else if (k == "adjoint_rows") {
    ctx->adjoint_rows = value;
  } else if (k == "regrid_host") {
    if (value > 1) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "regrid_host is 0 or 1");
    ctx->regrid_host = value;
  } else {
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "unknown tuning key " + k);
  }
  return LGD_OK;
}

int lgd_buf_alloc(lgd_ctx *ctx, size_t len, lgd_buf **out) {
  if (!ctx || !out) return LGD_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  lgd_buf *b = new (std::nothrow) lgd_buf();
  if (!b) return fail(ctx, LGD_ERR_OUT_OF_MEMORY, "host allocation failed");
  b->ctx = ctx;
  b->len = len;
  b->cap = len;
  if (len > 0) {
    hipError_t e = hipMalloc((void **)&b->data, len * sizeof(double));
    if (e != hipSuccess) {
      delete b;
      return hip_fail(ctx, e, "hipMalloc");
    }
    // The Allocator contract: storage starts zeroed (src/core/storage.rs:148-154).
    e = hipMemsetAsync(b->data, 0, len * sizeof(double), ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      (void)hipFree(b->data);
      delete b;
      return hip_fail(ctx, e, "zero-initialise");
    }
  }
  *out = b;
  return LGD_OK;
}

int lgd_buf_free(lgd_buf *buf) {
  if (!buf) return LGD_OK;
  if (buf->data) {
    (void)hipStreamSynchronize(buf->ctx->stream);
    (void)hipFree(buf->data);
  }
  delete buf;
  return LGD_OK;
}

size_t lgd_buf_len(const lgd_buf *buf) { return buf ? buf->len : 0; }

void *lgd_buf_device_ptr(lgd_buf *buf) { return buf ? (void *)buf->data : nullptr; }

int lgd_buf_upload(lgd_ctx *ctx, lgd_buf *buf, size_t offset, const double *host, size_t n) {
  if (!ctx || !buf || (!host && n)) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  if (offset > buf->len || n > buf->len - offset)
    return fail(ctx, LGD_ERR_SIZE_MISMATCH, "upload range exceeds the buffer");
  if (n == 0) return LGD_OK;
  hipError_t e = hipMemcpyAsync(buf->data + offset, host, n * sizeof(double),
                                hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "upload");
  return LGD_OK;
}

int lgd_buf_download(lgd_ctx *ctx, const lgd_buf *buf, size_t offset, double *host, size_t n) {
  if (!ctx || !buf || (!host && n)) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  if (offset > buf->len || n > buf->len - offset)
    return fail(ctx, LGD_ERR_SIZE_MISMATCH, "download range exceeds the buffer");
  if (n == 0) return LGD_OK;
  hipError_t e = hipMemcpyAsync(host, buf->data + offset, n * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "download");
  return LGD_OK;
}

int lgd_buf_copy(lgd_ctx *ctx, lgd_buf *dst, const lgd_buf *src) {
  if (!ctx || !dst || !src) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  if (dst->len != src->len) return fail(ctx, LGD_ERR_SIZE_MISMATCH, "copy between unequal buffers");
  if (dst->len == 0 || dst == src) return LGD_OK;
  hipError_t e = hipMemcpyAsync(dst->data, src->data, dst->len * sizeof(double),
                                hipMemcpyDeviceToDevice, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "copy");
  return LGD_OK;
}

int lgd_grid_create_cartesian(lgd_ctx *ctx, const lgd_cartesian_desc *desc, lgd_grid **out) {
  if (!ctx || !desc || !out) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  // CartesianGrid::new validation (src/geometry/cartesian.rs:48-82).
  if (desc->dim < 1 || desc->dim > 3)
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "dimension must be 1, 2 or 3");
  lgd_cart c{};
  c.dim = desc->dim;
  c.nblocks = 1;
  for (int d = 0; d < 3; ++d) {
    if (d < desc->dim) {
      if (!(desc->cells[d] > 0 && desc->block_cells[d] > 0))
        return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "empty dimension");
      if (desc->cells[d] % desc->block_cells[d] != 0)
        return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "cells not divisible by block_cells");
      if (!(desc->spacing[d] > 0.0) || !std::isfinite(desc->spacing[d]))
        return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "spacing must be positive and finite");
      c.n[d] = desc->block_cells[d];
      c.nb[d] = desc->cells[d] / desc->block_cells[d];
      c.h[d] = desc->spacing[d];
      c.origin[d] = desc->origin[d];
      c.periodic[d] = desc->periodic[d] ? 1 : 0;
    } else {
      c.n[d] = 1;
      c.nb[d] = 1;
      c.h[d] = 1.0;
      c.origin[d] = 0.0;
      c.periodic[d] = 0;
    }
    c.nblocks *= c.nb[d];
  }
  if (c.nblocks > (int64_t)UINT32_MAX)
    return fail(ctx, LGD_ERR_UNSUPPORTED, "block ids are u32 (src/geometry/grid.rs:36-46)");
  lgd_grid *g = new (std::nothrow) lgd_grid();
  if (!g) return fail(ctx, LGD_ERR_OUT_OF_MEMORY, "host allocation failed");
  g->ctx = ctx;
  g->cart = c;
  *out = g;
  return LGD_OK;
}

int lgd_grid_destroy(lgd_grid *grid) {
  delete grid;
  return LGD_OK;
}

int64_t lgd_grid_num_blocks(const lgd_grid *grid) { return grid ? grid->cart.nblocks : -1; }

int64_t lgd_grid_block_len(const lgd_grid *grid, uint32_t ghost) {
  return grid ? slab_len(grid->cart, ghost) : -1;
}

int64_t lgd_grid_face_neighbor(const lgd_grid *grid, int64_t block, int32_t dim, int32_t dir) {
  if (!grid || dim < 0 || dim >= grid->cart.dim || block < 0 || block >= grid->cart.nblocks)
    return -1;
  const lgd_cart &c = grid->cart;
  int64_t coords[3], rem = block;
  for (int d = 0; d < 3; ++d) {
    coords[d] = rem % c.nb[d];
    rem /= c.nb[d];
  }
  int64_t stepped = coords[dim] + (dir > 0 ? 1 : (dir < 0 ? -1 : 0));
  int64_t nb = c.nb[dim];
  int64_t next;
  if (stepped >= 0 && stepped < nb)
    next = stepped;
  else if (c.periodic[dim])
    next = ((stepped % nb) + nb) % nb;
  else
    return -1;
  int64_t id = 0, stride = 1;
  for (int d = 0; d < c.dim; ++d) {
    id += (d == dim ? next : coords[d]) * stride;
    stride *= c.nb[d];
  }
  return id;
}

int64_t lgd_grid_cell_key(const lgd_grid *grid, uint32_t ghost, int64_t offset) {
  if (!grid || offset < 0) return -1;
  const lgd_cart &c = grid->cart;
  int64_t g = ghost, rem = offset, key = 0, stride = 1;
  for (int d = 0; d < c.dim; ++d) {
    int64_t ghosted = c.n[d] + 2 * g;
    int64_t cc = rem % ghosted;
    rem /= ghosted;
    if (cc < g || cc >= c.n[d] + g) return -1;
    key += (cc - g) * stride;
    stride *= c.n[d];
  }
  return rem == 0 ? key : -1;
}

}  // extern "C"

hipError_t lgd_arena_take(lgd_ctx *ctx, size_t bytes, void **out, size_t *cap) {
  size_t best = ctx->arenas.size();
  for (size_t i = 0; i < ctx->arenas.size(); ++i)
    if (ctx->arenas[i].second >= bytes &&
        (best == ctx->arenas.size() || ctx->arenas[i].second < ctx->arenas[best].second))
      best = i;
  if (best < ctx->arenas.size()) {
    *out = ctx->arenas[best].first;
    *cap = ctx->arenas[best].second;
    ctx->arenas.erase(ctx->arenas.begin() + (std::ptrdiff_t)best);
    return hipSuccess;
  }
  const size_t want = bytes + bytes / 4;
  hipError_t e = hipMallocAsync(out, want, ctx->stream);
  if (e != hipSuccess) {
    *out = nullptr;
    *cap = 0;
    return e;
  }
  *cap = want;
  return hipSuccess;
}

void lgd_arena_give(lgd_ctx *ctx, void *p, size_t cap) {
  if (!p) return;
  ctx->arenas.emplace_back(p, cap);
  // Keep the three largest.
  if (ctx->arenas.size() > 3) {
    size_t small = 0;
    for (size_t i = 1; i < ctx->arenas.size(); ++i)
      if (ctx->arenas[i].second < ctx->arenas[small].second) small = i;
    (void)hipFreeAsync(ctx->arenas[small].first, ctx->stream);
    ctx->arenas.erase(ctx->arenas.begin() + (std::ptrdiff_t)small);
  }
}
