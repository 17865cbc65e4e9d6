// The observation seam: AsyncObserver's snapshot ring feeding ParquetObserver
// (src/core/observer.rs:60-182, src/io/parquet.rs:63-352), device-first.
//
// observe() runs on the solver's thread and never waits for the disk. At a snapshot
// step it takes a free ring slot (none free: the snapshot is dropped and counted),
// gathers the observed fields' interior cells into the slot's device staging area in
// the Parquet row order (block-major, dimension 0 fastest), and queues an
// asynchronous copy into the slot's pinned host buffer on a copy stream that waits
// for the gather only. A writer thread waits for each copy, writes the file
// (lgd_parquet.cpp) and returns the slot. After construction the pipeline allocates
// nothing on a fixed grid; on an AMR hierarchy a slot grows when an epoch has more
// rows than it holds.
//
// The per-epoch static file (coordinates, level, patch, static fields) is gathered on
// the device too and travels through the ring before the first snapshot of its
// epoch; a snapshot is never queued without its epoch's static file.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <cstdio>
#include <deque>
#include <filesystem>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "lgd_amr.h"
#include "lgd_internal.h"
#include "lgd_parquet.h"

namespace {

constexpr int kMaxCols = 16;
constexpr int kMaxLevels = 16;
constexpr int kTile = 256;
constexpr int64_t kRowsPerGroup = int64_t(1) << 22;  // ROWS_PER_GROUP, parquet.rs:55

// One block's rows: its first row, interior extent, lower corner in level cells,
// level, and (AMR) its ghost-1 slab offset.
struct ObsBlock {
  int64_t row0;
  int64_t off1;
  int64_t lo[3];
  int32_t n[3];
  int32_t level;
};

// A workgroup's rows: kTile rows of one block from `start`.
struct ObsTile {
  int32_t block;
  int32_t start;
};

struct ColumnSrc {
  const double *src;
  int32_t ghost;
  int64_t stride;  // Cartesian: elements per block slab; -1: the block's AMR offset
};

struct GatherArgs {
  const ObsBlock *blocks;
  const ObsTile *tiles;
  int64_t nrows;
  int32_t dim;
  int32_t ncol;
  int32_t coords;  // 1: write x, y[, z], level, patch first (the static file)
  ColumnSrc col[kMaxCols];
  double h[kMaxLevels][3];
  double origin[3];
  double *out;  // columns of nrows, in order
};

// Interior cells of every observed field, one row per thread (for_each_interior
// order, parquet.rs:190-230); with `coords`, the cell centre of
// CartesianGrid::cell_center / AmrGrid::cell_center ((lo + idx) + 0.5)·h + origin as
// a fused multiply-add (cartesian.rs:589-597, amr/grid.rs:388-395), then the
// block's level and id.
__global__ __launch_bounds__(kTile) void gather_kernel(GatherArgs a) {
  const ObsTile tile = a.tiles[blockIdx.x];
  const ObsBlock B = a.blocks[tile.block];
  // Block-local indices fit 32 bits (a block's cells are counted in int32).
  const uint32_t r = (uint32_t)tile.start + threadIdx.x;
  if (r >= (uint32_t)B.n[0] * (uint32_t)B.n[1] * (uint32_t)B.n[2]) return;
  const uint32_t n0 = (uint32_t)B.n[0], n1 = (uint32_t)B.n[1];
  const uint32_t rest = r / n0;
  const int64_t i = r - rest * n0;
  const uint32_t k32 = rest / n1;
  const int64_t j = rest - k32 * n1;
  const int64_t k = k32;
  const int64_t row = B.row0 + r;
  int c0 = 0;
  if (a.coords) {
    const int64_t idx[3] = {i, j, k};
    for (int d = 0; d < a.dim; ++d)
      a.out[(int64_t)d * a.nrows + row] =
          fma((double)(B.lo[d] + idx[d]) + 0.5, a.h[B.level][d], a.origin[d]);
    a.out[(int64_t)a.dim * a.nrows + row] = (double)B.level;
    a.out[(int64_t)(a.dim + 1) * a.nrows + row] = (double)tile.block;
    c0 = a.dim + 2;
  }
  for (int c = 0; c < a.ncol; ++c) {
    const ColumnSrc s = a.col[c];
    const int64_t g = s.ghost;
    const int64_t ex = B.n[0] + 2 * g;
    const int64_t ey = a.dim > 1 ? B.n[1] + 2 * g : 1;
    const int64_t jj = a.dim > 1 ? j + g : 0;
    const int64_t kk = a.dim > 2 ? k + g : 0;
    const int64_t base = s.stride >= 0 ? (int64_t)tile.block * s.stride : B.off1;
    a.out[(int64_t)(c0 + c) * a.nrows + row] = s.src[base + (kk * ey + jj) * ex + (i + g)];
  }
}

struct Slot {
  double *dev = nullptr;
  double *host = nullptr;
  size_t cap = 0;  // doubles
  hipEvent_t gathered = nullptr;
  hipEvent_t copied = nullptr;
};

struct Msg {
  int slot;
  bool is_static;
  uint64_t step;
  double t;
  uint64_t epoch;
  int64_t nrows;
  int dim;
  std::vector<int64_t> groups;
};

}  // namespace

struct lgd_observer {
  lgd_ctx *ctx = nullptr;
  std::string dir;
  uint64_t every = 1;
  std::vector<std::string> names, static_names;
  std::vector<int32_t> fields, statics;  // simulation field indices
  // Layout of the current epoch.
  const void *grid = nullptr;
  uint64_t epoch = 0;
  bool have_layout = false;
  bool static_queued = false;
  bool amr = false;
  int dim = 2;
  int64_t nrows = 0;
  std::vector<int64_t> groups;
  void *d_layout = nullptr;          // blocks then tiles
  ObsBlock *d_blocks = nullptr;
  ObsTile *d_tiles = nullptr;
  int64_t ntiles = 0;
  double h[kMaxLevels][3] = {};
  double origin[3] = {0.0, 0.0, 0.0};
  // Ring and writer.
  hipStream_t copy_stream = nullptr;
  std::vector<Slot> slots;  // slots[0] stages static files, the rest are the ring
  bool static_free = true;
  std::mutex m;
  std::condition_variable cv_work, cv_idle;
  std::deque<Msg> work;
  std::vector<int> free_slots;
  int busy = 0;
  bool stop = false, paused = false;
  bool blocking = false;  // wait for a free buffer instead of dropping
  std::thread writer;
  uint64_t written = 0, dropped = 0, failed = 0;
  double write_seconds = 0.0;
  std::string last_error;
};

using namespace lgd;

namespace {

bool due(const lgd_observer *o, uint64_t step) {
  return step == 1 || (o->every == 0 ? step == 0 : step % o->every == 0);
}

void write_msg(lgd_observer *o, const Msg &msg) {
  const Slot &s = o->slots[(size_t)msg.slot];
  std::vector<pq::Column> cols;
  char name[64];
  auto col = [&](const std::string &n, int c) {
    pq::Column k;
    k.name = n;
    k.type = pq::Type::kDouble;
    k.data = s.host + (size_t)c * (size_t)msg.nrows;
    cols.push_back(k);
  };
  if (msg.is_static) {
    static const char *kCoord[3] = {"x", "y", "z"};
    int c = 0;
    for (int d = 0; d < msg.dim; ++d) col(kCoord[d], c++);
    col("level", c++);
    col("patch", c++);
    for (const auto &n : o->static_names) col(n, c++);
    std::snprintf(name, sizeof name, "static_%07llu.parquet", (unsigned long long)msg.epoch);
  } else {
    pq::Column k;
    k.type = pq::Type::kInt64;
    k.name = "step";
    k.constant = msg.step;
    cols.push_back(k);
    k.type = pq::Type::kDouble;
    k.name = "t";
    std::memcpy(&k.constant, &msg.t, 8);
    cols.push_back(k);
    k.type = pq::Type::kInt64;
    k.name = "epoch";
    k.constant = msg.epoch;
    cols.push_back(k);
    for (size_t c = 0; c < o->names.size(); ++c) col(o->names[c], (int)c);
    std::snprintf(name, sizeof name, "snap_%07llu.parquet", (unsigned long long)msg.step);
  }
  std::string err;
  const auto t0 = std::chrono::steady_clock::now();
  const bool ok =
      pq::write_file(o->dir + "/" + name, cols, msg.nrows, msg.groups, &err) == pq::Status::kOk;
  const double sec =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::lock_guard<std::mutex> lk(o->m);
  o->write_seconds += sec;
  if (ok) {
    ++o->written;
  } else {
    ++o->failed;
    o->last_error = err;
  }
}

void writer_loop(lgd_observer *o) {
  (void)hipSetDevice(o->ctx->device);
  for (;;) {
    Msg msg;
    {
      std::unique_lock<std::mutex> lk(o->m);
      o->cv_work.wait(lk, [&] { return o->stop || (!o->paused && !o->work.empty()); });
      if (o->work.empty() || (o->paused && !o->stop)) {
        if (o->stop) return;
        continue;
      }
      msg = std::move(o->work.front());
      o->work.pop_front();
      ++o->busy;
    }
    if (hipEventSynchronize(o->slots[(size_t)msg.slot].copied) == hipSuccess) {
      write_msg(o, msg);
    } else {
      std::lock_guard<std::mutex> lk(o->m);
      ++o->failed;
      o->last_error = "snapshot copy failed";
    }
    {
      std::lock_guard<std::mutex> lk(o->m);
      if (msg.is_static)
        o->static_free = true;
      else
        o->free_slots.push_back(msg.slot);
      --o->busy;
    }
    o->cv_idle.notify_all();
  }
}

void free_layout(lgd_observer *o) {
  if (o->d_layout) (void)hipFree(o->d_layout);
  o->d_layout = nullptr;
  o->d_blocks = nullptr;
  o->d_tiles = nullptr;
  o->have_layout = false;
}

// Row groups as the reference's writer cuts them: whole blocks until at least
// ROWS_PER_GROUP rows, then a flush (parquet.rs:163-233).
int upload_layout(lgd_observer *o, const std::vector<ObsBlock> &blocks) {
  std::vector<ObsTile> tiles;
  o->groups.clear();
  int64_t acc = 0;
  for (size_t b = 0; b < blocks.size(); ++b) {
    const int64_t rows = (int64_t)blocks[b].n[0] * blocks[b].n[1] * blocks[b].n[2];
    if (rows > INT32_MAX) return fail(o->ctx, LGD_ERR_UNSUPPORTED, "a block of more than 2^31 cells");
    for (int64_t s = 0; s < rows; s += kTile) tiles.push_back({(int32_t)b, (int32_t)s});
    acc += rows;
    if (acc >= kRowsPerGroup) {
      o->groups.push_back(acc);
      acc = 0;
    }
  }
  if (acc > 0) o->groups.push_back(acc);
  free_layout(o);
  const size_t bb = blocks.size() * sizeof(ObsBlock), tb = tiles.size() * sizeof(ObsTile);
  if (bb + tb > 0) {
    hipError_t e = hipMalloc(&o->d_layout, bb + tb);
    if (e != hipSuccess) return hip_fail(o->ctx, e, "observer layout");
    o->d_blocks = (ObsBlock *)o->d_layout;
    o->d_tiles = (ObsTile *)((char *)o->d_layout + bb);
    e = hipMemcpy(o->d_blocks, blocks.data(), bb, hipMemcpyHostToDevice);
    if (e == hipSuccess && tb) e = hipMemcpy(o->d_tiles, tiles.data(), tb, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(o->ctx, e, "observer layout upload");
  }
  o->ntiles = (int64_t)tiles.size();
  o->have_layout = true;
  return LGD_OK;
}

int layout_cartesian(lgd_observer *o, const lgd_grid *grid, uint64_t epoch) {
  if (o->have_layout && !o->amr && o->grid == grid && o->epoch == epoch) return LGD_OK;
  const lgd_cart &c = grid->cart;
  std::vector<ObsBlock> blocks((size_t)c.nblocks);
  const int64_t ncell = c.n[0] * c.n[1] * c.n[2];
  for (int64_t b = 0; b < c.nblocks; ++b) {
    ObsBlock &B = blocks[(size_t)b];
    const int64_t coord[3] = {b % c.nb[0], (b / c.nb[0]) % c.nb[1], b / (c.nb[0] * c.nb[1])};
    B.row0 = b * ncell;
    B.off1 = 0;
    for (int d = 0; d < 3; ++d) {
      B.lo[d] = d < c.dim ? coord[d] * c.n[d] : 0;
      B.n[d] = (int32_t)(d < c.dim ? c.n[d] : 1);
    }
    B.level = 0;
  }
  o->amr = false;
  o->dim = c.dim;
  o->nrows = c.nblocks * ncell;
  for (int d = 0; d < 3; ++d) {
    o->h[0][d] = d < c.dim ? c.h[d] : 0.0;
    o->origin[d] = d < c.dim ? c.origin[d] : 0.0;
  }
  o->static_queued = false;  // a new layout starts a new static file
  o->grid = grid;
  o->epoch = epoch;
  return upload_layout(o, blocks);
}

int layout_amr(lgd_observer *o, const lgd_amr *g, uint64_t epoch) {
  if (o->have_layout && o->amr && o->grid == g && o->epoch == epoch) return LGD_OK;
  const int64_t nb = (int64_t)g->box.size();
  if (g->num_levels() > kMaxLevels)
    return fail(o->ctx, LGD_ERR_UNSUPPORTED, "the observer handles at most 16 levels");
  std::vector<ObsBlock> blocks((size_t)nb);
  int64_t row = 0;
  for (int64_t b = 0; b < nb; ++b) {
    ObsBlock &B = blocks[(size_t)b];
    const lgd_box &x = g->box[(size_t)b];
    B.row0 = row;
    B.off1 = g->off[(size_t)b];
    B.lo[0] = x.lo[0];
    B.lo[1] = x.lo[1];
    B.lo[2] = 0;
    B.n[0] = (int32_t)x.extent(0);
    B.n[1] = (int32_t)x.extent(1);
    B.n[2] = 1;
    B.level = g->level[(size_t)b];
    row += (int64_t)B.n[0] * B.n[1];
  }
  o->amr = true;
  o->dim = 2;
  o->nrows = row;
  for (int l = 0; l < g->num_levels(); ++l) {
    o->h[l][0] = g->spacing[(size_t)l][0];
    o->h[l][1] = g->spacing[(size_t)l][1];
    o->h[l][2] = 0.0;
  }
  o->origin[0] = g->base.origin[0];
  o->origin[1] = g->base.origin[1];
  o->origin[2] = 0.0;
  o->static_queued = false;
  o->grid = g;
  o->epoch = epoch;
  return upload_layout(o, blocks);
}

// Grow a free slot to `doubles` (allocation happens only when an epoch outgrows it).
int ensure_slot(lgd_observer *o, Slot &s, size_t doubles) {
  if (s.cap >= doubles) return LGD_OK;
  if (s.dev) (void)hipFree(s.dev);
  if (s.host) (void)hipHostFree(s.host);
  s.dev = nullptr;
  s.host = nullptr;
  s.cap = 0;
  hipError_t e = hipMalloc((void **)&s.dev, doubles * sizeof(double));
  if (e != hipSuccess) return hip_fail(o->ctx, e, "snapshot staging");
  e = hipHostMalloc((void **)&s.host, doubles * sizeof(double), hipHostMallocDefault);
  if (e != hipSuccess) return hip_fail(o->ctx, e, "snapshot ring (pinned host memory)");
  s.cap = doubles;
  return LGD_OK;
}

// Gather into `slot`, start its copy to the host and queue the message.
int enqueue(lgd_observer *o, int slot, bool is_static, uint64_t step, double t,
            const std::vector<ColumnSrc> &cols) {
  Slot &s = o->slots[(size_t)slot];
  const int ncoords = is_static ? o->dim + 2 : 0;
  const size_t doubles = (size_t)(ncoords + (int)cols.size()) * (size_t)o->nrows;
  int rc = ensure_slot(o, s, std::max<size_t>(doubles, 1));
  if (rc) return rc;
  lgd_ctx *ctx = o->ctx;
  if (o->ntiles > 0) {
    GatherArgs a{};
    a.blocks = o->d_blocks;
    a.tiles = o->d_tiles;
    a.nrows = o->nrows;
    a.dim = o->dim;
    a.ncol = (int32_t)cols.size();
    a.coords = is_static ? 1 : 0;
    for (size_t c = 0; c < cols.size(); ++c) a.col[c] = cols[c];
    std::memcpy(a.h, o->h, sizeof a.h);
    std::memcpy(a.origin, o->origin, sizeof a.origin);
    a.out = s.dev;
    gather_kernel<<<(unsigned)o->ntiles, kTile, 0, ctx->stream>>>(a);
    if ((rc = launch_check(ctx, "observer gather"))) return rc;
  }
  hipError_t e = hipEventRecord(s.gathered, ctx->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(o->copy_stream, s.gathered, 0);
  if (e == hipSuccess && doubles)
    e = hipMemcpyAsync(s.host, s.dev, doubles * sizeof(double), hipMemcpyDeviceToHost,
                       o->copy_stream);
  if (e == hipSuccess) e = hipEventRecord(s.copied, o->copy_stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "snapshot copy");
  Msg msg{slot, is_static, step, t, o->epoch, o->nrows, o->dim, o->groups};
  {
    std::lock_guard<std::mutex> lk(o->m);
    o->work.push_back(std::move(msg));
  }
  o->cv_work.notify_one();
  return LGD_OK;
}

// The snapshot itself, once the layout for (grid, epoch) is current.
int snapshot(lgd_observer *o, uint64_t step, double t, const std::vector<ColumnSrc> &dyn,
             const std::vector<ColumnSrc> &stat) {
  const bool need_static = !o->static_queued;
  int slot_d = -1;
  {
    std::unique_lock<std::mutex> lk(o->m);
    auto ready = [&] { return !o->free_slots.empty() && (!need_static || o->static_free); };
    // The synchronous ParquetObserver writes every snapshot (parquet.rs:318-333): in
    // blocking mode the solver waits for the writer to return a buffer.
    if (o->blocking) o->cv_idle.wait(lk, [&] { return ready() || o->paused; });
    if (!ready()) {
      ++o->dropped;  // AsyncObserver: the solver never waits (observer.rs:158-176)
      return LGD_OK;
    }
    if (need_static) o->static_free = false;
    slot_d = o->free_slots.back();
    o->free_slots.pop_back();
  }
  int rc = LGD_OK;
  auto give_back = [&](bool stat_slot, int s) {
    std::lock_guard<std::mutex> lk(o->m);
    if (stat_slot)
      o->static_free = true;
    else
      o->free_slots.push_back(s);
  };
  if (need_static) {
    if ((rc = enqueue(o, 0, true, step, t, stat))) {
      give_back(true, 0);
      give_back(false, slot_d);
      return rc;
    }
    o->static_queued = true;
  }
  if ((rc = enqueue(o, slot_d, false, step, t, dyn))) give_back(false, slot_d);
  return rc;
}

int check_cols(lgd_observer *o, const lgd_grid *grid, const lgd_buf *const *bufs,
               const uint32_t *ghost, size_t n, std::vector<ColumnSrc> &out) {
  out.clear();
  if (n && (!bufs || !ghost)) return fail(o->ctx, LGD_ERR_INVALID_ARGUMENT, "null field array");
  for (size_t c = 0; c < n; ++c) {
    int rc = check_field(o->ctx, grid, bufs[c], ghost[c], "observed field");
    if (rc) return rc;
    out.push_back({bufs[c]->data, (int32_t)ghost[c], slab_len(grid->cart, ghost[c])});
  }
  return LGD_OK;
}

}  // namespace

namespace lgd {

// Builds the layout of (grid or amr, epoch) and sizes the static buffer and every ring
// buffer for it, so a simulation's snapshots allocate nothing.
int observer_prepare(lgd_observer *o, const lgd_grid *grid, const lgd_amr *amr, uint64_t epoch) {
  int rc = amr ? layout_amr(o, amr, epoch) : layout_cartesian(o, grid, epoch);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(o->m);
  if (!o->static_free) return LGD_OK;  // in flight: sized when it is next used
  const size_t rows = (size_t)o->nrows;
  if ((rc = ensure_slot(o, o->slots[0], std::max<size_t>(1, (o->dim + 2 + o->static_names.size()) * rows))))
    return rc;
  for (int i : o->free_slots)
    if ((rc = ensure_slot(o, o->slots[(size_t)i], std::max<size_t>(1, o->names.size() * rows)))) return rc;
  return LGD_OK;
}

// Steps from `step` until the observer's next snapshot step (>= 1).
int64_t observer_steps_to_due(const lgd_observer *o, uint64_t step) {
  if (step == 0) return 1;  // step 1 always snapshots
  if (o->every == 0) return INT64_MAX;
  return (int64_t)((step / o->every + 1) * o->every - step);
}

bool observer_due(const lgd_observer *o, uint64_t step) { return due(o, step); }

const std::vector<int32_t> &observer_fields(const lgd_observer *o) { return o->fields; }
const std::vector<int32_t> &observer_statics(const lgd_observer *o) { return o->statics; }

}  // namespace lgd

extern "C" {

int lgd_observer_create(lgd_ctx *ctx, const char *dir, uint64_t every, size_t ring,
                        const lgd_observed_field *fields, size_t nfields,
                        const lgd_observed_field *statics, size_t nstatics, lgd_observer **out) {
  if (!ctx || !dir || !out) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  // ParquetObserver::new rejects an empty field list (parquet.rs:94-96).
  if (nfields == 0 || !fields) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "no fields to observe");
  if (nstatics && !statics) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null static field array");
  if (nfields > (size_t)kMaxCols || nstatics > (size_t)(kMaxCols))
    return fail(ctx, LGD_ERR_UNSUPPORTED, "at most 16 observed and 16 static fields");
  if (ring == 0) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "the snapshot ring needs a buffer");
  std::error_code ec;
  std::filesystem::create_directories(dir, ec);
  if (ec) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, std::string("cannot create ") + dir);
  lgd_observer *o = new (std::nothrow) lgd_observer();
  if (!o) return fail(ctx, LGD_ERR_OUT_OF_MEMORY, "host allocation failed");
  o->ctx = ctx;
  o->dir = dir;
  o->every = every;
  for (size_t i = 0; i < nfields; ++i) {
    if (!fields[i].name) {
      delete o;
      return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "unnamed field");
    }
    o->names.push_back(fields[i].name);
    o->fields.push_back(fields[i].field);
  }
  for (size_t i = 0; i < nstatics; ++i) {
    if (!statics[i].name) {
      delete o;
      return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "unnamed static field");
    }
    o->static_names.push_back(statics[i].name);
    o->statics.push_back(statics[i].field);
  }
  // slots[0] stages the per-epoch static file; slots[1..ring] are the snapshot ring.
  o->slots.resize(ring + 1);
  hipError_t e = hipStreamCreateWithFlags(&o->copy_stream, hipStreamNonBlocking);
  for (size_t i = 0; e == hipSuccess && i < o->slots.size(); ++i) {
    e = hipEventCreateWithFlags(&o->slots[i].gathered, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&o->slots[i].copied, hipEventDisableTiming);
  }
  for (size_t i = ring; i >= 1; --i) o->free_slots.push_back((int)i);
  if (e != hipSuccess) {
    int rc = hip_fail(ctx, e, "observer streams");
    for (auto &s : o->slots) {
      if (s.gathered) (void)hipEventDestroy(s.gathered);
      if (s.copied) (void)hipEventDestroy(s.copied);
    }
    if (o->copy_stream) (void)hipStreamDestroy(o->copy_stream);
    delete o;
    return rc;
  }
  o->writer = std::thread(writer_loop, o);
  *out = o;
  return LGD_OK;
}

int lgd_observer_flush(lgd_observer *obs) {
  if (!obs) return LGD_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(obs->m);
  obs->paused = false;
  obs->cv_work.notify_all();
  obs->cv_idle.wait(lk, [&] { return obs->work.empty() && obs->busy == 0; });
  return LGD_OK;
}

int lgd_observer_set_blocking(lgd_observer *obs, int32_t blocking) {
  if (!obs) return LGD_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(obs->m);
  obs->blocking = blocking != 0;
  return LGD_OK;
}

int lgd_observer_pause(lgd_observer *obs, int32_t paused) {
  if (!obs) return LGD_ERR_INVALID_ARGUMENT;
  {
    std::lock_guard<std::mutex> lk(obs->m);
    obs->paused = paused != 0;
  }
  obs->cv_work.notify_all();
  obs->cv_idle.notify_all();  // a blocked observe() drops instead of waiting on a held writer
  return LGD_OK;
}

int lgd_observer_destroy(lgd_observer *obs) {
  if (!obs) return LGD_OK;
  lgd_observer_flush(obs);
  {
    std::lock_guard<std::mutex> lk(obs->m);
    obs->stop = true;
  }
  obs->cv_work.notify_all();
  if (obs->writer.joinable()) obs->writer.join();
  (void)hipStreamSynchronize(obs->copy_stream);
  for (auto &s : obs->slots) {
    if (s.dev) (void)hipFree(s.dev);
    if (s.host) (void)hipHostFree(s.host);
    if (s.gathered) (void)hipEventDestroy(s.gathered);
    if (s.copied) (void)hipEventDestroy(s.copied);
  }
  (void)hipStreamDestroy(obs->copy_stream);
  free_layout(obs);
  delete obs;
  return LGD_OK;
}

int lgd_observer_stats(lgd_observer *obs, uint64_t *written, uint64_t *dropped, uint64_t *failed,
                       double *write_seconds) {
  if (!obs) return LGD_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(obs->m);
  if (written) *written = obs->written;
  if (dropped) *dropped = obs->dropped;
  if (failed) *failed = obs->failed;
  if (write_seconds) *write_seconds = obs->write_seconds;
  return LGD_OK;
}

int lgd_observer_last_error(lgd_observer *obs, char *buf, size_t len) {
  if (!obs || (!buf && len)) return LGD_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(obs->m);
  if (len) std::snprintf(buf, len, "%s", obs->last_error.c_str());
  return LGD_OK;
}

int lgd_observer_observe(lgd_observer *obs, const lgd_grid *grid, uint64_t step, double t,
                         uint64_t epoch, const lgd_buf *const *fields, const uint32_t *field_ghost,
                         const lgd_buf *const *statics, const uint32_t *static_ghost) {
  if (!obs || !grid) return LGD_ERR_INVALID_ARGUMENT;
  if (!due(obs, step)) {
    if (obs->have_layout && epoch != obs->epoch) obs->static_queued = false;
    return LGD_OK;
  }
  std::vector<ColumnSrc> dyn, stat;
  int rc;
  if ((rc = check_cols(obs, grid, fields, field_ghost, obs->names.size(), dyn))) return rc;
  if ((rc = check_cols(obs, grid, statics, static_ghost, obs->static_names.size(), stat))) return rc;
  if ((rc = layout_cartesian(obs, grid, epoch))) return rc;
  return snapshot(obs, step, t, dyn, stat);
}

int lgd_observer_observe_amr(lgd_observer *obs, const lgd_amr *amr, uint64_t step, double t,
                             uint64_t epoch, const lgd_buf *const *fields) {
  if (!obs || !amr) return LGD_ERR_INVALID_ARGUMENT;
  if (!obs->static_names.empty())
    return fail(obs->ctx, LGD_ERR_UNSUPPORTED, "static fields are observed on Cartesian grids");
  if (!due(obs, step)) {
    if (obs->have_layout && epoch != obs->epoch) obs->static_queued = false;
    return LGD_OK;
  }
  std::vector<ColumnSrc> dyn;
  const int64_t len = amr->off.back();
  for (size_t c = 0; c < obs->names.size(); ++c) {
    if (!fields || !fields[c])
      return fail(obs->ctx, LGD_ERR_INVALID_ARGUMENT, "null observed field");
    if ((int64_t)fields[c]->len != len)
      return fail(obs->ctx, LGD_ERR_SIZE_MISMATCH, "observed field is not a hierarchy field");
    dyn.push_back({fields[c]->data, 1, -1});
  }
  int rc;
  if ((rc = layout_amr(obs, amr, epoch))) return rc;
  return snapshot(obs, step, t, dyn, {});
}

}  // extern "C"
