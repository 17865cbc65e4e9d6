// A C++ caller of the C ABI (include/legendre_gpu.h) with no Python and no torch: the
// shape of the binding a Rust crate would write. It runs Model C twice from the same
// initial condition — the fused step and the per-intent sequence of the reference's
// EulerMaruyama::step — with a Parquet observer on the fused run, checks that the two
// agree bit for bit, and writes the fused run's phi and u (ghost-inclusive slabs,
// little-endian f64) for the calling test to compare with the oracle.
//
//   harness <out_dir> <steps>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "legendre_gpu.h"

namespace {

lgd_ctx *g_ctx = nullptr;

void check(int status, const char *what) {
  if (status == LGD_OK) return;
  std::fprintf(stderr, "%s failed: status %d (%s)\n", what, status,
               g_ctx ? lgd_ctx_last_error(g_ctx) : "");
  std::exit(1);
}

std::vector<double> download(lgd_sim *sim, int field) {
  lgd_buf *b = nullptr;
  check(lgd_sim_field(sim, field, &b), "lgd_sim_field");
  std::vector<double> out(lgd_buf_len(b));
  check(lgd_buf_download(g_ctx, b, 0, out.data(), out.size()), "lgd_buf_download");
  return out;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: harness <out_dir> <steps>\n");
    return 2;
  }
  const std::string out = argv[1];
  const long steps = std::atol(argv[2]);
  check(lgd_ctx_create(0, nullptr, &g_ctx), "lgd_ctx_create");
  lgd_cartesian_desc d{};
  d.dim = 2;
  d.cells[0] = 256;
  d.cells[1] = 192;
  d.block_cells[0] = 64;
  d.block_cells[1] = 64;
  d.spacing[0] = d.spacing[1] = 0.4;
  lgd_grid *grid = nullptr;
  check(lgd_grid_create_cartesian(g_ctx, &d, &grid), "lgd_grid_create_cartesian");
  const lgd_modelc_params p{0.05, 3.19, 2.0, 0.0, 1e-8, 1e-8};
  const double grains[] = {20.0, 30.0, 2.0, 70.0, 45.0, 2.0};
  lgd_sim *sims[2] = {nullptr, nullptr};
  const int modes[2] = {LGD_MODE_FUSED, LGD_MODE_PER_INTENT};
  for (int k = 0; k < 2; ++k) {
    check(lgd_sim_create_modelc(g_ctx, grid, &p, 0, 42, modes[k], &sims[k]), "lgd_sim_create_modelc");
    check(lgd_sim_initialize_grains(sims[k], grains, 2, 0.55), "lgd_sim_initialize_grains");
  }
  lgd_observer *obs = nullptr;
  const lgd_observed_field fields[2] = {{"phi", 0}, {"u", 1}};
  check(lgd_observer_create(g_ctx, (out + "/parquet").c_str(), 10, 2, fields, 2, nullptr, 0, &obs),
        "lgd_observer_create");
  check(lgd_observer_set_blocking(obs, 1), "lgd_observer_set_blocking");
  check(lgd_sim_attach_observer(sims[0], obs), "lgd_sim_attach_observer");
  for (int k = 0; k < 2; ++k) check(lgd_sim_step(sims[k], 0.008, steps), "lgd_sim_step");
  check(lgd_observer_flush(obs), "lgd_observer_flush");
  uint64_t written = 0, dropped = 0, failed = 0;
  check(lgd_observer_stats(obs, &written, &dropped, &failed, nullptr), "lgd_observer_stats");
  int same = 1;
  for (int f = 0; f < 2; ++f) {
    std::vector<double> a = download(sims[0], f), b = download(sims[1], f);
    // Interiors only: ghost entries are scratch after a step.
    const long sx = 66, sy = 66, nb = (long)a.size() / (sx * sy);
    for (long blk = 0; blk < nb; ++blk)
      for (long j = 1; j < sy - 1; ++j)
        for (long i = 1; i < sx - 1; ++i) {
          const long o = blk * sx * sy + j * sx + i;
          if (std::memcmp(&a[o], &b[o], 8) != 0) same = 0;
        }
    FILE *fp = std::fopen((out + (f == 0 ? "/phi.bin" : "/u.bin")).c_str(), "wb");
    if (!fp || std::fwrite(a.data(), 8, a.size(), fp) != a.size() || std::fclose(fp) != 0) {
      std::fprintf(stderr, "cannot write the fields\n");
      return 1;
    }
  }
  std::printf("{\"steps\": %ld, \"time\": %.17g, \"fused_equals_per_intent\": %d, \"written\": %llu, "
              "\"dropped\": %llu, \"failed\": %llu}\n",
              lgd_sim_step_index(sims[0]), lgd_sim_time(sims[0]), same, (unsigned long long)written,
              (unsigned long long)dropped, (unsigned long long)failed);
  check(lgd_sim_destroy(sims[0]), "lgd_sim_destroy");
  check(lgd_sim_destroy(sims[1]), "lgd_sim_destroy");
  check(lgd_observer_destroy(obs), "lgd_observer_destroy");
  check(lgd_grid_destroy(grid), "lgd_grid_destroy");
  check(lgd_ctx_destroy(g_ctx), "lgd_ctx_destroy");
  return 0;
}
