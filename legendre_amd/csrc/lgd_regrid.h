// Regrid of the device AMR backend (lgd_regrid.hip).
#pragma once

#include <vector>

#include "lgd_amr.h"

// Proposed refined levels from the ghost-filled tag field; returns 0 when the
// current hierarchy still serves (no regrid), 1 when `out` should replace it.
int lgd_regrid_propose(const lgd_amr &g, const std::vector<double> &tag_field,
                       const lgd_regrid_policy &policy, std::vector<std::vector<lgd_box>> &out);
// The same from per-level bitmasks of tagged, buffer-dilated cells (row-major over
// each level's domain, 64 cells per word; null = none).
int lgd_regrid_propose_bits(const lgd_amr &g, const std::vector<uint64_t *> &bits,
                            const lgd_regrid_policy &policy, std::vector<std::vector<lgd_box>> &out);
// A level's bitmask by its nonzero words in increasing order: word idx[i] (row
// idx[i] / wpr, cells 64·(idx[i] mod wpr) on) is bits[i]; every other word is zero.
struct lgd_flag_words {
  const int64_t *idx = nullptr;
  const uint64_t *bits = nullptr;
  size_t n = 0;
};
// The same from those words (a level past the vector's end has none).
int lgd_regrid_propose_words(const lgd_amr &g, const std::vector<lgd_flag_words> &flags,
                             const lgd_regrid_policy &policy, std::vector<std::vector<lgd_box>> &out);
// The bitmask of level `src` from a caller's tagged cells (n × [x, y], level cell
// coordinates, as TagCells::tag_level appends them), buffer-dilated and clipped to
// the level domain like the GradientTagger's.
void lgd_regrid_bits_from_cells(const lgd_amr &g, int src, const std::vector<int64_t> &cells,
                                int64_t buffer, std::vector<uint64_t> &bits);
// Migrate one field from the old hierarchy onto the new one: new_f receives the
// levels >= 1 of the new field (elements from new_g.off[level_start[1]] on); level 0
// is the same in both hierarchies and is copied on the device.
void lgd_regrid_migrate(const lgd_amr &old_g, const std::vector<double> &old_f,
                        const lgd_amr &new_g, std::vector<double> &new_f);

// Device scratch of the regrid's tagging pass (kept across regrids).
struct lgd_regrid_scratch {
  uint64_t *T = nullptr;       // tag bitmasks
  uint64_t *bits = nullptr;    // dilated bitmasks
  int64_t *idx = nullptr;      // the nonzero dilated words' indices
  uint64_t *words = nullptr;   // and values
  int32_t *counts = nullptr;   // nonzero words per workgroup
  int64_t *offsets = nullptr;  // their prefix sums, then the levels' starts
  uint64_t *host = nullptr;    // pinned host copy of the starts, indices and words
  size_t wcap = 0;             // words of each bitmask array
  size_t bcap = 0;             // workgroups of the word launches
  size_t last_total = 0;       // nonzero words at the last regrid
};
void lgd_regrid_scratch_free(lgd_regrid_scratch &sc);
// Whether the device path covers this hierarchy's depth.
bool lgd_regrid_device_supported(const lgd_amr &g);
// Tag and dilate every level that can be refined on the device (the tag field is
// the ghost-filled device field); flags[level] lists the nonzero words of its
// bitmask in the scratch's pinned host copy (valid until the next call).
int lgd_regrid_flag_words(lgd_ctx *ctx, const lgd_amr &g, const double *tag_field,
                          const lgd_regrid_policy &policy, lgd_regrid_scratch &sc,
                          std::vector<lgd_flag_words> &flags);
// Device migration: the interiors of new_g's levels >= 1 in new_f (a buffer of the
// new layout) sampled from old_f on old_g; level 0 and ghosts are not written.
int lgd_regrid_migrate_device(lgd_ctx *ctx, const lgd_amr &old_g, const double *const *old_f,
                              const lgd_amr &new_g, double *const *new_f, int nf);
