// Host-side regrid of the device AMR backend (lgd_regrid.hip).
#pragma once

#include <vector>

#include "lgd_amr.h"

// Proposed refined levels from the ghost-filled tag field; returns 0 when the
// current hierarchy still serves (no regrid), 1 when `out` should replace it.
int lgd_regrid_propose(const lgd_amr &g, const std::vector<double> &tag_field,
                       const lgd_regrid_policy &policy, std::vector<std::vector<lgd_box>> &out);
// Migrate one field from the old hierarchy onto the new one: new_f receives the
// levels >= 1 of the new field (elements from new_g.off[level_start[1]] on); level 0
// is the same in both hierarchies and is copied on the device.
void lgd_regrid_migrate(const lgd_amr &old_g, const std::vector<double> &old_f,
                        const lgd_amr &new_g, std::vector<double> &new_f);
