// plan_common.hpp -- definitions shared by the host planner (tree_plan.cpp) and the device
// planner (plan_kernels.hip), so both cut the same slot ranges.
#pragma once
#include <stdint.h>

#include "bwgpu_tree.h"

#ifdef __HIPCC__
#define BWT_HD __host__ __device__
#else
#define BWT_HD
#endif

namespace bwg {

// Slot ranges of the level-B work order.  The slots are split into BWT_XCDS parts exactly as the
// combine's slot pass splits them among the dies (whole tiles of BWT_COMBINE_TILE slots, part x
// = die x), and each part into ranges of at most range_slots slots; a range is numbered
// part * ranges-per-part + its index in the part.  Work items never cross a range, so the
// level-B results of one range form one run of the work order, read back by one die.
struct SlotRanges {
  int64_t part_slots, rs, nsub;
};

inline BWT_HD SlotRanges slot_ranges(int64_t M, int64_t range_slots) {
  const int64_t nt = (M + BWT_COMBINE_TILE - 1) / BWT_COMBINE_TILE;
  int64_t per_t = (nt + BWT_XCDS - 1) / BWT_XCDS;
  if (per_t < 1) per_t = 1;
  SlotRanges r;
  r.part_slots = per_t * BWT_COMBINE_TILE;
  r.rs = range_slots < 1 ? 1 : (range_slots < r.part_slots ? range_slots : r.part_slots);
  r.nsub = (r.part_slots + r.rs - 1) / r.rs;
  return r;
}

inline BWT_HD int64_t range_of(int64_t slot, const SlotRanges& r) {
  const int64_t p = slot / r.part_slots;
  return p * r.nsub + (slot - p * r.part_slots) / r.rs;
}

}  // namespace bwg
