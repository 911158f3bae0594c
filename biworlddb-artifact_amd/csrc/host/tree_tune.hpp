// tree_tune.hpp -- autotuner of the generalised batch path: picks the block size (rmax, the
// "chain-block size"), the key tile width (tile_shift), the level-B chunk size, the endpoints
// per chunk work item and the launch shapes by timing candidate configurations on the device
// over a sample of the actual query stream, and keeps the choice per workload form in a small
// JSON file ({"<form>": {"rmax": .., "tile_shift": .., ...}, ...}).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "tree_plan.hpp"
#include "tree_workload.hpp"

namespace bwg {

struct LaunchKnobs {
  int block_threads = 256;
  int chunk_waves = 4;
};

// Reads the knobs stored for `form` ("batched", "served", ...).  False when absent.
bool load_tuned(const char* path, const char* form, PlanKnobs& kn, LaunchKnobs& lk);
// Writes (or replaces) the entry of `form`, keeping the other entries of the file.
bool save_tuned(const char* path, const char* form, const PlanKnobs& kn, const LaunchKnobs& lk, const char* device,
                double ms);

// Knobs of the dense-carry form (include/bwgpu_dense.h): the block size (keys per block, which sets
// the carry's size), the persistent replay's workgroup threads and workgroups per CU, and how the
// answers reach their ids (BWD_STORE_*), and the queries per work item (0: whole blocks).  Stored
// as the "dense" entry of the same file.
struct DenseKnobs {
  int rmax = 4096;
  int threads = 1024;
  int wg_per_cu = 1;
  int store = 0;
  int chunk = 0;
};
bool load_dense_tuned(const char* path, DenseKnobs& dk);
bool save_dense_tuned(const char* path, const DenseKnobs& dk, const char* device, double ms);

struct TuneTrial {
  PlanKnobs kn;
  LaunchKnobs lk;
  double ms = 0;  // device time of one query launch sequence over the sample (< 0: failed)
};

// Coordinate descent from (kn, lk): every knob in turn takes each of its candidate values, the
// others fixed; a change is kept when it lowers the timed query pass over the sample (M queries
// of states qstate with coordinates [qlo, qhi)).  The index is re-planned and rebuilt only when
// an index knob changes.  Returns the best trial; every trial is appended to `log`.
TuneTrial autotune(const TreeWorkload& w, const uint32_t* kcoord, int64_t M, const int32_t* qstate,
                   const uint32_t* qlo, const uint32_t* qhi, PlanKnobs kn, LaunchKnobs lk, int repeat,
                   std::vector<TuneTrial>& log);

}  // namespace bwg
