// dense_engine.hpp -- host runtime of the dense-carry batched form (include/bwgpu_dense.h): the
// preorder-rank tables derived from a TreeEngine's walk, the dense carry rows built on the device,
// and a planned batch's workspace.  Every kernel goes through the bwd_* C-ABI.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "bwgpu_dense.h"
#include "dense_tables.hpp"
#include "tree_engine.hpp"

namespace bwg {

struct DenseEngine {
  BwdIndex x{};
  BwdLaunch launch{1024, 1, 0, 0};
  DeviceArena arena;   // the rank tables and the carry rows
  double build_ms = 0;  // device time of the carry build
  size_t carry_bytes = 0;

  // From an engine whose index weights are built (e.build_index: agg holds every block's net
  // weight per distinct key).  The engine must outlive this one (its arrays are shared).
  bool init(const TreeEngine& e, hipStream_t st, char* err, size_t errlen);
};

// A batch's workspace (plan + replay) for up to cap queries.
struct DenseBatch {
  int64_t cap = 0;
  void* work = nullptr;
  size_t work_bytes = 0;
  int64_t* flags = nullptr;
  DeviceArena arena;
};
bool alloc_dense_batch(const DenseEngine& d, int64_t cap, DenseBatch& b, char* err, size_t errlen);

}  // namespace bwg
