// dense_engine.hpp -- host runtime of the dense-carry batched form (include/bwgpu_dense.h): the
// preorder-rank tables derived from a TreeEngine's walk, the dense carry rows built on the device,
// and a planned batch's workspace.  Every kernel goes through the bwd_* C-ABI.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "bwgpu_dense.h"
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

// Host restatement of the rank tables (also used by the tests' emulation): for the walk `steps`
// cut at blk_step0, state_erank[v] = index of v's enter step among the enter steps, erank_step[e]
// = the step of the e-th enter, blk_e0[b] = enter steps before block b.  False if some state is
// not entered exactly once.
bool dense_rank_tables(int64_t S, const std::vector<int32_t>& steps, const std::vector<int32_t>& blk_step0,
                       std::vector<int32_t>& state_erank, std::vector<int32_t>& erank_step,
                       std::vector<int32_t>& blk_e0);

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
