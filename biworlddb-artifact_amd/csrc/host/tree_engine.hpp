// tree_engine.hpp -- host runtime of the generalised batch path: owns the device copy of the
// update-side index (BwtIndex) and of query batches (BwtBatch), built from the host plans.
// Nothing here touches the answers' arithmetic; every kernel goes through the bwt_* C-ABI.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "bwgpu_tree.h"
#include "tree_plan.hpp"

namespace bwg {

// Device allocations freed together.  Errors are reported through ok() / err.
class DeviceArena {
 public:
  DeviceArena() = default;
  DeviceArena(const DeviceArena&) = delete;
  DeviceArena& operator=(const DeviceArena&) = delete;
  ~DeviceArena() { release(); }
  void release();
  void* alloc(size_t bytes);
  template <typename T>
  T* scratch(size_t n) { return static_cast<T*>(alloc(std::max<size_t>(n, 1) * sizeof(T))); }
  template <typename T>
  T* upload(const T* src, size_t n) {
    T* d = scratch<T>(n);
    if (d && n) copy_in(d, src, n * sizeof(T));
    return d;
  }
  template <typename T>
  T* upload(const std::vector<T>& v) { return upload(v.data(), v.size()); }
  bool ok() const { return err == hipSuccess; }
  size_t bytes = 0;
  hipError_t err = hipSuccess;

 private:
  void copy_in(void* d, const void* s, size_t n);
  std::vector<void*> ptrs_;
};

struct TreeEngine {
  TreeIndexPlan plan;
  BwtIndex x{};
  BwtLaunch launch{256, 4, 0, 0};
  DeviceArena arena;
  BwtBatch all_blocks{};  // every block listed, no queries: the index-build batch

  // Host index plan + device upload of deltas and plan arrays; agg/tagg/cp are allocated, not
  // yet computed (build_index does that).  key holds coordinates in [0, NK).
  bool init(int64_t S, int U, uint64_t NK, const int32_t* parent, const uint32_t* key, const int64_t* delta,
            const PlanKnobs& knobs, const BwtLaunch& l, char* err, size_t errlen);
  // Block replay with the weight write over every block, chunk histograms, carry scan.
  int build_index(hipStream_t st);
};

// A query batch resident on the device.  qlo/qhi are device coordinate arrays owned by the caller
// (or uploaded here when given host arrays).
struct DeviceBatch {
  BatchPlan plan;
  BwtBatch b{};
  DeviceArena arena;
};

bool upload_batch(const TreeEngine& e, BatchPlan&& plan, const uint32_t* d_qlo, const uint32_t* d_qhi,
                  DeviceBatch& out, char* err, size_t errlen);

}  // namespace bwg
