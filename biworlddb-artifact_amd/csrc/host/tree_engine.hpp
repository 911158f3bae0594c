// tree_engine.hpp -- host runtime of the generalised batch path: owns the device copy of the
// update-side index (BwtIndex) and of query batches (BwtBatch), built from the host plans.
// Nothing here touches the answers' arithmetic; every kernel goes through the bwt_* C-ABI.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "bwgpu_index.h"
#include "bwgpu_plan.h"
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
  BwtPlanIndex px{};      // the device planner's tables (uploaded with the index)

  // Host index plan + device upload of deltas and plan arrays; agg/tagg/cp are allocated, not
  // yet computed (build_index does that).  key holds coordinates in [0, NK).
  bool init(int64_t S, int U, uint64_t NK, const int32_t* parent, const uint32_t* key, const int64_t* delta,
            const PlanKnobs& knobs, const BwtLaunch& l, char* err, size_t errlen);
  // The same engine with the index tables built on the device (include/bwgpu_index.h): the walk
  // and the block cut on the host (unless device_walk / device_cut), then the per-block tables and the
  // chains on the GPU from d_key (device keys; NULL: uploaded from key), then the chunk tables on
  // the host.  The host plan then holds the walk, blocks, offsets and chunk tables only (no bkey,
  // evrank, kitems or position tables): batches are planned on the device.
  // device_cut: the greedy block cut on the device too (bwt_index_cut_launch); device_walk: the
  // Euler walk too (bwt_index_walk_launch; implies device_cut) -- the host then only checks.
  // key may be NULL when d_key is given (keys only on the device; implies device_cut); d_delta:
  // the deltas already on the device (else uploaded from delta).
  bool init_device_index(int64_t S, int U, uint64_t NK, const int32_t* parent, const uint32_t* key,
                         const uint32_t* d_key, const int64_t* delta, const PlanKnobs& knobs, const BwtLaunch& l,
                         hipStream_t st, char* err, size_t errlen, bool device_cut = false, bool device_walk = false,
                         const int64_t* d_delta = nullptr);
  struct {
    double host_ms = 0, gpu_ms = 0, total_ms = 0;  // host checks (+ walk, cut); device stages; whole build
    int64_t nkeys = 0, nslots = 0;
  } device_index;
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

// The query sequence of bwt_query_launch with its two independent replays forked onto two
// streams (block replay on st, level-B chunk replay on side; combine on st after both), so their
// latency-bound phases overlap.  Capturable: fork and join are event record / wait.
int query_forked(const TreeEngine& e, const BwtBatch* b, hipStream_t st, hipStream_t side, hipEvent_t fork,
                 hipEvent_t join);

// Device arrays of a batch planned on the device, sized for up to cap queries: the plan's
// outputs, the staging/result buffers, the planner's workspace and counts.
struct DevicePlanSlot {
  int64_t cap = 0;
  BwtBatch b{};           // the output pointers (filled in by bwt_plan_launch per batch)
  int64_t* dcounts = nullptr;
  void* work = nullptr;
  size_t work_bytes = 0;
  DeviceArena arena;
};
bool alloc_plan_slot(const TreeEngine& e, int64_t cap, DevicePlanSlot& s, char* err, size_t errlen);
// Plan M <= cap queries (device qstate / qlo / qhi) on stream st; s.b is then ready to launch.
int device_plan(const TreeEngine& e, DevicePlanSlot& s, int64_t M, const int32_t* d_qstate, const uint32_t* d_qlo,
                const uint32_t* d_qhi, hipStream_t st);

}  // namespace bwg
