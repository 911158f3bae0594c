// tree_gen.hpp -- host runtime of the device workload generator (include/bwgpu_gen.h): the
// sequential parts (Zipf table, cfg-3 distinct keys, request scan) on the host, the per-element
// arrays on the device, ties settled with the host's exp.  The results equal tree_generate /
// stream_generate element for element.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "bwgpu_gen.h"
#include "tree_engine.hpp"
#include "tree_workload.hpp"

namespace bwg {

// Device Zipf inverse-CDF table (zipf_cumulative(m, s)) and its guide.
struct DeviceZipf {
  double* cum = nullptr;
  uint32_t* guide = nullptr;
  uint64_t m = 0;
  DeviceArena arena;
};
bool device_zipf(uint64_t m, double s, DeviceZipf& z, hipStream_t st, char* err, size_t errlen);

// The tree's per-element arrays on the device.  dkey: the host's sorted distinct keys (cfg 3).
struct DeviceTree {
  uint32_t* key = nullptr;
  int64_t* delta = nullptr;
  uint32_t *qlo = nullptr, *qhi = nullptr;
  uint32_t ties = 0;    // ranges settled by the host
  double gpu_ms = 0;    // device time of the launches (table upload excluded)
  DeviceArena arena;
};
bool device_generate_tree(const TreeSpec& spec, const std::vector<uint32_t>& dkey, DeviceTree& out,
                          hipStream_t st, char* err, size_t errlen);

// A request stream's per-element arrays on the device.  scan: the host's stream_scan of it.
struct DeviceStream {
  int32_t* qstate = nullptr;
  uint32_t *qlo = nullptr, *qhi = nullptr;
  int32_t* app_parent = nullptr;
  uint32_t* app_key = nullptr;
  int64_t* app_delta = nullptr;
  int64_t nq = 0, napp = 0;
  uint32_t ties = 0;
  double gpu_ms = 0;
  DeviceArena arena;
};
bool device_generate_stream(const TreeSpec& tree, const RequestStream& scan, DeviceStream& out, hipStream_t st,
                            char* err, size_t errlen,
                            const std::vector<uint32_t>& dkey = {});

// Elements of a device array that differ from the host vector (all of them if the copy fails).
template <typename T>
int64_t count_device_mismatch(const T* dev, const std::vector<T>& host) {
  std::vector<T> tmp(host.size());
  if (!host.empty() && hipMemcpy(tmp.data(), dev, host.size() * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
    return int64_t(host.size());
  int64_t n = 0;
  for (size_t i = 0; i < host.size(); ++i) n += tmp[i] != host[i];
  return n;
}

}  // namespace bwg
