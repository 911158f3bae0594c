// bw_floor.hip -- the device's random-access floor for the dense replay's two random streams, and its
// streaming bandwidth (one JSON line).
//
//   bw_floor [--gathers 205092024] [--scatters 102546012] [--rows 2048] [--copy-mb 2048]
//            [--workgroups W --visits V --per-visit P --share S]
//
// gather:  random 8-byte loads from rows of 2^20 int64 (8 MiB, one dense carry row), 4 independent
//          loads per thread -- what the replay does at a query's two endpoints;
// scatter: random 8-byte stores over an array of --scatters int64 -- the answers written by query id;
// copy:    a streaming device-to-device copy (16-byte loads and stores), the HBM reference rate;
// replay:  (with --workgroups) the random loads in the block replay's pattern: W persistent
//          512-thread workgroups take V row visits in turn, each visit P random loads from one row,
//          S consecutive visits on the same row (the work items of one block), so about W / S rows
//          are read at once.
// Each kernel runs once untimed, then three times between HIP events; the mean is reported.  On
// gfx950 every random 8-byte load fetches a whole 128-byte line from HBM, whatever the load's
// cache-policy bits or the memory kind (coarse, fine-grained or uncached), so `gather` is a floor
// for any design that reads one dense row entry per endpoint.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace {

constexpr int kRowShift = 20;  // 2^20 int64 per row
constexpr int kPer = 4;        // gathers per thread

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) fill_kernel(int64_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = i;
}

// nthreads * kPer gathers; rows is a power of two
__global__ void __launch_bounds__(256) gather_kernel(const int64_t* __restrict__ base, int64_t nthreads, uint32_t rows,
                                                     int64_t* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nthreads) return;
  int64_t s = 0;
#pragma unroll
  for (int g = 0; g < kPer; ++g) {
    const uint64_t h = mix64((uint64_t)t * kPer + g);
    const uint64_t row = h & (rows - 1);
    const uint64_t col = (h >> 32) & ((uint64_t(1) << kRowShift) - 1);
    s += base[(row << kRowShift) + col];
  }
  out[t] = s;
}

// visit v reads per_visit random entries of row (v / share) mod rows; workgroup g takes visits
// g, g + gridDim, ...
__global__ void __launch_bounds__(512) replay_gather_kernel(const int64_t* __restrict__ base, int64_t nvisits,
                                                            int per_visit, int share, uint32_t rows,
                                                            int64_t* __restrict__ out) {
  int64_t s = 0;
  for (int64_t v = blockIdx.x; v < nvisits; v += gridDim.x) {
    const uint64_t row = (uint64_t)((v / share) & (int64_t)(rows - 1)) << kRowShift;
#pragma unroll 4
    for (int j = (int)threadIdx.x; j < per_visit; j += 512) {
      const uint64_t h = mix64(((uint64_t)v << 32) ^ (uint64_t)j);
      s += base[row + ((h >> 32) & ((uint64_t(1) << kRowShift) - 1))];
    }
  }
  out[(int64_t)blockIdx.x * 512 + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) scatter_kernel(int64_t* __restrict__ a, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  a[mix64((uint64_t)t ^ 0x5bd1e995u) % (uint64_t)n] = t;
}

__global__ void __launch_bounds__(256) copy_kernel(const int4* __restrict__ src, int4* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}

#define CHECK(x)                                                                              \
  do {                                                                                        \
    const hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) {                                                                   \
      std::fprintf(stderr, "bw_floor: %s failed: %s\n", #x, hipGetErrorString(e_));          \
      std::exit(5);                                                                           \
    }                                                                                         \
  } while (0)

template <class F>
double time_ms(F launch) {
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  launch();
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  double tot = 0;
  for (int r = 0; r < 3; ++r) {
    CHECK(hipEventRecord(a, 0));
    launch();
    CHECK(hipEventRecord(b, 0));
    CHECK(hipEventSynchronize(b));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, a, b));
    tot += ms;
  }
  CHECK(hipEventDestroy(a));
  CHECK(hipEventDestroy(b));
  return tot / 3;
}

}  // namespace

int main(int argc, char** argv) {
  int64_t gathers = 205092024, scatters = 102546012, copy_mb = 2048, visits = 0;
  int workgroups = 0, per_visit = 0, share = 1;
  uint32_t rows = 2048;
  for (int i = 1; i < argc; ++i) {
    if (i + 1 >= argc) {
      std::fprintf(stderr, "usage: bw_floor [--gathers N] [--scatters N] [--rows R] [--copy-mb MB] "
                           "[--workgroups W --visits V --per-visit P --share S]\n");
      return 2;
    }
    const char* k = argv[i];
    const char* v = argv[++i];
    if (!std::strcmp(k, "--gathers")) gathers = std::atoll(v);
    else if (!std::strcmp(k, "--scatters")) scatters = std::atoll(v);
    else if (!std::strcmp(k, "--rows")) rows = (uint32_t)std::atoll(v);
    else if (!std::strcmp(k, "--copy-mb")) copy_mb = std::atoll(v);
    else if (!std::strcmp(k, "--workgroups")) workgroups = std::atoi(v);
    else if (!std::strcmp(k, "--visits")) visits = std::atoll(v);
    else if (!std::strcmp(k, "--per-visit")) per_visit = std::atoi(v);
    else if (!std::strcmp(k, "--share")) share = std::atoi(v);
    else {
      std::fprintf(stderr, "bw_floor: unknown argument %s\n", k);
      return 2;
    }
  }
  if (gathers < kPer || scatters < 1 || copy_mb < 1 || rows < 1 || (rows & (rows - 1)) != 0 || rows > (1u << 16) ||
      workgroups < 0 || (workgroups > 0 && (visits < 1 || per_visit < 1 || share < 1))) {
    std::fprintf(stderr, "bw_floor: bad sizes (rows must be a power of two <= 65536)\n");
    return 2;
  }
  const int64_t nthreads = gathers / kPer;
  const size_t span = (size_t)rows << kRowShift;  // int64 entries
  int64_t *rowsbuf = nullptr, *out = nullptr, *ans = nullptr;
  int4 *src = nullptr, *dst = nullptr;
  const int64_t n16 = copy_mb * (int64_t(1) << 20) / 16;
  CHECK(hipMalloc(&rowsbuf, span * 8));
  CHECK(hipMalloc(&out, (size_t)std::max<int64_t>(nthreads, (int64_t)workgroups * 512) * 8));
  CHECK(hipMalloc(&ans, (size_t)scatters * 8));
  CHECK(hipMalloc(&src, (size_t)n16 * 16));
  CHECK(hipMalloc(&dst, (size_t)n16 * 16));
  hipLaunchKernelGGL(fill_kernel, dim3(4096), dim3(256), 0, 0, rowsbuf, (int64_t)span);
  hipLaunchKernelGGL(fill_kernel, dim3(4096), dim3(256), 0, 0, reinterpret_cast<int64_t*>(src), n16 * 2);
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  const dim3 gg((unsigned)((nthreads + 255) / 256)), gs((unsigned)((scatters + 255) / 256));
  const double g_ms = time_ms([&] { hipLaunchKernelGGL(gather_kernel, gg, dim3(256), 0, 0, rowsbuf, nthreads, rows, out); });
  const double s_ms = time_ms([&] { hipLaunchKernelGGL(scatter_kernel, gs, dim3(256), 0, 0, ans, scatters); });
  const double c_ms = time_ms([&] { hipLaunchKernelGGL(copy_kernel, dim3(8192), dim3(256), 0, 0, src, dst, n16); });
  double r_ms = 0;
  if (workgroups > 0)
    r_ms = time_ms([&] {
      hipLaunchKernelGGL(replay_gather_kernel, dim3((unsigned)workgroups), dim3(512), 0, 0, rowsbuf, visits, per_visit,
                         share, rows, out);
    });
  hipDeviceProp_t prop{};
  CHECK(hipGetDeviceProperties(&prop, 0));
  const double ng = (double)nthreads * kPer;
  std::printf("{\"device\":\"%s\",\"gather\":{\"count\":%.0f,\"rows\":%u,\"row_bytes\":%d,\"ms\":%.4f,\"per_s\":%.1f},"
              "\"scatter\":{\"count\":%lld,\"span_bytes\":%lld,\"ms\":%.4f,\"per_s\":%.1f},"
              "\"copy\":{\"bytes\":%lld,\"ms\":%.4f,\"GBps\":%.1f}",
              prop.gcnArchName, ng, rows, 8 << kRowShift, g_ms, ng / (g_ms * 1e-3), (long long)scatters,
              (long long)scatters * 8, s_ms, (double)scatters / (s_ms * 1e-3), (long long)(n16 * 16), c_ms,
              2.0 * (double)n16 * 16 / (c_ms * 1e-3) / 1e9);
  if (workgroups > 0) {
    const double nr = (double)visits * per_visit;
    std::printf(",\"replay\":{\"count\":%.0f,\"workgroups\":%d,\"visits\":%lld,\"per_visit\":%d,\"share\":%d,"
                "\"rows_at_once\":%.1f,\"ms\":%.4f,\"per_s\":%.1f}",
                nr, workgroups, (long long)visits, per_visit, share, (double)workgroups / share, r_ms, nr / (r_ms * 1e-3));
  }
  std::printf("}\n");
  CHECK(hipFree(rowsbuf));
  CHECK(hipFree(out));
  CHECK(hipFree(ans));
  CHECK(hipFree(src));
  CHECK(hipFree(dst));
  return 0;
}
