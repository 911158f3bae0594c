// gen_kernels.hip -- the device workload generator (include/bwgpu_gen.h): one thread per array
// element, each element computed from its own counter-based draws (csrc/host/gen_draws.hpp, the
// definitions the host generator uses too), so the arrays come out bit-identical to the host's.
//
// Work per element is a few 64-bit multiplies, plus for Zipf keys a binary search inside a guide
// bracket of the cumulative table (most draws fall in the first ranks, whose brackets are a few
// entries) and for log-uniform ranges one exp.  The writes are coalesced streams; the kernels are
// bound by the integer pipeline and the table's random reads, far below HBM bandwidth.
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdint.h>
#include <vector>

#include "bwgpu_gen.h"
#include "gen_draws.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int kT = 256;
constexpr int64_t kMaxGrid = int64_t(1) << 16;

unsigned grid_for(int64_t n) { return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + kT - 1) / kT, kMaxGrid))); }

__device__ inline void tie(const BwtGenTies& t, int64_t i) {
  const uint32_t k = atomicAdd(t.count, 1u);
  if (k < t.cap) t.idx[k] = uint32_t(i);
}

struct DeviceExp {
  __device__ double operator()(double y) const { return exp(y); }
};

__global__ void __launch_bounds__(kT) gen_guide_kernel(const double* cum, uint64_t m, uint32_t* guide) {
  for (int64_t j = (int64_t)blockIdx.x * kT + threadIdx.x; j <= int64_t(bwg::kZipfGuide);
       j += (int64_t)gridDim.x * kT)
    guide[j] = bwg::zipf_guide_at(cum, m, uint32_t(j));
}

__global__ void __launch_bounds__(kT) gen_updates_kernel(const BwtGenSpec g, int64_t n, uint32_t* key,
                                                         int64_t* delta) {
  const bwg::GenStreams s(g.seed);
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    uint32_t k;
    int64_t d;
    bwg::gen_update_at(g, s, i, k, d);
    key[i] = k;
    delta[i] = d;
  }
}

__global__ void __launch_bounds__(kT) gen_zero_kernel(uint32_t* c) { *c = 0; }

__global__ void __launch_bounds__(kT) gen_state_queries_kernel(const BwtGenSpec g, int64_t n, uint64_t bound,
                                                               double log_n, uint32_t* qlo, uint32_t* qhi,
                                                               BwtGenTies t) {
  const bwg::GenStreams s(g.seed);
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    uint32_t lo, hi;
    if (bwg::gen_state_query_at(g, s, i, bound, log_n, DeviceExp{}, lo, hi)) tie(t, i);
    qlo[i] = lo;
    qhi[i] = hi;
  }
}

__global__ void __launch_bounds__(kT) gen_stream_kernel(const BwtGenSpec g, int form, const int64_t* req_q0,
                                                        const int64_t* req_app, int64_t nreq, int64_t nq,
                                                        double log_n, int32_t* qstate, uint32_t* qlo, uint32_t* qhi,
                                                        BwtGenTies t) {
  const bwg::GenStreams s(g.seed);
  for (int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x; q < nq; q += (int64_t)gridDim.x * kT) {
    int32_t st;
    uint32_t lo, hi;
    if (bwg::gen_stream_query_at(g, s, form, req_q0, req_app, nreq, q, log_n, DeviceExp{}, st, lo, hi)) tie(t, q);
    qstate[q] = st;
    qlo[q] = lo;
    qhi[q] = hi;
  }
}

__global__ void __launch_bounds__(kT) gen_appends_kernel(const BwtGenSpec g, int64_t napp, int32_t U,
                                                         int32_t* app_parent, uint32_t* app_key,
                                                         int64_t* app_delta) {
  const bwg::GenStreams s(g.seed);
  const int64_t n = napp * U;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    bwg::gen_append_at(g, s, napp, i, app_parent, app_key, app_delta);
}

double host_exp(double y) { return std::exp(y); }

bool zipf_ok(const BwtGenSpec* g) {
  return g->zipf_cum && g->zipf_guide && g->zipf_n >= 1 && g->zipf_n < (uint64_t(1) << 32);
}

// Range bound and draw streams of the log-uniform ranges of a query launch.
// ranges over a cfg-3 tree (its own queries and its streams') span the 32-bit domain
uint64_t range_bound(const BwtGenSpec* g, int what) {
  (void)what;
  return g->cfg == 3 ? 0xFFFFFFFFull : g->N;
}

}  // namespace

extern "C" int bwt_gen_guide_launch(const double* zipf_cum, uint64_t zipf_n, uint32_t* guide, void* stream) {
  if (!zipf_cum || !guide || zipf_n < 1 || zipf_n >= (uint64_t(1) << 32)) return int(hipErrorInvalidValue);
  hipLaunchKernelGGL(gen_guide_kernel, dim3(grid_for(bwg::kZipfGuide + 1)), dim3(kT), 0, hipStream_t(stream),
                     zipf_cum, zipf_n, guide);
  return int(hipGetLastError());
}

extern "C" int bwt_gen_updates_launch(const BwtGenSpec* g, uint32_t* key, int64_t* delta, void* stream) {
  if (!g || !key || !delta || g->S < 1 || g->U < 1 || g->cfg < 1 || g->cfg > 5 || g->N < 1)
    return int(hipErrorInvalidValue);
  if (g->cfg != 1 && (!zipf_ok(g) || (g->cfg == 3 && !g->dkey))) return int(hipErrorInvalidValue);
  const int64_t n = g->S * int64_t(g->U);
  hipLaunchKernelGGL(gen_updates_kernel, dim3(grid_for(n)), dim3(kT), 0, hipStream_t(stream), *g, n, key, delta);
  return int(hipGetLastError());
}

extern "C" int bwt_gen_state_queries_launch(const BwtGenSpec* g, uint32_t* qlo, uint32_t* qhi, BwtGenTies ties,
                                            void* stream) {
  if (!g || !qlo || !qhi || !ties.count || (ties.cap && !ties.idx) || g->S < 1 || g->Q < 1 || g->cfg < 1 ||
      g->cfg > 5 || g->N < 2)
    return int(hipErrorInvalidValue);
  const int64_t n = g->S * int64_t(g->Q);
  if (n >= (int64_t(1) << 32)) return int(hipErrorInvalidValue);
  const uint64_t bound = range_bound(g, 0);
  const double log_n = std::log(double(bound));
  hipStream_t st = hipStream_t(stream);
  hipLaunchKernelGGL(gen_zero_kernel, dim3(1), dim3(1), 0, st, ties.count);
  hipLaunchKernelGGL(gen_state_queries_kernel, dim3(grid_for(n)), dim3(kT), 0, st, *g, n, bound, log_n, qlo, qhi,
                     ties);
  return int(hipGetLastError());
}

extern "C" int bwt_gen_stream_launch(const BwtGenSpec* g, int32_t form, const int64_t* req_q0, const int64_t* req_app,
                                     int64_t nreq, int64_t nq, int32_t* qstate, uint32_t* qlo, uint32_t* qhi,
                                     BwtGenTies ties, void* stream) {
  if (!g || (form != 4 && form != 5) || nq < 0 || nreq < 0 || !ties.count || (ties.cap && !ties.idx) || g->S < 1 ||
      g->N < 2 || g->N > 0xFFFFFFFFull)
    return int(hipErrorInvalidValue);
  if (nq >= (int64_t(1) << 32) || (nq > 0 && (!qstate || !qlo || !qhi))) return int(hipErrorInvalidValue);
  if (form == 5 && (!req_q0 || !req_app || !zipf_ok(g))) return int(hipErrorInvalidValue);
  const double log_n = std::log(double(range_bound(g, 1)));
  hipStream_t st = hipStream_t(stream);
  hipLaunchKernelGGL(gen_zero_kernel, dim3(1), dim3(1), 0, st, ties.count);
  if (nq > 0)
    hipLaunchKernelGGL(gen_stream_kernel, dim3(grid_for(nq)), dim3(kT), 0, st, *g, int(form), req_q0,
                       form == 5 ? req_app : nullptr, nreq, nq, log_n, qstate, qlo, qhi, ties);
  return int(hipGetLastError());
}

extern "C" int bwt_gen_appends_launch(const BwtGenSpec* g, int64_t napp, int32_t append_U, int32_t* app_parent,
                                      uint32_t* app_key, int64_t* app_delta, void* stream) {
  if (!g || napp < 0 || append_U < 1 || g->S < 1 || !zipf_ok(g)) return int(hipErrorInvalidValue);
  if (napp == 0) return 0;
  if (!app_parent || !app_key || !app_delta) return int(hipErrorInvalidValue);
  const int64_t n = napp * append_U;
  hipLaunchKernelGGL(gen_appends_kernel, dim3(grid_for(n)), dim3(kT), 0, hipStream_t(stream), *g, napp, append_U,
                     app_parent, app_key, app_delta);
  return int(hipGetLastError());
}

extern "C" int bwt_gen_settle(const BwtGenSpec* g, int32_t what, uint32_t* qlo, uint32_t* qhi, BwtGenTies ties,
                              void* stream) {
  if (!g || (what != 0 && what != 1) || !qlo || !qhi || !ties.count) return int(hipErrorInvalidValue);
  hipStream_t st = hipStream_t(stream);
  uint32_t count = 0;
  hipError_t e = hipMemcpyAsync(&count, ties.count, sizeof(count), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return int(e);
  if (count > ties.cap) return int(hipErrorInvalidValue);
  if (count == 0) return 0;
  std::vector<uint32_t> idx(count), lo(count), hi(count);
  e = hipMemcpyAsync(idx.data(), ties.idx, count * sizeof(uint32_t), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return int(e);
  const uint64_t bound = range_bound(g, what);
  const double log_n = std::log(double(bound));
  const bwg::GenStreams s(g->seed);
  for (uint32_t k = 0; k < count; ++k) bwg::settle_range(s, what, bound, log_n, host_exp, idx[k], lo[k], hi[k]);
  for (uint32_t k = 0; k < count && e == hipSuccess; ++k) {
    e = hipMemcpyAsync(qlo + idx[k], &lo[k], sizeof(uint32_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(qhi + idx[k], &hi[k], sizeof(uint32_t), hipMemcpyHostToDevice, st);
  }
  const hipError_t sync = hipStreamSynchronize(st);
  return int(e != hipSuccess ? e : sync);
}
