// wave64.hpp -- exact int64 wavefront primitives for CDNA (64 lanes).
//
// Scans run on DPP row shifts and row broadcasts (GFX9 family, gfx950 included): Hillis-Steele
// over the four 16-lane rows (row_shr 1,2,4,8), then row_bcast:15 into rows 1 and 3 and
// row_bcast:31 into rows 2 and 3.  A 64-bit value moves as two 32-bit DPP moves with the same
// control; the add is an ordinary 64-bit add, so sums are exact (two's complement, no floats).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bw {

// DPP controls (GFX9 encoding)
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143;

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int64_t dpp_move64(int64_t x) {
  // lanes whose source is outside the row, or whose row is masked off, receive 0 (the `old`)
  const uint64_t u = (uint64_t)x;
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROW_MASK, 0xf, false);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}

// Inclusive prefix sum over the 64 lanes of the wavefront (all lanes must be active).
__device__ __forceinline__ int64_t wave_incl_scan(int64_t x) {
  x += dpp_move64<kRowShr1, 0xf>(x);
  x += dpp_move64<kRowShr2, 0xf>(x);
  x += dpp_move64<kRowShr4, 0xf>(x);
  x += dpp_move64<kRowShr8, 0xf>(x);
  x += dpp_move64<kRowBcast15, 0xa>(x);
  x += dpp_move64<kRowBcast31, 0xc>(x);
  return x;
}

__device__ __forceinline__ int64_t readlane64(int64_t x, int lane) {
  const uint64_t u = (uint64_t)x;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

// A value every active lane holds alike (e.g. loaded from one address by the whole wavefront),
// marked uniform so the compiler keeps it, and what is computed from it, in scalar registers.
__device__ __forceinline__ int64_t uniform64(int64_t x) {
  const uint64_t u = (uint64_t)x;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Orders this wavefront's LDS traffic: every LDS access issued before completes before any
// issued after (the kernels that use it keep all cross-lane LDS dependencies inside one wave).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void lds_add64(int64_t* p, int64_t v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}

// Workgroup-wide exclusive scan of n int64 values held in LDS, in place.  `tmp` is LDS scratch
// of at least blockDim.x/64 + 1 entries.  Every thread of the workgroup must call it.
__device__ inline void block_excl_scan_lds(int64_t* a, int n, int64_t* tmp) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
  const int per = (n + nt - 1) / nt;
  const int b0 = tid * per;
  const int b1 = (b0 + per < n) ? b0 + per : n;
  int64_t s = 0;
  for (int i = b0; i < b1; ++i) s += a[i];
  const int64_t inc = wave_incl_scan(s);
  if (lane == 63) tmp[wv] = inc;
  __syncthreads();
  if (tid == 0) {
    int64_t run = 0;
    for (int w = 0; w < nw; ++w) { const int64_t t = tmp[w]; tmp[w] = run; run += t; }
    tmp[nw] = run;
  }
  __syncthreads();
  int64_t run = tmp[wv] + inc - s;
  for (int i = b0; i < b1; ++i) { const int64_t t = a[i]; a[i] = run; run += t; }
  __syncthreads();
}

// Number of elements of the sorted array a[0..n) that are < x (lower bound).
template <typename T>
__device__ __forceinline__ int lower_bound_lds(const T* a, int n, T x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

}  // namespace bw
