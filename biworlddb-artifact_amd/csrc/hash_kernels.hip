// hash_kernels.hip -- FNV-1a 64 of a device buffer on the device (include/bwgpu_extract.h), by
// the block split of csrc/host/fnv_split.hpp:
//   1. maps     one workgroup per 4 KiB block runs the block's bytes through the low-byte chain
//               for all 256 entry values at once (two per thread, packed) -> a 256-entry map;
//   2. compose  maps are composed 64 at a time, level by level, until one group is left;
//   3. walk     from the start value's low byte, each level's group of 64 maps is walked to give
//               every map its entry low byte, down to the blocks;
//   4. poly     one thread per block replays its bytes from its entry low byte -> S_j;
//   5. finish   h = start * P^n + sum_j S_j * P^(n - end_j), one workgroup (a polynomial in
//               P^4096, by Horner over ranges of blocks).
// Step 1 does 256 chain steps per byte (two per VALU op sequence) and dominates: the work is
// fixed integer ALU per byte, spread over every CU, against one dependent multiply per byte for
// the sequential hash on one CPU core.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "bwgpu_extract.h"
#include "fnv_split.hpp"

namespace {

constexpr int kMapThreads = 128;  // two entry values per thread
constexpr int kGroup = 64;        // maps composed / walked per workgroup
constexpr size_t kWords = bwg::kFnvBlock / 4;

struct Levels {
  int n = 0;
  int64_t count[16];
  size_t map_off[16], start_off[16];  // byte offsets in the workspace
  size_t poly_off = 0, total = 0;
};

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

Levels levels_of(size_t n) {
  Levels L;
  const int64_t nblk = (int64_t)((n + bwg::kFnvBlock - 1) / bwg::kFnvBlock);
  size_t off = 0;
  int64_t c = nblk;
  while (true) {
    L.count[L.n] = c;
    L.map_off[L.n] = off;
    off = align256(off + (size_t)c * 256);
    L.start_off[L.n] = off;
    off = align256(off + (size_t)c);
    ++L.n;
    if (c <= kGroup) break;
    c = (c + kGroup - 1) / kGroup;
  }
  L.poly_off = off;
  L.total = align256(off + (size_t)nblk * 8);
  return L;
}

__global__ void __launch_bounds__(kMapThreads) fnv_map_kernel(const unsigned char* bytes, size_t n,
                                                              unsigned char* maps) {
  __shared__ uint32_t w[kWords];
  const int64_t blk = blockIdx.x;
  const size_t b0 = (size_t)blk * bwg::kFnvBlock;
  const size_t len = n - b0 < bwg::kFnvBlock ? n - b0 : bwg::kFnvBlock;
  const size_t full = len / 4;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(bytes + b0);
  for (size_t i = threadIdx.x; i < full; i += kMapThreads) w[i] = src[i];
  if (threadIdx.x == 0 && (len & 3)) {  // a last partial word, read byte by byte
    uint32_t v = 0;
    for (size_t k = 0; k < (len & 3); ++k) v |= (uint32_t)bytes[b0 + full * 4 + k] << (8 * k);
    w[full] = v;
  }
  __syncthreads();
  const uint32_t t = threadIdx.x;
  uint32_t x = t | ((t + kMapThreads) << 16);
  for (size_t i = 0; i < full; ++i) {
    const uint32_t v = w[i];
    x = bwg::fnv_low_step2(x, v & 0xFFu);
    x = bwg::fnv_low_step2(x, (v >> 8) & 0xFFu);
    x = bwg::fnv_low_step2(x, (v >> 16) & 0xFFu);
    x = bwg::fnv_low_step2(x, v >> 24);
  }
  if (len & 3) {
    const uint32_t v = w[full];
    for (size_t k = 0; k < (len & 3); ++k) x = bwg::fnv_low_step2(x, (v >> (8 * k)) & 0xFFu);
  }
  maps[blk * 256 + t] = (unsigned char)(x & 0xFFu);
  maps[blk * 256 + t + kMapThreads] = (unsigned char)(x >> 16);
}

// out[g] = in[g*64 + last] o ... o in[g*64]   (apply the group's maps in order)
__global__ void __launch_bounds__(256) fnv_compose_kernel(const unsigned char* in, int64_t n_in, unsigned char* out) {
  __shared__ unsigned char m[kGroup * 256];
  const int64_t g = blockIdx.x;
  const int64_t j0 = g * kGroup;
  const int cnt = (int)(n_in - j0 < kGroup ? n_in - j0 : kGroup);
  const uint32_t* src = reinterpret_cast<const uint32_t*>(in + j0 * 256);
  uint32_t* dst = reinterpret_cast<uint32_t*>(m);
  for (int i = threadIdx.x; i < cnt * 64; i += 256) dst[i] = src[i];
  __syncthreads();
  uint32_t x = threadIdx.x;
  for (int j = 0; j < cnt; ++j) x = m[j * 256 + x];
  out[g * 256 + threadIdx.x] = (unsigned char)x;
}

// starts[j] = entry low byte of map j: group g starts at parent[g] (or top when parent is NULL)
__global__ void __launch_bounds__(256) fnv_walk_kernel(const unsigned char* maps, int64_t n,
                                                       const unsigned char* parent, uint32_t top,
                                                       unsigned char* starts) {
  __shared__ unsigned char m[kGroup * 256];
  const int64_t g = blockIdx.x;
  const int64_t j0 = g * kGroup;
  const int cnt = (int)(n - j0 < kGroup ? n - j0 : kGroup);
  const uint32_t* src = reinterpret_cast<const uint32_t*>(maps + j0 * 256);
  uint32_t* dst = reinterpret_cast<uint32_t*>(m);
  for (int i = threadIdx.x; i < cnt * 64; i += 256) dst[i] = src[i];
  __syncthreads();
  if (threadIdx.x != 0) return;
  uint32_t x = parent ? parent[g] : top;
  for (int j = 0; j < cnt; ++j) {
    starts[j0 + j] = (unsigned char)x;
    x = m[j * 256 + x];
  }
}

__global__ void __launch_bounds__(256) fnv_poly_kernel(const unsigned char* bytes, size_t n, int64_t nblk,
                                                       const unsigned char* starts, bool a16, uint64_t* poly) {
  const int64_t blk = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (blk >= nblk) return;
  const size_t b0 = (size_t)blk * bwg::kFnvBlock;
  const size_t len = n - b0 < bwg::kFnvBlock ? n - b0 : bwg::kFnvBlock;
  const size_t full = len / 4;
  const uint4* src4 = reinterpret_cast<const uint4*>(bytes + b0);
  uint32_t l = starts[blk];
  uint64_t s = 0;
  size_t i = 0;
  for (; a16 && i + 4 <= full; i += 4) {  // 16-byte loads where the buffer is 16-byte aligned
    const uint4 v = src4[i / 4];
    bwg::fnv_poly_word(v.x, 4, l, s);
    bwg::fnv_poly_word(v.y, 4, l, s);
    bwg::fnv_poly_word(v.z, 4, l, s);
    bwg::fnv_poly_word(v.w, 4, l, s);
  }
  const uint32_t* src = reinterpret_cast<const uint32_t*>(bytes + b0);
  for (; i < full; ++i) bwg::fnv_poly_word(src[i], 4, l, s);
  for (size_t k = full * 4; k < len; ++k) bwg::fnv_poly_word(bytes[b0 + k], 1, l, s);
  poly[blk] = s;
}

// h = start * P^n + S_last + P^len_last * sum_{j < last} S_j Q^(last-1-j): each thread takes a
// range of the full blocks (Horner in Q), the ranges' shares are summed.
__global__ void __launch_bounds__(256) fnv_finish_kernel(const uint64_t* poly, int64_t nblk, size_t n, uint64_t start,
                                                         uint64_t* out) {
  __shared__ uint64_t part[256];
  const int64_t m = nblk > 0 ? nblk - 1 : 0;  // full blocks before the last one
  const int64_t per = (m + 255) / 256;
  const int64_t a = (int64_t)threadIdx.x * per, b = a + per < m ? a + per : m;
  part[threadIdx.x] = bwg::fnv_range_sum(poly, a, b, m);
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    uint64_t h = start * bwg::fnv_pow(n);
    if (nblk > 0) h += poly[m] + bwg::fnv_pow(n - (size_t)m * bwg::kFnvBlock) * part[0];
    *out = h;
  }
}

}  // namespace

extern "C" int bw_fnv1a_workspace(size_t n, size_t* bytes) {
  if (!bytes) return int(hipErrorInvalidValue);
  *bytes = n ? levels_of(n).total : 256;
  return 0;
}

extern "C" int bw_fnv1a_device(const void* d_bytes, size_t n, uint64_t start, void* work, size_t work_bytes,
                               uint64_t* d_result, void* stream) {
  if (!d_result || (n && (!d_bytes || (reinterpret_cast<uintptr_t>(d_bytes) & 3)))) return int(hipErrorInvalidValue);
  hipStream_t st = hipStream_t(stream);
  const unsigned char* bytes = static_cast<const unsigned char*>(d_bytes);
  if (n == 0) {
    hipLaunchKernelGGL(fnv_finish_kernel, dim3(1), dim3(256), 0, st, (const uint64_t*)nullptr, (int64_t)0, (size_t)0,
                       start, d_result);
    return int(hipGetLastError());
  }
  const Levels L = levels_of(n);
  if (!work || work_bytes < L.total || (reinterpret_cast<uintptr_t>(work) & 255)) return int(hipErrorInvalidValue);
  unsigned char* ws = static_cast<unsigned char*>(work);
  const int64_t nblk = L.count[0];
  hipLaunchKernelGGL(fnv_map_kernel, dim3((unsigned)nblk), dim3(kMapThreads), 0, st, bytes, n, ws + L.map_off[0]);
  for (int k = 1; k < L.n; ++k)
    hipLaunchKernelGGL(fnv_compose_kernel, dim3((unsigned)L.count[k]), dim3(256), 0, st, ws + L.map_off[k - 1],
                       L.count[k - 1], ws + L.map_off[k]);
  for (int k = L.n - 1; k >= 0; --k) {
    const int64_t groups = (L.count[k] + kGroup - 1) / kGroup;
    hipLaunchKernelGGL(fnv_walk_kernel, dim3((unsigned)groups), dim3(256), 0, st, ws + L.map_off[k], L.count[k],
                       k == L.n - 1 ? (const unsigned char*)nullptr : ws + L.start_off[k + 1],
                       (uint32_t)(start & 0xFFu), ws + L.start_off[k]);
  }
  uint64_t* poly = reinterpret_cast<uint64_t*>(ws + L.poly_off);
  hipLaunchKernelGGL(fnv_poly_kernel, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, st, bytes, n, nblk,
                     ws + L.start_off[0], (reinterpret_cast<uintptr_t>(d_bytes) & 15) == 0, poly);
  hipLaunchKernelGGL(fnv_finish_kernel, dim3(1), dim3(256), 0, st, (const uint64_t*)poly, nblk, n, start, d_result);
  return int(hipGetLastError());
}
