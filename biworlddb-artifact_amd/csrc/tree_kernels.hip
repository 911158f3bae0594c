// tree_kernels.hip -- CDNA4 (gfx950) kernels of the generalised batch path (bwgpu_tree.h).
//
// Stage map onto the reference's v7 algorithm (src/batch_engine_v7.hpp:832-968):
//   bwt_block_kernel      <- v7_block_kernel / run_blocks phase 1 (src/gpu_v7.cu:10-82): one
//                            workgroup replays one block from a zero local prefix; the class
//                            domain is the block's own distinct keys, so the Fenwick tree always
//                            fits LDS whatever the key space; steps are undo-aware (Euler walk).
//   bwt_chunk_hist_kernel <- the block's dense dV (src/gpu_v7.cu:81), per carry chunk
//   bwt_seg_*_kernel      <- the serial carry pass (src/batch_engine_v7.hpp:936-945), as a
//                            three-pass segmented column scan over chunk rows
//   bwt_chunk_replay_kernel <- lvals over the carry level (src/gpu_v7.cu:72-78), one
//                            wavefront per chunk with an LDS Fenwick tree
//   bwt_combine_kernel    <- gpu_combine_kernel (src/gpu_v7.cu:84-96)
//   bwt_rank_kernel       <- CsrBase::rank (src/sparse_engine.hpp:62-97), CSR form
// All arithmetic is int64 adds (LDS ds_add_u64, DPP scans): results are exact and independent
// of scheduling order.
#include <hip/hip_runtime.h>
#include <mutex>
#include <stdint.h>
#include <stdio.h>

#include "bwgpu_tree.h"
#include "wave64.hpp"

using namespace bw;

namespace {

__device__ __forceinline__ int64_t fen_prefix(const int64_t* fen, int y) {
  int64_t r = 0;
  for (int j = y; j > 0; j -= j & -j) r += fen[j];
  return r;
}

// ---------------------------------------------------------------- block replay
__global__ void __launch_bounds__(1024) bwt_block_kernel(const BwtPassArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int64_t* fen = reinterpret_cast<int64_t*>(smem_raw);          // rmax+1, 1-based
  uint32_t* sk = reinterpret_cast<uint32_t*>(fen + a.rmax + 1);  // rmax sorted keys
  const int b = blockIdx.x;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int64_t k0 = a.blk_key0[b];
  const int R = (int)(a.blk_key0[b + 1] - k0);
  for (int i = tid; i < R; i += nt) sk[i] = a.bkey[k0 + i];
  for (int i = tid; i <= R; i += nt) fen[i] = 0;
  __syncthreads();
  const int U = a.U, Q = a.Q;
  const int s1 = a.blk_step0[b + 1];
  for (int s = a.blk_step0[b]; s < s1; ++s) {
    const int st = a.steps[s];
    const int64_t v = st > 0 ? st : -st;
    const int64_t ub = (v - 1) * (int64_t)U;
    for (int i = tid; i < U; i += nt) {
      const uint32_t k = a.key[ub + i];
      const int64_t d = (st > 0) ? a.delta[ub + i] : -a.delta[ub + i];
      const int r = lower_bound_lds(sk, R, k);
      for (int j = r + 1; j <= R; j += j & -j) lds_add64(&fen[j], d);
    }
    if (st > 0) {  // queries of an entered state see every event before them in the block
      __syncthreads();
      const int64_t qb = (v - 1) * (int64_t)Q;
      for (int kq = tid; kq < Q; kq += nt) {
        const int ylo = lower_bound_lds(sk, R, a.qlo[qb + kq]);
        const int yhi = lower_bound_lds(sk, R, a.qhi[qb + kq]);
        a.ansr[qb + kq] = fen_prefix(fen, yhi) - fen_prefix(fen, ylo);
      }
      __syncthreads();
    }
  }
  __syncthreads();
  // net weight of the block per distinct key (carry level B) and per key tile (level C)
  for (int r = tid; r < R; r += nt) a.agg[k0 + r] = fen_prefix(fen, r + 1) - fen_prefix(fen, r);
  const int64_t t0 = a.blk_tile0[b];
  const int T = (int)(a.blk_tile0[b + 1] - t0);
  for (int s = tid; s < T; s += nt) {
    const int r0 = a.tile_rank0[t0 + s];
    const int r1 = (s + 1 < T) ? (int)a.tile_rank0[t0 + s + 1] : R;
    a.tagg[t0 + s] = fen_prefix(fen, r1) - fen_prefix(fen, r0);
  }
}

// ---------------------------------------------------------------- chunk histogram
// One wavefront per chunk: dense class histogram of the chunk's key items, written as its
// exclusive prefix over classes (cp row = sum of weights with class < y).
__global__ void __launch_bounds__(256) bwt_chunk_hist_kernel(const BwtPassArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * a.chunk_waves + wv;
  if (c >= a.nchunk) return;  // whole wavefront leaves; no workgroup barrier follows
  int64_t* h = reinterpret_cast<int64_t*>(smem_raw) + (int64_t)wv * (a.max_cls + 1);
  const int chain = a.chunk_chain[c];
  const int ncls = a.chain_ncls[chain];
  const int64_t* wsrc = a.chain_src[chain] ? a.tagg : a.agg;
  for (int i = lane; i < ncls; i += 64) h[i] = 0;
  wave_lds_sync();
  const int64_t i1 = a.chunk_item0[c + 1];
  for (int64_t it = a.chunk_item0[c] + lane; it < i1; it += 64) {
    const uint64_t x = a.items[it];
    if (!(x & BWT_ITEM_EP)) lds_add64(&h[(int)((x >> BWT_CLASS_SHIFT) & 0xFFF)], wsrc[x & BWT_REF_MASK]);
  }
  wave_lds_sync();
  int64_t* row = a.cp + a.chunk_cp0[c];
  int64_t carry = 0;
  for (int base = 0; base < ncls; base += 64) {
    const int y = base + lane;
    const int64_t v = (y < ncls) ? h[y] : 0;
    const int64_t inc = wave_incl_scan(v);
    if (y < ncls) row[y] = carry + inc - v;
    carry += readlane64(inc, 63);
  }
}

// ---------------------------------------------------------------- segmented carry scan
// cp rows of one chain form a matrix [chunks][ncls]; the carry of chunk r is the exclusive
// column sum over rows < r.  Pass A: segment sums; pass B: scan over segments; pass C: rows.
__global__ void __launch_bounds__(256) bwt_seg_sum_kernel(const BwtPassArgs a) {
  const int64_t s = blockIdx.x;
  const int y = blockIdx.y * blockDim.x + threadIdx.x;
  const int64_t c0 = a.seg_chunk0[s], c1 = a.seg_chunk0[s + 1];
  const int ncls = a.chain_ncls[a.chunk_chain[c0]];
  if (y >= ncls) return;
  int64_t acc = 0;
  for (int64_t c = c0; c < c1; ++c) acc += a.cp[a.chunk_cp0[c] + y];
  a.segsum[a.seg_cp0[s] + y] = acc;
}

__global__ void __launch_bounds__(256) bwt_seg_scan_kernel(const BwtPassArgs a) {
  const int chain = blockIdx.x;
  const int y = blockIdx.y * blockDim.x + threadIdx.x;
  if (y >= a.chain_ncls[chain]) return;
  int64_t run = 0;
  const int64_t e = a.chain_seg0[chain + 1];
  for (int64_t s = a.chain_seg0[chain]; s < e; ++s) {
    int64_t* p = a.segsum + a.seg_cp0[s] + y;
    const int64_t t = *p;
    *p = run;
    run += t;
  }
}

__global__ void __launch_bounds__(256) bwt_seg_apply_kernel(const BwtPassArgs a) {
  const int64_t s = blockIdx.x;
  const int y = blockIdx.y * blockDim.x + threadIdx.x;
  const int64_t c0 = a.seg_chunk0[s], c1 = a.seg_chunk0[s + 1];
  const int ncls = a.chain_ncls[a.chunk_chain[c0]];
  if (y >= ncls) return;
  int64_t run = a.segsum[a.seg_cp0[s] + y];
  for (int64_t c = c0; c < c1; ++c) {
    int64_t* p = a.cp + a.chunk_cp0[c] + y;
    const int64_t t = *p;
    *p = run;
    run += t;
  }
}

// ---------------------------------------------------------------- chunk replay
// One wavefront per chunk, 64 items per window.  An endpoint item's value is the chunk's carry
// at its class + the chunk-local Fenwick prefix (items of earlier windows) + the key items of
// its own window that precede it (broadcast lane by lane with readlane).  Key items are then
// added to the Fenwick tree.
__global__ void __launch_bounds__(256) bwt_chunk_replay_kernel(const BwtPassArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * a.chunk_waves + wv;
  if (c >= a.nchunk) return;
  int64_t* fen = reinterpret_cast<int64_t*>(smem_raw) + (int64_t)wv * (a.max_cls + 1);
  const int chain = a.chunk_chain[c];
  const int ncls = a.chain_ncls[chain];
  const int64_t* wsrc = a.chain_src[chain] ? a.tagg : a.agg;
  const int64_t* carry = a.cp + a.chunk_cp0[c];
  int64_t* gout = a.g + (int64_t)a.chain_out[chain] * (2 * a.S * (int64_t)a.Q);
  for (int i = lane; i <= ncls; i += 64) fen[i] = 0;
  wave_lds_sync();
  const int64_t i0 = a.chunk_item0[c], i1 = a.chunk_item0[c + 1];
  for (int64_t w0 = i0; w0 < i1; w0 += 64) {
    const int64_t it = w0 + lane;
    const bool valid = it < i1;
    const uint64_t x = valid ? a.items[it] : 0;
    const bool isep = valid && (x & BWT_ITEM_EP);
    const bool iskey = valid && !(x & BWT_ITEM_EP);
    const int cls = (int)((x >> BWT_CLASS_SHIFT) & 0xFFF);
    const int64_t w = iskey ? wsrc[x & BWT_REF_MASK] : 0;
    const uint64_t epmask = ballot64(isep);
    uint64_t keymask = ballot64(iskey);
    if (epmask) {
      int64_t val = isep ? carry[cls] + fen_prefix(fen, cls) : 0;
      // keys of this window that precede the endpoint: only lanes below the last endpoint matter
      const int last_ep = 63 - __builtin_clzll(epmask);
      keymask &= (last_ep == 63) ? ~0ull : ((1ull << last_ep) - 1);
      while (keymask) {
        const int j = __builtin_ctzll(keymask);
        keymask &= keymask - 1;
        const int kc = __builtin_amdgcn_readlane(cls, j);
        const int64_t kw = readlane64(w, j);
        if (j < lane && kc < cls) val += kw;
      }
      if (isep) gout[x & BWT_REF_MASK] = val;
    }
    if (iskey)
      for (int j = cls + 1; j <= ncls; j += j & -j) lds_add64(&fen[j], w);
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------- combine
__global__ void __launch_bounds__(256) bwt_combine_kernel(const BwtPassArgs a) {
  const int64_t n = a.S * (int64_t)a.Q;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t* gc = a.g;
  const int64_t* gb = a.g + 2 * n;
  a.answers[i] = a.ansr[i] + (gc[2 * i + 1] + gb[2 * i + 1]) - (gc[2 * i] + gb[2 * i]);
}

// ---------------------------------------------------------------- rank remap
__global__ void __launch_bounds__(256) bwt_rank_kernel(const BwtRankArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t x = a.in[i];
  const uint32_t h = x >> 16;
  const uint16_t l = (uint16_t)(x & 0xFFFFu);
  uint32_t lo = a.row_ptr[h], hi = a.row_ptr[h + 1];
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a.low16[mid] < l) lo = mid + 1; else hi = mid;
  }
  a.out[i] = lo;
}

// ---------------------------------------------------------------- launch helpers
size_t block_lds_bytes(const BwtPassArgs* a) {
  return (size_t)(a->rmax + 1) * sizeof(int64_t) + (size_t)a->rmax * sizeof(uint32_t);
}
size_t chunk_lds_bytes(const BwtPassArgs* a) {
  return (size_t)a->chunk_waves * (size_t)(a->max_cls + 1) * sizeof(int64_t);
}

int check_args(const BwtPassArgs* a) {
  if (!a || a->S < 1 || a->U < 1 || a->Q < 1 || a->nblk < 1) return (int)hipErrorInvalidValue;
  if (a->block_threads < 64 || a->block_threads > 1024 || (a->block_threads & 63)) return (int)hipErrorInvalidValue;
  if (a->rmax < 1 || a->rmax > 8192 || a->max_cls < 1 || a->max_cls > 4096) return (int)hipErrorInvalidValue;
  if (a->chunk_waves < 1 || a->chunk_waves > 4) return (int)hipErrorInvalidValue;
  if (block_lds_bytes(a) > 160 * 1024 || chunk_lds_bytes(a) > 160 * 1024) return (int)hipErrorInvalidValue;
  return 0;
}

int set_lds_limits() {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
    const int lim = 160 * 1024;
    hipError_t e1 = hipFuncSetAttribute(reinterpret_cast<const void*>(bwt_block_kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    hipError_t e2 = hipFuncSetAttribute(reinterpret_cast<const void*>(bwt_chunk_hist_kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    hipError_t e3 = hipFuncSetAttribute(reinterpret_cast<const void*>(bwt_chunk_replay_kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    err = e1 != hipSuccess ? e1 : (e2 != hipSuccess ? e2 : e3);
  });
  return (int)err;
}

int report(const char* what, hipError_t e) {
  if (e != hipSuccess) fprintf(stderr, "bwgpu: %s: %s\n", what, hipGetErrorString(e));
  return (int)e;
}

}  // namespace

extern "C" int bwt_block_replay_launch(const BwtPassArgs* a, void* stream) {
  int rc = check_args(a);
  if (rc) return rc;
  if ((rc = set_lds_limits())) return rc;
  hipLaunchKernelGGL(bwt_block_kernel, dim3(a->nblk), dim3(a->block_threads), block_lds_bytes(a),
                     (hipStream_t)stream, *a);
  return report("block replay", hipGetLastError());
}

extern "C" int bwt_chunk_hist_launch(const BwtPassArgs* a, void* stream) {
  int rc = check_args(a);
  if (rc) return rc;
  if ((rc = set_lds_limits())) return rc;
  if (a->nchunk == 0) return 0;
  const int64_t grid = (a->nchunk + a->chunk_waves - 1) / a->chunk_waves;
  hipLaunchKernelGGL(bwt_chunk_hist_kernel, dim3((unsigned)grid), dim3(64 * a->chunk_waves),
                     chunk_lds_bytes(a), (hipStream_t)stream, *a);
  return report("chunk histogram", hipGetLastError());
}

extern "C" int bwt_carry_scan_launch(const BwtPassArgs* a, void* stream) {
  int rc = check_args(a);
  if (rc) return rc;
  if (a->nseg == 0) return 0;
  const unsigned ytiles = (unsigned)((a->max_cls + 255) / 256);
  hipLaunchKernelGGL(bwt_seg_sum_kernel, dim3((unsigned)a->nseg, ytiles), dim3(256), 0, (hipStream_t)stream, *a);
  if ((rc = report("carry scan A", hipGetLastError()))) return rc;
  hipLaunchKernelGGL(bwt_seg_scan_kernel, dim3((unsigned)a->nchain, ytiles), dim3(256), 0, (hipStream_t)stream, *a);
  if ((rc = report("carry scan B", hipGetLastError()))) return rc;
  hipLaunchKernelGGL(bwt_seg_apply_kernel, dim3((unsigned)a->nseg, ytiles), dim3(256), 0, (hipStream_t)stream, *a);
  return report("carry scan C", hipGetLastError());
}

extern "C" int bwt_chunk_replay_launch(const BwtPassArgs* a, void* stream) {
  int rc = check_args(a);
  if (rc) return rc;
  if ((rc = set_lds_limits())) return rc;
  if (a->nchunk == 0) return 0;
  const int64_t grid = (a->nchunk + a->chunk_waves - 1) / a->chunk_waves;
  hipLaunchKernelGGL(bwt_chunk_replay_kernel, dim3((unsigned)grid), dim3(64 * a->chunk_waves),
                     chunk_lds_bytes(a), (hipStream_t)stream, *a);
  return report("chunk replay", hipGetLastError());
}

extern "C" int bwt_combine_launch(const BwtPassArgs* a, void* stream) {
  int rc = check_args(a);
  if (rc) return rc;
  const int64_t n = a->S * (int64_t)a->Q;
  hipLaunchKernelGGL(bwt_combine_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *a);
  return report("combine", hipGetLastError());
}

extern "C" int bwt_pass_launch(const BwtPassArgs* a, void* stream) {
  int rc;
  if ((rc = bwt_block_replay_launch(a, stream))) return rc;
  if ((rc = bwt_chunk_hist_launch(a, stream))) return rc;
  if ((rc = bwt_carry_scan_launch(a, stream))) return rc;
  if ((rc = bwt_chunk_replay_launch(a, stream))) return rc;
  return bwt_combine_launch(a, stream);
}

extern "C" int bwt_rank_remap_launch(const BwtRankArgs* a, void* stream) {
  if (!a || a->n < 0) return (int)hipErrorInvalidValue;
  if (a->n == 0) return 0;
  hipLaunchKernelGGL(bwt_rank_kernel, dim3((unsigned)((a->n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *a);
  return report("rank remap", hipGetLastError());
}
