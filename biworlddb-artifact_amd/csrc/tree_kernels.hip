// tree_kernels.hip -- CDNA4 (gfx950) kernels of the generalised batch path (bwgpu_tree.h).
//
// Stage map onto the reference's v7 algorithm (src/batch_engine_v7.hpp:832-968):
//   bwt_block_kernel        <- v7_block_kernel / run_blocks phase 1 (src/gpu_v7.cu:10-82): one
//                              workgroup replays one block from a zero local prefix over the
//                              block's own distinct keys (so the Fenwick tree always fits LDS),
//                              undo-aware (Euler steps); in-block prefixes at query endpoints;
//                              optionally the block's net weights (the dense dV, compacted)
//   bwt_chunk_hist_kernel   <- dV per carry chunk
//   bwt_seg_*_kernel        <- the serial carry pass (src/batch_engine_v7.hpp:936-945) as a
//                              three-pass segmented column scan over chunk rows
//   bwt_chunk_replay_kernel <- lvals on the carry levels (src/gpu_v7.cu:72-78): one wavefront
//                              per chunk merges its static key items with the batch's endpoint
//                              items, LDS Fenwick tree, in-window corrections by readlane
//   bwt_combine_kernel      <- gpu_combine_kernel (src/gpu_v7.cu:84-96)
//   bwt_rank_kernel         <- CsrBase::rank (src/sparse_engine.hpp:62-97), CSR form
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
// Events of step s: the U updates of state |steps[s]| with their precomputed in-block ranks.
// The next step's (rank, delta) pair is loaded before the current one is applied, so the global
// latency overlaps the LDS Fenwick updates; workgroup barriers only surround query evaluation.
__global__ void __launch_bounds__(1024) bwt_block_kernel(const BwtIndex x, const BwtBatch q, const BwtLaunch l) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int64_t* fen = reinterpret_cast<int64_t*>(smem_raw);  // rmax+1, 1-based
  const int b = q.qblk[blockIdx.x];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int64_t k0 = x.blk_key0[b];
  const int R = (int)(x.blk_key0[b + 1] - k0);
  const int64_t qa = q.qblk_q0[blockIdx.x], qe = q.qblk_q0[blockIdx.x + 1];
  for (int i = tid; i <= R; i += nt) fen[i] = 0;
  __syncthreads();
  const int U = x.U;
  const int s0 = x.blk_step0[b];
  // without the index write, nothing after the last queried step matters
  const int s1 = l.write_index ? x.blk_step0[b + 1] : (qa < qe ? q.qstep[qe - 1] + 1 : s0);
  int64_t qp = qa;
  int next_q = (qp < qe) ? q.qstep[qp] : -1;
  // one event per thread (nt >= U): the next step's (rank, signed delta) is loaded before the
  // current one enters the Fenwick tree; otherwise a strided loop
  const bool one = nt >= U;
  auto load_event = [&](int s, int i, int& r, int64_t& d) {
    const int st = x.steps[s];
    const int64_t v = st > 0 ? st : -st;
    r = x.evrank[(int64_t)s * U + i];
    const int64_t dv = x.delta[(v - 1) * (int64_t)U + i];
    d = st > 0 ? dv : -dv;
  };
  int r_cur = 0;
  int64_t d_cur = 0;
  if (one && tid < U && s0 < s1) load_event(s0, tid, r_cur, d_cur);
  for (int s = s0; s < s1; ++s) {
    if (one) {
      int r_nx = 0;
      int64_t d_nx = 0;
      if (tid < U && s + 1 < s1) load_event(s + 1, tid, r_nx, d_nx);
      if (tid < U)
        for (int f = r_cur + 1; f <= R; f += f & -f) lds_add64(&fen[f], d_cur);
      r_cur = r_nx;
      d_cur = d_nx;
    } else {
      for (int i = tid; i < U; i += nt) {
        int r;
        int64_t d;
        load_event(s, i, r, d);
        for (int f = r + 1; f <= R; f += f & -f) lds_add64(&fen[f], d);
      }
    }
    if (s == next_q) {  // queries of the state entered at step s (uniform branch)
      int64_t e = qp;
      while (e < qe && q.qstep[e] == s) ++e;
      __syncthreads();
      // one endpoint per lane: even lane = lo, odd lane = hi of the same query (nt is a
      // multiple of 64 and 2*qp is even, so a pair never straddles a wavefront)
      for (int64_t j = 2 * qp + tid; j < ((2 * e + 63) & ~63ll); j += nt) {
        const int64_t ent = j >> 1;
        const bool live = j < 2 * e;
        const uint32_t rk = live ? q.qrank[ent] : 0u;
        const int y = (j & 1) ? (int)(rk >> 16) : (int)(rk & 0xFFFFu);
        const int64_t pv = live ? fen_prefix(fen, y) : 0;
        const int64_t other = __shfl_xor(pv, 1);
        if (live && !(j & 1)) q.ansr[q.qorder[ent]] = other - pv;
      }
      __syncthreads();
      qp = e;
      next_q = (qp < qe) ? q.qstep[qp] : -1;
    }
  }
  if (!l.write_index) return;
  __syncthreads();
  // net weight of the block per distinct key (carry level B) and per key tile (level C)
  for (int r = tid; r < R; r += nt) x.agg[k0 + r] = fen_prefix(fen, r + 1) - fen_prefix(fen, r);
  const int64_t t0 = x.blk_tile0[b];
  const int T = (int)(x.blk_tile0[b + 1] - t0);
  for (int s = tid; s < T; s += nt) {
    const int r0 = x.tile_rank0[t0 + s];
    const int r1 = (s + 1 < T) ? (int)x.tile_rank0[t0 + s + 1] : R;
    x.tagg[t0 + s] = fen_prefix(fen, r1) - fen_prefix(fen, r0);
  }
}

// ---------------------------------------------------------------- chunk histogram
// One wavefront per chunk: dense class histogram of the chunk's key items, written as its
// exclusive prefix over classes (cp row = sum of weights with class < y).
__global__ void __launch_bounds__(256) bwt_chunk_hist_kernel(const BwtIndex x, const BwtLaunch l, int64_t c_begin,
                                                             int64_t c_end, int stride) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t c = c_begin + (int64_t)blockIdx.x * l.chunk_waves + wv;
  if (c >= c_end) return;  // whole wavefront leaves; no workgroup barrier follows
  int64_t* h = reinterpret_cast<int64_t*>(smem_raw) + (int64_t)wv * stride;
  const int chain = x.chunk_chain[c];
  const int ncls = x.chain_ncls[chain];
  const int64_t* wsrc = x.chain_src[chain] ? x.tagg : x.agg;
  for (int i = lane; i < ncls; i += 64) h[i] = 0;
  wave_lds_sync();
  const int64_t i1 = x.chunk_key0[c + 1];
  int64_t it = x.chunk_key0[c] + lane;
  uint64_t item = (it < i1) ? x.kitems[it] : 0;
  while (it < i1) {
    const int64_t nx = it + 64;
    const uint64_t nitem = (nx < i1) ? x.kitems[nx] : 0;  // prefetch
    lds_add64(&h[(int)((item >> BWT_CLASS_SHIFT) & 0xFFF)], wsrc[item & BWT_REF_MASK]);
    it = nx;
    item = nitem;
  }
  wave_lds_sync();
  int64_t* row = x.cp + x.chunk_cp0[c];
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
__global__ void __launch_bounds__(256) bwt_seg_sum_kernel(const BwtIndex x) {
  const int64_t s = blockIdx.x;
  const int y = blockIdx.y * blockDim.x + threadIdx.x;
  const int64_t c0 = x.seg_chunk0[s], c1 = x.seg_chunk0[s + 1];
  const int ncls = x.chain_ncls[x.chunk_chain[c0]];
  if (y >= ncls) return;
  int64_t acc = 0;
  for (int64_t c = c0; c < c1; ++c) acc += x.cp[x.chunk_cp0[c] + y];
  x.segsum[x.seg_cp0[s] + y] = acc;
}

__global__ void __launch_bounds__(256) bwt_seg_scan_kernel(const BwtIndex x) {
  const int chain = blockIdx.x;
  const int y = blockIdx.y * blockDim.x + threadIdx.x;
  if (y >= x.chain_ncls[chain]) return;
  int64_t run = 0;
  const int64_t e = x.chain_seg0[chain + 1];
  for (int64_t s = x.chain_seg0[chain]; s < e; ++s) {
    int64_t* p = x.segsum + x.seg_cp0[s] + y;
    const int64_t t = *p;
    *p = run;
    run += t;
  }
}

__global__ void __launch_bounds__(256) bwt_seg_apply_kernel(const BwtIndex x) {
  const int64_t s = blockIdx.x;
  const int y = blockIdx.y * blockDim.x + threadIdx.x;
  const int64_t c0 = x.seg_chunk0[s], c1 = x.seg_chunk0[s + 1];
  const int ncls = x.chain_ncls[x.chunk_chain[c0]];
  if (y >= ncls) return;
  int64_t run = x.segsum[x.seg_cp0[s] + y];
  for (int64_t c = c0; c < c1; ++c) {
    int64_t* p = x.cp + x.chunk_cp0[c] + y;
    const int64_t t = *p;
    *p = run;
    run += t;
  }
}

// ---------------------------------------------------------------- chunk replay
// One wavefront per listed chunk.  Key items are consumed in windows of 64 (one per lane); the
// endpoints whose position falls before the end of the window are evaluated first:
//   value = carry[cls] + Fenwick prefix(cls) over earlier windows
//         + sum of this window's keys that precede the endpoint's position with class < cls
// (keys broadcast lane by lane with readlane); then the window's keys enter the Fenwick tree.
__global__ void __launch_bounds__(256) bwt_chunk_replay_kernel(const BwtIndex x, const BwtBatch q, const BwtLaunch l,
                                                               int64_t w_begin, int64_t w_end, int stride) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t ci = w_begin + (int64_t)blockIdx.x * l.chunk_waves + wv;
  if (ci >= w_end) return;
  const int64_t c = q.wk_chunk[ci];
  int64_t* fen = reinterpret_cast<int64_t*>(smem_raw) + (int64_t)wv * stride;
  const int chain = x.chunk_chain[c];
  const int ncls = x.chain_ncls[chain];
  const int64_t* wsrc = x.chain_src[chain] ? x.tagg : x.agg;
  const int64_t* carry = x.cp + x.chunk_cp0[c];
  int64_t* gout = q.g + (int64_t)x.chain_out[chain] * (2 * q.M);
  for (int i = lane; i <= ncls; i += 64) fen[i] = 0;
  wave_lds_sync();
  int64_t ep = q.wk_e0[ci];
  const int64_t e1 = q.wk_e0[ci + 1];
  // keys after this work item's last endpoint (items are sorted by position) are not needed
  const int64_t k0 = x.chunk_key0[c];
  const int64_t k1 = min(x.chunk_key0[c + 1], k0 + (int64_t)(q.eitems[e1 - 1] >> BWT_EP_POS_SHIFT));
  // prefetched key window
  int64_t kit = k0 + lane;
  uint64_t kx = (kit < k1) ? x.kitems[kit] : 0;
  int64_t kw = (kit < k1) ? wsrc[kx & BWT_REF_MASK] : 0;
  for (int64_t w0 = k0;; w0 += 64) {
    const int64_t wend = (w0 + 64 < k1) ? w0 + 64 : k1;
    const bool last = wend == k1;
    const bool iskey = w0 + lane < k1;
    const int kcls = (int)((kx >> BWT_CLASS_SHIFT) & 0xFFF);
    const int64_t kwt = kw;
    // prefetch the next window while this one is processed
    const int64_t nkit = w0 + 64 + lane;
    const uint64_t nkx = (nkit < k1) ? x.kitems[nkit] : 0;
    const int64_t nkw = (nkit < k1) ? wsrc[nkx & BWT_REF_MASK] : 0;
    const uint64_t keymask = ballot64(iskey);
    // endpoints positioned inside this window (or anywhere, at the chunk's last window)
    const int64_t pend = wend - k0;
    while (ep < e1) {
      const int64_t e = ep + lane;
      const uint64_t it = (e < e1) ? q.eitems[e] : 0;
      const int64_t pos = (int64_t)(it >> BWT_EP_POS_SHIFT);
      const bool take = (e < e1) && (last || pos < pend);
      const uint64_t tm = ballot64(take);
      if (!tm) break;
      const int n = __popcll(tm);  // endpoints are sorted by position: a prefix of the lanes
      // every lane runs the loop (readlane stays in uniform control flow); lanes beyond the
      // batch hold item 0 (class 0), so every address below stays inside the chunk's rows
      const int cls = (int)((it >> BWT_EP_CLASS_SHIFT) & 0xFFF);
      int64_t val = carry[cls] + fen_prefix(fen, cls);
      const int64_t rel = pos - (w0 - k0);  // keys of this window before the endpoint
      uint64_t km = keymask;
      while (km) {
        const int j = __builtin_ctzll(km);
        km &= km - 1;
        const int jc = __builtin_amdgcn_readlane(kcls, j);
        const int64_t jw = readlane64(kwt, j);
        if (j < rel && jc < cls) val += jw;
      }
      if (take) gout[it & 0xFFFFFFFFull] = val;
      ep += n;
      if (n < 64) break;
    }
    if (iskey)
      for (int f = kcls + 1; f <= ncls; f += f & -f) lds_add64(&fen[f], kwt);
    wave_lds_sync();
    if (last) break;
    kx = nkx;
    kw = nkw;
  }
}

// ---------------------------------------------------------------- combine
__global__ void __launch_bounds__(256) bwt_combine_kernel(const BwtBatch q) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= q.M) return;
  const int64_t* gc = q.g;
  const int64_t* gb = q.g + 2 * q.M;
  q.answers[i] = q.ansr[i] + (gc[2 * i + 1] + gb[2 * i + 1]) - (gc[2 * i] + gb[2 * i]);
}

// ---------------------------------------------------------------- rank remap
__global__ void __launch_bounds__(256) bwt_rank_kernel(const BwtRankArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t v = a.in[i];
  const uint32_t h = v >> 16;
  const uint16_t lo16 = (uint16_t)(v & 0xFFFFu);
  uint32_t lo = a.row_ptr[h], hi = a.row_ptr[h + 1];
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a.low16[mid] < lo16) lo = mid + 1; else hi = mid;
  }
  a.out[i] = lo;
}

// ---------------------------------------------------------------- launch helpers
size_t block_lds_bytes(const BwtIndex* x) { return (size_t)(x->rmax + 1) * sizeof(int64_t); }
size_t chunk_lds_bytes(const BwtIndex* x, const BwtLaunch* l) {
  return (size_t)l->chunk_waves * (size_t)(x->max_cls + 1) * sizeof(int64_t);
}
// LDS stride (entries per wavefront) of a level: its classes + 1, rounded to an odd number of
// 8-byte words so consecutive wavefronts' trees start in different banks
int level_stride(int ncls) { return (ncls + 1) | 1; }

int check(const BwtIndex* x, const BwtLaunch* l) {
  if (!x || !l || x->S < 1 || x->U < 1 || x->nblk < 1) return (int)hipErrorInvalidValue;
  if (l->block_threads < 64 || l->block_threads > 1024 || (l->block_threads & 63)) return (int)hipErrorInvalidValue;
  if (x->rmax < 1 || x->rmax > 8192 || x->max_cls < 1 || x->max_cls > 4096) return (int)hipErrorInvalidValue;
  if (x->ncls_c < 1 || x->ncls_c > x->max_cls || x->ncls_b < 1 || x->ncls_b > x->max_cls) return (int)hipErrorInvalidValue;
  if (x->nchunk_c < 0 || x->nchunk_c > x->nchunk) return (int)hipErrorInvalidValue;
  if (l->chunk_waves < 1 || l->chunk_waves > 4) return (int)hipErrorInvalidValue;
  if (block_lds_bytes(x) > 160 * 1024 || chunk_lds_bytes(x, l) > 160 * 1024) return (int)hipErrorInvalidValue;
  return 0;
}

int set_lds_limits() {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
    const int lim = 160 * 1024;
    const void* ks[] = {reinterpret_cast<const void*>(bwt_block_kernel),
                        reinterpret_cast<const void*>(bwt_chunk_hist_kernel),
                        reinterpret_cast<const void*>(bwt_chunk_replay_kernel)};
    for (const void* k : ks) {
      const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
      if (e != hipSuccess && err == hipSuccess) err = e;
    }
  });
  return (int)err;
}

int report(const char* what, hipError_t e) {
  if (e != hipSuccess) fprintf(stderr, "bwgpu: %s: %s\n", what, hipGetErrorString(e));
  return (int)e;
}

}  // namespace

extern "C" int bwt_block_replay_launch(const BwtIndex* x, const BwtBatch* b, const BwtLaunch* l, void* stream) {
  int rc = check(x, l);
  if (rc) return rc;
  if (!b || b->nqblk < 0 || (l->write_index && b->nqblk != x->nblk)) return (int)hipErrorInvalidValue;
  if ((rc = set_lds_limits())) return rc;
  if (b->nqblk == 0) return 0;
  hipLaunchKernelGGL(bwt_block_kernel, dim3(b->nqblk), dim3(l->block_threads), block_lds_bytes(x),
                     (hipStream_t)stream, *x, *b, *l);
  return report("block replay", hipGetLastError());
}

extern "C" int bwt_chunk_hist_launch(const BwtIndex* x, const BwtLaunch* l, void* stream) {
  int rc = check(x, l);
  if (rc) return rc;
  if ((rc = set_lds_limits())) return rc;
  // level C chunks, then level B chunks, each with LDS sized to its own class count
  const int64_t rng[2][2] = {{0, x->nchunk_c}, {x->nchunk_c, x->nchunk}};
  const int ncls[2] = {x->ncls_c, x->ncls_b};
  for (int lv = 0; lv < 2; ++lv) {
    const int64_t n = rng[lv][1] - rng[lv][0];
    if (n <= 0) continue;
    const int stride = level_stride(ncls[lv]);
    const int64_t grid = (n + l->chunk_waves - 1) / l->chunk_waves;
    hipLaunchKernelGGL(bwt_chunk_hist_kernel, dim3((unsigned)grid), dim3(64 * l->chunk_waves),
                       (size_t)l->chunk_waves * stride * sizeof(int64_t), (hipStream_t)stream, *x, *l, rng[lv][0],
                       rng[lv][1], stride);
    if ((rc = report("chunk histogram", hipGetLastError()))) return rc;
  }
  return 0;
}

extern "C" int bwt_carry_scan_launch(const BwtIndex* x, const BwtLaunch* l, void* stream) {
  int rc = check(x, l);
  if (rc) return rc;
  if (x->nseg == 0) return 0;
  const unsigned ytiles = (unsigned)((x->max_cls + 255) / 256);
  hipLaunchKernelGGL(bwt_seg_sum_kernel, dim3((unsigned)x->nseg, ytiles), dim3(256), 0, (hipStream_t)stream, *x);
  if ((rc = report("carry scan A", hipGetLastError()))) return rc;
  hipLaunchKernelGGL(bwt_seg_scan_kernel, dim3((unsigned)x->nchain, ytiles), dim3(256), 0, (hipStream_t)stream, *x);
  if ((rc = report("carry scan B", hipGetLastError()))) return rc;
  hipLaunchKernelGGL(bwt_seg_apply_kernel, dim3((unsigned)x->nseg, ytiles), dim3(256), 0, (hipStream_t)stream, *x);
  return report("carry scan C", hipGetLastError());
}

extern "C" int bwt_chunk_replay_launch(const BwtIndex* x, const BwtBatch* b, const BwtLaunch* l, void* stream) {
  int rc = check(x, l);
  if (rc) return rc;
  if (!b || b->nwork < 0) return (int)hipErrorInvalidValue;
  if ((rc = set_lds_limits())) return rc;
  if (b->nwork_c < 0 || b->nwork_c > b->nwork) return (int)hipErrorInvalidValue;
  const int64_t rng[2][2] = {{0, b->nwork_c}, {b->nwork_c, b->nwork}};
  const int ncls[2] = {x->ncls_c, x->ncls_b};
  for (int lv = 0; lv < 2; ++lv) {
    const int64_t n = rng[lv][1] - rng[lv][0];
    if (n <= 0) continue;
    const int stride = level_stride(ncls[lv]);
    const int64_t grid = (n + l->chunk_waves - 1) / l->chunk_waves;
    hipLaunchKernelGGL(bwt_chunk_replay_kernel, dim3((unsigned)grid), dim3(64 * l->chunk_waves),
                       (size_t)l->chunk_waves * stride * sizeof(int64_t), (hipStream_t)stream, *x, *b, *l, rng[lv][0],
                       rng[lv][1], stride);
    if ((rc = report("chunk replay", hipGetLastError()))) return rc;
  }
  return 0;
}

extern "C" int bwt_combine_launch(const BwtIndex* x, const BwtBatch* b, const BwtLaunch* l, void* stream) {
  (void)x;
  (void)l;
  if (!b || b->M < 0) return (int)hipErrorInvalidValue;
  if (b->M == 0) return 0;
  hipLaunchKernelGGL(bwt_combine_kernel, dim3((unsigned)((b->M + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *b);
  return report("combine", hipGetLastError());
}

extern "C" int bwt_query_launch(const BwtIndex* x, const BwtBatch* b, const BwtLaunch* l, void* stream) {
  BwtLaunch q = *l;
  q.write_index = 0;
  int rc;
  if ((rc = bwt_block_replay_launch(x, b, &q, stream))) return rc;
  if ((rc = bwt_chunk_replay_launch(x, b, &q, stream))) return rc;
  return bwt_combine_launch(x, b, &q, stream);
}

extern "C" int bwt_pass_launch(const BwtIndex* x, const BwtBatch* b, const BwtLaunch* l, void* stream) {
  BwtLaunch w = *l;
  w.write_index = 1;
  int rc;
  if ((rc = bwt_block_replay_launch(x, b, &w, stream))) return rc;
  if ((rc = bwt_chunk_hist_launch(x, &w, stream))) return rc;
  if ((rc = bwt_carry_scan_launch(x, &w, stream))) return rc;
  if ((rc = bwt_chunk_replay_launch(x, b, &w, stream))) return rc;
  return bwt_combine_launch(x, b, &w, stream);
}

extern "C" int bwt_rank_remap_launch(const BwtRankArgs* a, void* stream) {
  if (!a || a->n < 0) return (int)hipErrorInvalidValue;
  if (a->n == 0) return 0;
  hipLaunchKernelGGL(bwt_rank_kernel, dim3((unsigned)((a->n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *a);
  return report("rank remap", hipGetLastError());
}
