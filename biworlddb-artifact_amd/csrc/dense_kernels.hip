// dense_kernels.hip -- the dense-carry batched form (include/bwgpu_dense.h) for gfx950.
//
// Stage map onto the reference's v7 algorithm (src/batch_engine_v7.hpp:832-968):
//   carry build      <- the carry pass (src/panel_batch_gpu.cpp:191-204): carrypref[b][k] =
//                       sum over blocks before b of their net weight of classes < k.  Here one
//                       dense row per block over every coordinate, built once: the blocks' net
//                       weights (agg) become per-block prefix functions D_b(x); rows are their
//                       running sum over blocks, computed per (coordinate range, block segment)
//                       with a scan over segments in between, so ~45k workgroups share the work.
//   plan             <- qlb/qhb (src/batch_engine_v7.hpp:666-698): queries grouped by their state
//                       id (a two-level bucket sort: high digit, then low digit inside each
//                       bucket), group offsets per preorder rank.
//   persistent block <- v7_block_kernel + gpu_combine_kernel (src/gpu_v7.cu:10-96): workgroups
//                       take blocks from a device queue; each replays its block's steps (events
//                       are ds_add_u64 into a dense per-rank array; at every enter step a DPP
//                       block scan publishes the in-block prefix), and answers the state's
//                       queries: in-block prefix at the endpoints' ranks (binary search in the
//                       block's keys, staged in LDS) plus the block's carry row at the endpoint
//                       coordinates (two global reads issued before the scan's barriers).
// All arithmetic is int64 adds: results are exact and independent of scheduling order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>
#include <stdint.h>
#include <stdio.h>

#include "bwgpu_dense.h"
#include "wave64.hpp"

using namespace bw;

// ---------------------------------------------------------------- the allocation-free entry's scratch
// bwd_answer_noalloc_launch plans and answers a batch in this array instead of a caller's
// workspace: every GPU the library runs on holds one copy, at its own address, until the process
// ends.  Host code must not pass `scratch` itself to a kernel (on the host the name is only a
// stand-in); its device address comes from hipGetSymbolAddress, once per device.
static constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_at[kMaxDevices] = {};

// ---------------------------------------------------------------- planned query layout
// In bucket order: key = state << 44 | query id << 17 | lo & 0x1FFFF;  val = hi | (lo >> 17) << 21
// (the state in 19 bits, the id in 27, the endpoints in 21 each).  Grouped by state: ends = lo |
// hi << 32 and, in a separate array, the query id -- the staged replay reads only the ends, the
// answers' partition only the ids.
constexpr int kStateShift = 44;
constexpr int kIdShift = 17;
constexpr uint64_t kIdMask = (uint64_t(1) << 27) - 1;
constexpr uint32_t kLoLowMask = (1u << 17) - 1;
constexpr uint32_t kHiMask = (1u << 21) - 1;

__device__ __forceinline__ uint64_t ends_of(uint64_t key, uint32_t val) {
  const uint32_t lo = ((uint32_t)key & kLoLowMask) | ((val >> 21) << 17);
  return (uint64_t)lo | ((uint64_t)(val & kHiMask) << 32);
}

// ---------------------------------------------------------------- carry build
constexpr int kCT = 256;          // threads of the carry-build workgroups
constexpr int kXPT = 8;           // coordinates per thread
constexpr int kW = kCT * kXPT;    // coordinates per range
constexpr int kChunk = 1024;      // block keys staged in LDS at a time
constexpr int kSegBlocks = 64;    // blocks per segment

// aggp[blk_key0[b] + b + r] = sum of agg of block b's r smallest keys, r in [0, R]
__global__ void __launch_bounds__(kCT) bwd_aggp_kernel(const BwdIndex x, const int64_t* agg, int64_t* aggp) {
  __shared__ int64_t a[BWD_MAX_RMAX + 1];
  __shared__ int64_t tmp[kCT / 64 + 1];
  const int b = blockIdx.x;
  const int64_t k0 = x.blk_key0[b];
  const int R = (int)(x.blk_key0[b + 1] - k0);
  for (int i = threadIdx.x; i <= R; i += kCT) a[i] = i < R ? agg[k0 + i] : 0;
  __syncthreads();
  block_excl_scan_lds(a, R + 1, tmp);
  for (int i = threadIdx.x; i <= R; i += kCT) aggp[k0 + b + i] = a[i];
}

// rb[b*(nr+1) + j] = number of block b's keys < j*kW
__global__ void __launch_bounds__(256) bwd_rb_kernel(const BwdIndex x, int nr, uint16_t* rb) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t n = (int64_t)x.nblk * (nr + 1);
  if (i >= n) return;
  const int b = (int)(i / (nr + 1)), j = (int)(i - (int64_t)b * (nr + 1));
  const int64_t k0 = x.blk_key0[b];
  const int R = (int)(x.blk_key0[b + 1] - k0);
  const uint64_t xv = (uint64_t)j * kW;
  const uint32_t t = xv > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)xv;
  int lo = 0, hi = R;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (x.bkey[k0 + mid] < t) lo = mid + 1; else hi = mid;
  }
  rb[i] = (uint16_t)lo;
}

// For the blocks of segment g and the coordinates of range j: with kRows, write every block's
// carry row (starting from the segment's carry-in T[g]) and advance; otherwise sum the blocks'
// prefix functions into T[g].  D_b(x) = aggp_b[rank_b(x)]: the block's keys of the range are
// staged kChunk at a time with their prefix values; a coordinate's rank is settled in the
// chunk where it falls below a key (or in the range's last chunk).
template <bool kRows>
__global__ void __launch_bounds__(kCT) bwd_carry_range_kernel(const BwdIndex x, const int64_t* aggp,
                                                              const uint16_t* rb, int nr, int64_t* T, int64_t ldT) {
  __shared__ uint32_t skey[kChunk];
  __shared__ int64_t sap[kChunk + 1];
  __shared__ int64_t srow[kW];  // the range's values, stored lane-contiguous from here
  const int j = blockIdx.x, g = blockIdx.y;
  const int64_t NKP = (int64_t)x.NK + 1;
  const int64_t xb = (int64_t)j * kW + (int64_t)threadIdx.x * kXPT;  // this thread's first coordinate
  const int b0 = g * kSegBlocks, b1 = min(x.nblk, b0 + kSegBlocks);
  int64_t acc[kXPT];
#pragma unroll
  for (int i = 0; i < kXPT; ++i) acc[i] = (kRows && xb + i < NKP) ? T[(int64_t)g * ldT + xb + i] : 0;
  // a thread's values cover kXPT consecutive coordinates; they go through LDS so that each store
  // instruction writes one contiguous run (a direct store would leave every 64-byte piece of the
  // row partly written per instruction)
  auto store_range = [&](int64_t* dst) {
    __syncthreads();  // the previous range's readers are done
#pragma unroll
    for (int i = 0; i < kXPT; ++i) srow[threadIdx.x * kXPT + i] = acc[i];
    __syncthreads();
    const int64_t x0 = (int64_t)j * kW;
#pragma unroll
    for (int i = 0; i < kXPT; ++i) {
      const int k = i * kCT + threadIdx.x;
      if (x0 + k < NKP) dst[x0 + k] = srow[k];
    }
  };
  for (int b = b0; b < b1; ++b) {
    if (kRows) store_range(x.carry + (int64_t)b * x.carry_ld);
    const int64_t k0 = x.blk_key0[b];
    const int64_t pbase = k0 + b;  // aggp offset of block b
    const int r0 = rb[(int64_t)b * (nr + 1) + j], r1 = rb[(int64_t)b * (nr + 1) + j + 1];
    int64_t d[kXPT];
    bool done[kXPT];
#pragma unroll
    for (int i = 0; i < kXPT; ++i) done[i] = false;
    for (int c0 = r0;; c0 += kChunk) {
      const int cm = min(kChunk, r1 - c0);  // keys in this chunk (may be 0 when the range has none)
      const bool last = c0 + kChunk >= r1;
      __syncthreads();  // the previous chunk's readers are done
      for (int i = threadIdx.x; i < cm; i += kCT) skey[i] = x.bkey[k0 + c0 + i];
      for (int i = threadIdx.x; i <= cm; i += kCT) sap[i] = aggp[pbase + c0 + i];
      __syncthreads();
      // local rank of the first coordinate, then walk the thread's consecutive coordinates
      const uint32_t xf = (uint32_t)min<int64_t>(xb, 0xFFFFFFFFll);
      int r = lower_bound_lds<uint32_t>(skey, cm, xf);
#pragma unroll
      for (int i = 0; i < kXPT; ++i) {
        const uint32_t xv = (uint32_t)min<int64_t>(xb + i, 0xFFFFFFFFll);
        while (r < cm && skey[r] < xv) ++r;
        if (!done[i] && (r < cm || last)) {
          d[i] = sap[r];
          done[i] = true;
        }
      }
      if (last) break;
    }
#pragma unroll
    for (int i = 0; i < kXPT; ++i) acc[i] += d[i];
  }
  if (!kRows) store_range(T + (int64_t)g * ldT);
}

// exclusive scan of T over segments, per coordinate
__global__ void __launch_bounds__(256) bwd_seg_scan_kernel(int64_t NKP, int nseg, int64_t* T, int64_t ldT) {
  const int64_t xv = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (xv >= NKP) return;
  int64_t run = 0;
  for (int g = 0; g < nseg; ++g) {
    int64_t* p = T + (int64_t)g * ldT + xv;
    const int64_t t = *p;
    *p = run;
    run += t;
  }
}

// ---------------------------------------------------------------- plan: queries grouped by state
// The replay needs the queries of each state together, in any order inside a state (answers are
// stored by sorted position and then routed to query ids, or stored by id), and per preorder rank
// of a state where its group lies.  The queries are grouped by state id (no lookup per query: a
// random 4-byte lookup costs one L2 request, ~0.4 ms per 10^8), and one small pass per block maps
// ranks to groups.  A two-level bucket sort -- the high digit v >> kLoBits of state v, then the low
// digit v & (kLoBins - 1) inside each high bucket -- so no pass needs a stable order and runs are
// placed by counts and cursors:
//   count    per tile of queries: LDS counts of the high digit, added to the global counts;
//   layout   one workgroup: bucket starts (also the buckets' cursors), the second level's tiles
//            per bucket (tile j of bucket h covers positions [hbase[h] + j*kPTile, ...) of it),
//            grp[S + 1] = M;
//   bucket   per tile of queries: the key packed (key = state << 44 | id << 17 | lo & 0x1FFFF,
//            val = hi | (lo >> 17) << 21), grouped by high digit in LDS, each bucket's run appended
//            at the bucket's cursor (16-byte loads when the query arrays allow);
//   tile     per second-level tile: LDS counts of the low digit;
//   offsets  per bucket: each tile's start inside every state of the bucket, and grp[v], the first
//            sorted position of state v;
//   state    per second-level tile: grouped by low digit in LDS, each state's run stored at grp[v]
//            plus the tile's start;
//   ranks    per block: the group [first, end) of the state of each of its ranks, the block's
//            last rank with queries (-1: none), which ends its replay, and its query count;
//   order    the blocks by descending query count (log-scale bins), the order in which the
//            persistent replay's workgroups take them: the heaviest first, so the replay ends on
//            light blocks and the compute units finish together.
// A run goes through LDS so that each store instruction writes consecutive positions.  Counts and
// placements come from the same rank of the same queries, so every position in [0, M) is written
// once; a position outside it (which the counts rule out) would set flags[0] to 2 and not be
// written.  An out-of-range query (state outside [1, S], an endpoint above NK) sets flags[0] to 1
// and is planned over [0, 0) (answer 0), as a query of state 1 when its state is out of range.
constexpr int kLoBits = 9, kLoBins = 1 << kLoBits;       // 512 states per bucket
constexpr int kMaxHi = (BWD_MAX_STATES >> kLoBits) + 1;  // 1,024 buckets
constexpr int kPT = 1024, kPE = 16, kPTile = kPT * kPE;  // 16,384 queries per tile, both levels
constexpr unsigned kCountGrid = 1024;                    // the count pass's grid bound (tiles a grid apart)
constexpr int kHiShift = kStateShift + kLoBits;           // key >> kHiShift = the high digit
static_assert(kMaxHi <= kPT && kLoBins <= kPT, "one thread per bucket or digit in the scans");

// Exclusive scan of one value per thread over the workgroup (every thread calls it); wtot is LDS
// scratch of blockDim.x / 64 entries.  *total receives the sum.
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t* wtot, uint32_t* total) {
  const int tid = threadIdx.x, wv = tid >> 6, nw = blockDim.x >> 6;
  const uint32_t inc = (uint32_t)wave_incl_scan((int64_t)v);
  if ((tid & 63) == 63) wtot[wv] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (int k = 0; k < nw; ++k) {
    const uint32_t t = wtot[k];
    off += k < wv ? t : 0u;
    tot += t;
  }
  __syncthreads();  // wtot is free for the next scan
  *total = tot;
  return off + inc - v;
}

__device__ __forceinline__ uint32_t plan_state(int64_t S, int32_t v) { return (v >= 1 && v <= S) ? (uint32_t)v : 1u; }

// The replay's block order: kCostBins bins of query counts, 4 per power of two, bin 0 the heaviest
// (a block without queries goes to the last bin).
constexpr int kCostBins = 128;
__device__ __forceinline__ int cost_bin(uint32_t q) {
  if (q == 0) return kCostBins - 1;
  const int lg = min(31 - __clz(q), 27);  // floor(log2 q); a batch holds at most 2^27 queries
  const int sub = lg >= 2 ? (int)((q >> (lg - 2)) & 3u) : 0;
  return kCostBins - 2 - (lg * 4 + sub);
}

__global__ void __launch_bounds__(1024) bwd_plan_init_kernel(int64_t* flags, uint32_t* hcount, int nh) {
  if (threadIdx.x == 0) flags[0] = 0;
  for (int i = threadIdx.x; i < nh; i += 1024) hcount[i] = 0;
}

__global__ void __launch_bounds__(kPT) bwd_count_kernel(int64_t S, int64_t M, const int32_t* __restrict__ qstate, int nh,
                                                        uint32_t* __restrict__ hcount) {
  __shared__ uint32_t c[kMaxHi];
  const int tid = threadIdx.x;
  for (int i = tid; i < nh; i += kPT) c[i] = 0;
  __syncthreads();  // counts are zero
  // tiles a grid apart into one set of LDS counts, flushed once: a workgroup per tile would add
  // every bucket's count to global memory once per tile, thousands of atomics on each address
  for (int64_t t0 = (int64_t)blockIdx.x * kPTile; t0 < M; t0 += (int64_t)gridDim.x * kPTile) {
    int32_t v[kPE];
#pragma unroll
    for (int k = 0; k < kPE; ++k) {
      const int64_t i = t0 + k * kPT + tid;
      v[k] = i < M ? qstate[i] : 1;
    }
#pragma unroll
    for (int k = 0; k < kPE; ++k)
      if (t0 + k * kPT + tid < M) atomicAdd(&c[plan_state(S, v[k]) >> kLoBits], 1u);
  }
  __syncthreads();
  for (int h = tid; h < nh; h += kPT)
    if (c[h]) atomicAdd(hcount + h, c[h]);
}

// hbase[h] (and the cursor hcur[h]) = first position of bucket h, hbase[nh] = M; tstart[h] = first
// second-level tile of bucket h, tstart[nh] = their number; tbkt[g] = the bucket of tile g;
// grp[S + 1] = M.
__global__ void __launch_bounds__(kPT) bwd_layout_kernel(int64_t S, int64_t M, int nh, const uint32_t* __restrict__ hcount,
                                                         uint32_t* __restrict__ hbase, uint32_t* __restrict__ hcur,
                                                         uint32_t* __restrict__ tstart, uint32_t* __restrict__ grp,
                                                         uint32_t* __restrict__ tbkt) {
  __shared__ uint32_t wtot[kPT / 64];
  const int h = threadIdx.x;
  const uint32_t c = h < nh ? hcount[h] : 0u;
  const uint32_t t = (c + kPTile - 1) / kPTile;
  uint32_t ctot, ttot;
  const uint32_t cb = block_excl_scan_u32(c, wtot, &ctot);
  const uint32_t tb = block_excl_scan_u32(t, wtot, &ttot);
  if (h < nh) {
    hbase[h] = cb;
    hcur[h] = cb;
    tstart[h] = tb;
    for (uint32_t k = 0; k < t; ++k) tbkt[tb + k] = (uint32_t)h;  // the bucket of each of its tiles
  }
  if (h == 0) {
    hbase[nh] = (uint32_t)M;
    tstart[nh] = ttot;
    grp[S + 1] = (uint32_t)M;
  }
}

// The bucket pass's tiles: kBE queries per thread, staged keys and values both in LDS (156 KiB)
constexpr int kBE = 12, kBTile = kPT * kBE;

template <bool kVec>
__global__ void __launch_bounds__(kPT) bwd_bucket_kernel(const BwdIndex x, int64_t M, const int32_t* __restrict__ qstate,
                                                         const uint32_t* __restrict__ qlo, const uint32_t* __restrict__ qhi,
                                                         int nh, uint32_t* __restrict__ hcur, uint64_t* __restrict__ keys,
                                                         uint32_t* __restrict__ vals, int64_t* flags) {
  __shared__ uint32_t cnt[kMaxHi], lstart[kMaxHi], gbase[kMaxHi];
  __shared__ uint32_t wtot[kPT / 64];
  __shared__ uint64_t skey[kBTile];
  __shared__ uint32_t sval[kBTile];
  const int tid = threadIdx.x;
  const int64_t t0 = (int64_t)blockIdx.x * kBTile;
  const int n = (int)min<int64_t>(kBTile, M - t0);
  for (int i = tid; i < nh; i += kPT) cnt[i] = 0;
  // the thread's k-th query: with kVec four consecutive ones per 16-byte load, in rounds a
  // workgroup of quadruples apart; otherwise one a workgroup apart
  auto qid = [&](int k) -> int64_t {
    return kVec ? t0 + (int64_t)(k >> 2) * (kPT * 4) + 4 * tid + (k & 3) : t0 + (int64_t)k * kPT + tid;
  };
  int32_t v[kBE];
  uint32_t lo[kBE], hi[kBE];
  if (kVec) {
#pragma unroll
    for (int r = 0; r < kBE / 4; ++r) {
      const int64_t b = qid(4 * r);
      if (b + 4 <= M) {
        const int4 a = *reinterpret_cast<const int4*>(qstate + b);
        const uint4 l = *reinterpret_cast<const uint4*>(qlo + b);
        const uint4 h = *reinterpret_cast<const uint4*>(qhi + b);
        v[4 * r] = a.x; v[4 * r + 1] = a.y; v[4 * r + 2] = a.z; v[4 * r + 3] = a.w;
        lo[4 * r] = l.x; lo[4 * r + 1] = l.y; lo[4 * r + 2] = l.z; lo[4 * r + 3] = l.w;
        hi[4 * r] = h.x; hi[4 * r + 1] = h.y; hi[4 * r + 2] = h.z; hi[4 * r + 3] = h.w;
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const bool in = b + c < M;
          v[4 * r + c] = in ? qstate[b + c] : 1;
          lo[4 * r + c] = in ? qlo[b + c] : 0;
          hi[4 * r + c] = in ? qhi[b + c] : 0;
        }
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < kBE; ++k) {
      const int64_t i = qid(k);
      const bool in = i < M;
      v[k] = in ? qstate[i] : 1;
      lo[k] = in ? qlo[i] : 0;
      hi[k] = in ? qhi[i] : 0;
    }
  }
  bool bad = false;
  uint64_t key[kBE];
  uint32_t val[kBE];
#pragma unroll
  for (int k = 0; k < kBE; ++k) {
    uint32_t l = lo[k], h = hi[k];
    if (qid(k) < M && (v[k] < 1 || v[k] > x.S || l > x.NK || h > x.NK)) {
      bad = true;
      l = h = 0;
    }
    key[k] = ((uint64_t)plan_state(x.S, v[k]) << kStateShift) | ((uint64_t)qid(k) << kIdShift) | (uint64_t)(l & kLoLowMask);
    val[k] = h | ((l >> 17) << 21);
  }
  __syncthreads();  // counts are zero
  uint32_t off[kBE];
#pragma unroll
  for (int k = 0; k < kBE; ++k) off[k] = qid(k) < M ? atomicAdd(&cnt[key[k] >> kHiShift], 1u) : 0u;
  __syncthreads();
  uint32_t tot;
  const uint32_t c = tid < nh ? cnt[tid] : 0u;
  const uint32_t ls = block_excl_scan_u32(c, wtot, &tot);
  if (tid < nh) {
    lstart[tid] = ls;
    if (c) gbase[tid] = atomicAdd(hcur + tid, c);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kBE; ++k) {
    if (qid(k) < M) {
      const uint32_t j = lstart[key[k] >> kHiShift] + off[k];
      skey[j] = key[k];
      sval[j] = val[k];
    }
  }
  __syncthreads();
  bool lost = false;
  for (int j = tid; j < n; j += kPT) {
    const uint64_t kk = skey[j];
    const int h = (int)(kk >> kHiShift);
    const uint32_t d = gbase[h] + (uint32_t)(j - (int)lstart[h]);
    if (d < (uint64_t)M) {
      keys[d] = kk;
      vals[d] = sval[j];
    } else {
      lost = true;
    }
  }
  if (bad) flags[0] = 1;
  if (lost) flags[0] = 2;
}

// The second-level tile of workgroup g: bucket h (tbkt[g], from the layout kernel: two rounds of
// dependent loads, where a search over the buckets' tile starts took one round per halving) and its
// positions [p0, p1).  False for the workgroups beyond the last tile (the grid is a bound, and tbkt
// has an entry for every workgroup of it, unwritten past the last tile).
__device__ __forceinline__ bool tile_of(int g, int nh, const uint32_t* tstart, const uint32_t* hbase,
                                        const uint32_t* tbkt, int& h, uint32_t& p0, uint32_t& p1) {
  const uint32_t ntiles = tstart[nh], hg = tbkt[g];
  if ((uint32_t)g >= ntiles) return false;
  h = (int)hg;
  p0 = hbase[h] + (uint32_t)(g - (int)tstart[h]) * kPTile;
  p1 = min(p0 + (uint32_t)kPTile, hbase[h + 1]);
  return true;
}

__global__ void __launch_bounds__(kPT) bwd_tile_count_kernel(int nh, const uint32_t* __restrict__ hbase,
                                                             const uint32_t* __restrict__ tstart,
                                                             const uint32_t* __restrict__ tbkt,
                                                             const uint64_t* __restrict__ keys, uint32_t* __restrict__ thist) {
  __shared__ uint32_t c[kLoBins];
  int h;
  uint32_t p0, p1;
  if (!tile_of(blockIdx.x, nh, tstart, hbase, tbkt, h, p0, p1)) return;  // the whole workgroup
  const int tid = threadIdx.x;
  if (tid < kLoBins) c[tid] = 0;
  uint32_t e[kPE];
#pragma unroll
  for (int k = 0; k < kPE; ++k) {
    const uint32_t p = p0 + k * kPT + tid;
    e[k] = p < p1 ? (uint32_t)(keys[p] >> kStateShift) : 0xFFFFFFFFu;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPE; ++k)
    if (e[k] != 0xFFFFFFFFu) atomicAdd(&c[e[k] & (kLoBins - 1)], 1u);
  __syncthreads();
  if (tid < kLoBins) thist[(size_t)blockIdx.x * kLoBins + tid] = c[tid];
}

// Per bucket h, thread l: thist[g][l] becomes tile g's start inside state h*kLoBins + l (the counts
// of the bucket's earlier tiles), and grp[v] = the state's first sorted position.
__global__ void __launch_bounds__(kLoBins) bwd_tile_offsets_kernel(int64_t S, const uint32_t* __restrict__ hbase,
                                                                   const uint32_t* __restrict__ tstart,
                                                                   uint32_t* __restrict__ thist, uint32_t* __restrict__ grp) {
  __shared__ uint32_t wtot[kLoBins / 64];
  const int h = blockIdx.x, l = threadIdx.x;
  const uint32_t g0 = tstart[h], g1 = tstart[h + 1];
  uint32_t s = 0;
  for (uint32_t g = g0; g < g1; ++g) {
    uint32_t* p = thist + (size_t)g * kLoBins + l;
    const uint32_t c = *p;
    *p = s;
    s += c;
  }
  uint32_t tot;
  const uint32_t base = block_excl_scan_u32(s, wtot, &tot);
  const int64_t e = (int64_t)h * kLoBins + l;
  if (e <= S) grp[e] = hbase[h] + base;
}

__global__ void __launch_bounds__(kPT) bwd_state_kernel(int64_t M, int nh, const uint32_t* __restrict__ hbase,
                                                       const uint32_t* __restrict__ tstart, const uint32_t* __restrict__ tbkt,
                                                       const uint32_t* __restrict__ thist,
                                                       const uint32_t* __restrict__ grp, const uint64_t* __restrict__ kin,
                                                       const uint32_t* __restrict__ vin, uint64_t* __restrict__ ends,
                                                       uint32_t* __restrict__ ids, int64_t* flags) {
  __shared__ uint32_t cnt[kLoBins], lstart[kLoBins], gdst[kLoBins];
  __shared__ uint32_t wtot[kPT / 64];
  __shared__ uint64_t stage[kPTile];
  int h;
  uint32_t p0, p1;
  if (!tile_of(blockIdx.x, nh, tstart, hbase, tbkt, h, p0, p1)) return;  // the whole workgroup
  const int tid = threadIdx.x;
  const int n = (int)(p1 - p0);
  if (tid < kLoBins) cnt[tid] = 0;
  // every load is issued before any is used (clamped to the tile's last position)
  uint64_t key[kPE];
  uint32_t val[kPE];
#pragma unroll
  for (int k = 0; k < kPE; ++k) {
    const uint32_t p = p0 + (uint32_t)min(k * kPT + tid, n - 1);
    key[k] = __builtin_nontemporal_load(kin + p);
    val[k] = __builtin_nontemporal_load(vin + p);
  }
  __syncthreads();  // counts are zero
  uint32_t off[kPE];
#pragma unroll
  for (int k = 0; k < kPE; ++k)
    off[k] = k * kPT + tid < n ? atomicAdd(&cnt[(key[k] >> kStateShift) & (kLoBins - 1)], 1u) : 0u;
  __syncthreads();
  uint32_t tot;
  const uint32_t c = tid < kLoBins ? cnt[tid] : 0u;
  const uint32_t ls = block_excl_scan_u32(c, wtot, &tot);
  if (tid < kLoBins) {
    lstart[tid] = ls;
    gdst[tid] = c ? grp[(int64_t)h * kLoBins + tid] + thist[(size_t)blockIdx.x * kLoBins + tid] : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPE; ++k)
    if (k * kPT + tid < n) stage[lstart[(key[k] >> kStateShift) & (kLoBins - 1)] + off[k]] = key[k];
  __syncthreads();
  uint32_t dst[kPE];
  bool lost = false;
#pragma unroll
  for (int i = 0; i < kPE; ++i) {
    const int j = tid + i * kPT;
    dst[i] = 0xFFFFFFFFu;
    if (j < n) {
      const uint64_t kk = stage[j];
      const int l = (int)((kk >> kStateShift) & (kLoBins - 1));
      const uint32_t d = gdst[l] + (uint32_t)(j - (int)lstart[l]);
      if (d < (uint64_t)M) {
        dst[i] = d;
        ids[d] = (uint32_t)((kk >> kIdShift) & kIdMask);
      } else {
        lost = true;
      }
    }
  }
  __syncthreads();  // the keys are out of the stage
#pragma unroll
  for (int k = 0; k < kPE; ++k)
    if (k * kPT + tid < n) stage[lstart[(key[k] >> kStateShift) & (kLoBins - 1)] + off[k]] = ends_of(key[k], val[k]);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kPE; ++i)
    if (dst[i] != 0xFFFFFFFFu) ends[dst[i]] = stage[tid + i * kPT];
  if (lost) flags[0] = 2;
}

// Per block, one workgroup: rg[e] = [first, end) of the sorted positions of the queries of the state
// of rank e, for the block's ranks, blast[b] = its last rank with queries (-1: none), and bq[b] =
// its number of queries.
__global__ void __launch_bounds__(256) bwd_rank_groups_kernel(const BwdIndex x, const uint32_t* __restrict__ grp,
                                                              uint2* __restrict__ rg, int32_t* __restrict__ blast,
                                                              uint32_t* __restrict__ bq) {
  __shared__ int last;
  __shared__ uint32_t nq;
  const int b = blockIdx.x;
  const int e0 = x.blk_e0[b], e1 = x.blk_e0[b + 1];
  if (threadIdx.x == 0) {
    last = -1;
    nq = 0;
  }
  __syncthreads();
  int mx = -1;
  uint32_t q = 0;
  for (int e = e0 + (int)threadIdx.x; e < e1; e += 256) {
    const int v = x.steps[x.erank_step[e]];
    const uint2 r = make_uint2(grp[v], grp[v + 1]);
    rg[e] = r;
    q += r.y - r.x;
    if (r.y > r.x) mx = e;
  }
  if (mx >= 0) atomicMax(&last, mx);
  if (q) atomicAdd(&nq, q);
  __syncthreads();
  if (threadIdx.x == 0) {
    blast[b] = last;
    bq[b] = nq;
  }
}

// One workgroup: order[p] = the block taken p-th, by cost bin (any order inside a bin): the bins'
// counts and cursors are LDS counters.
__global__ void __launch_bounds__(1024) bwd_order_kernel(int nblk, const uint32_t* __restrict__ bq,
                                                         int32_t* __restrict__ order) {
  __shared__ uint32_t cnt[kCostBins];
  __shared__ uint32_t wtot[1024 / 64];
  const int tid = threadIdx.x;
  if (tid < kCostBins) cnt[tid] = 0;
  __syncthreads();
  for (int b = tid; b < nblk; b += 1024) atomicAdd(&cnt[cost_bin(bq[b])], 1u);
  __syncthreads();
  uint32_t tot;
  const uint32_t start = block_excl_scan_u32(tid < kCostBins ? cnt[tid] : 0u, wtot, &tot);
  if (tid < kCostBins) cnt[tid] = start;
  __syncthreads();
  for (int b = tid; b < nblk; b += 1024) {
    const uint32_t p = atomicAdd(&cnt[cost_bin(bq[b])], 1u);
    if (p < (uint32_t)nblk) order[p] = b;
  }
}

// ---------------------------------------------------------------- persistent block replay
constexpr int kEPT = 4;   // events per thread per staged tile of steps
constexpr int kGCap = 256;  // group offsets staged in LDS
// dynamic LDS limit of the block replay (the device's 160 KiB less room for its static LDS)
constexpr int kBlockLdsLimit = 159 * 1024;

template <int RPT>
__device__ __forceinline__ int pslot(int r) {
  return r + r / RPT;  // one spare word per thread's run keeps strided run accesses on distinct banks
}

template <int RPT>
constexpr int span_of(int nt) {
  return nt * RPT + nt + 1;
}

// LDS of the block replay, byte offsets (each sub-array 16-byte aligned)
struct BlockLds {
  size_t w, wtot, skey, gq, sdel, srank, sstep, total;
};
__host__ __device__ constexpr size_t al16(size_t v) { return (v + 15) & ~size_t(15); }
__host__ __device__ inline BlockLds block_lds_layout(int span, int nt, int rmax, int U) {
  BlockLds L{};
  L.w = 0;
  L.wtot = al16(L.w + (size_t)span * 8);
  L.skey = al16(L.wtot + 16 * 8);
  L.gq = al16(L.skey + (size_t)rmax * 4);
  L.sdel = al16(L.gq + 2 * kGCap * 4);
  L.srank = al16(L.sdel + (size_t)kEPT * nt * 8);
  L.sstep = al16(L.srank + (size_t)kEPT * nt * 2);
  L.total = al16(L.sstep + (size_t)((kEPT * nt) / U + 1) * 4);  // one step id per step of a tile
  return L;
}

// kSorted: each answer goes to its query's sorted position (out[M], runs of consecutive positions
// per state; the unsort kernels below put them in query-id order); otherwise to out[query id].
template <int RPT, bool kSorted, int NT>
__device__ void dense_block(const BwdIndex& x, const uint64_t* __restrict__ ends, const uint32_t* __restrict__ qids,
                            const uint2* __restrict__ rg, const int32_t* __restrict__ blast,
                            int64_t* __restrict__ out, const int b, const int ea_item, const int off, const int cnt,
                            unsigned char* smem) {
  constexpr int nt = NT;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int span = span_of<RPT>(nt);
  const BlockLds lay = block_lds_layout(span, nt, x.rmax, x.U);
  int64_t* w = reinterpret_cast<int64_t*>(smem + lay.w);          // per rank: within-wave prefix + events since
  int64_t* wtot = reinterpret_cast<int64_t*>(smem + lay.wtot);    // 16 wavefront totals
  uint32_t* skey = reinterpret_cast<uint32_t*>(smem + lay.skey);  // the block's sorted keys
  uint32_t* gq = reinterpret_cast<uint32_t*>(smem + lay.gq);      // kGCap group starts, then ends
  int64_t* sdel = reinterpret_cast<int64_t*>(smem + lay.sdel);    // staged signed deltas
  uint16_t* srank = reinterpret_cast<uint16_t*>(smem + lay.srank);  // staged ranks
  int32_t* sstep = reinterpret_cast<int32_t*>(smem + lay.sstep);    // staged step ids (one per step)

  const int s0 = __builtin_amdgcn_readfirstlane(x.blk_step0[b]);
  const int e0 = __builtin_amdgcn_readfirstlane(x.blk_e0[b]);
  const int e1 = __builtin_amdgcn_readfirstlane(x.blk_e0[b + 1]);
  // replay up to the enter step of the last queried state (none: nothing to replay)
  const int el = __builtin_amdgcn_readfirstlane(blast[b]);
  if (el < 0) return;
  const int s1 = __builtin_amdgcn_readfirstlane(x.erank_step[el]) + 1;
  // the work item: its queries start at offset `off` into the group of rank ea and number cnt; an
  // item past the block's first rank starts its walk at ea's enter step ss, the events of the
  // steps before it applied first in any order (the adds commute)
  const int ea = ea_item < 0 ? e0 : ea_item;
  const int ss = ea == e0 ? s0 : __builtin_amdgcn_readfirstlane(x.erank_step[ea]);
  int left = cnt;
  const int64_t k0 = uniform64(x.blk_key0[b]);
  const int R = (int)(uniform64(x.blk_key0[b + 1]) - k0);
  const int U = x.U;
  const int64_t* crow = x.carry + (int64_t)b * x.carry_ld;

  for (int i = tid; i < span; i += nt) w[i] = 0;
  for (int i = tid; i < R; i += nt) skey[i] = x.bkey[k0 + i];
  if (ss > s0) {
    __syncthreads();  // w is zero
    const int np = (ss - s0) * U;
    const int64_t nu = x.S * (int64_t)U;
    for (int i = tid; i < np; i += nt) {
      const int s = s0 + i / U, k = i - (i / U) * U;
      const int sv = x.steps[s];
      const int64_t at = ((int64_t)(sv > 0 ? sv : -sv) - 1) * U + k;
      const int64_t d = (at >= 0 && at < nu) ? x.delta[at] : 0;
      lds_add64(&w[pslot<RPT>(x.evrank[(int64_t)s * U + k])], sv > 0 ? d : -d);
    }
  }
  int gwin = ea;  // gq[i], gq[kGCap + i] = the group of rank gwin + i
  auto stage_groups = [&](int g) {
    for (int i = tid; i < kGCap; i += nt) {
      const uint2 r = rg[min(g + i, e1 - 1)];
      gq[i] = r.x;
      gq[kGCap + i] = r.y;
    }
  };
  stage_groups(ea);
  int64_t tot[RPT];
#pragma unroll
  for (int j = 0; j < RPT; ++j) tot[j] = 0;

  // The prefix shares w: after a rebuild, the entry of rank y holds the sum of the totals of the
  // ranks < y owned by the same wavefront (its within-wave prefix), and the events add to it
  // afterwards.  The owning thread keeps its ranks' totals (tot) and its within-wave offset (wex),
  // so at the next rebuild it recovers the events added since (entry - what it wrote).  One
  // workgroup barrier; the prefix over all ranks adds the earlier wavefronts' totals (prefix()).
  int64_t wex = 0;
  auto rebuild = [&]() {
    const int base = tid * RPT;
    int64_t loc[RPT];
    int64_t run = 0, was = wex;  // was: what the last rebuild wrote for rank base + j
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int pi = pslot<RPT>(base + j);
      const int64_t d = w[pi] - was;  // the events of rank base + j since then
      was += tot[j];
      tot[j] += d;
      loc[j] = run;
      run += tot[j];
    }
    const int64_t inc = wave_incl_scan(run);
    wex = inc - run;
#pragma unroll
    for (int j = 0; j < RPT; ++j) w[pslot<RPT>(base + j)] = wex + loc[j];
    if (tid == nt - 1) w[pslot<RPT>(nt * RPT)] = inc;  // past the last rank (no event reaches it)
    if (lane == 63) wtot[wv] = inc;
    __syncthreads();
  };
  // sum of the totals of ranks < y, after a rebuild and before the next events
  auto prefix = [&](int y) -> int64_t {
    const int ow = min(y / RPT, nt - 1) >> 6;  // the wavefront owning rank y
    int64_t v = w[pslot<RPT>(y)];
    for (int k = 0; k < ow; ++k) v += wtot[k];
    return v;
  };

  // The queries of the state of rank e (entered at the current step).  The first round's ends
  // (en_next) were loaded by the previous call (or before the walk): its carry-row reads go out at
  // once, and the scan's barriers hide their latency; then the next rank's ends are loaded into the
  // same register, which is read before it is loaded again.
  uint64_t en_next = 0;
  auto first_qa = [&](int e) { return gq[e - gwin] + (e == ea ? (uint32_t)off : 0u); };
  auto query = [&](int e) {
    if (e + 1 - gwin >= kGCap) {  // uniform: ranks e and e + 1 must be staged
      __syncthreads();
      gwin = e;
      stage_groups(e);
      __syncthreads();
    }
    const uint32_t qa = first_qa(e), qb = gq[kGCap + e - gwin];
    const int nq = min((int)(qb - qa), left);
    left -= max(nq, 0);
    const bool live = tid < nq;
    const uint32_t lo = (uint32_t)en_next, hi = (uint32_t)(en_next >> 32);
    int64_t clo = 0, chi = 0;
    if (live) {
      clo = crow[lo];
      chi = crow[hi];
    }
    {  // the next rank's first round: every lane loads (clamped into the group, or to a valid
       // position when it is empty or past the block), so no loaded value is merged with another
      const uint32_t a = gq[e + 1 - gwin], z = gq[kGCap + e + 1 - gwin];
      en_next = ends[min(a + (uint32_t)tid, max(z, 1u) - 1u)];
    }
    if (nq <= 0) return;
    __syncthreads();  // every event of the steps so far is in w
    rebuild();
    if (live) {
      const int rl = lower_bound_lds<uint32_t>(skey, R, lo), rh = lower_bound_lds<uint32_t>(skey, R, hi);
      out[kSorted ? qa + tid : qids[qa + tid]] = (chi - clo) + (prefix(rh) - prefix(rl));
    }
    for (int j = tid + nt; j < nq; j += nt) {  // groups larger than the workgroup
      const uint64_t en = ends[qa + j];
      const uint32_t lo = (uint32_t)en, hi = (uint32_t)(en >> 32);
      const int rl = lower_bound_lds<uint32_t>(skey, R, lo), rh = lower_bound_lds<uint32_t>(skey, R, hi);
      out[kSorted ? qa + j : qids[qa + j]] = (crow[hi] - crow[lo]) + (prefix(rh) - prefix(rl));
    }
    __syncthreads();  // every prefix is read before the next events add to w
  };

  const int T = (kEPT * nt) / U;  // steps per tile (U <= kEPT * nt is checked at launch)
  const int TE = T * U;
  const int64_t nu = x.S * (int64_t)U;
  // events of the tile at ts, given its step ids: ranks and raw deltas are loads in flight
  auto events = [&](int ts, const int* st, int* r, int64_t* d, bool* neg) {
#pragma unroll
    for (int k = 0; k < kEPT; ++k) {
      const int ev = tid + k * nt;
      const int s = ts + ev / U, i = ev - (ev / U) * U;
      const bool ok = ev < TE && s < s1;
      const int64_t v = st[k] > 0 ? st[k] : -st[k];
      r[k] = x.evrank[ok ? (int64_t)s * U + i : 0];
      const int64_t at = (v - 1) * (int64_t)U + i;
      d[k] = x.delta[(ok && at >= 0 && at < nu) ? at : 0];
      neg[k] = st[k] < 0;
    }
  };
  auto ids = [&](int ts, int* st) {
#pragma unroll
    for (int k = 0; k < kEPT; ++k) {
      const int ev = tid + k * nt;
      const int s = ts + ev / U;
      st[k] = x.steps[(ev < TE && s < s1) ? s : s0];
    }
  };
  auto stage_events = [&](int ts, const int* st, const int* r, const int64_t* d, const bool* neg) {
#pragma unroll
    for (int k = 0; k < kEPT; ++k) {
      const int ev = tid + k * nt;
      if (ev < TE) {
        const int s = ts + ev / U;
        const bool ok = s < s1;
        srank[ev] = (uint16_t)(ok ? r[k] : 0);
        sdel[ev] = ok ? (neg[k] ? -d[k] : d[k]) : 0;
        if (ev - (ev / U) * U == 0) sstep[ev / U] = ok ? st[k] : 0;
      }
    }
  };
  int st[kEPT], stn[kEPT], r_n[kEPT];
  int64_t d_n[kEPT];
  bool neg_n[kEPT];
  // prologue: tile 0 staged, tile 1's events in flight, tile 2's step ids in flight
  ids(ss, st);
  events(ss, st, r_n, d_n, neg_n);
  stage_events(ss, st, r_n, d_n, neg_n);
  ids(ss + T, stn);
  events(ss + T, stn, r_n, d_n, neg_n);
  ids(ss + 2 * T, st);
  __syncthreads();
  {  // the first rank's first round (its group from offset off), clamped as in query()
    const uint32_t a = first_qa(ea), z = gq[kGCap];
    en_next = ends[min(a + (uint32_t)tid, max(z, 1u) - 1u)];
  }
  int e = ea;  // rank of the next entered state
  bool done = false;  // the item's queries are answered
  for (int t0 = ss; t0 < s1; t0 += T) {
    const int t1 = min(t0 + T, s1);
    for (int s = t0; s < t1;) {
      // the steps up to the next enter step (or the tile's end) in one pass: the adds commute,
      // and only a query needs the steps before it complete
      int se = s;
      while (se < t1 && sstep[se - t0] <= 0) ++se;
      const int last = se < t1 ? se : t1 - 1;
      const int base = (s - t0) * U, nev = (last - s + 1) * U;
      for (int i = tid; i < nev; i += nt) lds_add64(&w[pslot<RPT>(srank[base + i])], sdel[base + i]);
      if (se < t1) {
        query(e);
        ++e;
        if (left == 0) {  // uniform
          done = true;
          break;
        }
      }
      s = last + 1;
    }
    if (done || t1 >= s1) break;
    __syncthreads();  // the stage is free
    stage_events(t0 + T, stn, r_n, d_n, neg_n);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kEPT; ++k) stn[k] = st[k];
    events(t0 + 2 * T, stn, r_n, d_n, neg_n);
    ids(t0 + 3 * T, st);
  }
}

// Work queue: whole blocks in `order` (items null), or the work items (block, first rank, offset
// into its group, query count), *nitems of them.
template <int RPT, bool kSorted, int NT>
__global__ void __launch_bounds__(NT) bwd_block_kernel(const BwdIndex x, const uint64_t* ends, const uint32_t* ids,
                                                         const uint2* rg, const int32_t* blast, const int32_t* order,
                                                         const int4* items, const uint32_t* nitems, int64_t* out,
                                                         uint32_t* counter) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int4 next;
  for (;;) {
    if (threadIdx.x == 0) {
      const uint32_t i = atomicAdd(counter, 1u);
      if (items) next = i < *nitems ? items[i] : make_int4(x.nblk, 0, 0, 0);
      else next = make_int4(i < (uint32_t)x.nblk ? order[i] : x.nblk, -1, 0, 0x7fffffff);
    }
    __syncthreads();
    const int b = __builtin_amdgcn_readfirstlane(next.x), ea = __builtin_amdgcn_readfirstlane(next.y);
    const int off = __builtin_amdgcn_readfirstlane(next.z), cnt = __builtin_amdgcn_readfirstlane(next.w);
    __syncthreads();  // every thread has read `next` before it is taken again
    if (b >= x.nblk) break;
    dense_block<RPT, kSorted, NT>(x, ends, ids, rg, blast, out, b, ea, off, cnt, smem_raw);
    __syncthreads();  // the next block reuses the LDS
  }
}

// One workgroup: block order[p] gets ceil(bq / chunk) work items (none without queries), the first
// at ibase[p]; ipos[b] = b's place in the order; ibase[nblk] = the item count.
__global__ void __launch_bounds__(1024) bwd_items_base_kernel(int nblk, uint32_t chunk, const int32_t* __restrict__ order,
                                                              const uint32_t* __restrict__ bq, int32_t* __restrict__ ipos,
                                                              uint32_t* __restrict__ ibase) {
  __shared__ uint32_t wtot[1024 / 64];
  uint32_t carry = 0;
  for (int p0 = 0; p0 < nblk; p0 += 1024) {
    const int p = p0 + (int)threadIdx.x;
    uint32_t k = 0;
    if (p < nblk) {
      const int b = order[p];
      ipos[b] = p;
      k = (bq[b] + chunk - 1) / chunk;
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan_u32(k, wtot, &tot);
    if (p < nblk) ibase[p] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) ibase[nblk] = carry;
}

// One wavefront per block: its ranks in walk order and their queries' running count; item j of the
// block starts at the rank holding the block's query j * chunk (a rank's queries in group order),
// at that offset into the rank's group, and holds min(chunk, the rest) queries.
__global__ void __launch_bounds__(256) bwd_items_kernel(const BwdIndex x, uint32_t chunk, const uint2* __restrict__ rg,
                                                        const uint32_t* __restrict__ bq, const int32_t* __restrict__ ipos,
                                                        const uint32_t* __restrict__ ibase, int4* __restrict__ items) {
  const int lane = threadIdx.x & 63;
  const int b = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (b >= x.nblk) return;
  const uint32_t nqb = bq[b];
  if (nqb == 0) return;
  const uint32_t base = ibase[ipos[b]];
  const int e0 = x.blk_e0[b], e1 = x.blk_e0[b + 1];
  uint32_t cs = 0;  // queries of the ranks before this round of 64
  for (int r0 = e0; r0 < e1; r0 += 64) {
    const int e = r0 + lane;
    uint32_t c = 0;
    if (e < e1) {
      const uint2 r = rg[e];
      c = r.y - r.x;
    }
    const uint32_t inc = (uint32_t)wave_incl_scan((int64_t)c);
    const uint32_t q = cs + inc - c;  // the block's ordinal of this rank's first query
    for (uint32_t j = (q + chunk - 1) / chunk; j * chunk < q + c; ++j)
      items[base + j] = make_int4(b, e, (int)(j * chunk - q), (int)min(chunk, nqb - j * chunk));
    cs += (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
  }
}

// ---------------------------------------------------------------- unsort: answers to query-id order
// A random 8-byte store costs a 32-byte write request and a partial line in some die's cache, so
// the replay (kSorted) stores its answers in sorted order, and two streaming passes put them in
// query-id order through windows of 2^kWinBits ids (2 MiB of answers each):
//   partition -- a tile of sorted positions is grouped by window in LDS, and each window's run is
//                appended to that window's region of a staging buffer (region w = staging positions
//                [w << kWinBits, ...): every window but the last holds exactly 2^kWinBits ids, so the
//                staging buffer is dense over [0, M)); a per-window cursor hands out the runs, so
//                the order inside a region is free (the next pass stores by id);
//   scatter   -- the staging buffer is cut into kDies contiguous parts, part d taken by the
//                workgroups that run on die d (workgroup g runs on die g mod kDies), so each
//                window's stores come from one die and its lines fill in that die's L2 before they
//                are written back.  The launch reserves kScatLds bytes of LDS it does not use, so
//                one workgroup runs per compute unit: with a quarter window in flight per die the
//                partly written lines stay in its L2 until complete (measured over 102.5M answers:
//                0.69 ms and 2.1 GB written, against 1.12 ms and 3.5 GB at full occupancy).
constexpr int kWinBits = 18;
constexpr int kMaxWin = (int)(BWD_MAX_QUERIES >> kWinBits);  // 512 windows cover every query id
// the partition's tiles: 12,288 answers (147 KiB staged in LDS, one workgroup per CU); tiles of
// 8,192 / 4,096 (two workgroups per CU) measured 0.73 / 0.83 ms against 0.68 ms over 102.5M answers:
// a bigger tile takes fewer cursor atomics and stores longer runs per window
constexpr int kPartT = 1024, kPartE = 12, kPartTile = kPartT * kPartE;
constexpr int kScatT = 256, kScatE = 8, kScatTile = kScatT * kScatE;
constexpr int kDies = 8;
constexpr int kScatLds = 100 * 1024;
static_assert(kMaxWin <= kPartT && kMaxWin % 64 == 0, "one thread per window in the partition's scan");

__global__ void __launch_bounds__(256) bwd_replay_init_kernel(uint32_t* counter, uint32_t* cursor, int nw) {
  if (threadIdx.x == 0) counter[0] = 0;
  if (cursor)
    for (int i = threadIdx.x; i < nw; i += 256) cursor[i] = 0;
}

__global__ void __launch_bounds__(kPartT) bwd_part_kernel(int64_t M, const uint32_t* __restrict__ ids,
                                                          const int64_t* __restrict__ sans, uint32_t* __restrict__ cursor,
                                                          int64_t* __restrict__ stage_ans, uint32_t* __restrict__ stage_id) {
  __shared__ uint32_t cnt[kMaxWin], lstart[kMaxWin], gbase[kMaxWin];
  __shared__ uint32_t wtot[kMaxWin / 64];
  __shared__ int64_t sv[kPartTile];
  __shared__ uint32_t si[kPartTile];
  const int tid = threadIdx.x;
  const int64_t t0 = (int64_t)blockIdx.x * kPartTile;
  const int n = (int)min<int64_t>(kPartTile, M - t0);
  if (tid < kMaxWin) cnt[tid] = 0;
  // every load is issued before any is used (clamped to the tile's last position)
  uint32_t id[kPartE];
  int64_t a[kPartE];
#pragma unroll
  for (int k = 0; k < kPartE; ++k) {
    const int64_t p = t0 + min(k * kPartT + tid, n - 1);
    id[k] = __builtin_nontemporal_load(ids + p);
    a[k] = __builtin_nontemporal_load(sans + p);
  }
  __syncthreads();  // counts are zero
  uint32_t off[kPartE];
#pragma unroll
  for (int k = 0; k < kPartE; ++k) off[k] = k * kPartT + tid < n ? atomicAdd(&cnt[id[k] >> kWinBits], 1u) : 0u;
  __syncthreads();
  // thread w: window w's start in the tile (exclusive scan of the counts) and its run's place in
  // the staging region
  if (tid < kMaxWin) {
    const uint32_t c = cnt[tid];
    const uint32_t inc = (uint32_t)wave_incl_scan((int64_t)c);
    if ((tid & 63) == 63) wtot[tid >> 6] = inc;
    lstart[tid] = inc - c;
    if (c) gbase[tid] = atomicAdd(cursor + tid, c);
  }
  __syncthreads();
  if (tid < kMaxWin) {
    uint32_t o = 0;
    for (int w = 0; w < (tid >> 6); ++w) o += wtot[w];
    lstart[tid] += o;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPartE; ++k) {
    if (k * kPartT + tid < n) {
      const int j = (int)(lstart[id[k] >> kWinBits] + off[k]);
      sv[j] = a[k];
      si[j] = id[k];
    }
  }
  __syncthreads();
  // consecutive tile places of one window go to consecutive staging positions
  for (int j = tid; j < n; j += kPartT) {
    const uint32_t q = si[j];
    const int w = (int)(q >> kWinBits);
    const int64_t g = ((int64_t)w << kWinBits) + gbase[w] + (uint32_t)(j - (int)lstart[w]);
    stage_ans[g] = sv[j];
    stage_id[g] = q;
  }
}

__global__ void __launch_bounds__(kScatT) bwd_scatter_kernel(int64_t M, int64_t per, const int64_t* __restrict__ stage_ans,
                                                             const uint32_t* __restrict__ stage_id,
                                                             int64_t* __restrict__ answers) {
  const int64_t i0 = ((int64_t)(blockIdx.x % kDies) * per + blockIdx.x / kDies) * kScatTile + threadIdx.x;
  uint32_t q[kScatE];
  int64_t v[kScatE];
#pragma unroll
  for (int k = 0; k < kScatE; ++k) {
    const int64_t i = min(i0 + k * kScatT, M - 1);
    q[k] = __builtin_nontemporal_load(stage_id + i);
    v[k] = __builtin_nontemporal_load(stage_ans + i);
  }
#pragma unroll
  for (int k = 0; k < kScatE; ++k)
    if (i0 + k * kScatT < M) answers[q[k]] = v[k];
}

// ---------------------------------------------------------------- host helpers
int report(const char* what, hipError_t e) {
  if (e != hipSuccess) fprintf(stderr, "bwd: %s: %s\n", what, hipGetErrorString(e));
  return (int)e;
}

bool index_ok(const BwdIndex* x) {
  return x && x->S >= 1 && x->S <= BWD_MAX_STATES && x->U >= 1 && x->nblk >= 1 && x->rmax >= 1 &&
         x->rmax <= BWD_MAX_RMAX && x->NK >= 1 && x->NK <= BWD_MAX_NK && x->carry_ld >= (int64_t)x->NK + 1 &&
         x->carry_ld % 64 == 0 && x->nsteps >= 1 && x->delta && x->steps && x->blk_step0 && x->evrank &&
         x->blk_key0 && x->bkey && x->blk_e0 && x->state_erank && x->erank_step && x->carry;
}

struct CarryLayout {
  int nr, nseg;
  int64_t ldT;
  size_t aggp, rb, T, total;
};

CarryLayout carry_layout(const BwdIndex* x, int64_t nkeys) {
  CarryLayout L;
  const int64_t NKP = (int64_t)x->NK + 1;
  L.nr = (int)((NKP + kW - 1) / kW);
  L.nseg = (x->nblk + kSegBlocks - 1) / kSegBlocks;
  L.ldT = ((NKP + 63) / 64) * 64;
  size_t o = 0;
  auto take = [&](size_t n) {
    const size_t a = o;
    o += (n + 255) & ~size_t(255);
    return a;
  };
  L.aggp = take((size_t)(nkeys + x->nblk) * 8);
  L.rb = take((size_t)x->nblk * (L.nr + 1) * 2);
  L.T = take((size_t)L.nseg * L.ldT * 8);
  L.total = o;
  return L;
}

// The planned batch: keys/vals (the bucket pass's output: queries grouped by high rank digit; then,
// once the plan has run, the unsort's staging buffer: stage_ans = kin, stage_id = vin), ends/ids
// (queries grouped by state: endpoints, query ids), grp, the bucket tables, per rank its state's group (rg) and per block
// its last queried rank (blast); sans holds the replay's answers in sorted order.
struct AnswerLayout {
  size_t kin, ends, vin, ids, grp, rg, blast, bq, order, counter, cursor, sans, hcount, hcur, hbase, tstart, tbkt,
      thist, items, ipos, ibase, total;
  int nh;
  unsigned tiles;  // bound on the second level's tiles (its grid)
};

hipError_t answer_layout(const BwdIndex* x, int64_t M, AnswerLayout& L) {
  L.nh = (int)((x->S >> kLoBits) + 1);
  L.tiles = (unsigned)((M + kPTile - 1) / kPTile + L.nh);
  size_t o = 0;
  auto take = [&](size_t n) {
    const size_t a = o;
    o += (n + 255) & ~size_t(255);
    return a;
  };
  L.kin = take((size_t)M * 8);
  L.ends = take((size_t)M * 8);
  L.vin = take((size_t)M * 4);
  L.ids = take((size_t)M * 4);
  L.grp = take((size_t)(x->S + 2) * 4);
  L.rg = take((size_t)x->S * 8);
  L.blast = take((size_t)x->nblk * 4);
  L.bq = take((size_t)x->nblk * 4);
  L.order = take((size_t)x->nblk * 4);
  L.counter = take(4);
  L.cursor = take((size_t)kMaxWin * 4);
  L.sans = take((size_t)M * 8);
  L.hcount = take((size_t)kMaxHi * 4);
  L.hcur = take((size_t)kMaxHi * 4);
  L.hbase = take((size_t)(kMaxHi + 1) * 4);
  L.tstart = take((size_t)(kMaxHi + 1) * 4);
  L.tbkt = take((size_t)L.tiles * 4);
  L.thist = take((size_t)L.tiles * kLoBins * 4);
  // work items: at most one per block plus one per BWD_MIN_CHUNK queries
  L.items = take(((size_t)x->nblk + (size_t)(M / BWD_MIN_CHUNK) + 1) * 16);
  L.ipos = take((size_t)x->nblk * 4);
  L.ibase = take((size_t)(x->nblk + 1) * 4);
  L.total = o;
  return hipSuccess;
}

template <class T>
T* at(void* base, size_t off) {
  return reinterpret_cast<T*>(static_cast<unsigned char*>(base) + off);
}

// the block replay's instances: [sorted][ranks per thread 4 / 8 / 16][threads 256 / 512 / 1024],
// each compiled for its thread count (its register budget)
using BlockKernel = void (*)(const BwdIndex, const uint64_t*, const uint32_t*, const uint2*, const int32_t*,
                             const int32_t*, const int4*, const uint32_t*, int64_t*, uint32_t*);
template <bool kS, int RPT>
constexpr BlockKernel kBlockRow[3] = {bwd_block_kernel<RPT, kS, 256>, bwd_block_kernel<RPT, kS, 512>,
                                      bwd_block_kernel<RPT, kS, 1024>};
BlockKernel block_kernel(bool sorted, int rpt, int nt) {
  const int r = rpt == 4 ? 0 : rpt == 8 ? 1 : 2, t = nt == 256 ? 0 : nt == 512 ? 1 : 2;
  static const BlockKernel* const tab[2][3] = {{kBlockRow<false, 4>, kBlockRow<false, 8>, kBlockRow<false, 16>},
                                               {kBlockRow<true, 4>, kBlockRow<true, 8>, kBlockRow<true, 16>}};
  return tab[sorted ? 1 : 0][r][t];
}

int set_lds_limits() {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
    std::vector<const void*> ks;
    for (int sorted = 0; sorted < 2; ++sorted)
      for (int rpt : {4, 8, 16})
        for (int nt : {256, 512, 1024}) ks.push_back(reinterpret_cast<const void*>(block_kernel(sorted, rpt, nt)));
    for (const void* k : ks) {
      const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, kBlockLdsLimit);
      if (e != hipSuccess && err == hipSuccess) err = e;
    }
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(bwd_scatter_kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, kScatLds);
    if (e != hipSuccess && err == hipSuccess) err = e;
  });
  return (int)err;
}

}  // namespace

extern "C" int bwd_carry_workspace(const BwdIndex* x, size_t* bytes) {
  if (!bytes || !index_ok(x)) return (int)hipErrorInvalidValue;
  int64_t nkeys = 0;
  if (hipMemcpy(&nkeys, x->blk_key0 + x->nblk, sizeof(nkeys), hipMemcpyDeviceToHost) != hipSuccess)
    return (int)hipErrorInvalidValue;
  *bytes = carry_layout(x, nkeys).total;
  return 0;
}

extern "C" int bwd_carry_build_launch(const BwdIndex* x, const int64_t* agg, void* work, size_t work_bytes,
                                      void* stream) {
  if (!index_ok(x) || !agg || !work) return (int)hipErrorInvalidValue;
  int64_t nkeys = 0;
  if (hipMemcpy(&nkeys, x->blk_key0 + x->nblk, sizeof(nkeys), hipMemcpyDeviceToHost) != hipSuccess)
    return (int)hipErrorInvalidValue;
  const CarryLayout L = carry_layout(x, nkeys);
  if (work_bytes < L.total) return (int)hipErrorInvalidValue;
  const hipStream_t st = (hipStream_t)stream;
  int64_t* aggp = at<int64_t>(work, L.aggp);
  uint16_t* rb = at<uint16_t>(work, L.rb);
  int64_t* T = at<int64_t>(work, L.T);
  const int64_t NKP = (int64_t)x->NK + 1;
  int rc;
  hipLaunchKernelGGL(bwd_aggp_kernel, dim3((unsigned)x->nblk), dim3(kCT), 0, st, *x, agg, aggp);
  if ((rc = report("carry prefix", hipGetLastError()))) return rc;
  const int64_t nrb = (int64_t)x->nblk * (L.nr + 1);
  hipLaunchKernelGGL(bwd_rb_kernel, dim3((unsigned)((nrb + 255) / 256)), dim3(256), 0, st, *x, L.nr, rb);
  if ((rc = report("carry range ranks", hipGetLastError()))) return rc;
  const dim3 grid((unsigned)L.nr, (unsigned)L.nseg);
  hipLaunchKernelGGL(bwd_carry_range_kernel<false>, grid, dim3(kCT), 0, st, *x, aggp, rb, L.nr, T, L.ldT);
  if ((rc = report("carry segment totals", hipGetLastError()))) return rc;
  hipLaunchKernelGGL(bwd_seg_scan_kernel, dim3((unsigned)((NKP + 255) / 256)), dim3(256), 0, st, NKP, L.nseg, T, L.ldT);
  if ((rc = report("carry segment scan", hipGetLastError()))) return rc;
  hipLaunchKernelGGL(bwd_carry_range_kernel<true>, grid, dim3(kCT), 0, st, *x, aggp, rb, L.nr, T, L.ldT);
  return report("carry rows", hipGetLastError());
}

extern "C" int bwd_answer_workspace(const BwdIndex* x, int64_t M, size_t* bytes) {
  if (!bytes || !index_ok(x) || M < 1 || M > BWD_MAX_QUERIES) return (int)hipErrorInvalidValue;
  AnswerLayout L;
  const hipError_t e = answer_layout(x, M, L);
  if (e) return report("answer workspace", e);
  *bytes = L.total;
  return 0;
}

extern "C" int bwd_plan_launch(const BwdIndex* x, int64_t M, const int32_t* qstate, const uint32_t* qlo,
                               const uint32_t* qhi, int64_t* flags, void* work, size_t work_bytes, void* stream) {
  if (!index_ok(x) || M < 1 || M > BWD_MAX_QUERIES || !qstate || !qlo || !qhi || !flags || !work)
    return (int)hipErrorInvalidValue;
  AnswerLayout L;
  hipError_t e = answer_layout(x, M, L);
  if (e) return report("answer workspace", e);
  if (work_bytes < L.total) return (int)hipErrorInvalidValue;
  const hipStream_t st = (hipStream_t)stream;
  uint64_t* kin = at<uint64_t>(work, L.kin);
  uint64_t* ends = at<uint64_t>(work, L.ends);
  uint32_t* vin = at<uint32_t>(work, L.vin);
  uint32_t* ids = at<uint32_t>(work, L.ids);
  uint32_t* grp = at<uint32_t>(work, L.grp);
  uint32_t* hcount = at<uint32_t>(work, L.hcount);
  uint32_t* hcur = at<uint32_t>(work, L.hcur);
  uint32_t* hbase = at<uint32_t>(work, L.hbase);
  uint32_t* tstart = at<uint32_t>(work, L.tstart);
  uint32_t* tbkt = at<uint32_t>(work, L.tbkt);
  uint32_t* thist = at<uint32_t>(work, L.thist);
  const int nh = L.nh;
  const unsigned qgrid = (unsigned)((M + kPTile - 1) / kPTile), bgrid = (unsigned)((M + kBTile - 1) / kBTile);
  int rc;
  hipLaunchKernelGGL(bwd_plan_init_kernel, dim3(1), dim3(1024), 0, st, flags, hcount, nh);
  hipLaunchKernelGGL(bwd_count_kernel, dim3(std::min(qgrid, kCountGrid)), dim3(kPT), 0, st, x->S, M, qstate, nh, hcount);
  hipLaunchKernelGGL(bwd_layout_kernel, dim3(1), dim3(kPT), 0, st, x->S, M, nh, hcount, hbase, hcur, tstart, grp, tbkt);
  if ((rc = report("plan counts", hipGetLastError()))) return rc;
  // the 16-byte form when the caller's query arrays allow it
  if ((((uintptr_t)qstate | (uintptr_t)qlo | (uintptr_t)qhi) & 15) == 0)
    hipLaunchKernelGGL(bwd_bucket_kernel<true>, dim3(bgrid), dim3(kPT), 0, st, *x, M, qstate, qlo, qhi, nh, hcur, kin,
                       vin, flags);
  else
    hipLaunchKernelGGL(bwd_bucket_kernel<false>, dim3(bgrid), dim3(kPT), 0, st, *x, M, qstate, qlo, qhi, nh, hcur, kin,
                       vin, flags);
  if ((rc = report("plan buckets", hipGetLastError()))) return rc;
  hipLaunchKernelGGL(bwd_tile_count_kernel, dim3(L.tiles), dim3(kPT), 0, st, nh, hbase, tstart, tbkt, kin, thist);
  hipLaunchKernelGGL(bwd_tile_offsets_kernel, dim3((unsigned)nh), dim3(kLoBins), 0, st, x->S, hbase, tstart, thist, grp);
  hipLaunchKernelGGL(bwd_state_kernel, dim3(L.tiles), dim3(kPT), 0, st, M, nh, hbase, tstart, tbkt, thist, grp, kin, vin, ends,
                     ids, flags);
  if ((rc = report("plan states", hipGetLastError()))) return rc;
  hipLaunchKernelGGL(bwd_rank_groups_kernel, dim3((unsigned)x->nblk), dim3(256), 0, st, *x, grp, at<uint2>(work, L.rg),
                     at<int32_t>(work, L.blast), at<uint32_t>(work, L.bq));
  if ((rc = report("plan rank groups", hipGetLastError()))) return rc;
  hipLaunchKernelGGL(bwd_order_kernel, dim3(1), dim3(1024), 0, st, x->nblk, at<const uint32_t>(work, L.bq),
                     at<int32_t>(work, L.order));
  return report("plan block order", hipGetLastError());
}

// The block replay's launch for a batch of M queries: its instance, grid, workgroups per CU that
// reside (as asked, at most as many as the instance's registers and LDS allow: a workgroup beyond
// them would only start once the queue is empty) and LDS per workgroup.
struct ReplayShape {
  bool staged;
  int nt, rpt, wg_per_cu, workgroups;
  size_t lds;
  BlockKernel kernel;
};
static int replay_shape(const BwdIndex* x, int64_t M, const BwdLaunch* l, ReplayShape& sh) {
  if (!index_ok(x) || M < 1 || M > BWD_MAX_QUERIES || !l) return (int)hipErrorInvalidValue;
  const int nt = l->threads;
  if (nt != 256 && nt != 512 && nt != 1024) return (int)hipErrorInvalidValue;
  if (l->wg_per_cu < 1 || l->wg_per_cu > 8 || l->max_workgroups < 0 || l->store < BWD_STORE_AUTO ||
      l->store > BWD_STORE_STAGED || l->chunk < 0 || (l->chunk > 0 && l->chunk < BWD_MIN_CHUNK))
    return (int)hipErrorInvalidValue;
  // ranks per thread: the smallest of 4, 8, 16 with rpt * nt >= rmax
  const int need = (x->rmax + nt - 1) / nt;
  const int rpt = need <= 4 ? 4 : need <= 8 ? 8 : 16;
  if (need > 16 || x->U > kEPT * nt) return (int)hipErrorInvalidValue;
  const int spn = rpt == 4 ? span_of<4>(nt) : rpt == 8 ? span_of<8>(nt) : span_of<16>(nt);
  const size_t lds = block_lds_layout(spn, nt, x->rmax, x->U).total;
  if (lds > (size_t)kBlockLdsLimit) return (int)hipErrorInvalidValue;
  int rc = set_lds_limits();
  if (rc) return report("replay LDS limit", (hipError_t)rc);
  hipError_t e;
  int dev = 0, ncu = 0;
  if ((e = hipGetDevice(&dev)) || (e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)))
    return report("device query", e);
  // staged stores pay off once the ids span more than one window
  sh.staged = l->store == BWD_STORE_STAGED || (l->store == BWD_STORE_AUTO && M > (int64_t(1) << kWinBits));
  sh.nt = nt;
  sh.rpt = rpt;
  sh.lds = lds;
  sh.kernel = block_kernel(sh.staged, rpt, nt);
  int resident = 0;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, reinterpret_cast<const void*>(sh.kernel), nt, lds)))
    return report("replay occupancy", e);
  sh.wg_per_cu = std::max(1, std::min(l->wg_per_cu, resident));
  sh.workgroups = ncu * sh.wg_per_cu;
  if (l->max_workgroups > 0 && l->max_workgroups < sh.workgroups) sh.workgroups = l->max_workgroups;
  return 0;
}

// The replay's launches; with ev (four events) they are recorded at the phase boundaries: before
// the block replay, after it, after the partition, after the scatter.
static int replay_impl(const BwdIndex* x, int64_t M, int64_t* answers, const BwdLaunch* l, void* work,
                       size_t work_bytes, hipStream_t st, hipEvent_t* ev) {
  if (!answers || !work) return (int)hipErrorInvalidValue;
  ReplayShape sh;
  int rc = replay_shape(x, M, l, sh);
  if (rc) return rc;
  AnswerLayout L;
  hipError_t e = answer_layout(x, M, L);
  if (e) return report("answer workspace", e);
  if (work_bytes < L.total) return (int)hipErrorInvalidValue;
  const bool staged = sh.staged;
  const int nt = sh.nt;
  const size_t lds = sh.lds;
  const BlockKernel bk = sh.kernel;
  const int nwin = (int)((M + (int64_t(1) << kWinBits) - 1) >> kWinBits);
  uint32_t* counter = at<uint32_t>(work, L.counter);
  uint32_t* cursor = at<uint32_t>(work, L.cursor);
  hipLaunchKernelGGL(bwd_replay_init_kernel, dim3(1), dim3(256), 0, st, counter, staged ? cursor : nullptr, nwin);
  const dim3 grid((unsigned)sh.workgroups);
  const uint64_t* ends = at<uint64_t>(work, L.ends);
  const uint32_t* ids = at<uint32_t>(work, L.ids);
  const uint2* rg = at<uint2>(work, L.rg);
  const int32_t* blast = at<int32_t>(work, L.blast);
  const int32_t* order = at<int32_t>(work, L.order);
  int64_t* out = staged ? at<int64_t>(work, L.sans) : answers;
  auto mark = [&](int i) { return ev ? hipEventRecord(ev[i], st) : hipSuccess; };
  if ((e = mark(0))) return report("replay event", e);
  const int4* items = nullptr;
  const uint32_t* nitems = nullptr;
  if (l->chunk > 0) {  // the work items, in the blocks' order
    const uint32_t* bq = at<const uint32_t>(work, L.bq);
    int32_t* ipos = at<int32_t>(work, L.ipos);
    uint32_t* ibase = at<uint32_t>(work, L.ibase);
    hipLaunchKernelGGL(bwd_items_base_kernel, dim3(1), dim3(1024), 0, st, x->nblk, (uint32_t)l->chunk, order, bq, ipos,
                       ibase);
    hipLaunchKernelGGL(bwd_items_kernel, dim3((unsigned)((x->nblk + 3) / 4)), dim3(256), 0, st, *x,
                       (uint32_t)l->chunk, rg, bq, ipos, ibase, at<int4>(work, L.items));
    if ((rc = report("replay work items", hipGetLastError()))) return rc;
    items = at<const int4>(work, L.items);
    nitems = ibase + x->nblk;
  }
  auto launch = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, grid, dim3(nt), lds, st, *x, ends, ids, rg, blast, order, items, nitems, out, counter);
  };
  launch(bk);
  if ((rc = report("dense block replay", hipGetLastError()))) return rc;
  if ((e = mark(1))) return report("replay event", e);
  if (!staged) {
    if ((e = mark(2)) || (e = mark(3))) return report("replay event", e);
    return 0;
  }
  hipLaunchKernelGGL(bwd_part_kernel, dim3((unsigned)((M + kPartTile - 1) / kPartTile)), dim3(kPartT), 0, st, M, ids,
                     at<const int64_t>(work, L.sans), cursor, at<int64_t>(work, L.kin), at<uint32_t>(work, L.vin));
  if ((rc = report("answer partition", hipGetLastError()))) return rc;
  if ((e = mark(2))) return report("replay event", e);
  const int64_t per = ((M + kScatTile - 1) / kScatTile + kDies - 1) / kDies;
  hipLaunchKernelGGL(bwd_scatter_kernel, dim3((unsigned)(per * kDies)), dim3(kScatT), kScatLds, st, M, per,
                     at<const int64_t>(work, L.kin), at<const uint32_t>(work, L.vin), answers);
  if ((rc = report("answer scatter", hipGetLastError()))) return rc;
  if ((e = mark(3))) return report("replay event", e);
  return 0;
}

extern "C" int bwd_replay_shape(const BwdIndex* x, int64_t M, const BwdLaunch* l, int32_t* workgroups,
                                int32_t* wg_per_cu, size_t* lds_bytes) {
  if (!workgroups || !wg_per_cu || !lds_bytes) return (int)hipErrorInvalidValue;
  ReplayShape sh;
  const int rc = replay_shape(x, M, l, sh);
  if (rc) return rc;
  *workgroups = sh.workgroups;
  *wg_per_cu = sh.wg_per_cu;
  *lds_bytes = sh.lds;
  return 0;
}

extern "C" int bwd_replay_work_items(const BwdIndex* x, int64_t M, const void* work, size_t work_bytes,
                                     uint32_t* items) {
  if (!index_ok(x) || M < 1 || M > BWD_MAX_QUERIES || !work || !items) return (int)hipErrorInvalidValue;
  AnswerLayout L;
  hipError_t e = answer_layout(x, M, L);
  if (e) return report("answer workspace", e);
  if (work_bytes < L.total) return (int)hipErrorInvalidValue;
  const uint32_t* n = reinterpret_cast<const uint32_t*>(static_cast<const unsigned char*>(work) + L.ibase) + x->nblk;
  if ((e = hipMemcpy(items, n, sizeof(uint32_t), hipMemcpyDeviceToHost))) return report("work item count", e);
  return 0;
}

extern "C" int bwd_replay_launch(const BwdIndex* x, int64_t M, int64_t* answers, const BwdLaunch* l, void* work,
                                 size_t work_bytes, void* stream) {
  return replay_impl(x, M, answers, l, work, work_bytes, (hipStream_t)stream, nullptr);
}

extern "C" int bwd_replay_timed(const BwdIndex* x, int64_t M, int64_t* answers, const BwdLaunch* l, void* work,
                                size_t work_bytes, void* stream, float* ms) {
  if (!ms) return (int)hipErrorInvalidValue;
  hipEvent_t ev[4] = {};
  hipError_t e = hipSuccess;
  for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipEventCreate(&ev[i]);
  int rc = e ? report("replay events", e) : replay_impl(x, M, answers, l, work, work_bytes, (hipStream_t)stream, ev);
  if (!rc && (e = hipEventSynchronize(ev[3]))) rc = report("replay", e);
  for (int i = 0; i < 3 && !rc; ++i)
    if ((e = hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]))) rc = report("replay event time", e);
  for (int i = 0; i < 4; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  return rc;
}

// the current device's address of `scratch` (looked up on the first call on each device)
static hipError_t scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_at[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e) return e;
    g_scratch_at[dev] = p;
  }
  *out = g_scratch_at[dev];
  return hipSuccess;
}

extern "C" int bwd_answer_launch(const BwdIndex* x, int64_t M, const int32_t* qstate, const uint32_t* qlo,
                                 const uint32_t* qhi, int64_t* answers, int64_t* flags, const BwdLaunch* l, void* work,
                                 size_t work_bytes, void* stream) {
  const int rc = bwd_plan_launch(x, M, qstate, qlo, qhi, flags, work, work_bytes, stream);
  if (rc) return rc;
  return bwd_replay_launch(x, M, answers, l, work, work_bytes, stream);
}

extern "C" int bwd_answer_noalloc_launch(const BwdIndex* x, int64_t M, const int32_t* qstate, const uint32_t* qlo,
                                         const uint32_t* qhi, int64_t* answers, int64_t* flags, const BwdLaunch* l,
                                         void* stream) {
  if (!index_ok(x) || M < 1 || M > BWD_MAX_QUERIES) return (int)hipErrorInvalidValue;
  AnswerLayout L;
  hipError_t e = answer_layout(x, M, L);
  if (e) return report("answer workspace", e);
  if (L.total > kScratchBytes) {
    fprintf(stderr, "bwd_answer_noalloc_launch: a batch of %lld queries over %lld states needs %zu bytes of scratch; "
                    "this entry point holds %zu (use bwd_answer_launch with a workspace)\n",
            (long long)M, (long long)x->S, L.total, kScratchBytes);
    return (int)hipErrorInvalidValue;
  }
  void* work = nullptr;
  if ((e = scratch_address(&work))) return report("scratch address", e);
  return bwd_answer_launch(x, M, qstate, qlo, qhi, answers, flags, l, work, kScratchBytes, stream);
}

extern "C" int bwd_answer_noalloc_capacity(const BwdIndex* x, int64_t M, size_t* bytes_needed, size_t* bytes_held) {
  if (!bytes_needed || !bytes_held || !index_ok(x) || M < 1 || M > BWD_MAX_QUERIES) return (int)hipErrorInvalidValue;
  AnswerLayout L;
  const hipError_t e = answer_layout(x, M, L);
  if (e) return report("answer workspace", e);
  *bytes_needed = L.total;
  *bytes_held = kScratchBytes;
  return 0;
}
