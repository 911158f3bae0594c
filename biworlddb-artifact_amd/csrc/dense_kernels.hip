// dense_kernels.hip -- the dense-carry batched form (include/bwgpu_dense.h) for gfx950.
//
// Stage map onto the reference's v7 algorithm (src/batch_engine_v7.hpp:832-968):
//   carry build      <- the carry pass (src/panel_batch_gpu.cpp:191-204): carrypref[b][k] =
//                       sum over blocks before b of their net weight of classes < k.  Here one
//                       dense row per block over every coordinate, built once: the blocks' net
//                       weights (agg) become per-block prefix functions D_b(x); rows are their
//                       running sum over blocks, computed per (coordinate range, block segment)
//                       with a scan over segments in between, so ~45k workgroups share the work.
//   plan             <- qlb/qhb (src/batch_engine_v7.hpp:666-698): queries sorted by the preorder
//                       rank of their state (rocPRIM radix sort in two 10-bit passes), group offsets
//                       per rank.
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
#include <rocprim/device/device_radix_sort.hpp>
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

// ---------------------------------------------------------------- sort key layout
// key = rank << 44 | query id << 17 | lo & 0x1FFFF;  val = hi | (lo >> 17) << 21.  The sorted bits
// [44, 44 + 19) end below bit 64: rocPRIM's path for inputs under 2^20 items (merge sort) left keys
// sorted on bits [45, 64) out of order here, while its onesweep path sorted them.
constexpr int kRankShift = 44;
constexpr int kIdShift = 17;
constexpr uint64_t kIdMask = (uint64_t(1) << 27) - 1;
constexpr uint32_t kLoLowMask = (1u << 17) - 1;
constexpr uint32_t kHiMask = (1u << 21) - 1;

__device__ __forceinline__ void unpack(uint64_t key, uint32_t val, uint32_t& qid, uint32_t& lo, uint32_t& hi) {
  qid = (uint32_t)((key >> kIdShift) & kIdMask);
  lo = ((uint32_t)key & kLoLowMask) | ((val >> 21) << 17);
  hi = val & kHiMask;
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

// ---------------------------------------------------------------- plan
__global__ void bwd_init_kernel(int64_t* flags, uint32_t* counter) {
  if (threadIdx.x == 0) {
    if (flags) flags[0] = 0;
    if (counter) counter[0] = 0;
  }
}

// The sort's input: key = rank << 44 | query id << 17 | lo & 0x1FFFF, val = hi | (lo >> 17) << 21.
// An out-of-range query sets flags[0] and is planned as a query of state 1 over [0, 0).  Each
// thread packs kKeysPer queries: with kVec four consecutive ones, read and written with 16-byte
// accesses (the caller's arrays 16-byte aligned), else four a workgroup apart; either way all of a
// thread's loads are issued before its rank gathers.
constexpr int kKeysPer = 4;
template <bool kVec>
__global__ void __launch_bounds__(256) bwd_keys_kernel(const BwdIndex x, int64_t M, const int32_t* __restrict__ qstate,
                                                       const uint32_t* __restrict__ qlo,
                                                       const uint32_t* __restrict__ qhi, uint64_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals, int64_t* flags) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t qv = t * kKeysPer;  // kVec: the thread's first query
  const bool whole = kVec && qv + kKeysPer <= M;
  int64_t q[kKeysPer];
  int32_t v[kKeysPer];
  uint32_t lo[kKeysPer], hi[kKeysPer];
#pragma unroll
  for (int k = 0; k < kKeysPer; ++k) q[k] = kVec ? qv + k : (int64_t)blockIdx.x * (256 * kKeysPer) + threadIdx.x + k * 256;
  if (whole) {
    const int4 a = *reinterpret_cast<const int4*>(qstate + qv);
    const uint4 l = *reinterpret_cast<const uint4*>(qlo + qv);
    const uint4 h = *reinterpret_cast<const uint4*>(qhi + qv);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    lo[0] = l.x; lo[1] = l.y; lo[2] = l.z; lo[3] = l.w;
    hi[0] = h.x; hi[1] = h.y; hi[2] = h.z; hi[3] = h.w;
  } else {
#pragma unroll
    for (int k = 0; k < kKeysPer; ++k) {
      const bool in = q[k] < M;
      v[k] = in ? qstate[q[k]] : 1;
      lo[k] = in ? qlo[q[k]] : 0;
      hi[k] = in ? qhi[q[k]] : 0;
    }
  }
  bool bad = false;
#pragma unroll
  for (int k = 0; k < kKeysPer; ++k) {
    if (v[k] < 1 || v[k] > x.S || lo[k] > x.NK || hi[k] > x.NK) {
      bad = true;
      v[k] = 1;
      lo[k] = hi[k] = 0;
    }
  }
  uint64_t key[kKeysPer];
  uint32_t val[kKeysPer];
#pragma unroll
  for (int k = 0; k < kKeysPer; ++k) {
    const uint64_t e = (uint64_t)(uint32_t)x.state_erank[v[k]];
    key[k] = (e << kRankShift) | ((uint64_t)q[k] << kIdShift) | (uint64_t)(lo[k] & kLoLowMask);
    val[k] = hi[k] | ((lo[k] >> 17) << 21);
  }
  if (whole) {
    ulonglong2* kp = reinterpret_cast<ulonglong2*>(keys + qv);
    kp[0] = make_ulonglong2(key[0], key[1]);
    kp[1] = make_ulonglong2(key[2], key[3]);
    *reinterpret_cast<uint4*>(vals + qv) = make_uint4(val[0], val[1], val[2], val[3]);
  } else {
#pragma unroll
    for (int k = 0; k < kKeysPer; ++k) {
      if (q[k] < M) {
        keys[q[k]] = key[k];
        vals[q[k]] = val[k];
      }
    }
  }
  if (bad) flags[0] = 1;
}

// grp[e] = first sorted position whose rank is >= e, for e in [0, S]; grp[S] = M
__global__ void __launch_bounds__(256) bwd_groups_kernel(int64_t S, int64_t M, const uint64_t* keys, uint32_t* grp) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= M) return;
  const int64_t e = (int64_t)(keys[j] >> kRankShift);
  const int64_t pe = j > 0 ? (int64_t)(keys[j - 1] >> kRankShift) : -1;
  for (int64_t t = pe + 1; t <= e; ++t) grp[t] = (uint32_t)j;
  if (j == M - 1)
    for (int64_t t = e + 1; t <= S; ++t) grp[t] = (uint32_t)M;
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
  size_t w, P, wtot, skey, gq, sdel, srank, sstep, total;
};
__host__ __device__ constexpr size_t al16(size_t v) { return (v + 15) & ~size_t(15); }
__host__ __device__ inline BlockLds block_lds_layout(int span, int nt, int rmax) {
  BlockLds L{};
  L.w = 0;
  L.P = al16(L.w + (size_t)span * 8);
  L.wtot = al16(L.P + (size_t)span * 8);
  L.skey = al16(L.wtot + 16 * 8);
  L.gq = al16(L.skey + (size_t)rmax * 4);
  L.sdel = al16(L.gq + (kGCap + 1) * 4);
  L.srank = al16(L.sdel + (size_t)kEPT * nt * 8);
  L.sstep = al16(L.srank + (size_t)kEPT * nt * 2);
  L.total = al16(L.sstep + (size_t)kEPT * nt * 4);
  return L;
}

// kSorted: each answer goes to its query's sorted position (out[M], runs of consecutive positions
// per state; the unsort kernels below put them in query-id order); otherwise to out[query id].
template <int RPT, bool kSorted>
__device__ void dense_block(const BwdIndex& x, const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                            const uint32_t* __restrict__ grp, int64_t* __restrict__ out, const int b,
                            unsigned char* smem) {
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wv = tid >> 6;
  const int span = span_of<RPT>(nt);
  const BlockLds lay = block_lds_layout(span, nt, x.rmax);
  int64_t* w = reinterpret_cast<int64_t*>(smem + lay.w);          // deltas since the last scan, per rank
  int64_t* P = reinterpret_cast<int64_t*>(smem + lay.P);          // exclusive prefix over ranks
  int64_t* wtot = reinterpret_cast<int64_t*>(smem + lay.wtot);    // 16 wavefront totals
  uint32_t* skey = reinterpret_cast<uint32_t*>(smem + lay.skey);  // the block's sorted keys
  uint32_t* gq = reinterpret_cast<uint32_t*>(smem + lay.gq);      // kGCap + 1 group offsets
  int64_t* sdel = reinterpret_cast<int64_t*>(smem + lay.sdel);    // staged signed deltas
  uint16_t* srank = reinterpret_cast<uint16_t*>(smem + lay.srank);  // staged ranks
  int32_t* sstep = reinterpret_cast<int32_t*>(smem + lay.sstep);    // staged step ids (one per step)

  const int s0 = __builtin_amdgcn_readfirstlane(x.blk_step0[b]);
  const int e0 = __builtin_amdgcn_readfirstlane(x.blk_e0[b]);
  const int e1 = __builtin_amdgcn_readfirstlane(x.blk_e0[b + 1]);
  const uint32_t q0 = __builtin_amdgcn_readfirstlane(grp[e0]), q1 = __builtin_amdgcn_readfirstlane(grp[e1]);
  if (q1 == q0) return;  // no query in this block: nothing to replay
  // replay up to the enter step of the last queried state
  const int el = (int)(uniform64((int64_t)(keys[q1 - 1] >> kRankShift)));
  const int s1 = __builtin_amdgcn_readfirstlane(x.erank_step[el]) + 1;
  const int64_t k0 = uniform64(x.blk_key0[b]);
  const int R = (int)(uniform64(x.blk_key0[b + 1]) - k0);
  const int U = x.U;
  const int64_t* crow = x.carry + (int64_t)b * x.carry_ld;

  for (int i = tid; i < span; i += nt) w[i] = 0;
  for (int i = tid; i < R; i += nt) skey[i] = x.bkey[k0 + i];
  int gwin = e0;  // gq[i] = grp[gwin + i]
  auto stage_groups = [&](int g) {
    for (int i = tid; i <= kGCap; i += nt) gq[i] = grp[min(g + i, e1)];
  };
  stage_groups(e0);
  int64_t tot[RPT];
#pragma unroll
  for (int j = 0; j < RPT; ++j) tot[j] = 0;

  // w -> rank totals; P[y] = sum of totals of ranks < y.  Two workgroup barriers.
  auto rebuild = [&]() {
    const int base = tid * RPT;
    int64_t loc[RPT];
    int64_t run = 0;
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int pi = pslot<RPT>(base + j);
      tot[j] += w[pi];
      w[pi] = 0;
      loc[j] = run;
      run += tot[j];
    }
    const int64_t inc = wave_incl_scan(run);
    if (lane == 63) wtot[wv] = inc;
    __syncthreads();
    int64_t off = inc - run;
    for (int k = 0; k < wv; ++k) off += wtot[k];
#pragma unroll
    for (int j = 0; j < RPT; ++j) P[pslot<RPT>(base + j)] = off + loc[j];
    if (tid == nt - 1) P[pslot<RPT>(nt * RPT)] = off + run;
    __syncthreads();
  };

  // the queries of the state of rank e (entered at the current step)
  auto query = [&](int e) {
    if (e - gwin >= kGCap) {  // uniform
      __syncthreads();
      gwin = e;
      stage_groups(e);
      __syncthreads();
    }
    const uint32_t qa = gq[e - gwin], qb = gq[e - gwin + 1];
    const int nq = (int)(qb - qa);
    if (nq == 0) return;
    // first round: the queries' items and carry-row reads are issued before the scan, whose
    // barriers hide their latency
    uint32_t qid = 0, lo = 0, hi = 0;
    int64_t clo = 0, chi = 0;
    const bool live = tid < nq;
    if (live) {
      unpack(keys[qa + tid], vals[qa + tid], qid, lo, hi);
      clo = crow[lo];
      chi = crow[hi];
    }
    __syncthreads();  // every event of the steps so far is in w; previous readers of P are done
    rebuild();
    if (live) {
      const int rl = lower_bound_lds<uint32_t>(skey, R, lo), rh = lower_bound_lds<uint32_t>(skey, R, hi);
      out[kSorted ? qa + tid : qid] = (chi - clo) + (P[pslot<RPT>(rh)] - P[pslot<RPT>(rl)]);
    }
    for (int j = tid + nt; j < nq; j += nt) {  // groups larger than the workgroup
      unpack(keys[qa + j], vals[qa + j], qid, lo, hi);
      const int rl = lower_bound_lds<uint32_t>(skey, R, lo), rh = lower_bound_lds<uint32_t>(skey, R, hi);
      out[kSorted ? qa + j : qid] = (crow[hi] - crow[lo]) + (P[pslot<RPT>(rh)] - P[pslot<RPT>(rl)]);
    }
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
  ids(s0, st);
  events(s0, st, r_n, d_n, neg_n);
  stage_events(s0, st, r_n, d_n, neg_n);
  ids(s0 + T, stn);
  events(s0 + T, stn, r_n, d_n, neg_n);
  ids(s0 + 2 * T, st);
  __syncthreads();
  int e = e0;  // rank of the next entered state
  for (int t0 = s0; t0 < s1; t0 += T) {
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
      }
      s = last + 1;
    }
    if (t1 >= s1) break;
    __syncthreads();  // the stage is free
    stage_events(t0 + T, stn, r_n, d_n, neg_n);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kEPT; ++k) stn[k] = st[k];
    events(t0 + 2 * T, stn, r_n, d_n, neg_n);
    ids(t0 + 3 * T, st);
  }
}

template <int RPT, bool kSorted>
__global__ void __launch_bounds__(1024) bwd_block_kernel(const BwdIndex x, const uint64_t* keys, const uint32_t* vals,
                                                         const uint32_t* grp, int64_t* out, uint32_t* counter) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int next;
  for (;;) {
    if (threadIdx.x == 0) next = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int b = __builtin_amdgcn_readfirstlane(next);
    __syncthreads();  // every thread has read `next` before it is taken again
    if (b >= x.nblk) break;
    dense_block<RPT, kSorted>(x, keys, vals, grp, out, b, smem_raw);
    __syncthreads();  // the next block reuses the LDS
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
//                are written back.
constexpr int kWinBits = 18;
constexpr int kMaxWin = (int)(BWD_MAX_QUERIES >> kWinBits);  // 512 windows cover every query id
constexpr int kPartT = 1024, kPartE = 8, kPartTile = kPartT * kPartE;
constexpr int kScatT = 256, kScatE = 8, kScatTile = kScatT * kScatE;
constexpr int kDies = 8;
static_assert(kMaxWin <= kPartT && kMaxWin % 64 == 0, "one thread per window in the partition's scan");

__global__ void __launch_bounds__(512) bwd_replay_init_kernel(uint32_t* counter, uint32_t* cursor, int nw) {
  if (threadIdx.x == 0) counter[0] = 0;
  if (cursor && (int)threadIdx.x < nw) cursor[threadIdx.x] = 0;
}

__global__ void __launch_bounds__(kPartT) bwd_part_kernel(int64_t M, const uint64_t* __restrict__ keys,
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
    id[k] = (uint32_t)((__builtin_nontemporal_load(keys + p) >> kIdShift) & kIdMask);
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

unsigned bits_for(uint64_t v) {
  unsigned b = 1;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
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

// The plan's sort: 10-bit digits, so the 19 rank bits of a 300,000-state tree take two passes
// (rocPRIM's default for 64-bit keys takes three 8-bit passes); 1,024-thread workgroups of 12 items.
using PlanSortConfig = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 16>, rocprim::kernel_config<1024, 12>, 10,
                                        rocprim::block_radix_rank_algorithm::match>>;

hipError_t plan_sort(void* temp, size_t& temp_bytes, const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
                     uint32_t* vout, int64_t M, unsigned end_bit, hipStream_t st) {
  return rocprim::radix_sort_pairs<PlanSortConfig>(temp, temp_bytes, kin, kout, vin, vout, (size_t)M,
                                                   (unsigned)kRankShift, end_bit, st);
}

// kin / vin (the sort's input) are free once the plan has run: the unsort's staging buffer reuses
// them (stage_ans = kin, stage_id = vin); sans holds the replay's answers in sorted order.
struct AnswerLayout {
  size_t kin, kout, vin, vout, grp, counter, cursor, sans, temp, temp_bytes, total;
  unsigned end_bit;
};

hipError_t answer_layout(const BwdIndex* x, int64_t M, AnswerLayout& L) {
  L.end_bit = kRankShift + bits_for((uint64_t)(x->S > 1 ? x->S - 1 : 1));
  L.temp_bytes = 0;
  const hipError_t e = plan_sort(nullptr, L.temp_bytes, nullptr, nullptr, nullptr, nullptr, M, L.end_bit, nullptr);
  if (e) return e;
  size_t o = 0;
  auto take = [&](size_t n) {
    const size_t a = o;
    o += (n + 255) & ~size_t(255);
    return a;
  };
  L.kin = take((size_t)M * 8);
  L.kout = take((size_t)M * 8);
  L.vin = take((size_t)M * 4);
  L.vout = take((size_t)M * 4);
  L.grp = take((size_t)(x->S + 1) * 4);
  L.counter = take(4);
  L.cursor = take((size_t)kMaxWin * 4);
  L.sans = take((size_t)M * 8);
  L.temp = take(std::max<size_t>(L.temp_bytes, 1));
  L.total = o;
  return hipSuccess;
}

template <class T>
T* at(void* base, size_t off) {
  return reinterpret_cast<T*>(static_cast<unsigned char*>(base) + off);
}

int set_lds_limits() {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
    const void* ks[] = {reinterpret_cast<const void*>(bwd_block_kernel<4, false>),
                        reinterpret_cast<const void*>(bwd_block_kernel<8, false>),
                        reinterpret_cast<const void*>(bwd_block_kernel<16, false>),
                        reinterpret_cast<const void*>(bwd_block_kernel<4, true>),
                        reinterpret_cast<const void*>(bwd_block_kernel<8, true>),
                        reinterpret_cast<const void*>(bwd_block_kernel<16, true>)};
    for (const void* k : ks) {
      const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, kBlockLdsLimit);
      if (e != hipSuccess && err == hipSuccess) err = e;
    }
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
  uint64_t* kout = at<uint64_t>(work, L.kout);
  uint32_t* vin = at<uint32_t>(work, L.vin);
  uint32_t* vout = at<uint32_t>(work, L.vout);
  uint32_t* grp = at<uint32_t>(work, L.grp);
  int rc;
  hipLaunchKernelGGL(bwd_init_kernel, dim3(1), dim3(64), 0, st, flags, (uint32_t*)nullptr);
  const unsigned kgrid = (unsigned)((M + 256 * kKeysPer - 1) / (256 * kKeysPer));
  // the 16-byte form when the caller's query arrays allow it (the workspace arrays are 256-byte aligned)
  if ((((uintptr_t)qstate | (uintptr_t)qlo | (uintptr_t)qhi) & 15) == 0)
    hipLaunchKernelGGL(bwd_keys_kernel<true>, dim3(kgrid), dim3(256), 0, st, *x, M, qstate, qlo, qhi, kin, vin, flags);
  else
    hipLaunchKernelGGL(bwd_keys_kernel<false>, dim3(kgrid), dim3(256), 0, st, *x, M, qstate, qlo, qhi, kin, vin, flags);
  if ((rc = report("plan keys", hipGetLastError()))) return rc;
  size_t tb = L.temp_bytes;
  if ((rc = report("plan sort", plan_sort(at<void>(work, L.temp), tb, kin, kout, vin, vout, M, L.end_bit, st))))
    return rc;
  const unsigned grid = (unsigned)((M + 255) / 256);
  hipLaunchKernelGGL(bwd_groups_kernel, dim3(grid), dim3(256), 0, st, x->S, M, kout, grp);
  return report("plan groups", hipGetLastError());
}

extern "C" int bwd_replay_launch(const BwdIndex* x, int64_t M, int64_t* answers, const BwdLaunch* l, void* work,
                                 size_t work_bytes, void* stream) {
  if (!index_ok(x) || M < 1 || M > BWD_MAX_QUERIES || !answers || !l || !work) return (int)hipErrorInvalidValue;
  const int nt = l->threads;
  if (nt != 256 && nt != 512 && nt != 1024) return (int)hipErrorInvalidValue;
  if (l->wg_per_cu < 1 || l->wg_per_cu > 8 || l->max_workgroups < 0 || l->store < BWD_STORE_AUTO ||
      l->store > BWD_STORE_STAGED)
    return (int)hipErrorInvalidValue;
  // ranks per thread: the smallest of 4, 8, 16 with rpt * nt >= rmax
  const int need = (x->rmax + nt - 1) / nt;
  const int rpt = need <= 4 ? 4 : need <= 8 ? 8 : 16;
  if (need > 16 || x->U > kEPT * nt) return (int)hipErrorInvalidValue;
  const int spn = rpt == 4 ? span_of<4>(nt) : rpt == 8 ? span_of<8>(nt) : span_of<16>(nt);
  if (block_lds_layout(spn, nt, x->rmax).total > (size_t)kBlockLdsLimit) return (int)hipErrorInvalidValue;
  AnswerLayout L;
  hipError_t e = answer_layout(x, M, L);
  if (e) return report("answer workspace", e);
  if (work_bytes < L.total) return (int)hipErrorInvalidValue;
  int rc = set_lds_limits();
  if (rc) return report("replay LDS limit", (hipError_t)rc);
  int dev = 0, ncu = 0;
  if ((e = hipGetDevice(&dev)) || (e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)))
    return report("device query", e);
  const hipStream_t st = (hipStream_t)stream;
  // staged stores pay off once the ids span more than one window
  const bool staged = l->store == BWD_STORE_STAGED || (l->store == BWD_STORE_AUTO && M > (int64_t(1) << kWinBits));
  const int nwin = (int)((M + (int64_t(1) << kWinBits) - 1) >> kWinBits);
  uint32_t* counter = at<uint32_t>(work, L.counter);
  uint32_t* cursor = at<uint32_t>(work, L.cursor);
  hipLaunchKernelGGL(bwd_replay_init_kernel, dim3(1), dim3(512), 0, st, counter, staged ? cursor : nullptr, nwin);
  int nwg = ncu * l->wg_per_cu;
  if (l->max_workgroups > 0 && l->max_workgroups < nwg) nwg = l->max_workgroups;
  const dim3 grid((unsigned)nwg);
  const uint64_t* kout = at<uint64_t>(work, L.kout);
  const uint32_t* vout = at<uint32_t>(work, L.vout);
  const uint32_t* grp = at<uint32_t>(work, L.grp);
  int64_t* out = staged ? at<int64_t>(work, L.sans) : answers;
  const size_t lds = block_lds_layout(spn, nt, x->rmax).total;
  auto launch = [&](auto kernel) { hipLaunchKernelGGL(kernel, grid, dim3(nt), lds, st, *x, kout, vout, grp, out, counter); };
  if (staged) {
    if (rpt == 4) launch(bwd_block_kernel<4, true>);
    else if (rpt == 8) launch(bwd_block_kernel<8, true>);
    else launch(bwd_block_kernel<16, true>);
  } else {
    if (rpt == 4) launch(bwd_block_kernel<4, false>);
    else if (rpt == 8) launch(bwd_block_kernel<8, false>);
    else launch(bwd_block_kernel<16, false>);
  }
  if ((rc = report("dense block replay", hipGetLastError()))) return rc;
  if (!staged) return 0;
  hipLaunchKernelGGL(bwd_part_kernel, dim3((unsigned)((M + kPartTile - 1) / kPartTile)), dim3(kPartT), 0, st, M, kout,
                     at<const int64_t>(work, L.sans), cursor, at<int64_t>(work, L.kin), at<uint32_t>(work, L.vin));
  if ((rc = report("answer partition", hipGetLastError()))) return rc;
  const int64_t per = ((M + kScatTile - 1) / kScatTile + kDies - 1) / kDies;
  hipLaunchKernelGGL(bwd_scatter_kernel, dim3((unsigned)(per * kDies)), dim3(kScatT), 0, st, M, per,
                     at<const int64_t>(work, L.kin), at<const uint32_t>(work, L.vin), answers);
  return report("answer scatter", hipGetLastError());
}

// the current device's address of `scratch` (looked up on the first call on each device)
hipError_t scratch_address(void** out) {
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
