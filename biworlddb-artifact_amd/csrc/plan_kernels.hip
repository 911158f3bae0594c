// plan_kernels.hip -- the device batch planner (include/bwgpu_plan.h): a query batch becomes
// the launch plan of the replay kernels through four stable radix sorts (rocPRIM), three scans
// and a handful of one-thread-per-element kernels, all on the caller's stream.
//
// The reference plans on the host before each launch (qlb/qhb bounds per query,
// src/batch_engine_v7.hpp:666-698; radix-sort-driven plan build, :1069-1348).  The steps here
// produce exactly the plan of the host planner build_batch_plan (csrc/host/tree_plan.cpp):
//   1. slots      sort (enter step, query id), carrying the endpoints -> qorder; per slot its
//                 block, in-block ranks (binary search in the block's sorted keys), key tiles;
//   2. groups     scan of "new step" / "new block" flags -> grp_step/grp_q0, qblk/qblk_g0;
//   3. endpoints  per endpoint its level-B chain position (a dense (block, tile) table, or a
//                 binary search over the blocks that hold keys of its tile), chunk and item;
//                 sort (slot range, chunk, position) keeping endpoint order;
//   4. work       run starts by a max-scan, cuts at every ep_per_work-th item of a run (moved
//                 on past a slot's pair), scan -> wk_chunk/wk_e0, pair count; each endpoint's
//                 index in the work order -> epos;
//   5. staging    sort slots by (query-id window, die part) -> ipos; sort those positions by
//                 slot tile -> spos (ascending per tile) and srank.
// Every sort carries what the next step needs, so no step gathers through the permutation.
// Out-of-range queries are planned as state 1 over [0, 0) and flagged in dcounts[4].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <stdint.h>
#include <stdio.h>

#include "bwgpu_plan.h"
#include "bwgpu_tree.h"
#include "plan_common.hpp"

using bwg::SlotRanges;

namespace {

constexpr int kT = 256;

// The sorts' configuration: up to 4,096 items sort in one single-block launch (512 threads x 8
// items) instead of a block sort and a merge pass per doubling -- a served batch of a few thousand
// queries then plans in a few launches; larger sorts use rocPRIM's own choices.  Every sort is
// stable, so its result does not depend on the algorithm.
using PlanSort = rocprim::radix_sort_config<rocprim::kernel_config<512, 8>>;

// values < 2^bits for every v' <= v (at least one bit, so no sort runs over an empty key)
unsigned bits_for(uint64_t v) {
  unsigned b = 1;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

struct Shape {
  int64_t M, nE, ntile, per_t, nwin;
  unsigned sbits, rb, cb, pb, gbits, tbits;
  int64_t G, CH;
  SlotRanges sr;
  int sh, ts;
};

bool shape_of(const BwtIndex* x, const BwtPlanIndex* p, int64_t M, Shape& s) {
  if (!x || !p || M < 1 || M > BWT_MAX_BATCH || p->nsteps < 1 || p->chunk_items < 1 || p->ep_per_work < 64 ||
      p->range_slots < 1 || x->tile_shift < 1 || x->tile_shift > 12)
    return false;
  s.M = M;
  s.nE = 2 * M;
  s.ntile = (M + BWT_COMBINE_TILE - 1) / BWT_COMBINE_TILE;
  s.per_t = std::max<int64_t>(1, (s.ntile + BWT_XCDS - 1) / BWT_XCDS);
  s.sh = std::min(31, std::max(8, (int)p->stage_shift));
  s.nwin = (M >> s.sh) + 1;
  s.ts = x->tile_shift;
  s.G = p->ep_per_work;
  s.sr = bwg::slot_ranges(M, p->range_slots);
  s.CH = p->chunk_items;
  s.sbits = bits_for(uint64_t(p->nsteps - 1));
  s.rb = bits_for(uint64_t(bwg::range_of(M - 1, s.sr)));
  s.cb = bits_for(uint64_t(x->nchunk - 1));
  // a position within a chunk is at most the chunk length (an endpoint after a chain's last
  // key sits at the end of the chain's last chunk, which is at most one chunk long)
  s.pb = bits_for(uint64_t(s.CH));
  s.gbits = bits_for(uint64_t(s.nwin * BWT_XCDS - 1));
  s.tbits = bits_for(uint64_t(s.ntile - 1));
  return s.rb + s.cb + s.pb <= 64 && s.pb <= 19;
}

struct Layout {
  size_t skey_in, skey_out, sval_in, sval_out, flags, fscan;           // slots
  size_t ekey_in, ekey_out, eval_in, wflag, wscan, a0in, a0out;        // endpoints
  size_t gkey_in, gkey_out, gval_in, gval_out, tkey_in, tkey_out, tval_in, tval_out;  // staging
  size_t temp, temp_bytes, total;
};

template <class T>
T* at(void* base, size_t off) {
  return reinterpret_cast<T*>(static_cast<unsigned char*>(base) + off);
}

// Temporary-storage bytes of every rocPRIM call the planner makes (size queries only).
hipError_t temp_bytes(const Shape& s, size_t& need) {
  need = 0;
  size_t b = 0;
  hipError_t e;
  if ((e = rocprim::radix_sort_pairs<PlanSort>(nullptr, b, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                     (const uint64_t*)nullptr, (uint64_t*)nullptr, size_t(s.M), 32u, 32u + s.sbits)))
    return e;
  need = std::max(need, b);
  if ((e = rocprim::inclusive_scan(nullptr, b, (const uint64_t*)nullptr, (uint64_t*)nullptr, size_t(s.M),
                                   rocprim::plus<uint64_t>())))
    return e;
  need = std::max(need, b);
  if ((e = rocprim::radix_sort_pairs<PlanSort>(nullptr, b, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                     (const uint64_t*)nullptr, (uint64_t*)nullptr, size_t(s.nE), 0u,
                                     s.rb + s.cb + s.pb)))
    return e;
  need = std::max(need, b);
  if ((e = rocprim::inclusive_scan(nullptr, b, (const uint32_t*)nullptr, (uint32_t*)nullptr, size_t(s.nE),
                                   rocprim::maximum<uint32_t>())))
    return e;
  need = std::max(need, b);
  if ((e = rocprim::inclusive_scan(nullptr, b, (const uint64_t*)nullptr, (uint64_t*)nullptr, size_t(s.nE),
                                   rocprim::plus<uint64_t>())))
    return e;
  need = std::max(need, b);
  for (const unsigned gb : {s.gbits, s.tbits}) {
    if ((e = rocprim::radix_sort_pairs<PlanSort>(nullptr, b, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                       (const uint64_t*)nullptr, (uint64_t*)nullptr, size_t(s.M), 0u, gb)))
      return e;
    need = std::max(need, b);
  }
  return hipSuccess;
}

hipError_t layout_of(const Shape& s, Layout& L) {
  size_t o = 0;
  auto take = [&](size_t n) {
    const size_t a = o;
    o += (n + 255) & ~size_t(255);
    return a;
  };
  const size_t M = size_t(s.M), E = size_t(s.nE);
  L.skey_in = take(M * 8);
  L.skey_out = take(M * 8);
  L.sval_in = take(M * 8);
  L.sval_out = take(M * 8);
  L.flags = take(M * 8);
  L.fscan = take(M * 8);
  L.ekey_in = take(E * 8);
  L.ekey_out = take(E * 8);
  L.eval_in = take(E * 8);
  L.wflag = take(E * 8);
  L.wscan = take(E * 8);
  L.a0in = take(E * 4);
  L.a0out = take(E * 4);
  L.gkey_in = take(M * 4);
  L.gkey_out = take(M * 4);
  L.gval_in = take(M * 8);
  L.gval_out = take(M * 8);
  L.tkey_in = take(M * 4);
  L.tkey_out = take(M * 4);
  L.tval_in = take(M * 8);
  L.tval_out = take(M * 8);
  const hipError_t e = temp_bytes(s, L.temp_bytes);
  if (e) return e;
  L.temp = take(std::max<size_t>(L.temp_bytes, 1));
  L.total = o;
  return hipSuccess;
}

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* a, int64_t n, uint32_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return (uint32_t)lo;
}

__device__ __forceinline__ int64_t lower_bound_i32(const int32_t* a, int64_t lo, int64_t hi, int32_t v) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__global__ void plan_init_kernel(int64_t* dcounts) {
  if (threadIdx.x < 5) dcounts[threadIdx.x] = 0;
}

// 1. sort keys of the slots: enter step << 32 | query id; the values carry the endpoints
//    (an out-of-range query becomes state 1 over [0, 0))
__global__ void __launch_bounds__(kT) plan_slot_keys_kernel(const BwtIndex x, const BwtPlanIndex p,
                                                            const int32_t* qstate, const BwtBatch b, int64_t M,
                                                            uint64_t* skey, uint64_t* sval, int64_t* dcounts) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (q >= M) return;
  const int32_t v = qstate[q];
  uint32_t lo = b.qlo[q], hi = b.qhi[q];
  const bool bad = v < 1 || v > x.S || lo > p.NK || hi > p.NK;
  if (bad) {
    dcounts[4] = 1;
    lo = hi = 0;
  }
  skey[q] = ((uint64_t)(uint32_t)p.state_step[bad ? 1 : v] << 32) | (uint64_t)q;
  sval[q] = (uint64_t)lo | ((uint64_t)hi << 32);
}

// per slot: query id, block, in-block ranks, key tiles; flags "new step group" | "new block" << 32
__global__ void __launch_bounds__(kT) plan_slot_info_kernel(const BwtIndex x, const BwtPlanIndex p,
                                                            const BwtBatch b, int64_t M, const uint64_t* skey,
                                                            const uint64_t* sval, uint64_t* flags) {
  const int64_t j = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (j >= M) return;
  const uint64_t key = skey[j], lh = sval[j];
  const uint32_t s = (uint32_t)(key >> 32), lo = (uint32_t)lh, hi = (uint32_t)(lh >> 32);
  const int32_t blk = p.step_block[s];
  const int64_t k0 = x.blk_key0[blk], n = x.blk_key0[blk + 1] - k0;
  const uint32_t rl = lower_bound_u32(x.bkey + k0, n, lo), rh = lower_bound_u32(x.bkey + k0, n, hi);
  const_cast<uint32_t*>(b.qorder)[j] = (uint32_t)key;
  const_cast<int32_t*>(b.sblk)[j] = blk;
  const_cast<uint32_t*>(b.qrank)[j] = rl | (rh << 16);
  const_cast<uint32_t*>(b.qtile)[j] = (lo >> x.tile_shift) | ((hi >> x.tile_shift) << 16);
  const uint32_t ps = j > 0 ? (uint32_t)(skey[j - 1] >> 32) : 0u;
  const bool g = j == 0 || ps != s;
  const bool nb = g && (j == 0 || (int64_t)ps < (int64_t)x.blk_step0[blk]);
  flags[j] = (uint64_t)g | ((uint64_t)nb << 32);
}

// 2. step groups and queried blocks from the scanned flags
__global__ void __launch_bounds__(kT) plan_groups_kernel(const BwtBatch b, int64_t M, const uint64_t* skey,
                                                         const uint64_t* flags, const uint64_t* fscan,
                                                         int64_t* dcounts) {
  const int64_t j = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (j >= M) return;
  const uint64_t f = flags[j], c = fscan[j];
  const int64_t gi = (int64_t)(uint32_t)c - 1, bi = (int64_t)(c >> 32) - 1;
  if (f & 1u) {
    const_cast<int32_t*>(b.grp_step)[gi] = (int32_t)(skey[j] >> 32);
    const_cast<int64_t*>(b.grp_q0)[gi] = j;
  }
  if (f >> 32) {
    const_cast<int32_t*>(b.qblk)[bi] = b.sblk[j];
    const_cast<int64_t*>(b.qblk_g0)[bi] = gi;
  }
  if (j == M - 1) {
    const int64_t ngrp = (int64_t)(uint32_t)c, nqb = (int64_t)(c >> 32);
    const_cast<int64_t*>(b.grp_q0)[ngrp] = M;
    const_cast<int64_t*>(b.qblk_g0)[nqb] = ngrp;
    dcounts[0] = nqb;
    dcounts[1] = ngrp;
  }
}

// 3. endpoint items and their sort keys (slot range, chunk, position)
__global__ void __launch_bounds__(kT) plan_endpoints_kernel(const BwtIndex x, const BwtPlanIndex p, const BwtBatch b,
                                                            int64_t nE, SlotRanges sr, int64_t CH, unsigned cb,
                                                            unsigned pb, const uint64_t* sval, uint64_t* ekey,
                                                            uint64_t* eval) {
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (e >= nE) return;
  const int64_t j = e >> 1;
  const uint64_t lh = sval[j];
  const uint32_t xk = (e & 1) ? (uint32_t)(lh >> 32) : (uint32_t)lh;
  const int32_t blk = b.sblk[j];
  // the keys of tile t in blocks < blk precede the endpoint on chain 1 + t
  const uint32_t t = xk >> x.tile_shift;
  int64_t pos;
  if (p.bt_pos) {
    pos = p.bt_pos[(int64_t)blk * p.KT + t];
  } else {
    const int64_t lo = p.tile_tb0[t], hi = p.tile_tb0[t + 1];
    const int64_t idx = lower_bound_i32(p.tb_block, lo, hi, blk);
    pos = idx < hi ? p.tb_pos[idx] : p.chain_key0[t + 2] - p.chain_key0[t + 1];
  }
  const int64_t c0 = x.chain_chunk0[1 + t], n = x.chain_chunk0[2 + t] - c0;
  int64_t k = pos / CH;
  if (k >= n) k = n - 1;
  const uint64_t pin = (uint64_t)(pos - k * CH);
  const uint64_t mask = (1u << x.tile_shift) - 1;
  eval[e] = (pin << BWT_EP_POS_SHIFT) | ((uint64_t)(xk & mask) << BWT_EP_CLASS_SHIFT) | ((uint64_t)e & BWT_EP_ID_MASK);
  ekey[e] = ((uint64_t)bwg::range_of(j, sr) << (cb + pb)) | ((uint64_t)(c0 + k) << pb) | pin;
}

// 4. work items: run starts (for a max-scan) ...
__global__ void __launch_bounds__(kT) plan_run_start_kernel(int64_t nE, unsigned pb, const uint64_t* ekey,
                                                            uint32_t* a0) {
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (e >= nE) return;
  const bool start = e == 0 || (ekey[e] >> pb) != (ekey[e - 1] >> pb);
  a0[e] = start ? (uint32_t)e : 0u;
}

// ... cuts ("starts a work item" | "second endpoint of a pair" << 32) ...
__global__ void __launch_bounds__(kT) plan_work_flags_kernel(const BwtBatch b, int64_t nE, int64_t G,
                                                             const uint32_t* a0, uint64_t* wflag) {
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (e >= nE) return;
  const int64_t off = e - (int64_t)a0[e];
  auto slot = [&](int64_t i) { return BWT_EP_ID(b.eitems[i]) >> 1; };
  bool start = off == 0;
  if (off > 0 && off % G == 0) start = slot(e) != slot(e - 1);
  else if (off > G && off % G == 1) start = slot(e - 1) == slot(e - 2);
  const bool pair = off > 0 && slot(e) == slot(e - 1);
  wflag[e] = (uint64_t)start | ((uint64_t)pair << 32);
}

// ... and the work list from the scanned flags
__global__ void __launch_bounds__(kT) plan_work_kernel(const BwtBatch b, int64_t nE, unsigned cb, unsigned pb,
                                                       const uint64_t* ekey, const uint64_t* wflag,
                                                       const uint64_t* wscan, int64_t* dcounts) {
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (e >= nE) return;
  const uint64_t c = wscan[e];
  const_cast<uint32_t*>(b.epos)[BWT_EP_ID(b.eitems[e])] = (uint32_t)e;  // the endpoint's place in the work order
  if (wflag[e] & 1u) {
    const int64_t w = (int64_t)(uint32_t)c - 1;
    const_cast<int64_t*>(b.wk_chunk)[w] = (int64_t)((ekey[e] >> pb) & ((uint64_t(1) << cb) - 1));
    const_cast<int64_t*>(b.wk_e0)[w] = e;
  }
  if (e == nE - 1) {
    const int64_t nw = (int64_t)(uint32_t)c;
    const_cast<int64_t*>(b.wk_e0)[nw] = nE;
    dcounts[2] = nw;
    dcounts[3] = (int64_t)(c >> 32);
  }
}

// 5. staging: sort keys (query-id window, die part of the slot's tile); values slot | id << 32 ...
__global__ void __launch_bounds__(kT) plan_stage_keys_kernel(const BwtBatch b, int64_t M, int sh, int64_t per_t,
                                                             uint32_t* gkey, uint64_t* gval) {
  const int64_t j = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (j >= M) return;
  const uint32_t q = b.qorder[j];
  gkey[j] = (uint32_t)(((int64_t)(q >> sh)) * BWT_XCDS + (j / BWT_COMBINE_TILE) / per_t);
  gval[j] = (uint64_t)j | ((uint64_t)q << 32);
}

// ... position p of each slot -> ipos of its query; then sort the positions by slot tile ...
__global__ void __launch_bounds__(kT) plan_stage_pos_kernel(const BwtBatch b, int64_t M, const uint64_t* gval,
                                                            uint32_t* tkey, uint64_t* tval) {
  const int64_t p = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (p >= M) return;
  const uint64_t v = gval[p];
  const uint32_t j = (uint32_t)v;
  const_cast<uint32_t*>(b.ipos)[v >> 32] = (uint32_t)p;
  tkey[p] = j / BWT_COMBINE_TILE;
  tval[p] = (uint64_t)p | ((uint64_t)j << 32);
}

// ... every tile's positions now ascending: spos, and each slot's index in its tile's list
__global__ void __launch_bounds__(kT) plan_stage_rank_kernel(const BwtBatch b, int64_t M, const uint64_t* tval) {
  const int64_t o = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (o >= M) return;
  const uint64_t v = tval[o];
  const uint32_t j = (uint32_t)(v >> 32);
  const_cast<uint32_t*>(b.spos)[o] = (uint32_t)v;
  const_cast<uint16_t*>(b.srank)[j] = (uint16_t)(o - (int64_t)(j / BWT_COMBINE_TILE) * BWT_COMBINE_TILE);
}

unsigned grid_of(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

// every block listed (a pass that also writes the index replays every block): qblk[b] = b and
// qblk_g0[b] = the first step group whose step is in block b or later (grp_step ascending)
__global__ void __launch_bounds__(kT) plan_all_blocks_kernel(const BwtIndex x, const BwtBatch b, int64_t* dcounts) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i > x.nblk) return;
  const int64_t ngrp = dcounts[1];
  if (i == x.nblk) {
    const_cast<int64_t*>(b.qblk_g0)[i] = ngrp;
    dcounts[0] = x.nblk;
    return;
  }
  const_cast<int32_t*>(b.qblk)[i] = (int32_t)i;
  const_cast<int64_t*>(b.qblk_g0)[i] = lower_bound_i32(b.grp_step, 0, ngrp, x.blk_step0[i]);
}


int report(const char* what, hipError_t e) {
  if (e != hipSuccess) std::fprintf(stderr, "bwt_plan: %s: %s\n", what, hipGetErrorString(e));
  return (int)e;
}

}  // namespace

extern "C" int bwt_plan_workspace(const BwtIndex* x, const BwtPlanIndex* p, int64_t M, size_t* bytes) {
  Shape s;
  if (!bytes || !shape_of(x, p, M, s)) return (int)hipErrorInvalidValue;
  Layout L;
  const hipError_t e = layout_of(s, L);
  if (e) return report("workspace", e);
  *bytes = L.total;
  return 0;
}

extern "C" int bwt_plan_launch(const BwtIndex* x, const BwtPlanIndex* p, const int32_t* qstate, int64_t M,
                               BwtBatch* b, int64_t* dcounts, void* work, size_t work_bytes, void* stream) {
  Shape s;
  if (!b || !qstate || !dcounts || !work || !shape_of(x, p, M, s)) return (int)hipErrorInvalidValue;
  if (!b->qlo || !b->qhi || !b->qblk || !b->qblk_g0 || !b->grp_step || !b->grp_q0 || !b->qorder || !b->qrank ||
      !b->sblk || !b->qtile || !b->wk_chunk || !b->wk_e0 || !b->eitems || !b->epos || !b->spos || !b->srank ||
      !b->ipos)
    return (int)hipErrorInvalidValue;
  Layout L;
  hipError_t e = layout_of(s, L);
  if (e) return report("workspace", e);
  if (work_bytes < L.total) return (int)hipErrorInvalidValue;
  const hipStream_t st = (hipStream_t)stream;
  b->M = M;
  b->dcounts = dcounts;
  b->nqblk = (int32_t)std::min<int64_t>(x->nblk, M);
  b->ngrp = M;
  b->nwork = s.nE;
  const BwtBatch B = *b;
  void* temp = at<void>(work, L.temp);
  size_t tb = L.temp_bytes;
  uint64_t* skey_in = at<uint64_t>(work, L.skey_in);
  uint64_t* skey_out = at<uint64_t>(work, L.skey_out);
  uint64_t* sval_in = at<uint64_t>(work, L.sval_in);
  uint64_t* sval_out = at<uint64_t>(work, L.sval_out);
  uint64_t* flags = at<uint64_t>(work, L.flags);
  uint64_t* fscan = at<uint64_t>(work, L.fscan);
  uint64_t* ekey_in = at<uint64_t>(work, L.ekey_in);
  uint64_t* ekey_out = at<uint64_t>(work, L.ekey_out);
  uint64_t* eval_in = at<uint64_t>(work, L.eval_in);
  uint64_t* wflag = at<uint64_t>(work, L.wflag);
  uint64_t* wscan = at<uint64_t>(work, L.wscan);
  uint32_t* a0in = at<uint32_t>(work, L.a0in);
  uint32_t* a0out = at<uint32_t>(work, L.a0out);
  uint32_t* gkey_in = at<uint32_t>(work, L.gkey_in);
  uint32_t* gkey_out = at<uint32_t>(work, L.gkey_out);
  uint64_t* gval_in = at<uint64_t>(work, L.gval_in);
  uint64_t* gval_out = at<uint64_t>(work, L.gval_out);
  uint32_t* tkey_in = at<uint32_t>(work, L.tkey_in);
  uint32_t* tkey_out = at<uint32_t>(work, L.tkey_out);
  uint64_t* tval_in = at<uint64_t>(work, L.tval_in);
  uint64_t* tval_out = at<uint64_t>(work, L.tval_out);
  int rc;
  const dim3 blk(kT);
  hipLaunchKernelGGL(plan_init_kernel, dim3(1), dim3(64), 0, st, dcounts);
  // 1. slots
  hipLaunchKernelGGL(plan_slot_keys_kernel, dim3(grid_of(M)), blk, 0, st, *x, *p, qstate, B, M, skey_in, sval_in,
                     dcounts);
  if ((rc = report("slot keys", hipGetLastError()))) return rc;
  if ((rc = report("slot sort", rocprim::radix_sort_pairs<PlanSort>(temp, tb, skey_in, skey_out, sval_in, sval_out, size_t(M),
                                                          32u, 32u + s.sbits, st))))
    return rc;
  hipLaunchKernelGGL(plan_slot_info_kernel, dim3(grid_of(M)), blk, 0, st, *x, *p, B, M, skey_out, sval_out, flags);
  if ((rc = report("slot info", hipGetLastError()))) return rc;
  // 2. groups
  tb = L.temp_bytes;
  if ((rc = report("group scan",
                   rocprim::inclusive_scan(temp, tb, flags, fscan, size_t(M), rocprim::plus<uint64_t>(), st))))
    return rc;
  hipLaunchKernelGGL(plan_groups_kernel, dim3(grid_of(M)), blk, 0, st, B, M, skey_out, flags, fscan, dcounts);
  if ((rc = report("groups", hipGetLastError()))) return rc;
  // 3. endpoints
  hipLaunchKernelGGL(plan_endpoints_kernel, dim3(grid_of(s.nE)), blk, 0, st, *x, *p, B, s.nE, s.sr, s.CH, s.cb, s.pb,
                     sval_out, ekey_in, eval_in);
  if ((rc = report("endpoints", hipGetLastError()))) return rc;
  tb = L.temp_bytes;
  if ((rc = report("endpoint sort", rocprim::radix_sort_pairs<PlanSort>(temp, tb, ekey_in, ekey_out, eval_in,
                                                              const_cast<uint64_t*>(B.eitems), size_t(s.nE), 0u,
                                                              s.rb + s.cb + s.pb, st))))
    return rc;
  // 4. work items
  hipLaunchKernelGGL(plan_run_start_kernel, dim3(grid_of(s.nE)), blk, 0, st, s.nE, s.pb, ekey_out, a0in);
  if ((rc = report("run starts", hipGetLastError()))) return rc;
  tb = L.temp_bytes;
  if ((rc = report("run scan", rocprim::inclusive_scan(temp, tb, a0in, a0out, size_t(s.nE),
                                                       rocprim::maximum<uint32_t>(), st))))
    return rc;
  hipLaunchKernelGGL(plan_work_flags_kernel, dim3(grid_of(s.nE)), blk, 0, st, B, s.nE, s.G, a0out, wflag);
  if ((rc = report("work flags", hipGetLastError()))) return rc;
  tb = L.temp_bytes;
  if ((rc = report("work scan",
                   rocprim::inclusive_scan(temp, tb, wflag, wscan, size_t(s.nE), rocprim::plus<uint64_t>(), st))))
    return rc;
  hipLaunchKernelGGL(plan_work_kernel, dim3(grid_of(s.nE)), blk, 0, st, B, s.nE, s.cb, s.pb, ekey_out, wflag, wscan,
                     dcounts);
  if ((rc = report("work", hipGetLastError()))) return rc;
  // 5. staging
  hipLaunchKernelGGL(plan_stage_keys_kernel, dim3(grid_of(M)), blk, 0, st, B, M, s.sh, s.per_t, gkey_in, gval_in);
  if ((rc = report("stage keys", hipGetLastError()))) return rc;
  tb = L.temp_bytes;
  if ((rc = report("stage sort", rocprim::radix_sort_pairs<PlanSort>(temp, tb, gkey_in, gkey_out, gval_in, gval_out, size_t(M),
                                                           0u, s.gbits, st))))
    return rc;
  hipLaunchKernelGGL(plan_stage_pos_kernel, dim3(grid_of(M)), blk, 0, st, B, M, gval_out, tkey_in, tval_in);
  if ((rc = report("stage positions", hipGetLastError()))) return rc;
  tb = L.temp_bytes;
  if ((rc = report("tile sort", rocprim::radix_sort_pairs<PlanSort>(temp, tb, tkey_in, tkey_out, tval_in, tval_out, size_t(M),
                                                          0u, s.tbits, st))))
    return rc;
  hipLaunchKernelGGL(plan_stage_rank_kernel, dim3(grid_of(M)), blk, 0, st, B, M, tval_out);
  return report("tile ranks", hipGetLastError());
}

extern "C" int bwt_plan_all_blocks_launch(const BwtIndex* x, BwtBatch* b, void* stream) {
  if (!x || !b || x->nblk < 1 || !b->dcounts || !b->qblk || !b->qblk_g0 || !b->grp_step) return (int)hipErrorInvalidValue;
  // dcounts[1] (step groups) is read on the device: this launch follows bwt_plan_launch on the stream
  hipLaunchKernelGGL(plan_all_blocks_kernel, dim3(grid_of((int64_t)x->nblk + 1)), dim3(kT), 0, (hipStream_t)stream, *x,
                     *b, const_cast<int64_t*>(b->dcounts));
  const int rc = report("all blocks", hipGetLastError());
  if (rc == 0) b->nqblk = x->nblk;
  return rc;
}
