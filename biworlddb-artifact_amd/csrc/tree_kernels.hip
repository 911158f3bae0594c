// tree_kernels.hip -- CDNA4 (gfx950) kernels of the generalised batch path (bwgpu_tree.h).
//
// Stage map onto the reference's v7 algorithm (src/batch_engine_v7.hpp:832-968):
//   bwt_block_kernel        <- v7_block_kernel / run_blocks phase 1 (src/gpu_v7.cu:10-82): one
//                              workgroup replays one block from a zero local prefix over the
//                              block's own distinct keys, undo-aware (Euler steps).  Instead of
//                              a Fenwick tree (whose upper nodes serialise LDS atomics) every
//                              event is one ds_add_u64 into a dense per-rank delta array; at a
//                              query step a block scan folds the deltas into per-thread rank
//                              totals (registers) and publishes their exclusive prefix P, which
//                              the step's endpoints read once each.  Optionally the block's net
//                              weights (the dense dV, compacted); on the query side it also adds
//                              the level-C carry (the block's dense row at each endpoint's tile)
//   bwt_chunk_hist_kernel   <- dV per carry chunk (level C: one chunk per block)
//   bwt_seg_*_kernel        <- the serial carry pass (src/batch_engine_v7.hpp:936-945) as a
//                              three-pass segmented column scan over chunk rows
//   bwt_chunk_replay_kernel <- lvals on carry level B (src/gpu_v7.cu:72-78): one wavefront per
//                              work item; the chunk's key items and the batch's endpoint items
//                              are consumed in merged groups of 64; a Fenwick tree seeded from
//                              the chunk's carry row; pairs inside a group by LDS broadcast;
//                              results stored in work order
//   bwt_combine_*_kernel    <- gpu_combine_kernel (src/gpu_v7.cu:84-96): a slot pass gathers each
//                              slot's level-B results (die-local) and adds them to its partial
//                              result (plus the level-C carry when the block replay did not add
//                              it), a staged id pass stores the answers in query order
//   bwt_rank_kernel         <- CsrBase::rank (src/sparse_engine.hpp:62-97), CSR form
// All arithmetic is int64 adds (LDS ds_add_u64, DPP scans): results are exact and independent
// of scheduling order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <stdint.h>
#include <stdio.h>

#include "bwgpu_tree.h"
#include "wave64.hpp"

using namespace bw;

namespace {

constexpr int kQCap = 1024;  // query slots staged in LDS by the block replay
constexpr int kGCap = 64;    // query step groups staged in LDS by the block replay
constexpr int kEPT = 4;      // block replay: events per thread per staged tile of steps
// grid cap of the replay kernels when the counts are on the device: enough workgroups that the
// dispatcher, not a fixed share per workgroup, balances uneven blocks / work items; workgroups
// past the count leave at once
constexpr int64_t kDeviceCountGrid = int64_t(1) << 16;

// Padded LDS slot of in-block rank r when each thread owns RPT consecutive ranks: one spare
// word per run keeps a wavefront's strided run accesses on distinct banks.
template <int RPT>
__device__ __forceinline__ int pslot(int r) {
  return RPT > 1 ? r + r / RPT : r;
}

template <int RPT>
constexpr int block_span(int nt) {
  return (RPT > 1 ? nt * RPT + nt : nt * RPT) + 1;
}

// ---------------------------------------------------------------- block replay
// Events of step s: the U updates of state |steps[s]| with their precomputed in-block ranks.
// Steps are replayed in tiles of T = kEPT*nt/U steps whose events (rank, signed delta) are staged
// in LDS: while tile k replays, every thread holds its kEPT events of tile k+1 in flight (loaded
// with step ids fetched one tile earlier, every load unconditional with clamped indices) and the
// step ids of tile k+2; between tiles the arrived events go to the stage.  Query ranks and query
// step groups are staged in LDS too, so a query step touches no global memory before its answers
// are stored.
template <int RPT>
__device__ __forceinline__ void block_replay_one(const BwtIndex& x, const BwtBatch& q, const BwtLaunch& l,
                                                 const int64_t bi, unsigned char* smem_raw) {
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wv = tid >> 6;
  const int span = block_span<RPT>(nt);
  int64_t* w = reinterpret_cast<int64_t*>(smem_raw);          // deltas since the last scan
  int64_t* P = w + span;                                       // exclusive prefix over ranks
  int64_t* wtot = P + span;                                    // 16 wavefront totals
  int64_t* gq0 = wtot + 16;                                    // kGCap + 1 slot offsets
  int32_t* gstep = reinterpret_cast<int32_t*>(gq0 + kGCap + 1);  // kGCap steps
  uint32_t* qbuf = reinterpret_cast<uint32_t*>(gstep + kGCap);   // kQCap packed ranks
  int64_t* sdel = reinterpret_cast<int64_t*>(qbuf + kQCap);      // kEPT*nt staged signed deltas
  uint16_t* srank = reinterpret_cast<uint16_t*>(sdel + kEPT * nt);  // kEPT*nt staged ranks
  uint32_t* qtbuf = reinterpret_cast<uint32_t*>(srank + kEPT * nt);  // kQCap key tiles (carry_in_block)
  const bool add_carry = l.carry_in_block && !l.write_index;
  // the block's dense level-C row (level-C chunk b = block b); its entries stay in cache
  const int64_t* crow = add_carry ? x.cp + uniform64(x.chunk_cp0[q.qblk[bi]]) : nullptr;
  // key tiles of slot s's endpoints, lo | hi << 16
  auto tiles_of = [&](int64_t s) -> uint32_t {
    if (q.qtile) return q.qtile[s];
    const uint32_t id = q.qorder[s];
    return (q.qlo[id] >> x.tile_shift) | ((q.qhi[id] >> x.tile_shift) << 16);
  };
  // the block's fields are the same in every thread: marked uniform, they stay in scalar registers
  const int b = __builtin_amdgcn_readfirstlane(q.qblk[bi]);
  const int64_t k0 = uniform64(x.blk_key0[b]);
  const int R = (int)(uniform64(x.blk_key0[b + 1]) - k0);
  const int U = x.U;
  const int s0 = __builtin_amdgcn_readfirstlane(x.blk_step0[b]);
  const int64_t gi0 = uniform64(q.qblk_g0[bi]), ge = uniform64(q.qblk_g0[bi + 1]);
  // without the index write, nothing after the last queried step matters
  const int s1 = __builtin_amdgcn_readfirstlane(
      l.write_index ? x.blk_step0[b + 1] : (gi0 < ge ? q.grp_step[ge - 1] + 1 : s0));
  const int64_t qend = uniform64(q.grp_q0[ge]);
  auto stage_queries = [&](int64_t a) {
    for (int i = tid; i < kQCap; i += nt) {
      const bool in = a + i < qend;
      qbuf[i] = in ? q.qrank[a + i] : 0u;
      if (add_carry) qtbuf[i] = in ? tiles_of(a + i) : 0u;
    }
  };
  auto stage_groups = [&](int64_t g) {
    for (int i = tid; i <= kGCap; i += nt) {
      const int64_t gg = g + i;
      if (i < kGCap) gstep[i] = gg < ge ? q.grp_step[gg] : -1;
      gq0[i] = gg <= ge ? q.grp_q0[gg] : qend;
    }
  };
  for (int i = tid; i < span; i += nt) w[i] = 0;
  stage_groups(gi0);
  const int64_t qbeg = uniform64(q.grp_q0[gi0]);
  stage_queries(qbeg);
  int64_t tot[RPT];
#pragma unroll
  for (int j = 0; j < RPT; ++j) tot[j] = 0;
  int64_t gi = gi0, gwin = gi0, qwin = qbeg;
  __syncthreads();
  int next_q = __builtin_amdgcn_readfirstlane(gi < ge ? gstep[0] : -1);

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

  // answers of the query group gi (the state entered at the current step)
  auto query = [&]() {
    __syncthreads();  // every event of the steps so far is in w; previous readers are done
    const int64_t qa = uniform64(gq0[gi - gwin]), qb = uniform64(gq0[gi - gwin + 1]);
    const int64_t nq = qb - qa;
    const bool staged = nq <= kQCap;
    const bool fresh = staged && qb > qwin + kQCap;  // uniform
    if (fresh) {
      qwin = qa;
      stage_queries(qa);
    }
    // one endpoint per thread: even = lo, odd = hi of the same slot (nt is a multiple of 64,
    // so a pair never straddles a wavefront); kUnroll rounds at a time, so the carry-row loads
    // of all of them are in flight together.  When the group's tiles were staged by an earlier
    // step, the first round's carry-row loads (which do not depend on the replay) are issued
    // before the rebuild, whose barriers then hide them.  (A restaging step would have to load
    // its tiles from global memory first: that chain, waited on before the barriers, costs more
    // than it hides.)
    constexpr int kUnroll = 4;
    const bool pre = add_carry && staged && !fresh;  // uniform
    int64_t cpre[kUnroll];
    if (pre) {
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int64_t j = tid + (int64_t)u * nt;
        const bool live = j < 2 * nq;
        const uint32_t tl = live ? qtbuf[qa - qwin + (j >> 1)] : 0u;
        cpre[u] = live ? crow[(j & 1) ? (tl >> 16) : (tl & 0xFFFFu)] : 0;
      }
    }
    rebuild();  // its barriers also publish qbuf
    const int64_t jn = (2 * nq + 63) & ~(int64_t)63;
    for (int64_t j0 = tid; j0 < jn; j0 += (int64_t)kUnroll * nt) {
      int64_t pv[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int64_t j = j0 + (int64_t)u * nt;
        const int64_t sl = j >> 1;
        const bool live = j < 2 * nq;
        uint32_t rk = 0;
        if (live) rk = staged ? qbuf[qa - qwin + sl] : q.qrank[qa + sl];
        const int y = (j & 1) ? (int)(rk >> 16) : (int)(rk & 0xFFFFu);
        pv[u] = live ? P[pslot<RPT>(y)] : 0;
        if (add_carry && live) {  // plus the carry of the earlier blocks' keys in lower tiles
          if (pre && j0 == tid) {
            pv[u] += cpre[u];
          } else {
            const uint32_t tl = staged ? qtbuf[qa - qwin + sl] : tiles_of(qa + sl);
            pv[u] += crow[(j & 1) ? (tl >> 16) : (tl & 0xFFFFu)];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int64_t j = j0 + (int64_t)u * nt;
        const int64_t other = __shfl_xor(pv[u], 1);
        if (j < 2 * nq && !(j & 1)) q.ansr[qa + (j >> 1)] = other - pv[u];
      }
    }
    ++gi;
    if (gi < ge && gi - gwin >= kGCap) {  // uniform
      __syncthreads();
      gwin = gi;
      stage_groups(gi);
      __syncthreads();
    }
    next_q = __builtin_amdgcn_readfirstlane(gi < ge ? gstep[gi - gwin] : -1);
  };

  if (U <= kEPT * nt) {
    const int T = (kEPT * nt) / U;  // steps per tile
    const int TE = T * U;           // events per tile
    const int64_t nu = x.S * (int64_t)U;
    // events of the tile at ts, given its step ids: ranks and raw deltas are loads in flight,
    // the signs are taken from the step ids at once (so the ids registers are free again)
    auto events = [&](int ts, const int* st, int* r, int64_t* d, bool* neg) {
#pragma unroll
      for (int k = 0; k < kEPT; ++k) {
        const int e = tid + k * nt;
        const int s = ts + e / U, i = e - (e / U) * U;
        const bool ok = e < TE && s < s1;
        const int64_t v = st[k] > 0 ? st[k] : -st[k];
        r[k] = x.evrank[ok ? (int64_t)s * U + i : 0];
        const int64_t at = (v - 1) * (int64_t)U + i;
        d[k] = x.delta[(ok && at >= 0 && at < nu) ? at : 0];
        neg[k] = st[k] < 0;
      }
    };
    // step id of this thread's k-th event slot in the tile starting at step ts (clamped)
    auto ids = [&](int ts, int* st) {
#pragma unroll
      for (int k = 0; k < kEPT; ++k) {
        const int e = tid + k * nt;
        const int s = ts + e / U;
        st[k] = x.steps[(e < TE && s < s1) ? s : s0];
      }
    };
    auto stage_events = [&](int ts, const int* r, const int64_t* d, const bool* neg) {
#pragma unroll
      for (int k = 0; k < kEPT; ++k) {
        const int e = tid + k * nt;
        if (e < TE) {
          const bool ok = ts + e / U < s1;
          srank[e] = (uint16_t)(ok ? r[k] : 0);
          sdel[e] = ok ? (neg[k] ? -d[k] : d[k]) : 0;
        }
      }
    };
    int st[kEPT], r_n[kEPT];
    int64_t d_n[kEPT];
    bool neg_n[kEPT];
    // prologue: tile 0 staged, tile 1's events in flight, tile 2's step ids in flight
    ids(s0, st);
    events(s0, st, r_n, d_n, neg_n);
    stage_events(s0, r_n, d_n, neg_n);
    ids(s0 + T, st);
    events(s0 + T, st, r_n, d_n, neg_n);
    ids(s0 + 2 * T, st);
    __syncthreads();
    for (int t0 = s0; t0 < s1; t0 += T) {
      const int t1 = min(t0 + T, s1);
      // the staged events of every step up to the next query step (or the tile's end) in one
      // pass: the adds commute, and only a query needs the steps before it complete
      for (int s = t0; s < t1;) {
        const int se = (next_q >= s && next_q < t1) ? next_q + 1 : t1;
        const int base = (s - t0) * U, nev = (se - s) * U;
        for (int i = tid; i < nev; i += nt) lds_add64(&w[pslot<RPT>(srank[base + i])], sdel[base + i]);
        s = se;
        if (se - 1 == next_q) query();
      }
      if (t1 >= s1) break;
      __syncthreads();  // the stage is free
      stage_events(t0 + T, r_n, d_n, neg_n);
      __syncthreads();
      events(t0 + 2 * T, st, r_n, d_n, neg_n);
      ids(t0 + 3 * T, st);
    }
  } else {
    for (int s = s0; s < s1; ++s) {
      const int st = x.steps[s];
      const int64_t v = st > 0 ? st : -st;
      for (int i = tid; i < U; i += nt) {
        const int r = x.evrank[(int64_t)s * U + i];
        const int64_t dv = x.delta[(v - 1) * (int64_t)U + i];
        lds_add64(&w[pslot<RPT>(r)], st > 0 ? dv : -dv);
      }
      if (s == next_q) query();
    }
  }
  if (!l.write_index) return;
  __syncthreads();
  rebuild();
  // net weight of the block per distinct key (carry level B) and per key tile (level C)
  for (int r = tid; r < R; r += nt) x.agg[k0 + r] = P[pslot<RPT>(r + 1)] - P[pslot<RPT>(r)];
  const int64_t t0 = x.blk_tile0[b];
  const int T = (int)(x.blk_tile0[b + 1] - t0);
  for (int s = tid; s < T; s += nt) {
    const int r0 = x.tile_rank0[t0 + s];
    const int r1 = (s + 1 < T) ? (int)x.tile_rank0[t0 + s + 1] : R;
    x.tagg[t0 + s] = P[pslot<RPT>(r1)] - P[pslot<RPT>(r0)];
  }
}

// One workgroup per listed block; with device counts (device planner) the grid is bounded and
// each workgroup takes every gridDim-th listed block.
template <int RPT>
__global__ void __launch_bounds__(1024) bwt_block_kernel(const BwtIndex x, const BwtBatch q, const BwtLaunch l) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int64_t n = q.dcounts ? q.dcounts[0] : (int64_t)q.nqblk;
  for (int64_t bi = blockIdx.x; bi < n; bi += gridDim.x) {
    block_replay_one<RPT>(x, q, l, bi, smem_raw);
    __syncthreads();  // the next block reuses the LDS
  }
}

// ---------------------------------------------------------------- chunk histogram
constexpr int kHistBatch = 8;  // key items per lane whose gathers are in flight at once
// One wavefront per chunk: dense class histogram of the chunk's key items, written as its
// exclusive prefix over classes (cp row = sum of weights with class < y).  With write_kw the
// gathered weight of every item is also stored in chain order (kweight), so the chunk replay
// streams weights instead of gathering them.
__global__ void __launch_bounds__(256) bwt_chunk_hist_kernel(const BwtIndex x, const BwtLaunch l, int64_t c_begin,
                                                             int64_t c_end, int stride, int write_kw) {
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
  // kHistBatch items per lane at a time: all their loads, then all their weight gathers, are
  // issued before any is used (offsets past the chunk are clamped to its last item, their
  // values ignored), so a wavefront has that many random gathers in flight instead of one
  const int64_t i0 = x.chunk_key0[c], i1 = x.chunk_key0[c + 1];
  for (int64_t base = i0 + lane; base < i1; base += 64 * kHistBatch) {
    uint64_t item[kHistBatch];
    int64_t wt[kHistBatch];
#pragma unroll
    for (int k = 0; k < kHistBatch; ++k) item[k] = x.kitems[min(base + 64 * k, i1 - 1)];
#pragma unroll
    for (int k = 0; k < kHistBatch; ++k) wt[k] = wsrc[item[k] & BWT_REF_MASK];
#pragma unroll
    for (int k = 0; k < kHistBatch; ++k) {
      const int64_t it = base + 64 * k;
      if (it < i1) {
        if (write_kw) x.kweight[it] = wt[k];
        lds_add64(&h[(int)BWT_KEY_CLASS(item[k])], wt[k]);
      }
    }
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

// Level C, one workgroup per block: the block's tile slots are its distinct key tiles in
// ascending order (one key item each, weight tagg[slot]), so its row is a step function over the
// tiles: row[y] = sum of the weights of its slots with tile < y.  The slots' tiles and inclusive
// weight prefix go to LDS; each row entry is then one short search and one coalesced store.  (The
// per-wavefront histogram below would scan the whole row in one wavefront.)
constexpr int kRowCThreads = 256;
__global__ void __launch_bounds__(kRowCThreads) bwt_row_c_kernel(const BwtIndex x, int ncap) {
  // dynamic LDS only (the launch raises its limit to the whole 160 KiB): wave totals, then
  // ncap inclusive prefix values, then ncap tiles
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int64_t* wtot = reinterpret_cast<int64_t*>(smem_raw);
  int64_t* pre = wtot + kRowCThreads / 64;
  uint32_t* tl = reinterpret_cast<uint32_t*>(pre + ncap);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t c = blockIdx.x;  // level-C chunk c = block c
  const int64_t i0 = x.chunk_key0[c];
  // a block has at most rmax distinct keys, so at most ncap = rmax tile slots
  const int n = min((int)(x.chunk_key0[c + 1] - i0), ncap);
  const int ncls = x.chain_ncls[0];
  const int64_t* wsrc = x.chain_src[0] ? x.tagg : x.agg;
  // per thread a run of consecutive slots: its sum, a block scan of the sums, then the prefix
  const int per = (n + kRowCThreads - 1) / kRowCThreads;
  const int a = min(tid * per, n), b = min(a + per, n);
  int64_t sum = 0;
  for (int i = a; i < b; ++i) {
    const uint64_t item = x.kitems[i0 + i];
    tl[i] = BWT_KEY_CLASS(item);
    sum += wsrc[item & BWT_REF_MASK];
    pre[i] = sum;
  }
  const int64_t inc = wave_incl_scan(sum);
  if (lane == 63) wtot[wv] = inc;
  __syncthreads();
  int64_t off = inc - sum;
  for (int k = 0; k < wv; ++k) off += wtot[k];
  for (int i = a; i < b; ++i) pre[i] += off;
  __syncthreads();
  int64_t* row = x.cp + x.chunk_cp0[c];
  for (int y = tid; y < ncls; y += kRowCThreads) {
    int lo = 0, hi = n;  // slots with tile < y
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (tl[mid] < (uint32_t)y) lo = mid + 1; else hi = mid;
    }
    row[y] = lo > 0 ? pre[lo - 1] : 0;
  }
}

// ---------------------------------------------------------------- segmented carry scan
constexpr int kSegBatch = 8;  // rows whose loads a thread has in flight at once
// cp rows of one chain form a matrix [chunks][ncls]; the carry of chunk r is the exclusive
// column sum over rows < r.  Pass A: segment sums; pass B: scan over segments; pass C: rows.
__global__ void __launch_bounds__(256) bwt_seg_sum_kernel(const BwtIndex x) {
  const int64_t s = blockIdx.x;
  const int y = blockIdx.y * blockDim.x + threadIdx.x;
  const int64_t c0 = x.seg_chunk0[s], c1 = x.seg_chunk0[s + 1];
  const int ncls = x.chain_ncls[x.chunk_chain[c0]];
  if (y >= ncls) return;
  // kSegBatch rows' loads in flight at a time (their offsets are uniform: scalar loads)
  int64_t acc = 0;
  int64_t c = c0;
  for (; c + kSegBatch <= c1; c += kSegBatch) {
    int64_t v[kSegBatch];
#pragma unroll
    for (int k = 0; k < kSegBatch; ++k) v[k] = x.cp[x.chunk_cp0[c + k] + y];
#pragma unroll
    for (int k = 0; k < kSegBatch; ++k) acc += v[k];
  }
  for (; c < c1; ++c) acc += x.cp[x.chunk_cp0[c] + y];
  x.segsum[x.seg_cp0[s] + y] = acc;
}

__global__ void __launch_bounds__(256) bwt_seg_scan_kernel(const BwtIndex x) {
  const int chain = blockIdx.x;
  const int y = blockIdx.y * blockDim.x + threadIdx.x;
  if (y >= x.chain_ncls[chain]) return;
  int64_t run = 0;
  const int64_t e = x.chain_seg0[chain + 1];
  int64_t s = x.chain_seg0[chain];
  for (; s + kSegBatch <= e; s += kSegBatch) {  // the batch's loads before its stores
    int64_t v[kSegBatch];
#pragma unroll
    for (int k = 0; k < kSegBatch; ++k) v[k] = x.segsum[x.seg_cp0[s + k] + y];
#pragma unroll
    for (int k = 0; k < kSegBatch; ++k) {
      x.segsum[x.seg_cp0[s + k] + y] = run;
      run += v[k];
    }
  }
  for (; s < e; ++s) {
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
  int64_t c = c0;
  for (; c + kSegBatch <= c1; c += kSegBatch) {  // the batch's loads before its stores
    int64_t v[kSegBatch];
#pragma unroll
    for (int k = 0; k < kSegBatch; ++k) v[k] = x.cp[x.chunk_cp0[c + k] + y];
#pragma unroll
    for (int k = 0; k < kSegBatch; ++k) {
      x.cp[x.chunk_cp0[c + k] + y] = run;
      run += v[k];
    }
  }
  for (; c < c1; ++c) {
    int64_t* p = x.cp + x.chunk_cp0[c] + y;
    const int64_t t = *p;
    *p = run;
    run += t;
  }
}

// ---------------------------------------------------------------- chunk replay (level B)
// Per-wavefront LDS of the chunk replay: 64 group key weights, 64 group key classes, the Fenwick
// tree (its node slots rounded up to whole runs of 32, see fen_slot); every wavefront's area is a
// multiple of 16 bytes.
__host__ __device__ constexpr int replay_wave_words(int stride) { return 96 + ((stride + 31) & ~31); }

// LDS slot of Fenwick node f: f with its low 5 bits xor-ed with bits 5-9, a permutation of every
// aligned run of 32 nodes.  The nodes one update touches at level t are odd multiples of 2^t;
// stored in place, every level t >= 4 falls on one bank pair (64-bit words), so the lanes of an
// update pass (several keys x every level) and late prefix steps serialise on it.  Levels 5-9
// then spread over 32 slots of their runs (levels >= 10 have a few nodes each).
__device__ __forceinline__ int fen_slot(int f) { return f ^ ((f >> 5) & 31); }

// Fenwick prefix over classes < y for y < 2^kBits (at most kBits set bits); node 0 stays 0, so
// the fixed trip count needs no branch and the LDS reads issue back to back.
template <int kBits>
__device__ __forceinline__ int64_t fen_prefix(const int64_t* fen, int y) {
  int64_t r = 0;
#pragma unroll
  for (int k = 0; k < kBits; ++k) {
    r += fen[fen_slot(y)];
    y &= y - 1;
  }
  return r;
}

// The value lane `src` holds (ds_bpermute: every lane active, src in [0, 64)).
__device__ __forceinline__ int lane_read(int v, int src) { return __builtin_amdgcn_ds_bpermute(src << 2, v); }
__device__ __forceinline__ int64_t lane_read64(int64_t v, int src) {
  const uint64_t u = (uint64_t)v;
  const uint32_t lo = (uint32_t)lane_read((int)(uint32_t)u, src);
  const uint32_t hi = (uint32_t)lane_read((int)(uint32_t)(u >> 32), src);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// One wavefront per work item (a chunk and a run of its endpoints, sorted by position).  The
// Fenwick tree starts as the chunk's carry row (fen[i] = row[i] - row[i - lsb(i)]), so a prefix
// read is the whole carry.  Each group takes the next 64 items of the merged order (endpoint
// items before key items at equal position): the group's keys go to LDS, every endpoint adds
// the Fenwick prefix plus the group's keys that precede it with a smaller class (summed by a
// segment of lanes per endpoint), then the keys enter the tree.  The next group's items are loaded while this one runs
// (only the high word of a key item, which holds its class).  Results are stored in work order
// (g[i] for eitems[i]): a group's taken lanes write one contiguous run.  Classes < 2^kBits.
template <int kBits>
__device__ __forceinline__ void chunk_replay_one(const BwtIndex& x, const BwtBatch& q, const int stride,
                                                 const int64_t ci, int64_t* kws) {
  const int lane = threadIdx.x & 63;
  // the work item's fields are the same in every lane: marked uniform, they and the group
  // cursors below stay in scalar registers
  const int64_t c = uniform64(q.wk_chunk[ci]);
  int32_t* kcs = reinterpret_cast<int32_t*>(kws + 64);  // 64 group key classes
  int64_t* fen = kws + 96;
  const int ncls = __builtin_amdgcn_readfirstlane(x.chain_ncls[x.chunk_chain[c]]);
  const int64_t* row = x.cp + x.chunk_cp0[c];
  for (int i = lane; i < stride; i += 64) {
    int64_t v = 0;
    if (i >= 1 && i < ncls) v = row[i] - row[i - (i & -i)];
    fen[fen_slot(i)] = v;
  }
  // the work item's endpoints e0 .. e0+ne_all-1 and keys k0 .. k0+nk_all-1 (keys after its last
  // endpoint are not needed); the cursors below are 32-bit offsets from them
  const int64_t e0 = uniform64(q.wk_e0[ci]);
  const int ne_all = (int)(uniform64(q.wk_e0[ci + 1]) - e0);
  const int64_t k0 = uniform64(x.chunk_key0[c]);
  const int nk_all = (int)(uniform64(min(x.chunk_key0[c + 1], k0 + (int64_t)BWT_EP_POS(q.eitems[e0 + ne_all - 1]))) - k0);
  const uint64_t* eit = q.eitems + e0;
  // offset 0 must stay readable when the work item needs no key (its chunk may end the array)
  const int64_t kb = nk_all > 0 ? k0 : 0;
  const uint32_t* khi = reinterpret_cast<const uint32_t*>(x.kitems + kb) + 1;  // little endian
  const int64_t* kwt = x.kweight + kb;
  int64_t* gout = q.g + e0;
  // the items of the group starting at (e, k): every load is issued (an offset past the work
  // item's range is clamped to a valid one and the value ignored), so the number of loads in
  // flight is the same on every path and a use waits only for its own group's loads
  auto load_group = [&](int e, int k, uint64_t& it_, uint32_t& kx_, int64_t& kw_) {
    // unsigned 32-bit offsets from uniform bases: scalar-base loads, no 64-bit address math
    const uint32_t ei = (uint32_t)min(e + lane, ne_all - 1);
    const uint32_t kk = k + lane < nk_all ? (uint32_t)(k + lane) : 0u;
    it_ = eit[ei];
    kx_ = khi[2u * kk];
    kw_ = kwt[kk];
  };
  int ep = 0, ki = 0;
  // the first Fenwick level of key lane `lane` in the tree update below
  constexpr int kLevels = kBits + 1;
  const int kLev0 = lane % kLevels;
  // One group of 64 merged items whose items are in (it, kx, kw); the next group's items are
  // loaded into (it2, kx2, kw2) as soon as its start is known.  Returns false after the last group.
  auto group = [&](const uint64_t it, const uint32_t kx, const int64_t kw, uint64_t& it2, uint32_t& kx2,
                   int64_t& kw2) -> bool {
    const bool ev = ep + lane < ne_all;
    // keys of this chunk still ahead of the endpoint (>= 0: earlier ones were consumed)
    const int rel = ev ? (int)BWT_EP_POS(it) - ki : 0;
    const bool take = ev && lane + rel < 64;  // its merged index falls in this group
    const int ne = __popcll(ballot64(take));  // positions are sorted: a prefix of the lanes
    const int nk = min(64 - ne, nk_all - ki);
    const int ep2 = __builtin_amdgcn_readfirstlane(ep + ne), ki2 = __builtin_amdgcn_readfirstlane(ki + nk);
    load_group(ep2, ki2, it2, kx2, kw2);
    const int kcls = (int)((kx >> (BWT_CLASS_SHIFT - 32)) & BWT_CLASS_MASK);
    if (lane < nk) {
      kcs[lane] = kcls;
      kws[lane] = kw;
    }
    wave_lds_sync();
    if (ne > 0) {
      const int cls = (int)BWT_EP_CLASS(it);
      // a taken endpoint needs the group's first `rel` keys (rel <= nk); 0 for the other lanes
      const int rl = take ? rel : 0;
      int64_t val = fen_prefix<kBits>(fen, take ? cls : 0);
      if (ne > 32) {
        // Endpoint-heavy group (at most 31 keys): every endpoint lane walks the keys before it,
        // four per 16-byte broadcast read (entries past nk are excluded by m < rl <= nk)
        const int mmax = __builtin_amdgcn_readlane(rel, ne - 1);  // rel grows with the lane
        for (int m = 0; m < mmax; m += 4) {
          const int4 c4 = *reinterpret_cast<const int4*>(kcs + m);
          const longlong2 w01 = *reinterpret_cast<const longlong2*>(kws + m);
          const longlong2 w23 = *reinterpret_cast<const longlong2*>(kws + m + 2);
          if (m < rl && c4.x < cls) val += w01.x;
          if (m + 1 < rl && c4.y < cls) val += w01.y;
          if (m + 2 < rl && c4.z < cls) val += w23.x;
          if (m + 3 < rl && c4.w < cls) val += w23.y;
        }
      } else {
        // The group's keys before endpoint j with a smaller class, summed by the lanes: the
        // wavefront splits into segments of L = 64 / 2^ceil(log2 ne) >= 2 lanes, segment j serving
        // endpoint j, its lane s taking the key pairs (2s, 2s + 1), (2s + 2L, 2s + 2L + 1), ...
        // below rel_j (one 8-byte and one 16-byte read per pair, consecutive within a segment;
        // segments past ne have rel 0).  rel_j <= 63, so a pair's second entry is inside the 64
        // slots (a slot at or past rel_j may hold an earlier group's key and is excluded).  One
        // wavefront scan then gives every segment's total as the difference of two of its
        // inclusive sums.
        const int lg = 6 - (ne > 1 ? 32 - __builtin_clz((unsigned)(ne - 1)) : 0);  // log2 L, uniform
        const int L = 1 << lg;
        const int sj = lane >> lg;
        const int rj = lane_read(rl, sj), cj = lane_read(cls, sj);
        int64_t part = 0;
        for (int m = 2 * (lane & (L - 1)); m < rj; m += 2 * L) {
          const int2 c2 = *reinterpret_cast<const int2*>(kcs + m);
          const longlong2 w2 = *reinterpret_cast<const longlong2*>(kws + m);
          // selects, not branches: both weights are read by one access whatever the classes
          part += (c2.x < cj ? w2.x : 0) + (((m + 1 < rj) & (c2.y < cj)) ? w2.y : 0);
        }
        const int64_t inc = wave_incl_scan(part);
        // endpoint lane e (< ne): its segment spans lanes [e L, e L + L), inside the wavefront
        const int first = (lane << lg) & 63;
        const int64_t seg_end = lane_read64(inc, (first + L - 1) & 63);
        const int64_t seg_before = lane_read64(inc, (first + 63) & 63);  // lane first - 1
        val += seg_end - (first == 0 ? 0 : seg_before);
      }
      // results in work order: the taken lanes store one contiguous run
      if (take) gout[(uint32_t)(ep + lane)] = val;
    }
    if (ep2 >= ne_all) return false;
    // the group's keys enter the tree: the update path of class c is (c | (2^t - 1)) + 1 for every
    // level t with bit t of c clear (up to ncls).  Key lane m (its class and weight in registers)
    // walks the levels from m mod kLevels on, so the lanes of one atomic spread over the levels
    // (the top levels have a few nodes each: all keys at one level would serialise on them)
    if (lane < nk) {
      int lev = kLev0;
#pragma unroll
      for (int t = 0; t < kLevels; ++t) {
        const int f = (kcls | ((1 << lev) - 1)) + 1;
        if (!((kcls >> lev) & 1) && f <= ncls) lds_add64(&fen[fen_slot(f)], kw);
        lev = lev == kLevels - 1 ? 0 : lev + 1;
      }
    }
    wave_lds_sync();
    ep = ep2;
    ki = ki2;
    return true;
  };
  // two register sets, alternating, so no loaded value is copied before it is used
  uint64_t itA, itB;
  uint32_t kxA, kxB;
  int64_t kwA, kwB;
  load_group(ep, ki, itA, kxA, kwA);
  wave_lds_sync();
  while (group(itA, kxA, kwA, itB, kxB, kwB) && group(itB, kxB, kwB, itA, kxA, kwA)) {
  }
}

// One wavefront per work item; with device counts (device planner) the grid is bounded and each
// wavefront takes every (gridDim * chunk_waves)-th work item.
template <int kBits>
__global__ void __launch_bounds__(256) bwt_chunk_replay_kernel(const BwtIndex x, const BwtBatch q, const BwtLaunch l,
                                                               int stride) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // the wavefront's index, marked uniform: the work item and everything derived from it (its
  // chunk, cursors and bounds) then live in scalar registers and load through scalar loads
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int64_t* kws = reinterpret_cast<int64_t*>(smem_raw) + (int64_t)wv * replay_wave_words(stride);
  const int64_t n = q.dcounts ? q.dcounts[2] : q.nwork;
  const int64_t step = (int64_t)gridDim.x * l.chunk_waves;
  for (int64_t ci = (int64_t)blockIdx.x * l.chunk_waves + wv; ci < n; ci += step) {
    chunk_replay_one<kBits>(x, q, stride, ci, kws);
    wave_lds_sync();  // the next work item reuses this wavefront's LDS
  }
}

// ---------------------------------------------------------------- combine
// Two passes.  Slots (block order) add the level-C carry, read from the dense carry row of the
// slot's block (level-C chunk b = block b) at the endpoints' key tiles, and the level-B carries,
// which the chunk replay stored in work order (the results of a die's slots lie in that die's
// ranges of it), to the in-block part, and store the value at the slot's staging position.  The staging buffer is
// grouped by windows of query ids (then by die part, then slot), so these stores fill each
// window's lines in order.  Then query ids, each die taking its part, load their value from
// their window (a few MiB, resident in that die's cache) and store the answers in order.
// XCD-major block of this workgroup (see BWT_XCDS): grid = BWT_XCDS * per workgroups
__device__ __forceinline__ int64_t xcd_major_block(int64_t per) {
  return (int64_t)(blockIdx.x % BWT_XCDS) * per + blockIdx.x / BWT_XCDS;
}

// One tile of BWT_COMBINE_TILE slots per workgroup: the finished values are put in LDS in the
// order of their staging positions, then stored in that order, so the lanes of a store
// instruction write runs of consecutive positions (one run per query-id window).  With
// kCarry the level-C carry is added here (read from the block's dense row); otherwise the block
// replay already added it.
template <bool kCarry>
__global__ void __launch_bounds__(BWT_COMBINE_THREADS) bwt_combine_slots_kernel(const BwtIndex x, const BwtBatch q,
                                                                                int64_t per) {
  constexpr int kPer = BWT_COMBINE_TILE / BWT_COMBINE_THREADS;  // slots per thread
  __shared__ int64_t sv[BWT_COMBINE_TILE];
  const int64_t t0 = xcd_major_block(per) * BWT_COMBINE_TILE;
  if (t0 >= q.M) return;  // the whole workgroup
  // every load of the thread's slots is issued before any is used (clamped, unconditional)
  const int64_t last = q.M - 1;
  int32_t blk[kPer];
  uint32_t tl[kPer];
  int64_t part[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int64_t s = min(t0 + k * BWT_COMBINE_THREADS + (int64_t)threadIdx.x, last);
    if (kCarry) {
      blk[k] = q.sblk[s];
      if (q.qtile) {
        tl[k] = q.qtile[s];
      } else {  // tiles of the endpoint coordinates (e.g. ranks the device just remapped)
        const uint32_t id = q.qorder[s];
        tl[k] = (q.qlo[id] >> x.tile_shift) | ((q.qhi[id] >> x.tile_shift) << 16);
      }
    }
    // level B: the slot's two results, stored in work order (one die's range of it)
    // epos and ansr are read once: non-temporal, so they do not evict the gathered g lines
    const uint64_t epp = __builtin_nontemporal_load(reinterpret_cast<const uint64_t*>(q.epos) + s);
    part[k] = __builtin_nontemporal_load(q.ansr + s) + (q.g[(uint32_t)(epp >> 32)] - q.g[(uint32_t)epp]);
  }
  if (kCarry) {
    int64_t r0[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) r0[k] = x.chunk_cp0[blk[k]];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int64_t* row = x.cp + r0[k];
      part[k] += row[tl[k] >> 16] - row[tl[k] & 0xFFFFu];
    }
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int64_t s = t0 + k * BWT_COMBINE_THREADS + threadIdx.x;
    if (s <= last) sv[__builtin_nontemporal_load(q.srank + s)] = part[k];
  }
  __syncthreads();
  for (int k = threadIdx.x; k < BWT_COMBINE_TILE; k += BWT_COMBINE_THREADS) {
    const int64_t p = t0 + k;
    if (p >= q.M) break;
    q.stage[__builtin_nontemporal_load(q.spos + p)] = sv[k];
  }
}

// query ids per thread of the id pass: all their loads are issued before any is used
constexpr int kIdsPer = 8;

__global__ void __launch_bounds__(BWT_COMBINE_THREADS) bwt_combine_ids_kernel(const BwtBatch q, int64_t per) {
  const int64_t i0 = xcd_major_block(per) * (BWT_COMBINE_THREADS * kIdsPer) + threadIdx.x;
  const int64_t last = q.M - 1;
  uint32_t p[kIdsPer];
#pragma unroll
  for (int k = 0; k < kIdsPer; ++k) p[k] = __builtin_nontemporal_load(q.ipos + min(i0 + k * BWT_COMBINE_THREADS, last));
  int64_t v[kIdsPer];
#pragma unroll
  for (int k = 0; k < kIdsPer; ++k) v[k] = q.stage[p[k]];
#pragma unroll
  for (int k = 0; k < kIdsPer; ++k)
    if (i0 + k * BWT_COMBINE_THREADS <= last) __builtin_nontemporal_store(v[k], q.answers + i0 + k * BWT_COMBINE_THREADS);
}

// ---------------------------------------------------------------- overlay (appended states)
// One wavefront per overlay query: walk from the appended state up to its first indexed
// ancestor, every lane summing its share of each appended state's updates in [lo, hi).
__global__ void __launch_bounds__(256) bwt_overlay_kernel(const BwtOverlay o) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + wv;
  if (i >= o.n) return;
  const uint32_t id = o.qid[i];
  const uint32_t lo = o.qlo[id], hi = o.qhi[id];
  int64_t acc = 0;
  for (int64_t v = o.qstate[i]; v > o.S; v = o.app_parent[v - o.S - 1]) {
    const int64_t base = (v - o.S - 1) * (int64_t)o.U;
    for (int u = lane; u < o.U; u += 64) {
      const uint32_t k = o.app_key[base + u];
      if (k >= lo && k < hi) acc += o.app_delta[base + u];
    }
  }
  const int64_t tot = readlane64(wave_incl_scan(acc), 63);
  if (lane == 0) o.answers[id] += tot;
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

// The same remap with the LUT's row offsets staged in LDS: each workgroup stages the 65,537 row
// offsets as a 32-bit base per 64 rows plus a 16-bit offset per row (132 KiB), then takes
// endpoints a grid apart, so a lookup reads its row's bounds from LDS and only its row's keys
// (one or two 128-byte lines) from global memory.  A workgroup whose rows do not fit 16-bit
// offsets (more than 65,535 keys under 64 consecutive rows) reads the bounds from global memory.
constexpr int kRankLdsThreads = 1024;
constexpr int64_t kRankLdsMin = int64_t(1) << 20;  // endpoints from which the staged form is used
constexpr int kRankRows = 65536, kRankGroup = 64;
constexpr size_t kRankLdsBytes = 16 + (kRankRows / kRankGroup + 1) * 4 + (kRankRows + 1) * 2;
__global__ void __launch_bounds__(kRankLdsThreads) bwt_rank_lds_kernel(const BwtRankArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int* overflow = reinterpret_cast<int*>(smem_raw);
  uint32_t* base = reinterpret_cast<uint32_t*>(smem_raw + 16);
  uint16_t* off16 = reinterpret_cast<uint16_t*>(base + kRankRows / kRankGroup + 1);
  const int tid = threadIdx.x;
  if (tid == 0) overflow[0] = 0;
  for (int g = tid; g <= kRankRows / kRankGroup; g += kRankLdsThreads) base[g] = a.row_ptr[g * kRankGroup];
  __syncthreads();
  for (int h = tid; h <= kRankRows; h += kRankLdsThreads) {
    const uint32_t d = a.row_ptr[h] - base[h / kRankGroup];
    if (d > 0xFFFFu) atomicOr(overflow, 1);
    off16[h] = (uint16_t)d;
  }
  __syncthreads();
  const bool staged = overflow[0] == 0;  // uniform
  for (int64_t i = (int64_t)blockIdx.x * kRankLdsThreads + tid; i < a.n; i += (int64_t)gridDim.x * kRankLdsThreads) {
    const uint32_t v = a.in[i];
    const uint32_t h = v >> 16;
    const uint16_t lo16 = (uint16_t)(v & 0xFFFFu);
    uint32_t lo, hi;
    if (staged) {
      lo = base[h / kRankGroup] + off16[h];
      hi = base[(h + 1) / kRankGroup] + off16[h + 1];
    } else {
      lo = a.row_ptr[h];
      hi = a.row_ptr[h + 1];
    }
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (a.low16[mid] < lo16) lo = mid + 1; else hi = mid;
    }
    a.out[i] = lo;
  }
}

// ---------------------------------------------------------------- launch helpers
// ranks per thread of the block replay: the smallest power of two with nt * RPT >= rmax
int block_rpt(const BwtIndex* x, int nt) {
  const int need = (x->rmax + nt - 1) / nt;
  int r = 1;
  while (r < need) r <<= 1;
  return r;
}

// with_tiles: the query tiles are staged too (the replay adds the level-C carry)
size_t block_lds_bytes(int rpt, int nt, bool with_tiles) {
  const size_t span = (size_t)(rpt > 1 ? nt * rpt + nt : nt * rpt) + 1;
  return (2 * span + 16 + kGCap + 1) * sizeof(int64_t) + kGCap * sizeof(int32_t) + kQCap * sizeof(uint32_t) +
         (size_t)kEPT * nt * (sizeof(int64_t) + sizeof(uint16_t)) + (with_tiles ? kQCap * sizeof(uint32_t) : 0);
}

bool block_adds_carry(const BwtLaunch* l) { return l->carry_in_block && !l->write_index; }


// LDS stride (entries per wavefront) of a level: its classes + 1, rounded to an odd number of
// 8-byte words so consecutive wavefronts' rows start in different banks
int level_stride(int ncls) { return (ncls + 1) | 1; }

size_t replay_lds_bytes(const BwtIndex* x, const BwtLaunch* l) {
  return (size_t)l->chunk_waves * (size_t)replay_wave_words(level_stride(x->ncls_b)) * sizeof(int64_t);
}

// wavefronts per workgroup of the chunk histogram for a level of ncls classes: the launch's, fewer
// when that many wavefronts' class rows would not fit the workgroup's LDS
int hist_waves(int ncls, const BwtLaunch* l) {
  int w = l->chunk_waves;
  while (w > 1 && (size_t)w * (size_t)level_stride(ncls) * sizeof(int64_t) > 160 * 1024) --w;
  return w;
}

int check(const BwtIndex* x, const BwtLaunch* l) {
  if (!x || !l || x->S < 1 || x->U < 1 || x->nblk < 1) return (int)hipErrorInvalidValue;
  if (l->block_threads < 64 || l->block_threads > 1024 || (l->block_threads & 63)) return (int)hipErrorInvalidValue;
  if (x->rmax < 1 || x->rmax > 8192 || x->max_cls < 1 || x->max_cls > BWT_MAX_CLASSES) return (int)hipErrorInvalidValue;
  if (x->ncls_c < 1 || x->ncls_c > x->max_cls || x->ncls_b < 1 || x->ncls_b > x->max_cls) return (int)hipErrorInvalidValue;
  if (x->ncls_b > 4096) return (int)hipErrorInvalidValue;  // level-B classes are tile offsets (tile_shift <= 12)
  if (x->nchunk_c != x->nblk || x->nchunk_c > x->nchunk) return (int)hipErrorInvalidValue;
  if (x->tile_shift < 1 || x->tile_shift > 12) return (int)hipErrorInvalidValue;
  if (l->chunk_waves < 1 || l->chunk_waves > 4) return (int)hipErrorInvalidValue;
  const int rpt = block_rpt(x, l->block_threads);
  if (rpt > 16 || block_lds_bytes(rpt, l->block_threads, block_adds_carry(l)) > 160 * 1024) return (int)hipErrorInvalidValue;
  if (replay_lds_bytes(x, l) > 160 * 1024 || (size_t)level_stride(x->max_cls) * sizeof(int64_t) > 160 * 1024)
    return (int)hipErrorInvalidValue;
  return 0;
}

int set_lds_limits() {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
    const int lim = 160 * 1024;
    const void* ks[] = {reinterpret_cast<const void*>(bwt_block_kernel<1>),
                        reinterpret_cast<const void*>(bwt_block_kernel<2>),
                        reinterpret_cast<const void*>(bwt_block_kernel<4>),
                        reinterpret_cast<const void*>(bwt_block_kernel<8>),
                        reinterpret_cast<const void*>(bwt_block_kernel<16>),
                        reinterpret_cast<const void*>(bwt_chunk_hist_kernel),
                        reinterpret_cast<const void*>(bwt_row_c_kernel),
                        reinterpret_cast<const void*>(bwt_rank_lds_kernel),
                        reinterpret_cast<const void*>(bwt_chunk_replay_kernel<9>),
                        reinterpret_cast<const void*>(bwt_chunk_replay_kernel<10>),
                        reinterpret_cast<const void*>(bwt_chunk_replay_kernel<12>)};
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

template <int RPT>
void launch_block(const BwtIndex* x, const BwtBatch* b, const BwtLaunch* l, hipStream_t st) {
  const int64_t grid = b->dcounts ? std::min<int64_t>(b->nqblk, kDeviceCountGrid) : b->nqblk;
  hipLaunchKernelGGL(bwt_block_kernel<RPT>, dim3((unsigned)grid), dim3(l->block_threads),
                     block_lds_bytes(RPT, l->block_threads, block_adds_carry(l)), st, *x, *b, *l);
}

}  // namespace

extern "C" int bwt_block_replay_launch(const BwtIndex* x, const BwtBatch* b, const BwtLaunch* l, void* stream) {
  int rc = check(x, l);
  if (rc) return rc;
  if (!b || b->nqblk < 0 || (l->write_index && b->nqblk != x->nblk)) return (int)hipErrorInvalidValue;
  if ((rc = set_lds_limits())) return rc;
  if (b->nqblk == 0) return 0;
  const hipStream_t st = (hipStream_t)stream;
  switch (block_rpt(x, l->block_threads)) {
    case 1: launch_block<1>(x, b, l, st); break;
    case 2: launch_block<2>(x, b, l, st); break;
    case 4: launch_block<4>(x, b, l, st); break;
    case 8: launch_block<8>(x, b, l, st); break;
    case 16: launch_block<16>(x, b, l, st); break;
    default: return (int)hipErrorInvalidValue;
  }
  return report("block replay", hipGetLastError());
}

extern "C" int bwt_chunk_hist_launch(const BwtIndex* x, const BwtLaunch* l, void* stream) {
  int rc = check(x, l);
  if (rc) return rc;
  if ((rc = set_lds_limits())) return rc;
  // level C: one workgroup per block writes its row from the block's sorted tile slots
  if (x->nchunk_c > 0) {
    hipLaunchKernelGGL(bwt_row_c_kernel, dim3((unsigned)x->nchunk_c), dim3(kRowCThreads),
                       (kRowCThreads / 64) * sizeof(int64_t) + (size_t)x->rmax * (sizeof(int64_t) + sizeof(uint32_t)),
                       (hipStream_t)stream, *x, x->rmax);
    if ((rc = report("level-C rows", hipGetLastError()))) return rc;
  }
  // level B chunks (which also store their item weights), LDS sized to the class count
  const int64_t rng[2][2] = {{0, x->nchunk_c}, {x->nchunk_c, x->nchunk}};
  const int ncls[2] = {x->ncls_c, x->ncls_b};
  for (int lv = 1; lv < 2; ++lv) {
    const int64_t n = rng[lv][1] - rng[lv][0];
    if (n <= 0) continue;
    const int stride = level_stride(ncls[lv]);
    BwtLaunch lh = *l;
    lh.chunk_waves = hist_waves(ncls[lv], l);
    const int64_t grid = (n + lh.chunk_waves - 1) / lh.chunk_waves;
    hipLaunchKernelGGL(bwt_chunk_hist_kernel, dim3((unsigned)grid), dim3(64 * lh.chunk_waves),
                       (size_t)lh.chunk_waves * stride * sizeof(int64_t), (hipStream_t)stream, *x, lh, rng[lv][0],
                       rng[lv][1], stride, lv);
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
  if (b->nwork == 0) return 0;
  const int stride = level_stride(x->ncls_b);
  int64_t grid = (b->nwork + l->chunk_waves - 1) / l->chunk_waves;
  if (b->dcounts) grid = std::min<int64_t>(grid, kDeviceCountGrid);
  // classes of a level-B chain: key offsets within a tile, < ncls_b <= 4096
  if (x->ncls_b <= 512)
    hipLaunchKernelGGL(bwt_chunk_replay_kernel<9>, dim3((unsigned)grid), dim3(64 * l->chunk_waves),
                       replay_lds_bytes(x, l), (hipStream_t)stream, *x, *b, *l, stride);
  else if (x->ncls_b <= 1024)
    hipLaunchKernelGGL(bwt_chunk_replay_kernel<10>, dim3((unsigned)grid), dim3(64 * l->chunk_waves),
                       replay_lds_bytes(x, l), (hipStream_t)stream, *x, *b, *l, stride);
  else
    hipLaunchKernelGGL(bwt_chunk_replay_kernel<12>, dim3((unsigned)grid), dim3(64 * l->chunk_waves),
                       replay_lds_bytes(x, l), (hipStream_t)stream, *x, *b, *l, stride);
  return report("chunk replay", hipGetLastError());
}

extern "C" int bwt_combine_launch(const BwtIndex* x, const BwtBatch* b, const BwtLaunch* l, void* stream) {
  if (!x || !b || b->M < 0) return (int)hipErrorInvalidValue;
  if (b->M == 0) return 0;
  if (!b->spos || !b->srank || !b->ipos || !b->stage) return (int)hipErrorInvalidValue;
  const int64_t nt = (b->M + BWT_COMBINE_TILE - 1) / BWT_COMBINE_TILE;
  const int64_t per_t = (nt + BWT_XCDS - 1) / BWT_XCDS;
  const bool carry = !(l && l->carry_in_block && !l->write_index);  // else the block replay added it
  if (carry)
    hipLaunchKernelGGL(bwt_combine_slots_kernel<true>, dim3((unsigned)(per_t * BWT_XCDS)), dim3(BWT_COMBINE_THREADS), 0,
                       (hipStream_t)stream, *x, *b, per_t);
  else
    hipLaunchKernelGGL(bwt_combine_slots_kernel<false>, dim3((unsigned)(per_t * BWT_XCDS)), dim3(BWT_COMBINE_THREADS),
                       0, (hipStream_t)stream, *x, *b, per_t);
  int rc = report("combine (slots)", hipGetLastError());
  if (rc) return rc;
  const int64_t nb = (b->M + BWT_COMBINE_THREADS * kIdsPer - 1) / (BWT_COMBINE_THREADS * kIdsPer);
  const int64_t per = (nb + BWT_XCDS - 1) / BWT_XCDS;
  hipLaunchKernelGGL(bwt_combine_ids_kernel, dim3((unsigned)(per * BWT_XCDS)), dim3(BWT_COMBINE_THREADS), 0,
                     (hipStream_t)stream, *b, per);
  return report("combine (ids)", hipGetLastError());
}

extern "C" int bwt_query_launch(const BwtIndex* x, const BwtBatch* b, const BwtLaunch* l, void* stream) {
  BwtLaunch q = *l;
  q.write_index = 0;
  q.carry_in_block = 1;
  int rc;
  if ((rc = bwt_block_replay_launch(x, b, &q, stream))) return rc;
  if ((rc = bwt_chunk_replay_launch(x, b, &q, stream))) return rc;
  return bwt_combine_launch(x, b, &q, stream);
}

extern "C" int bwt_pass_launch(const BwtIndex* x, const BwtBatch* b, const BwtLaunch* l, void* stream) {
  BwtLaunch w = *l;
  w.write_index = 1;
  w.carry_in_block = 0;
  int rc;
  if ((rc = bwt_block_replay_launch(x, b, &w, stream))) return rc;
  if ((rc = bwt_chunk_hist_launch(x, &w, stream))) return rc;
  if ((rc = bwt_carry_scan_launch(x, &w, stream))) return rc;
  if ((rc = bwt_chunk_replay_launch(x, b, &w, stream))) return rc;
  return bwt_combine_launch(x, b, &w, stream);
}

extern "C" int bwt_overlay_launch(const BwtOverlay* o, void* stream) {
  if (!o || o->n < 0 || o->U < 1 || o->S < 1) return (int)hipErrorInvalidValue;
  if (o->n == 0) return 0;
  hipLaunchKernelGGL(bwt_overlay_kernel, dim3((unsigned)((o->n + 3) / 4)), dim3(256), 0, (hipStream_t)stream, *o);
  return report("overlay", hipGetLastError());
}

extern "C" int bwt_rank_remap_launch(const BwtRankArgs* a, void* stream) {
  if (!a || a->n < 0) return (int)hipErrorInvalidValue;
  if (a->n == 0) return 0;
  // large remaps stage the row offsets in LDS (one workgroup per CU, each staging 256 KiB of
  // offsets once); small ones (a served batch's endpoints) read them from global memory
  if (a->n >= kRankLdsMin) {
    int rc = set_lds_limits();
    if (rc) return rc;
    int dev = 0, ncu = 0;
    hipError_t e;
    if ((e = hipGetDevice(&dev)) || (e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)))
      return report("device query", e);
    const int64_t want = (a->n + kRankLdsThreads - 1) / kRankLdsThreads;
    const unsigned grid = (unsigned)std::min<int64_t>(ncu, want);
    hipLaunchKernelGGL(bwt_rank_lds_kernel, dim3(grid), dim3(kRankLdsThreads), kRankLdsBytes, (hipStream_t)stream, *a);
  } else {
    hipLaunchKernelGGL(bwt_rank_kernel, dim3((unsigned)((a->n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *a);
  }
  return report("rank remap", hipGetLastError());
}
