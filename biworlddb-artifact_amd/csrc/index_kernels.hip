// index_kernels.hip -- the device index build (include/bwgpu_index.h): the per-block and
// per-chain tables of the Euler-walk index from the walk and its block cut.
//   stage 1  idx_keys_kernel     one workgroup per block: LDS hash set of its events' keys, the
//                                distinct keys sorted (bitonic, in LDS), event ranks by binary
//                                search; step -> block, state -> enter step; key / tile counts
//   stage 2  idx_write_kernel    per block: its keys into bkey, its tile slots (first rank, tile,
//                                block, key count) and their level-C items
//            radix sort          slots by tile, stable (rocPRIM), then their key counts scanned
//            idx_tiles_kernel    chain_key0 and tile_tb0 by binary search over the sorted tiles
//            idx_items_kernel    one thread per sorted slot: its level-B items, tb_block, tb_pos
//            idx_btpos_kernel    the dense (block, tile) table of level-B positions
// Everything is exact integer work; the results equal the host build array for array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <stdint.h>

#include "bwgpu_index.h"
#include "bwgpu_tree.h"
#include "wave64.hpp"

namespace {

constexpr int kT = 256;

__device__ __forceinline__ uint64_t key_item(uint32_t cls, uint64_t ref) {
  return ((uint64_t)cls << BWT_CLASS_SHIFT) | (ref & BWT_REF_MASK);
}

// the LDS hash capacity of a block: a power of two >= 2 * rmax (load at most one half)
int cap_for(int rmax) {
  int c = 64;
  while (c < 2 * rmax) c <<= 1;
  return c;
}

unsigned bits_for(uint64_t v) {
  unsigned b = 1;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

template <int kCap>
__global__ void __launch_bounds__(kT) idx_keys_kernel(const BwtIndexBuild in, uint16_t* evrank, int32_t* state_step,
                                                      int32_t* step_block, uint32_t* tmp, int32_t* nkeys,
                                                      int32_t* ntiles) {
  __shared__ uint32_t tab[kCap];  // key + 1, 0 = empty
  __shared__ uint32_t srt[kCap];
  __shared__ int cnt, tiles;
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  const int s0 = in.blk_step0[b], s1 = in.blk_step0[b + 1];
  const int U = in.U;
  const int64_t ne = (int64_t)(s1 - s0) * U;
  for (int i = tid; i < kCap; i += kT) {
    tab[i] = 0;
    srt[i] = 0xFFFFFFFFu;
  }
  if (tid == 0) {
    cnt = 0;
    tiles = 0;
    if (b == 0) state_step[0] = -1;
  }
  __syncthreads();
  for (int64_t e = tid; e < ne; e += kT) {
    const int s = s0 + (int)(e / U), i = (int)(e - (int64_t)(e / U) * U);
    const int st = in.steps[s];
    const int64_t v = st > 0 ? st : -st;
    if (i == 0) {
      step_block[s] = b;
      if (st > 0) state_step[v] = s;
    }
    const uint32_t k = in.key[(v - 1) * (int64_t)U + i];
    if ((k >> in.tile_shift) >= (uint32_t)in.KT) {  // a key outside the tiles: reported as an overfull block
      atomicAdd(&cnt, kCap);
      continue;
    }
    uint32_t h = (k * 0x9E3779B1u) >> (32 - __builtin_ctz((unsigned)kCap));
    // at most kCap probes: a full table (more distinct keys than the cut allows) ends the insert
    for (int probe = 0; probe < kCap; ++probe) {
      const uint32_t old = atomicCAS(&tab[h], 0u, k + 1u);
      if (old == 0u) {
        const int at = atomicAdd(&cnt, 1);
        if (at < kCap / 2) srt[at] = k;
        break;
      }
      if (old == k + 1u) break;
      h = (h + 1) & (kCap - 1);
      if (probe == kCap - 1) atomicAdd(&cnt, kCap);  // counted as too many keys
    }
  }
  __syncthreads();
  const int R = cnt;
  if (R > in.rmax) {  // cannot come from the host cut; reported through the count
    if (tid == 0) {
      nkeys[b] = R;
      ntiles[b] = 0;
    }
    return;
  }
  int P = 1;
  while (P < R) P <<= 1;
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < P; i += kT) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const uint32_t a = srt[i], c = srt[ixj];
          if ((a > c) == ((i & k) == 0)) {
            srt[i] = c;
            srt[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  const int ts = in.tile_shift;
  int tl = 0;
  for (int i = tid; i < R; i += kT) {
    tmp[(int64_t)b * in.rmax + i] = srt[i];
    tl += (i == 0 || (srt[i] >> ts) != (srt[i - 1] >> ts)) ? 1 : 0;
  }
  if (tl) atomicAdd(&tiles, tl);
  for (int64_t e = tid; e < ne; e += kT) {
    const int s = s0 + (int)(e / U), i = (int)(e - (int64_t)(e / U) * U);
    const int st = in.steps[s];
    const int64_t v = st > 0 ? st : -st;
    const uint32_t k = in.key[(v - 1) * (int64_t)U + i];
    evrank[(int64_t)s * U + i] = (uint16_t)bw::lower_bound_lds(srt, R, k);
  }
  __syncthreads();
  if (tid == 0) {
    nkeys[b] = R;
    ntiles[b] = tiles;
  }
}

// per block: bkey, and its tile slots: first rank, tile, block, key count, level-C item
__global__ void __launch_bounds__(kT) idx_write_kernel(const BwtIndexBuild in, const uint32_t* tmp, BwtIndexTables t,
                                                       uint32_t* slot_tile, int32_t* slot_block, int64_t* slot_n,
                                                       uint32_t* slot_id) {
  __shared__ int64_t wsum[kT / 64 + 1];
  extern __shared__ uint32_t chg[];  // rank of each tile's first key (+ R at the end)
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int b = blockIdx.x;
  const int64_t k0 = t.blk_key0[b], j0 = t.blk_tile0[b];
  const int R = (int)(t.blk_key0[b + 1] - k0);
  const int T = (int)(t.blk_tile0[b + 1] - j0);
  const int ts = in.tile_shift;
  const uint32_t* ks = tmp + (int64_t)b * in.rmax;
  // contiguous ranges per thread, so the change positions come out in order
  const int per = (R + kT - 1) / kT;
  const int a0 = tid * per, a1 = min(a0 + per, R);
  int64_t c = 0;
  for (int i = a0; i < a1; ++i) {
    const uint32_t k = ks[i];
    t.bkey[k0 + i] = k;
    c += (i == 0 || (k >> ts) != (ks[i - 1] >> ts)) ? 1 : 0;
  }
  const int64_t inc = bw::wave_incl_scan(c);
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  int64_t off = inc - c;
  for (int w = 0; w < wv; ++w) off += wsum[w];
  for (int i = a0; i < a1; ++i) {
    const uint32_t k = ks[i];
    if (i == 0 || (k >> ts) != (ks[i - 1] >> ts)) chg[off++] = (uint32_t)i;
  }
  if (tid == 0) chg[T] = (uint32_t)R;
  __syncthreads();
  for (int j = tid; j < T; j += kT) {
    const int64_t s = j0 + j;
    const uint32_t r = chg[j];
    const uint32_t tile = ks[r] >> ts;
    t.tile_rank0[s] = (uint16_t)r;
    t.kitems[s] = key_item(tile, (uint64_t)s);
    slot_tile[s] = tile;
    slot_block[s] = b;
    slot_n[s] = (int64_t)(chg[j + 1] - r);
    slot_id[s] = (uint32_t)s;
  }
}

__global__ void __launch_bounds__(kT) idx_gather_n_kernel(int64_t nslots, const uint32_t* sorted_id,
                                                          const int64_t* slot_n, int64_t* n_sorted) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < nslots) n_sorted[i] = slot_n[sorted_id[i]];
}

// chain_key0 and tile_tb0: the first sorted slot of each tile, by binary search
__global__ void __launch_bounds__(kT) idx_tiles_kernel(int KT, int64_t nslots, int64_t nC, const uint32_t* sorted_tile,
                                                       const int64_t* start, const int64_t* n_sorted,
                                                       int64_t* chain_key0, int64_t* tile_tb0) {
  const int t = blockIdx.x * kT + threadIdx.x;
  if (t > KT) return;
  int64_t lo = 0, hi = nslots;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (sorted_tile[mid] < (uint32_t)t) lo = mid + 1;
    else hi = mid;
  }
  const int64_t totalB = start[nslots - 1] + n_sorted[nslots - 1];
  tile_tb0[t] = lo;
  chain_key0[1 + t] = nC + (lo < nslots ? start[lo] : totalB);
  if (t == 0) chain_key0[0] = 0;
}

// one thread per sorted slot (a slot holds a few keys): its keys' level-B items, tb_block, tb_pos
__global__ void __launch_bounds__(kT) idx_items_kernel(const BwtIndexBuild in, BwtIndexTables t, int64_t nslots,
                                                       int64_t nC, const uint32_t* sorted_tile,
                                                       const uint32_t* sorted_id, const int32_t* slot_block,
                                                       const int64_t* n_sorted, const int64_t* start) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= nslots) return;
  const uint32_t tile = sorted_tile[i];
  const uint32_t slot = sorted_id[i];
  const int b = slot_block[slot];
  const int64_t r0 = t.blk_key0[b] + t.tile_rank0[slot];
  const int64_t n = n_sorted[i], pos0 = nC + start[i];
  const uint32_t mask = (1u << in.tile_shift) - 1;
  for (int64_t j = 0; j < n; ++j) t.kitems[pos0 + j] = key_item(t.bkey[r0 + j] & mask, (uint64_t)(r0 + j));
  t.tb_block[i] = b;
  t.tb_pos[i] = start[i] - (t.chain_key0[1 + tile] - nC);
}

// bt_pos[b*KT + t]: the level-B position of block b's first key of tile t (or of the next block
// holding the tile; the chain length past the last)
__global__ void __launch_bounds__(kT) idx_btpos_kernel(int nblk, int KT, BwtIndexTables t) {
  const int64_t x = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (x >= (int64_t)nblk * KT) return;
  const int b = (int)(x / KT), tl = (int)(x - (int64_t)b * KT);
  int64_t lo = t.tile_tb0[tl], hi = t.tile_tb0[tl + 1];
  const int64_t e = hi;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (t.tb_block[mid] < b) lo = mid + 1;
    else hi = mid;
  }
  t.bt_pos[x] = lo < e ? (int32_t)t.tb_pos[lo] : (int32_t)(t.chain_key0[tl + 2] - t.chain_key0[tl + 1]);
}

// ---- the Euler walk: children lists by a stable sort of (parent, state); every tree edge gets
// its successor in the preorder walk (down an edge to the first child, or back up; up an edge to
// the next sibling, or further up); list ranking by pointer jumping gives each edge its place.
// Edge 2(c-1) enters state c, edge 2(c-1)+1 leaves it; edge 2S ends the walk.
__global__ void __launch_bounds__(kT) walk_keys_kernel(int64_t S, const int32_t* parent, uint32_t* key,
                                                       uint32_t* val) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= S) return;
  key[i] = (uint32_t)parent[i + 1];
  val[i] = (uint32_t)(i + 1);
}

__global__ void __launch_bounds__(kT) walk_koff_kernel(int64_t S, const uint32_t* sorted_parent,
                                                       const uint32_t* kids, int64_t* koff, int32_t* pos_of) {
  const int64_t p = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (p <= S + 1) {  // koff[p]: first child slot of state p (koff[S+1] = S)
    int64_t lo = 0, hi = S;
    while (lo < hi) {
      const int64_t mid = lo + (hi - lo) / 2;
      if ((int64_t)sorted_parent[mid] < p) lo = mid + 1;
      else hi = mid;
    }
    koff[p] = lo;
  }
  if (p < S) pos_of[kids[p]] = (int32_t)p;
}

__global__ void __launch_bounds__(kT) walk_succ_kernel(int64_t S, const int32_t* parent, const uint32_t* kids,
                                                       const int64_t* koff, const int32_t* pos_of, int64_t* nxt,
                                                       int64_t* d) {
  const int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x + 1;
  if (c > S + 1) return;
  const int64_t END = 2 * S;
  if (c == S + 1) {
    nxt[END] = END;
    d[END] = 0;
    return;
  }
  const int64_t dn = 2 * (c - 1), up = dn + 1;
  const int64_t p = parent[c];
  const int64_t i = pos_of[c];
  nxt[dn] = koff[c] < koff[c + 1] ? 2 * ((int64_t)kids[koff[c]] - 1) : up;
  nxt[up] = i + 1 < koff[p + 1] ? 2 * ((int64_t)kids[i + 1] - 1) : (p == 0 ? END : 2 * (p - 1) + 1);
  d[dn] = 1;
  d[up] = 1;
}

__global__ void __launch_bounds__(kT) walk_rank_kernel(int64_t n, const int64_t* nxt, const int64_t* d, int64_t* nxt2,
                                                       int64_t* d2) {
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (e > n) return;
  const int64_t x = nxt[e];
  d2[e] = d[e] + d[x];
  nxt2[e] = nxt[x];
}

// place every edge at 2S - (its distance to the end); the walk stops after the last entry
__global__ void __launch_bounds__(kT) walk_place_kernel(int64_t S, const int64_t* d, int32_t* steps,
                                                        unsigned long long* last) {
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (e >= 2 * S) return;
  const int64_t pos = 2 * S - d[e];
  const int32_t c = (int32_t)(e / 2 + 1);
  steps[pos] = (e & 1) ? -c : c;
  if (!(e & 1)) atomicMax(last, (unsigned long long)pos);
}

__global__ void __launch_bounds__(kT) walk_zero_kernel(unsigned long long* last) { *last = 0; }
__global__ void __launch_bounds__(kT) walk_nsteps_kernel(const unsigned long long* last, int64_t* nsteps) {
  *nsteps = (int64_t)*last + 1;
}

struct WalkLayout {
  size_t key, key_s, val, kids, koff, pos_of, nxt, d, nxt2, d2, last, temp, temp_bytes, total;
};

hipError_t walk_layout_of(int64_t S, WalkLayout& L) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (bytes + 255) & ~size_t(255);
    return at;
  };
  const size_t n = (size_t)S, ne = 2 * (size_t)S + 1;
  L.key = take(n * 4);
  L.key_s = take(n * 4);
  L.val = take(n * 4);
  L.kids = take(n * 4);
  L.koff = take((n + 2) * 8);
  L.pos_of = take((n + 1) * 4);
  L.nxt = take(ne * 8);
  L.d = take(ne * 8);
  L.nxt2 = take(ne * 8);
  L.d2 = take(ne * 8);
  L.last = take(8);
  size_t b = 0;
  const hipError_t e = rocprim::radix_sort_pairs(nullptr, b, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                 (const uint32_t*)nullptr, (uint32_t*)nullptr, n, 0u,
                                                 bits_for((uint64_t)S));
  if (e != hipSuccess) return e;
  L.temp_bytes = std::max<size_t>(b, 1);
  L.temp = take(L.temp_bytes);
  L.total = off;
  return hipSuccess;
}

// ---- stage 0: the block cut.  e(s): where a block starting at step s would end -- the first
// step t > s whose keys take the block's distinct keys past rmax, or at max_block_steps.
constexpr int kCutT = 128;

template <int kCap>
__global__ void __launch_bounds__(kCutT) idx_cut_end_kernel(const BwtIndexBuild in, int32_t mbs, int32_t* e_out) {
  __shared__ uint32_t tab[kCap];  // key + 1, 0 = empty
  __shared__ int total;
  const int tid = threadIdx.x;
  const int64_t s0 = blockIdx.x;
  for (int i = tid; i < kCap; i += kCutT) tab[i] = 0;
  if (tid == 0) total = 0;
  __syncthreads();
  const int U = in.U;
  int64_t e = in.nsteps;
  for (int64_t t = s0; t < in.nsteps; ++t) {
    if (t > s0 && t - s0 >= mbs) {
      e = t;
      break;
    }
    const int st = in.steps[t];
    const int64_t v = st > 0 ? st : -st;
    for (int i = tid; i < U; i += kCutT) {
      const uint32_t k = in.key[(v - 1) * (int64_t)U + i];
      uint32_t h = (k * 0x9E3779B1u) >> (32 - __builtin_ctz((unsigned)kCap));
      for (int probe = 0; probe < kCap; ++probe) {
        const uint32_t old = atomicCAS(&tab[h], 0u, k + 1u);
        if (old == 0u) {
          atomicAdd(&total, 1);
          break;
        }
        if (old == k + 1u) break;
        h = (h + 1) & (kCap - 1);
      }
    }
    __syncthreads();
    const int d = total;  // distinct keys of steps s0..t
    __syncthreads();      // every thread has read it before the next step adds to it
    if (t > s0 && d > in.rmax) {
      e = t;
      break;
    }
  }
  if (tid == 0) e_out[s0] = (int32_t)e;
}

__global__ void __launch_bounds__(kT) idx_cut_init_kernel(int64_t n, int32_t* J, int32_t* mark) {
  const int64_t s = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (s > n) return;
  if (s == n) J[s] = (int32_t)n;  // the end maps to itself
  mark[s] = s == 0 ? 1 : 0;
}

// one round of pointer doubling: marks J(s) for every marked s, and J <- J o J
__global__ void __launch_bounds__(kT) idx_cut_mark_kernel(int64_t n, const int32_t* J, int32_t* mark) {
  const int64_t s = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (s <= n && mark[s]) mark[J[s]] = 1;
}
__global__ void __launch_bounds__(kT) idx_cut_jump_kernel(int64_t n, const int32_t* J, int32_t* J2) {
  const int64_t s = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (s <= n) J2[s] = J[J[s]];
}

__global__ void __launch_bounds__(kT) idx_cut_scatter_kernel(int64_t n, const int32_t* mark, const int32_t* pos,
                                                             int32_t* blk_step0, int32_t* nblk) {
  const int64_t s = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (s < n && mark[s]) blk_step0[pos[s]] = (int32_t)s;
  if (s == n - 1) {
    const int32_t nb = pos[s] + mark[s];
    blk_step0[nb] = (int32_t)n;
    *nblk = nb;
  }
}

struct CutLayout {
  size_t J, J2, mark, pos, temp, temp_bytes, total;
};

hipError_t cut_layout_of(const BwtIndexBuild* in, CutLayout& L) {
  const size_t n = (size_t)in->nsteps + 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (bytes + 255) & ~size_t(255);
    return at;
  };
  L.J = take(n * 4);
  L.J2 = take(n * 4);
  L.mark = take(n * 4);
  L.pos = take(n * 4);
  size_t b = 0;
  const hipError_t e = rocprim::exclusive_scan(nullptr, b, (const int32_t*)nullptr, (int32_t*)nullptr, (int32_t)0,
                                               (size_t)in->nsteps, rocprim::plus<int32_t>());
  if (e != hipSuccess) return e;
  L.temp_bytes = std::max<size_t>(b, 1);
  L.temp = take(L.temp_bytes);
  L.total = off;
  return hipSuccess;
}

// workspace layout
struct WsLayout {
  size_t tmp, slot_tile, slot_tile_s, slot_id, slot_id_s, slot_block, slot_n, n_sorted, start, temp, temp_bytes,
      total;
};

size_t slot_cap(const BwtIndexBuild* in) {
  return (size_t)in->nblk * (size_t)std::min<int64_t>(in->rmax, in->KT);
}

hipError_t layout_of(const BwtIndexBuild* in, WsLayout& L) {
  const size_t nb = (size_t)in->nblk, cap = slot_cap(in);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (bytes + 255) & ~size_t(255);
    return at;
  };
  L.tmp = take(nb * (size_t)in->rmax * 4);
  L.slot_tile = take(cap * 4);
  L.slot_tile_s = take(cap * 4);
  L.slot_id = take(cap * 4);
  L.slot_id_s = take(cap * 4);
  L.slot_block = take(cap * 4);
  L.slot_n = take(cap * 8);
  L.n_sorted = take(cap * 8);
  L.start = take(cap * 8);
  size_t a = 0, b = 0;
  hipError_t e = rocprim::radix_sort_pairs(nullptr, a, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (const uint32_t*)nullptr, (uint32_t*)nullptr, cap, 0u,
                                           bits_for((uint64_t)in->KT));
  if (e != hipSuccess) return e;
  e = rocprim::exclusive_scan(nullptr, b, (const int64_t*)nullptr, (int64_t*)nullptr, (int64_t)0, cap,
                              rocprim::plus<int64_t>());
  if (e != hipSuccess) return e;
  L.temp_bytes = std::max<size_t>(std::max(a, b), 1);
  L.temp = take(L.temp_bytes);
  L.total = off;
  return hipSuccess;
}

bool build_ok(const BwtIndexBuild* in) {
  return in && in->S >= 1 && in->U >= 1 && in->rmax >= in->U && in->rmax >= 8 && in->rmax <= 8192 &&
         in->tile_shift >= 1 && in->tile_shift <= 12 && in->KT >= 1 && in->KT < BWT_MAX_CLASSES && in->nsteps >= 1 &&
         in->nblk >= 1 && in->steps && in->blk_step0 && in->key;
}

template <typename T>
T* at(void* base, size_t off) {
  return reinterpret_cast<T*>(static_cast<unsigned char*>(base) + off);
}

}  // namespace

extern "C" int bwt_index_walk_workspace(int64_t S, size_t* bytes) {
  if (S < 1 || S >= (int64_t(1) << 30) || !bytes) return int(hipErrorInvalidValue);
  WalkLayout L;
  const hipError_t e = walk_layout_of(S, L);
  if (e != hipSuccess) return int(e);
  *bytes = L.total;
  return 0;
}

extern "C" int bwt_index_walk_launch(int64_t S, const int32_t* parent, int32_t* steps, int64_t* nsteps, void* work,
                                     size_t work_bytes, void* stream) {
  if (S < 1 || S >= (int64_t(1) << 30) || !parent || !steps || !nsteps) return int(hipErrorInvalidValue);
  WalkLayout L;
  hipError_t e = walk_layout_of(S, L);
  if (e != hipSuccess) return int(e);
  if (!work || work_bytes < L.total) return int(hipErrorInvalidValue);
  hipStream_t st = hipStream_t(stream);
  uint32_t* key = at<uint32_t>(work, L.key);
  uint32_t* key_s = at<uint32_t>(work, L.key_s);
  uint32_t* val = at<uint32_t>(work, L.val);
  uint32_t* kids = at<uint32_t>(work, L.kids);
  int64_t* koff = at<int64_t>(work, L.koff);
  int32_t* pos_of = at<int32_t>(work, L.pos_of);
  int64_t* nxt = at<int64_t>(work, L.nxt);
  int64_t* d = at<int64_t>(work, L.d);
  int64_t* nxt2 = at<int64_t>(work, L.nxt2);
  int64_t* d2 = at<int64_t>(work, L.d2);
  unsigned long long* last = at<unsigned long long>(work, L.last);
  const dim3 blk(kT);
  hipLaunchKernelGGL(walk_keys_kernel, dim3((unsigned)((S + kT - 1) / kT)), blk, 0, st, S, parent, key, val);
  size_t tb = L.temp_bytes;
  if ((e = rocprim::radix_sort_pairs(at<void>(work, L.temp), tb, key, key_s, val, kids, (size_t)S, 0u,
                                     bits_for((uint64_t)S), st)) != hipSuccess)
    return int(e);
  hipLaunchKernelGGL(walk_koff_kernel, dim3((unsigned)((S + 2 + kT - 1) / kT)), blk, 0, st, S, (const uint32_t*)key_s,
                     (const uint32_t*)kids, koff, pos_of);
  hipLaunchKernelGGL(walk_succ_kernel, dim3((unsigned)((S + 1 + kT - 1) / kT)), blk, 0, st, S, parent,
                     (const uint32_t*)kids, (const int64_t*)koff, (const int32_t*)pos_of, nxt, d);
  const int64_t n = 2 * S;  // edges 0..2S-1 and the end 2S
  const dim3 ge((unsigned)((n + 1 + kT - 1) / kT));
  for (int64_t reach = 1; reach <= n; reach <<= 1) {
    hipLaunchKernelGGL(walk_rank_kernel, ge, blk, 0, st, n, (const int64_t*)nxt, (const int64_t*)d, nxt2, d2);
    std::swap(nxt, nxt2);
    std::swap(d, d2);
  }
  hipLaunchKernelGGL(walk_zero_kernel, dim3(1), dim3(1), 0, st, last);
  hipLaunchKernelGGL(walk_place_kernel, dim3((unsigned)((n + kT - 1) / kT)), blk, 0, st, S, (const int64_t*)d, steps,
                     last);
  if ((e = hipGetLastError()) != hipSuccess) return int(e);
  // nsteps = last entry + 1
  hipLaunchKernelGGL(walk_nsteps_kernel, dim3(1), dim3(1), 0, st, (const unsigned long long*)last, nsteps);
  return int(hipGetLastError());
}

extern "C" int bwt_index_cut_workspace(const BwtIndexBuild* in, size_t* bytes) {
  if (!in || !bytes || in->nsteps < 1 || in->nsteps >= (int64_t(1) << 31) - 1) return int(hipErrorInvalidValue);
  CutLayout L;
  const hipError_t e = cut_layout_of(in, L);
  if (e != hipSuccess) return int(e);
  *bytes = L.total;
  return 0;
}

extern "C" int bwt_index_cut_launch(const BwtIndexBuild* in, int32_t max_block_steps, int32_t* blk_step0, int32_t* nblk,
                                    void* work, size_t work_bytes, void* stream) {
  if (!in || !in->steps || !in->key || in->U < 1 || in->rmax < in->U || in->rmax > 8192 || in->nsteps < 1 ||
      in->nsteps >= (int64_t(1) << 31) - 1 || max_block_steps < 1 || !blk_step0 || !nblk)
    return int(hipErrorInvalidValue);
  CutLayout L;
  hipError_t e = cut_layout_of(in, L);
  if (e != hipSuccess) return int(e);
  if (!work || work_bytes < L.total) return int(hipErrorInvalidValue);
  hipStream_t st = hipStream_t(stream);
  int32_t* J = at<int32_t>(work, L.J);
  int32_t* J2 = at<int32_t>(work, L.J2);
  int32_t* mark = at<int32_t>(work, L.mark);
  int32_t* pos = at<int32_t>(work, L.pos);
  const int64_t n = in->nsteps;
  const dim3 gn((unsigned)((n + 1 + kT - 1) / kT)), blk(kT);
  hipLaunchKernelGGL(idx_cut_init_kernel, gn, blk, 0, st, n, J, mark);
#define BWT_IDX_CUT(C) \
  hipLaunchKernelGGL(idx_cut_end_kernel<C>, dim3((unsigned)n), dim3(kCutT), 0, st, *in, max_block_steps, J)
  int cap = 64;
  while (cap < 2 * (in->rmax + in->U)) cap <<= 1;
  switch (cap) {
    case 64: BWT_IDX_CUT(64); break;
    case 128: BWT_IDX_CUT(128); break;
    case 256: BWT_IDX_CUT(256); break;
    case 512: BWT_IDX_CUT(512); break;
    case 1024: BWT_IDX_CUT(1024); break;
    case 2048: BWT_IDX_CUT(2048); break;
    case 4096: BWT_IDX_CUT(4096); break;
    case 8192: BWT_IDX_CUT(8192); break;
    case 16384: BWT_IDX_CUT(16384); break;
    default: BWT_IDX_CUT(32768); break;
  }
#undef BWT_IDX_CUT
  if ((e = hipGetLastError()) != hipSuccess) return int(e);
  // pointer doubling: after k rounds every e^i(0), i < 2^k, is marked
  for (int64_t reach = 1; reach <= n; reach <<= 1) {
    hipLaunchKernelGGL(idx_cut_mark_kernel, gn, blk, 0, st, n, (const int32_t*)J, mark);
    hipLaunchKernelGGL(idx_cut_jump_kernel, gn, blk, 0, st, n, (const int32_t*)J, J2);
    std::swap(J, J2);
  }
  hipLaunchKernelGGL(idx_cut_mark_kernel, gn, blk, 0, st, n, (const int32_t*)J, mark);
  size_t tb = L.temp_bytes;
  if ((e = rocprim::exclusive_scan(at<void>(work, L.temp), tb, (const int32_t*)mark, pos, (int32_t)0, (size_t)n,
                                   rocprim::plus<int32_t>(), st)) != hipSuccess)
    return int(e);
  hipLaunchKernelGGL(idx_cut_scatter_kernel, dim3((unsigned)((n + kT - 1) / kT)), blk, 0, st, n, (const int32_t*)mark,
                     (const int32_t*)pos, blk_step0, nblk);
  return int(hipGetLastError());
}

extern "C" int bwt_index_workspace(const BwtIndexBuild* in, size_t* bytes) {
  if (!build_ok(in) || !bytes) return int(hipErrorInvalidValue);
  WsLayout L;
  const hipError_t e = layout_of(in, L);
  if (e != hipSuccess) return int(e);
  *bytes = L.total;
  return 0;
}

extern "C" int bwt_index_keys_launch(const BwtIndexBuild* in, BwtIndexTables* t, int32_t* nkeys, int32_t* ntiles,
                                     void* work, size_t work_bytes, void* stream) {
  if (!build_ok(in) || !t || !t->evrank || !t->state_step || !t->step_block || !nkeys || !ntiles)
    return int(hipErrorInvalidValue);
  WsLayout L;
  hipError_t e = layout_of(in, L);
  if (e != hipSuccess) return int(e);
  if (!work || work_bytes < L.total) return int(hipErrorInvalidValue);
  hipStream_t st = hipStream_t(stream);
  uint32_t* tmp = at<uint32_t>(work, L.tmp);
  const dim3 g((unsigned)in->nblk), blk(kT);
#define BWT_IDX_KEYS(C)                                                                                        \
  hipLaunchKernelGGL(idx_keys_kernel<C>, g, blk, 0, st, *in, t->evrank, t->state_step, t->step_block, tmp, nkeys, \
                     ntiles)
  switch (cap_for(in->rmax)) {
    case 64: BWT_IDX_KEYS(64); break;
    case 128: BWT_IDX_KEYS(128); break;
    case 256: BWT_IDX_KEYS(256); break;
    case 512: BWT_IDX_KEYS(512); break;
    case 1024: BWT_IDX_KEYS(1024); break;
    case 2048: BWT_IDX_KEYS(2048); break;
    case 4096: BWT_IDX_KEYS(4096); break;
    case 8192: BWT_IDX_KEYS(8192); break;
    default: BWT_IDX_KEYS(16384); break;
  }
#undef BWT_IDX_KEYS
  return int(hipGetLastError());
}

extern "C" int bwt_index_chains_launch(const BwtIndexBuild* in, BwtIndexTables* t, int64_t nkeys_total, int64_t nslots,
                                       void* work, size_t work_bytes, void* stream) {
  if (!build_ok(in) || !t || !t->blk_key0 || !t->blk_tile0 || !t->bkey || !t->tile_rank0 || !t->kitems ||
      !t->chain_key0 || !t->tile_tb0 || !t->tb_block || !t->tb_pos || nkeys_total < 1 || nslots < 1 ||
      (size_t)nslots > slot_cap(in))
    return int(hipErrorInvalidValue);
  WsLayout L;
  hipError_t e = layout_of(in, L);
  if (e != hipSuccess) return int(e);
  if (!work || work_bytes < L.total) return int(hipErrorInvalidValue);
  hipStream_t st = hipStream_t(stream);
  uint32_t* tmp = at<uint32_t>(work, L.tmp);
  uint32_t* slot_tile = at<uint32_t>(work, L.slot_tile);
  uint32_t* slot_tile_s = at<uint32_t>(work, L.slot_tile_s);
  uint32_t* slot_id = at<uint32_t>(work, L.slot_id);
  uint32_t* slot_id_s = at<uint32_t>(work, L.slot_id_s);
  int32_t* slot_block = at<int32_t>(work, L.slot_block);
  int64_t* slot_n = at<int64_t>(work, L.slot_n);
  int64_t* n_sorted = at<int64_t>(work, L.n_sorted);
  int64_t* start = at<int64_t>(work, L.start);
  void* temp = at<void>(work, L.temp);
  const int chg_cap = std::min(in->rmax, in->KT) + 1;
  hipLaunchKernelGGL(idx_write_kernel, dim3((unsigned)in->nblk), dim3(kT), (size_t)chg_cap * sizeof(uint32_t), st, *in,
                     (const uint32_t*)tmp, *t, slot_tile, slot_block, slot_n, slot_id);
  if ((e = hipGetLastError()) != hipSuccess) return int(e);
  size_t tb = L.temp_bytes;
  if ((e = rocprim::radix_sort_pairs(temp, tb, slot_tile, slot_tile_s, slot_id, slot_id_s, (size_t)nslots, 0u,
                                     bits_for((uint64_t)in->KT), st)) != hipSuccess)
    return int(e);
  const unsigned gs = (unsigned)((nslots + kT - 1) / kT);
  hipLaunchKernelGGL(idx_gather_n_kernel, dim3(gs), dim3(kT), 0, st, nslots, (const uint32_t*)slot_id_s,
                     (const int64_t*)slot_n, n_sorted);
  tb = L.temp_bytes;
  if ((e = rocprim::exclusive_scan(temp, tb, (const int64_t*)n_sorted, start, (int64_t)0, (size_t)nslots,
                                   rocprim::plus<int64_t>(), st)) != hipSuccess)
    return int(e);
  int64_t nC = nslots;  // level-C items: one per slot
  hipLaunchKernelGGL(idx_tiles_kernel, dim3((unsigned)((in->KT + 1 + kT - 1) / kT)), dim3(kT), 0, st, in->KT, nslots,
                     nC, (const uint32_t*)slot_tile_s, (const int64_t*)start, (const int64_t*)n_sorted,
                     t->chain_key0, t->tile_tb0);
  hipLaunchKernelGGL(idx_items_kernel, dim3((unsigned)((nslots + kT - 1) / kT)), dim3(kT), 0, st, *in, *t,
                     nslots, nC, (const uint32_t*)slot_tile_s, (const uint32_t*)slot_id_s,
                     (const int32_t*)slot_block, (const int64_t*)n_sorted, (const int64_t*)start);
  if (t->bt_pos) {
    const int64_t nbt = (int64_t)in->nblk * in->KT;
    hipLaunchKernelGGL(idx_btpos_kernel, dim3((unsigned)((nbt + kT - 1) / kT)), dim3(kT), 0, st, in->nblk, in->KT, *t);
  }
  (void)nkeys_total;
  return int(hipGetLastError());
}
