// tree_plan.cpp -- host plan of the generalised batch path (see tree_plan.hpp).
#include "tree_plan.hpp"

#include <omp.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "bwgpu_tree.h"

namespace bwg {

size_t TreePlan::host_bytes() const {
  return steps.size() * 4 + blk_step0.size() * 4 + blk_key0.size() * 8 + blk_tile0.size() * 8 + bkey.size() * 4 +
         tile_rank0.size() * 2 + items.size() * 8 + chunk_item0.size() * 8 + chunk_cp0.size() * 8 +
         chunk_chain.size() * 4 + seg_chunk0.size() * 8 + seg_cp0.size() * 8;
}

namespace {
inline uint64_t make_item(bool ep, uint32_t cls, uint64_t ref) {
  return (ep ? BWT_ITEM_EP : 0ull) | (uint64_t(cls) << BWT_CLASS_SHIFT) | (ref & BWT_REF_MASK);
}
}  // namespace

bool build_tree_plan(int64_t S, int U, int Q, uint64_t NK, const int32_t* parent, const uint32_t* key,
                     const uint32_t* qlo, const uint32_t* qhi, const PlanKnobs& kn, TreePlan& P, char* err,
                     size_t errlen) {
  auto fail = [&](const char* m) {
    if (err && errlen) std::snprintf(err, errlen, "%s", m);
    return false;
  };
  const int ts = kn.tile_shift;
  if (S < 1 || U < 1 || Q < 1 || NK < 1 || NK > 0xFFFFFFFFull) return fail("bad sizes");
  if (ts < 1 || ts > 12) return fail("tile_shift must be in [1,12]");
  const uint64_t KT = (NK >> ts) + 1;
  if (KT > 4096) return fail("key space too large for the tile classes (NK >> tile_shift must be < 4096)");
  if (kn.rmax < U || kn.rmax > 8192) return fail("rmax must be in [U, 8192]");
  if (kn.chunk_items < 64 || kn.seg_chunks < 1) return fail("bad chunking");
  if (uint64_t(2 * S * Q) > BWT_REF_MASK) return fail("too many endpoints");
  for (int64_t v = 1; v <= S; ++v)
    if (parent[v] < 0 || parent[v] >= v) return fail("parent[v] must be in [0, v)");
  const size_t nu = size_t(S) * size_t(U), nq = size_t(S) * size_t(Q);
  for (size_t i = 0; i < nu; ++i)
    if (key[i] >= NK) return fail("key coordinate out of range");
  for (size_t i = 0; i < nq; ++i)
    if (qlo[i] > NK || qhi[i] > NK) return fail("query coordinate out of range");

  P = TreePlan();
  P.S = S; P.U = U; P.Q = Q; P.NK = NK; P.knobs = kn;
  const uint32_t W = 1u << ts, mask = W - 1;

  // ---- 1. children (ascending) and the Euler walk: +v on entry, -v after the subtree
  std::vector<int64_t> koff(size_t(S) + 2, 0);
  std::vector<int32_t> kids(static_cast<size_t>(S));
  for (int64_t v = 1; v <= S; ++v) koff[size_t(parent[v]) + 1]++;
  for (int64_t v = 0; v <= S; ++v) koff[size_t(v) + 1] += koff[size_t(v)];
  {
    std::vector<int64_t> cur(koff.begin(), koff.end() - 1);
    for (int64_t v = 1; v <= S; ++v) kids[size_t(cur[size_t(parent[v])]++)] = int32_t(v);
  }
  P.steps.reserve(size_t(2 * S));
  {
    std::vector<std::pair<int32_t, int64_t>> st;
    st.reserve(1024);
    st.push_back({0, koff[0]});
    while (!st.empty()) {
      auto& top = st.back();
      if (top.second < koff[size_t(top.first) + 1]) {
        const int32_t c = kids[size_t(top.second++)];
        P.steps.push_back(c);
        st.push_back({c, koff[size_t(c)]});
      } else {
        if (top.first != 0) P.steps.push_back(-top.first);
        st.pop_back();
      }
    }
    while (!P.steps.empty() && P.steps.back() < 0) P.steps.pop_back();  // undo after the last query
  }
  const int64_t nsteps = int64_t(P.steps.size());

  // ---- 2. blocks: greedy, at most rmax distinct keys per block
  {
    std::vector<int32_t> seen_blk(NK, -1), seen_step(NK, -1);
    P.blk_step0.push_back(0);
    int32_t b = 0;
    int64_t distinct = 0, in_blk = 0;
    for (int64_t s = 0; s < nsteps; ++s) {
      const int64_t v = P.steps[size_t(s)] > 0 ? P.steps[size_t(s)] : -P.steps[size_t(s)];
      const uint32_t* kk = key + size_t(v - 1) * size_t(U);
      int64_t uniq = 0, fresh = 0;
      for (int i = 0; i < U; ++i) {
        const uint32_t k = kk[i];
        if (seen_step[k] != int32_t(s)) {
          seen_step[k] = int32_t(s);
          ++uniq;
          if (seen_blk[k] != b) ++fresh;
        }
      }
      if (in_blk > 0 && (distinct + fresh > kn.rmax || in_blk >= kn.max_block_steps)) {
        P.blk_step0.push_back(int32_t(s));
        ++b;
        distinct = 0;
        in_blk = 0;
        fresh = uniq;
      }
      for (int i = 0; i < U; ++i) seen_blk[kk[i]] = b;
      distinct += fresh;
      ++in_blk;
    }
    P.blk_step0.push_back(int32_t(nsteps));
  }
  const int32_t nblk = P.nblk();

  // ---- 3. per-block sorted distinct keys, key tiles, enter-step counts
  std::vector<std::vector<uint32_t>> bk(static_cast<size_t>(nblk));
  std::vector<std::vector<uint16_t>> bt(static_cast<size_t>(nblk));
  std::vector<int64_t> nenter(static_cast<size_t>(nblk), 0);
  P.state_block.assign(size_t(S) + 1, -1);
#pragma omp parallel for schedule(dynamic, 16)
  for (int32_t b = 0; b < nblk; ++b) {
    auto& ks = bk[size_t(b)];
    for (int32_t s = P.blk_step0[size_t(b)]; s < P.blk_step0[size_t(b) + 1]; ++s) {
      const int32_t st = P.steps[size_t(s)];
      const int64_t v = st > 0 ? st : -st;
      if (st > 0) { ++nenter[size_t(b)]; P.state_block[size_t(v)] = b; }
      const uint32_t* kk = key + size_t(v - 1) * size_t(U);
      ks.insert(ks.end(), kk, kk + U);
    }
    std::sort(ks.begin(), ks.end());
    ks.erase(std::unique(ks.begin(), ks.end()), ks.end());
    auto& tr = bt[size_t(b)];
    for (size_t r = 0; r < ks.size(); ++r)
      if (r == 0 || (ks[r] >> ts) != (ks[r - 1] >> ts)) tr.push_back(uint16_t(r));
  }
  P.blk_key0.assign(size_t(nblk) + 1, 0);
  P.blk_tile0.assign(size_t(nblk) + 1, 0);
  for (int32_t b = 0; b < nblk; ++b) {
    if (int64_t(bk[size_t(b)].size()) > kn.rmax) return fail("internal: block exceeds rmax");
    P.blk_key0[size_t(b) + 1] = P.blk_key0[size_t(b)] + int64_t(bk[size_t(b)].size());
    P.blk_tile0[size_t(b) + 1] = P.blk_tile0[size_t(b)] + int64_t(bt[size_t(b)].size());
  }
  P.bkey.resize(size_t(P.blk_key0[size_t(nblk)]));
  P.tile_rank0.resize(size_t(P.blk_tile0[size_t(nblk)]));
#pragma omp parallel for schedule(static)
  for (int32_t b = 0; b < nblk; ++b) {
    std::copy(bk[size_t(b)].begin(), bk[size_t(b)].end(), P.bkey.begin() + P.blk_key0[size_t(b)]);
    std::copy(bt[size_t(b)].begin(), bt[size_t(b)].end(), P.tile_rank0.begin() + P.blk_tile0[size_t(b)]);
  }
  bt.clear();
  bt.shrink_to_fit();

  // per block: its endpoints in step order; endpoint id 2*((v-1)*Q+k)+side
  auto for_block_eps = [&](int32_t b, auto&& fn) {
    for (int32_t s = P.blk_step0[size_t(b)]; s < P.blk_step0[size_t(b) + 1]; ++s) {
      const int32_t st = P.steps[size_t(s)];
      if (st <= 0) continue;
      const size_t qb = size_t(st - 1) * size_t(Q);
      for (int k = 0; k < Q; ++k) {
        fn(uint64_t(2 * (qb + size_t(k))), qlo[qb + size_t(k)]);
        fn(uint64_t(2 * (qb + size_t(k)) + 1), qhi[qb + size_t(k)]);
      }
    }
  };

  // ---- 4. level C: one chain over all blocks, class = key tile
  std::vector<int64_t> cstart(size_t(nblk) + 1, 0);
  for (int32_t b = 0; b < nblk; ++b)
    cstart[size_t(b) + 1] = cstart[size_t(b)] + 2 * int64_t(Q) * nenter[size_t(b)] +
                            (P.blk_tile0[size_t(b) + 1] - P.blk_tile0[size_t(b)]);
  const int64_t nC = cstart[size_t(nblk)];
  // ---- 5. level B: one chain per tile, stable counting sort by tile of
  //          [block 0 endpoints, block 0 keys, block 1 endpoints, block 1 keys, ...]
  const int nth = std::max(1, omp_get_max_threads());
  std::vector<int64_t> hist(size_t(nth) * KT, 0);
  std::vector<int32_t> tb0(size_t(nth) + 1);
  for (int t = 0; t <= nth; ++t) tb0[size_t(t)] = int32_t(int64_t(nblk) * t / nth);
#pragma omp parallel num_threads(nth)
  {
    const int t = omp_get_thread_num();
    int64_t* h = &hist[size_t(t) * KT];
    for (int32_t b = tb0[size_t(t)]; b < tb0[size_t(t) + 1]; ++b) {
      for_block_eps(b, [&](uint64_t, uint32_t x) { h[x >> ts]++; });
      for (int64_t r = P.blk_key0[size_t(b)]; r < P.blk_key0[size_t(b) + 1]; ++r) h[P.bkey[size_t(r)] >> ts]++;
    }
  }
  std::vector<int64_t> tile_tot(KT, 0);
  for (uint64_t tile = 0; tile < KT; ++tile)
    for (int t = 0; t < nth; ++t) tile_tot[tile] += hist[size_t(t) * KT + tile];
  // chain ids: 0 = level C; then every tile with items
  P.chain_ncls.push_back(int32_t(KT));
  P.chain_src.push_back(1);
  P.chain_out.push_back(0);
  std::vector<int64_t> tile_base(KT, -1);
  int64_t nB = 0;
  std::vector<std::pair<int64_t, int64_t>> chain_rng;  // item ranges
  chain_rng.push_back({0, nC});
  for (uint64_t tile = 0; tile < KT; ++tile) {
    if (tile_tot[tile] == 0) continue;
    tile_base[tile] = nC + nB;
    chain_rng.push_back({nC + nB, nC + nB + tile_tot[tile]});
    nB += tile_tot[tile];
    P.chain_ncls.push_back(int32_t(W));
    P.chain_src.push_back(0);
    P.chain_out.push_back(1);
  }
  // per (thread, tile) write cursors
  std::vector<int64_t> cur(size_t(nth) * KT, 0);
  for (uint64_t tile = 0; tile < KT; ++tile) {
    int64_t o = tile_base[tile];
    for (int t = 0; t < nth; ++t) {
      cur[size_t(t) * KT + tile] = o;
      o += hist[size_t(t) * KT + tile];
    }
  }
  P.items.resize(size_t(nC + nB));
  uint64_t* items = P.items.data();
#pragma omp parallel num_threads(nth)
  {
    const int t = omp_get_thread_num();
    int64_t* c = &cur[size_t(t) * KT];
    for (int32_t b = tb0[size_t(t)]; b < tb0[size_t(t) + 1]; ++b) {
      int64_t oc = cstart[size_t(b)];
      for_block_eps(b, [&](uint64_t ep, uint32_t x) {
        items[oc++] = make_item(true, x >> ts, ep);
        items[c[x >> ts]++] = make_item(true, x & mask, ep);
      });
      for (int64_t s = P.blk_tile0[size_t(b)]; s < P.blk_tile0[size_t(b) + 1]; ++s) {
        const uint32_t tile = P.bkey[size_t(P.blk_key0[size_t(b)] + P.tile_rank0[size_t(s)])] >> ts;
        items[oc++] = make_item(false, tile, uint64_t(s));
      }
      for (int64_t r = P.blk_key0[size_t(b)]; r < P.blk_key0[size_t(b) + 1]; ++r) {
        const uint32_t k = P.bkey[size_t(r)];
        items[c[k >> ts]++] = make_item(false, k & mask, uint64_t(r));
      }
    }
  }

  // ---- 6. chunks and scan segments
  P.chain_chunk0.push_back(0);
  P.chain_seg0.push_back(0);
  P.seg_chunk0.push_back(0);
  int64_t cp = 0, sg = 0;
  for (int32_t ch = 0; ch < int32_t(chain_rng.size()); ++ch) {
    const int64_t a = chain_rng[size_t(ch)].first, e = chain_rng[size_t(ch)].second;
    const int ncls = P.chain_ncls[size_t(ch)];
    int64_t nck = 0;
    for (int64_t i = a; i < e || (nck == 0 && i == a); i += kn.chunk_items) {
      P.chunk_item0.push_back(i);
      P.chunk_chain.push_back(ch);
      P.chunk_cp0.push_back(cp);
      cp += ncls;
      ++nck;
      if (nck % kn.seg_chunks == 0) {
        P.seg_chunk0.push_back(int64_t(P.chunk_chain.size()));
        P.seg_cp0.push_back(sg);
        sg += ncls;
      }
    }
    if (nck % kn.seg_chunks != 0) {
      P.seg_chunk0.push_back(int64_t(P.chunk_chain.size()));
      P.seg_cp0.push_back(sg);
      sg += ncls;
    }
    P.chain_chunk0.push_back(int64_t(P.chunk_chain.size()));
    P.chain_seg0.push_back(int64_t(P.seg_cp0.size()));
  }
  P.chunk_item0.push_back(nC + nB);
  P.n_cp = cp;
  P.n_segsum = sg;
  P.max_cls = int(std::max<uint64_t>(KT, W));
  return true;
}

}  // namespace bwg
