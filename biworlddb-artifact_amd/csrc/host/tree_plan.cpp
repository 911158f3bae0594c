// tree_plan.cpp -- host plans of the generalised batch path (see tree_plan.hpp).
#include "tree_plan.hpp"

#include <omp.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "bwgpu_tree.h"
#include "plan_common.hpp"

namespace bwg {

namespace {
inline uint64_t key_item(uint32_t cls, uint64_t ref) { return (uint64_t(cls) << BWT_CLASS_SHIFT) | (ref & BWT_REF_MASK); }
inline uint64_t ep_item(uint64_t pos, uint32_t cls, uint64_t ep) {
  return (pos << BWT_EP_POS_SHIFT) | (uint64_t(cls) << BWT_EP_CLASS_SHIFT) | (ep & BWT_EP_ID_MASK);
}
}  // namespace

bool plan_walk_blocks(int64_t S, int U, uint64_t NK, const int32_t* parent, const uint32_t* key,
                      const PlanKnobs& kn, TreeIndexPlan& P, char* err, size_t errlen, bool cut, bool walk) {
  auto fail = [&](const char* m) {
    if (err && errlen) std::snprintf(err, errlen, "%s", m);
    return false;
  };
  if (S < 1 || U < 1 || NK < 1 || NK > 0xFFFFFFFFull) return fail("bad sizes");
  int ts = kn.tile_shift;
  if (ts <= 0) {  // balanced split: the smallest max(level-C classes, level-B classes)
    uint64_t best = ~0ull;
    for (int t = 1; t <= 12; ++t) {
      const uint64_t m = std::max<uint64_t>((NK >> t) + 1, uint64_t(1) << t);
      if (m < best) { best = m; ts = t; }
    }
  }
  if (ts < 1 || ts > 12) return fail("tile_shift must be in [1,12]");
  const uint64_t KT = (NK >> ts) + 1;
  if (KT > BWT_MAX_CLASSES - 1)
    return fail("key space too large for the tile classes (NK >> tile_shift must be < BWT_MAX_CLASSES - 2)");
  if (kn.rmax < U || kn.rmax > 8192) return fail("rmax must be in [U, 8192]");
  if ((kn.chunk_items != 0 && (kn.chunk_items < 64 || kn.chunk_items >= (int64_t(1) << 19))) ||
      kn.seg_chunks < 1 || kn.ep_per_work < 64)
    return fail("bad chunking");
  for (int64_t v = 1; v <= S; ++v)
    if (parent[v] < 0 || parent[v] >= v) return fail("parent[v] must be in [0, v)");
  const size_t nu = size_t(S) * size_t(U);
  if (key)  // (keys held only on the device are checked there)
    for (size_t i = 0; i < nu; ++i)
      if (key[i] >= NK) return fail("key coordinate out of range");

  P = TreeIndexPlan();
  P.S = S; P.U = U; P.NK = NK; P.knobs = kn; P.knobs.tile_shift = ts; P.KT = int(KT);
  if (!walk) return true;  // the checks only

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
    while (!P.steps.empty() && P.steps.back() < 0) P.steps.pop_back();  // undo after the last state
  }
  const int64_t nsteps = int64_t(P.steps.size());
  if (nsteps >= (int64_t(1) << 31) / U) return fail("too many steps");

  // ---- 2. blocks: greedy, at most rmax distinct keys per block
  if (cut) {
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
  return true;
}

bool plan_block_tables(const uint32_t* key, TreeIndexPlan& P, char* err, size_t errlen) {
  auto fail = [&](const char* m) {
    if (err && errlen) std::snprintf(err, errlen, "%s", m);
    return false;
  };
  const int64_t S = P.S;
  const int U = P.U;
  const PlanKnobs& kn = P.knobs;
  const int ts = kn.tile_shift;
  const uint64_t KT = uint64_t(P.KT);
  const uint32_t W = 1u << ts, mask = W - 1;
  const int64_t nsteps = int64_t(P.steps.size());
  const int32_t nblk = P.nblk();

  // ---- 3. per block: sorted distinct keys, key tiles, in-block ranks of every event.  A block's
  //          events go through an open-addressing table (tagged with the block, so it is never
  //          cleared): its distinct keys are collected, sorted, and their ranks stored back in the
  //          table, from which every event reads its rank.
  std::vector<std::vector<uint32_t>> bk(static_cast<size_t>(nblk));
  std::vector<std::vector<uint16_t>> bt(static_cast<size_t>(nblk));
  P.state_block.assign(size_t(S) + 1, -1);
  P.state_step.assign(size_t(S) + 1, -1);
  P.evrank.resize(size_t(nsteps) * size_t(U));
  int tbits = 1;
  while ((1 << tbits) < 2 * kn.rmax) ++tbits;
#pragma omp parallel
  {
    const size_t cap = size_t(1) << tbits;
    std::vector<uint32_t> tkey(cap);
    std::vector<int32_t> ttag(cap, -1);
    std::vector<uint16_t> trank(cap);
    auto slot_of = [&](uint32_t k, int32_t tag, bool& found) {
      size_t h = size_t((k * 0x9E3779B1u) >> (32 - tbits));
      while (ttag[h] == tag) {
        if (tkey[h] == k) {
          found = true;
          return h;
        }
        h = (h + 1) & (cap - 1);
      }
      found = false;
      return h;
    };
#pragma omp for schedule(dynamic, 16)
    for (int32_t b = 0; b < nblk; ++b) {
      auto& ks = bk[size_t(b)];
      bool ok = true;
      for (int32_t s = P.blk_step0[size_t(b)]; s < P.blk_step0[size_t(b) + 1] && ok; ++s) {
        const int32_t st = P.steps[size_t(s)];
        const int64_t v = st > 0 ? st : -st;
        if (st > 0) { P.state_block[size_t(v)] = b; P.state_step[size_t(v)] = s; }
        const uint32_t* kk = key + size_t(v - 1) * size_t(U);
        for (int i = 0; i < U; ++i) {
          bool found;
          const size_t h = slot_of(kk[i], b, found);
          if (found) continue;
          if (int64_t(ks.size()) >= kn.rmax) { ok = false; break; }  // cannot happen: the cut bounds it
          tkey[h] = kk[i];
          ttag[h] = b;
          ks.push_back(kk[i]);
        }
      }
      if (!ok) { ks.assign(size_t(kn.rmax) + 1, 0); continue; }  // reported below
      std::sort(ks.begin(), ks.end());
      for (size_t r = 0; r < ks.size(); ++r) {
        bool found;
        trank[slot_of(ks[r], b, found)] = uint16_t(r);
      }
      for (int32_t s = P.blk_step0[size_t(b)]; s < P.blk_step0[size_t(b) + 1]; ++s) {
        const int32_t st = P.steps[size_t(s)];
        const int64_t v = st > 0 ? st : -st;
        const uint32_t* kk = key + size_t(v - 1) * size_t(U);
        uint16_t* er = &P.evrank[size_t(s) * size_t(U)];
        for (int i = 0; i < U; ++i) {
          bool found;
          er[i] = trank[slot_of(kk[i], b, found)];
        }
      }
      auto& tr = bt[size_t(b)];
      for (size_t r = 0; r < ks.size(); ++r)
        if (r == 0 || (ks[r] >> ts) != (ks[r - 1] >> ts)) tr.push_back(uint16_t(r));
    }
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
  bk.clear();
  bk.shrink_to_fit();
  bt.clear();
  bt.shrink_to_fit();

  // ---- 4. chains.  Level C (chain 0): every block's tile slots in block order, class = tile.
  //          Level B (chain 1+t): every block's keys of tile t in block order, class = key & mask.
  const int64_t nC = P.blk_tile0[size_t(nblk)];
  const int nth = std::max(1, omp_get_max_threads());
  std::vector<int64_t> hist(size_t(nth) * KT, 0);
  std::vector<int32_t> tb0(size_t(nth) + 1);
  for (int t = 0; t <= nth; ++t) tb0[size_t(t)] = int32_t(int64_t(nblk) * t / nth);
#pragma omp parallel num_threads(nth)
  {
    const int t = omp_get_thread_num();
    int64_t* h = &hist[size_t(t) * KT];
    for (int32_t b = tb0[size_t(t)]; b < tb0[size_t(t) + 1]; ++b)
      for (int64_t r = P.blk_key0[size_t(b)]; r < P.blk_key0[size_t(b) + 1]; ++r) h[P.bkey[size_t(r)] >> ts]++;
  }
  std::vector<int64_t> tile_tot(KT, 0);
  for (uint64_t tile = 0; tile < KT; ++tile)
    for (int t = 0; t < nth; ++t) tile_tot[tile] += hist[size_t(t) * KT + tile];
  P.chain_key0.assign(KT + 2, 0);
  P.chain_key0[1] = nC;
  for (uint64_t tile = 0; tile < KT; ++tile) P.chain_key0[tile + 2] = P.chain_key0[tile + 1] + tile_tot[tile];
  std::vector<int64_t> cur(size_t(nth) * KT, 0);
  for (uint64_t tile = 0; tile < KT; ++tile) {
    int64_t o = P.chain_key0[tile + 1];
    for (int t = 0; t < nth; ++t) {
      cur[size_t(t) * KT + tile] = o;
      o += hist[size_t(t) * KT + tile];
    }
  }
  P.kitems.resize(size_t(P.chain_key0[KT + 1]));
  uint64_t* items = P.kitems.data();
#pragma omp parallel num_threads(nth)
  {
    const int t = omp_get_thread_num();
    int64_t* c = &cur[size_t(t) * KT];
    for (int32_t b = tb0[size_t(t)]; b < tb0[size_t(t) + 1]; ++b) {
      for (int64_t s = P.blk_tile0[size_t(b)]; s < P.blk_tile0[size_t(b) + 1]; ++s) {
        const uint32_t tile = P.bkey[size_t(P.blk_key0[size_t(b)] + P.tile_rank0[size_t(s)])] >> ts;
        items[s] = key_item(tile, uint64_t(s));
      }
      for (int64_t r = P.blk_key0[size_t(b)]; r < P.blk_key0[size_t(b) + 1]; ++r) {
        const uint32_t k = P.bkey[size_t(r)];
        items[c[k >> ts]++] = key_item(k & mask, uint64_t(r));
      }
    }
  }
  // level-B positions: per tile, the blocks holding its keys (in block order) and where their
  // keys start on the tile's chain.  The blocks are split among the threads as for the items
  // above; a thread's slots of a tile follow those of the threads before it.
  {
    std::vector<int64_t> scnt(size_t(nth) * KT, 0);
    auto tile_of_slot = [&](int32_t b, int64_t s) {
      return P.bkey[size_t(P.blk_key0[size_t(b)] + P.tile_rank0[size_t(s)])] >> ts;
    };
#pragma omp parallel num_threads(nth)
    {
      const int t = omp_get_thread_num();
      int64_t* c = &scnt[size_t(t) * KT];
      for (int32_t b = tb0[size_t(t)]; b < tb0[size_t(t) + 1]; ++b)
        for (int64_t s = P.blk_tile0[size_t(b)]; s < P.blk_tile0[size_t(b) + 1]; ++s) c[tile_of_slot(b, s)]++;
    }
    P.tile_tb0.assign(KT + 1, 0);
    for (uint64_t tile = 0; tile < KT; ++tile) {
      int64_t n = 0;
      for (int t = 0; t < nth; ++t) n += scnt[size_t(t) * KT + tile];
      P.tile_tb0[tile + 1] = P.tile_tb0[tile] + n;
    }
    std::vector<int64_t> fo(size_t(nth) * KT), ro(size_t(nth) * KT);
    for (uint64_t tile = 0; tile < KT; ++tile) {
      int64_t f = P.tile_tb0[tile], r = 0;
      for (int t = 0; t < nth; ++t) {
        fo[size_t(t) * KT + tile] = f;
        ro[size_t(t) * KT + tile] = r;
        f += scnt[size_t(t) * KT + tile];
        r += hist[size_t(t) * KT + tile];
      }
    }
    P.tb_block.resize(size_t(P.tile_tb0[KT]));
    P.tb_pos.resize(size_t(P.tile_tb0[KT]));
#pragma omp parallel num_threads(nth)
    {
      const int t = omp_get_thread_num();
      int64_t* f = &fo[size_t(t) * KT];
      int64_t* r = &ro[size_t(t) * KT];
      for (int32_t b = tb0[size_t(t)]; b < tb0[size_t(t) + 1]; ++b) {
        const int64_t t0 = P.blk_tile0[size_t(b)], t1 = P.blk_tile0[size_t(b) + 1];
        const int64_t R = P.blk_key0[size_t(b) + 1] - P.blk_key0[size_t(b)];
        for (int64_t s = t0; s < t1; ++s) {
          const uint32_t tile = tile_of_slot(b, s);
          const int64_t n = ((s + 1 < t1) ? P.tile_rank0[size_t(s) + 1] : R) - P.tile_rank0[size_t(s)];
          P.tb_block[size_t(f[tile])] = b;
          P.tb_pos[size_t(f[tile]++)] = r[tile];
          r[tile] += n;
        }
      }
    }
  }
  (void)fail;
  return true;
}

void plan_chunks(TreeIndexPlan& P) {
  const PlanKnobs& kn = P.knobs;
  const uint64_t KT = uint64_t(P.KT);
  const uint32_t W = 1u << kn.tile_shift;
  const int32_t nblk = P.nblk();
  const int64_t nC = P.blk_tile0[size_t(nblk)];
  // the chains: 0 = level C (classes = tiles), 1 + t = level B of tile t (classes = keys in a tile)
  P.chain_ncls.assign(1, int32_t(KT));
  P.chain_src.assign(1, 1);
  P.chain_out.assign(1, 0);
  for (uint64_t tile = 0; tile < KT; ++tile) {
    P.chain_ncls.push_back(int32_t(W));
    P.chain_src.push_back(0);
    P.chain_out.push_back(1);
  }
  // ---- 5. chunks (at least one per chain) and carry-scan segments.  Level C: one chunk per
  //          block (its tile slots), so every level-C endpoint sits at a chunk start and its
  //          carry is a plain read of the block's dense row.  Level B: chunks of chunk_items key
  //          items; the auto size (knob 0) keeps the dense rows near 64 MiB: the smallest power
  //          of two >= 1024 with (key items / chunk) * classes * 8 B <= 64 MiB.
  auto auto_chunk = [](int64_t items, int64_t ncls) {
    int64_t c = 1024;
    while (c < (int64_t(1) << 18) && (items / c) * ncls * 8 > (int64_t(64) << 20)) c <<= 1;
    return c;
  };
  if (P.knobs.chunk_items == 0) P.knobs.chunk_items = auto_chunk(P.chain_key0[KT + 1] - nC, int64_t(W));
  P.chain_chunk0.push_back(0);
  P.chain_seg0.push_back(0);
  P.seg_chunk0.push_back(0);
  int64_t cp = 0, sg = 0;
  for (int32_t ch = 0; ch < P.nchain(); ++ch) {
    const int64_t CH = P.knobs.chunk_items;
    const int64_t a = P.chain_key0[size_t(ch)], e = P.chain_key0[size_t(ch) + 1];
    const int ncls = P.chain_ncls[size_t(ch)];
    const int64_t nck = ch == 0 ? int64_t(nblk) : std::max<int64_t>(1, (e - a + CH - 1) / CH);
    for (int64_t k = 0; k < nck; ++k) {
      P.chunk_key0.push_back(ch == 0 ? P.blk_tile0[size_t(k)] : a + k * CH);
      P.chunk_chain.push_back(ch);
      P.chunk_cp0.push_back(cp);
      cp += ncls;
      if ((k + 1) % kn.seg_chunks == 0 || k + 1 == nck) {
        P.seg_chunk0.push_back(int64_t(P.chunk_chain.size()));
        P.seg_cp0.push_back(sg);
        sg += ncls;
      }
    }
    P.chain_chunk0.push_back(int64_t(P.chunk_chain.size()));
    P.chain_seg0.push_back(int64_t(P.seg_cp0.size()));
  }
  P.chunk_key0.push_back(P.chain_key0.back());
  P.n_cp = cp;
  P.n_segsum = sg;
  P.max_cls = int(std::max<uint64_t>(KT, W));
}

bool build_tree_index(int64_t S, int U, uint64_t NK, const int32_t* parent, const uint32_t* key,
                      const PlanKnobs& knobs, TreeIndexPlan& P, char* err, size_t errlen) {
  if (!plan_walk_blocks(S, U, NK, parent, key, knobs, P, err, errlen)) return false;
  if (!plan_block_tables(key, P, err, errlen)) return false;
  plan_chunks(P);
  return true;
}

bool build_batch_plan(const TreeIndexPlan& X, int64_t M, const int32_t* qstate, const uint32_t* qlo,
                      const uint32_t* qhi, bool all_blocks, BatchPlan& B, char* err, size_t errlen) {
  auto fail = [&](const char* m) {
    if (err && errlen) std::snprintf(err, errlen, "%s", m);
    return false;
  };
  if (M < 0 || M > BWT_MAX_BATCH) return fail("bad batch size");
  for (int64_t q = 0; q < M; ++q)
    if (qstate[q] < 1 || qstate[q] > X.S || qlo[q] > X.NK || qhi[q] > X.NK) return fail("query out of range");
  B = BatchPlan();
  B.M = M;
  const int32_t nblk = X.nblk();
  const int64_t nsteps = int64_t(X.steps.size());
  const int64_t nchunk = X.nchunk();
  const int ts = X.knobs.tile_shift;
  const uint32_t mask = (1u << ts) - 1;
  const int nth = std::max(1, omp_get_max_threads());
  // Small batches (a few queries per thousand steps) are sorted; large ones counting-sorted over
  // every step and chunk.  Both give the same plan: slots by (enter step, query id), endpoint
  // items by (chunk, position, query order).
  const bool sparse = M * 8 < nsteps;
  // ---- slots: queries ordered by enter step (stable), hence grouped by block
  B.qorder.resize(size_t(M));
  B.islot.resize(size_t(M));
  if (sparse) {
    std::vector<uint64_t> key(static_cast<size_t>(M));
    for (int64_t q = 0; q < M; ++q) key[size_t(q)] = (uint64_t(X.state_step[size_t(qstate[q])]) << 32) | uint64_t(q);
    std::sort(key.begin(), key.end());
    for (int64_t at = 0; at < M; ++at) {
      const uint32_t q = uint32_t(key[size_t(at)] & 0xFFFFFFFFu);
      B.qorder[size_t(at)] = q;
      B.islot[size_t(q)] = uint32_t(at);
    }
  } else {
    std::vector<int64_t> scnt(size_t(nsteps) + 1, 0);
    for (int64_t q = 0; q < M; ++q) scnt[size_t(X.state_step[size_t(qstate[q])]) + 1]++;
    for (int64_t s = 0; s < nsteps; ++s) scnt[size_t(s) + 1] += scnt[size_t(s)];
    std::vector<int64_t> f(scnt.begin(), scnt.end() - 1);
    for (int64_t q = 0; q < M; ++q) {
      const int64_t at = f[size_t(X.state_step[size_t(qstate[q])])]++;
      B.qorder[size_t(at)] = uint32_t(q);
      B.islot[size_t(q)] = uint32_t(at);
    }
  }
  // ---- staging positions of the combine: grouped by window of query ids, then by the die part
  //      of the slot (the combine's XCD-major workgroup order), then slot
  {
    const int64_t nt = (M + BWT_COMBINE_TILE - 1) / BWT_COMBINE_TILE;
    const int64_t per = std::max<int64_t>(1, (nt + BWT_XCDS - 1) / BWT_XCDS);
    const int sh = std::min(31, std::max(8, X.knobs.stage_shift));
    const int64_t nwin = (M >> sh) + 1;
    auto seg = [&](int64_t j) {
      return (int64_t(B.qorder[size_t(j)]) >> sh) * BWT_XCDS + (j / BWT_COMBINE_TILE) / per;
    };
    std::vector<int64_t> base(size_t(nwin * BWT_XCDS) + 1, 0);
    for (int64_t j = 0; j < M; ++j) base[size_t(seg(j)) + 1]++;
    for (size_t k = 1; k < base.size(); ++k) base[k] += base[k - 1];
    std::vector<uint32_t> at(static_cast<size_t>(M));
    B.ipos.resize(size_t(M));
    for (int64_t j = 0; j < M; ++j) {
      at[size_t(j)] = uint32_t(base[size_t(seg(j))]++);
      B.ipos[size_t(B.qorder[size_t(j)])] = at[size_t(j)];
    }
    // each tile's positions ascending, and every slot's index in that list
    B.spos.resize(size_t(M));
    B.srank.resize(size_t(M));
#pragma omp parallel for schedule(static) if (nt > 64)
    for (int64_t t = 0; t < nt; ++t) {
      const int64_t t0 = t * BWT_COMBINE_TILE, n = std::min<int64_t>(BWT_COMBINE_TILE, M - t0);
      std::vector<uint64_t> key(static_cast<size_t>(n));
      for (int64_t k = 0; k < n; ++k) key[size_t(k)] = (uint64_t(at[size_t(t0 + k)]) << 16) | uint64_t(k);
      std::sort(key.begin(), key.end());
      for (int64_t r = 0; r < n; ++r) {
        B.spos[size_t(t0 + r)] = uint32_t(key[size_t(r)] >> 16);
        B.srank[size_t(t0 + int64_t(key[size_t(r)] & 0xFFFFu))] = uint16_t(r);
      }
    }
  }
  // ---- listed blocks and their step groups (the queried steps, ascending); with all_blocks,
  //      blocks without queries are listed too (with no groups)
  B.qblk_g0.push_back(0);
  B.grp_q0.push_back(0);
  {
    int64_t j = 0;  // next slot
    int32_t b = 0;
    while (true) {
      const int32_t nb = j < M ? X.state_block[size_t(qstate[B.qorder[size_t(j)]])] : nblk;
      if (all_blocks)
        for (; b < nb; ++b) {
          B.qblk.push_back(b);
          B.qblk_g0.push_back(int64_t(B.grp_step.size()));
        }
      if (j >= M) break;
      b = nb;
      B.qblk.push_back(b);
      const int32_t sa = X.blk_step0[size_t(b)];
      int32_t last = sa;
      while (j < M && X.state_block[size_t(qstate[B.qorder[size_t(j)]])] == b) {
        const int32_t s = X.state_step[size_t(qstate[B.qorder[size_t(j)]])];
        while (j < M && X.state_step[size_t(qstate[B.qorder[size_t(j)]])] == s) ++j;
        B.grp_step.push_back(s);
        B.grp_q0.push_back(j);
        last = s;
      }
      B.qblk_g0.push_back(int64_t(B.grp_step.size()));
      B.steps_run += last + 1 - sa;
      ++b;
    }
  }
  // ---- per slot: block, in-block ranks and key tiles of its endpoints
  B.qrank.resize(size_t(M));
  B.sblk.resize(size_t(M));
  B.qtile.resize(size_t(M));
#pragma omp parallel for schedule(static) if (M > 4096)
  for (int64_t j = 0; j < M; ++j) {
    const uint32_t q = B.qorder[size_t(j)];
    const int32_t b = X.state_block[size_t(qstate[q])];
    const uint32_t* k0 = X.bkey.data() + X.blk_key0[size_t(b)];
    const uint32_t* k1 = X.bkey.data() + X.blk_key0[size_t(b) + 1];
    const uint32_t rl = uint32_t(std::lower_bound(k0, k1, qlo[q]) - k0);
    const uint32_t rh = uint32_t(std::lower_bound(k0, k1, qhi[q]) - k0);
    B.qrank[size_t(j)] = rl | (rh << 16);
    B.sblk[size_t(j)] = b;
    B.qtile[size_t(j)] = (qlo[q] >> ts) | ((qhi[q] >> ts) << 16);
  }
  // ---- level-B endpoint items: (chunk, item) for every endpoint, in slot order
  const int64_t nE = 2 * M;
  std::vector<int64_t> ech(static_cast<size_t>(nE));
  std::vector<uint64_t> eit(static_cast<size_t>(nE));
  const int64_t CH = X.knobs.chunk_items;
#pragma omp parallel for schedule(static) if (M > 4096)
  for (int64_t j = 0; j < M; ++j) {
    const uint32_t q = B.qorder[size_t(j)];
    const int32_t b = B.sblk[size_t(j)];
    for (int side = 0; side < 2; ++side) {
      const uint32_t x = side ? qhi[q] : qlo[q];
      // the keys of tile t in blocks < b precede the endpoint on chain 1 + t
      const uint32_t t = x >> ts;
      const int64_t lo = X.tile_tb0[t], hi = X.tile_tb0[size_t(t) + 1];
      const int64_t idx = std::lower_bound(X.tb_block.begin() + lo, X.tb_block.begin() + hi, b) - X.tb_block.begin();
      const int64_t pos = idx < hi ? X.tb_pos[size_t(idx)] : X.chain_key0[size_t(t) + 2] - X.chain_key0[size_t(t) + 1];
      const int32_t chain = 1 + int32_t(t);
      const int64_t n = X.chain_chunk0[size_t(chain) + 1] - X.chain_chunk0[size_t(chain)];
      int64_t k = pos / CH;
      if (k >= n) k = n - 1;
      ech[size_t(2 * j + side)] = X.chain_chunk0[size_t(chain)] + k;
      eit[size_t(2 * j + side)] = ep_item(uint64_t(pos - k * CH), x & mask, uint64_t(2 * j) + uint64_t(side));
    }
  }
  // ---- group by chunk, each chunk's items by position (ties keep slot order)
  std::vector<int64_t> cfirst;  // chunks holding endpoints, ascending, and their item ranges
  std::vector<int64_t> cstart;
  B.eitems.resize(size_t(nE));
  if (sparse) {
    std::vector<uint32_t> ord(static_cast<size_t>(nE));
    for (int64_t i = 0; i < nE; ++i) ord[size_t(i)] = uint32_t(i);
    std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) {
      if (ech[a] != ech[b]) return ech[a] < ech[b];
      const uint64_t pa = BWT_EP_POS(eit[a]), pb = BWT_EP_POS(eit[b]);
      return pa != pb ? pa < pb : a < b;
    });
    for (int64_t i = 0; i < nE; ++i) {
      const int64_t c = ech[ord[size_t(i)]];
      if (cfirst.empty() || cfirst.back() != c) {
        cfirst.push_back(c);
        cstart.push_back(i);
      }
      B.eitems[size_t(i)] = eit[ord[size_t(i)]];
    }
    cstart.push_back(nE);
  } else {
    std::vector<int64_t> ccnt(size_t(nth) * size_t(nchunk), 0);
    std::vector<int64_t> e0(size_t(nth) + 1);
    for (int t = 0; t <= nth; ++t) e0[size_t(t)] = nE * t / nth;
#pragma omp parallel num_threads(nth)
    {
      const int t = omp_get_thread_num();
      int64_t* h = &ccnt[size_t(t) * size_t(nchunk)];
      for (int64_t i = e0[size_t(t)]; i < e0[size_t(t) + 1]; ++i) h[size_t(ech[size_t(i)])]++;
    }
    std::vector<int64_t> all(size_t(nchunk) + 1, 0);
    for (int64_t c = 0; c < nchunk; ++c) {
      int64_t s = 0;
      for (int t = 0; t < nth; ++t) s += ccnt[size_t(t) * size_t(nchunk) + size_t(c)];
      all[size_t(c) + 1] = all[size_t(c)] + s;
    }
    {
      std::vector<int64_t> pos(all.begin(), all.end() - 1);
      for (int t = 0; t < nth; ++t)
        for (int64_t c = 0; c < nchunk; ++c) {
          int64_t& slot = ccnt[size_t(t) * size_t(nchunk) + size_t(c)];
          const int64_t n = slot;
          slot = pos[size_t(c)];
          pos[size_t(c)] += n;
        }
    }
#pragma omp parallel num_threads(nth)
    {
      const int t = omp_get_thread_num();
      int64_t* c = &ccnt[size_t(t) * size_t(nchunk)];
      for (int64_t i = e0[size_t(t)]; i < e0[size_t(t) + 1]; ++i)
        B.eitems[size_t(c[size_t(ech[size_t(i)])]++)] = eit[size_t(i)];
    }
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t c = 0; c < nchunk; ++c)
      std::stable_sort(B.eitems.begin() + all[size_t(c)], B.eitems.begin() + all[size_t(c) + 1],
                       [](uint64_t a, uint64_t b) { return BWT_EP_POS(a) < BWT_EP_POS(b); });
    for (int64_t c = 0; c < nchunk; ++c)
      if (all[size_t(c) + 1] > all[size_t(c)]) {
        cfirst.push_back(c);
        cstart.push_back(all[size_t(c)]);
      }
    cstart.push_back(nE);
  }
  // work items: the endpoints of each (slot range, chunk) run -- a chunk's items are in slot
  // order (position follows the block, ties keep slot order), so a range is one run of them --
  // cut at every ep_per_work-th item from the run's start (a work item replays the chunk's keys
  // only up to its own last endpoint), a cut that would part a slot's two endpoints moved one
  // item on; ordered by slot range, then chunk.  The device planner makes the same cuts.
  const int64_t G = std::max<int64_t>(2, X.knobs.ep_per_work);
  const SlotRanges SR = slot_ranges(M, X.knobs.range_slots);
  auto slot_of = [&](int64_t e) { return int64_t(BWT_EP_ID(B.eitems[size_t(e)])) >> 1; };
  struct Work {
    int64_t range, chunk, e0, e1;
  };
  std::vector<Work> wk;
  for (size_t i = 0; i < cfirst.size(); ++i) {
    int64_t a0 = cstart[i];
    const int64_t ce = cstart[i + 1];
    while (a0 < ce) {
      const int64_t r = range_of(slot_of(a0), SR);
      const int64_t e = std::partition_point(B.eitems.begin() + a0, B.eitems.begin() + ce, [&](uint64_t it) {
                          return range_of(int64_t(BWT_EP_ID(it)) >> 1, SR) == r;
                        }) - B.eitems.begin();
      int64_t a = a0;
      for (int64_t k = 1; a < e; ++k) {
        const int64_t c = a0 + k * G;
        const int64_t b = c >= e ? e : c + (slot_of(c) == slot_of(c - 1) ? 1 : 0);
        wk.push_back(Work{r, cfirst[i], a, b});
        a = b;
      }
      a0 = e;
    }
  }
  for (const Work& w : wk)  // pairs: a slot's two endpoints on one chain (same tile), adjacent
    for (int64_t i = w.e0; i + 1 < w.e1; ++i) {
      const uint32_t id = BWT_EP_ID(B.eitems[size_t(i)]);
      if (!(id & 1u) && BWT_EP_ID(B.eitems[size_t(i) + 1]) == id + 1) ++B.npairs;
    }
  const bool ranged = !wk.empty() && wk.back().range != wk.front().range;
  if (ranged) std::stable_sort(wk.begin(), wk.end(), [](const Work& p, const Work& q) { return p.range < q.range; });
  B.wk_chunk.resize(wk.size());
  B.wk_e0.resize(wk.size() + 1);
  B.wk_e0[0] = 0;
  for (size_t i = 0; i < wk.size(); ++i) {
    B.wk_chunk[i] = wk[i].chunk;
    B.wk_e0[i + 1] = B.wk_e0[i] + (wk[i].e1 - wk[i].e0);
  }
  if (ranged) {  // lay the items out in work order
    std::vector<uint64_t> out(B.eitems.size());
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < int64_t(wk.size()); ++i)
      std::copy(B.eitems.begin() + wk[size_t(i)].e0, B.eitems.begin() + wk[size_t(i)].e1,
                out.begin() + B.wk_e0[size_t(i)]);
    B.eitems.swap(out);
  }
  // each endpoint's place in the work order (its level-B result is stored there)
  B.epos.resize(size_t(nE));
#pragma omp parallel for schedule(static) if (nE > 8192)
  for (int64_t i = 0; i < nE; ++i) B.epos[size_t(BWT_EP_ID(B.eitems[size_t(i)]))] = uint32_t(i);
  return true;
}

}  // namespace bwg
