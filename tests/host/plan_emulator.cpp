// plan_emulator.cpp -- test infrastructure: a CPU emulation of the bwt_* kernels
// (csrc/tree_kernels.hip) driven by the product's host plans (csrc/host/tree_plan.cpp), lane by
// lane for the chunk replay's 64-item groups, with a bounds check on every index.  It answers the
// declared workload of a tree config and prints the SHA-256 of the answers, which the tests compare
// with the oracle's golden hashes: a plan whose slots, groups, endpoint items, work cuts, pairs or
// staging positions were wrong gives a different hash (or fails a check) without any GPU.
//   plan_emulator CFG [--S --U --Q --N --seed] [--rmax --tile-shift --chunk --seg --ep-per-work
//                 --range-slots --stage-shift --threads] [--all 0|1] [--subset K]
// --all 0: the query-side replay lists only the blocks with queries; --subset K keeps every K-th
// query (small batches take the planner's sorted path); --answers FILE writes the int64 answers.
#include <openssl/sha.h>

#include <algorithm>
#include <cassert>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "bwgpu_tree.h"
#include "tree_plan.hpp"
#include "tree_workload.hpp"
using namespace bwg;

#define CHECK(c)                                                       \
  do {                                                                 \
    if (!(c)) {                                                        \
      std::fprintf(stderr, "CHECK failed %s at line %d\n", #c, __LINE__); \
      std::exit(1);                                                    \
    }                                                                  \
  } while (0)

static const int kQCap = 1024, kGCap = 64;

int main(int argc, char** argv) {
  int cfg = atoi(argv[1]);
  TreeSpec s = tree_defaults(cfg);
  PlanKnobs kn;
  int nt = 256, all = 1;
  int64_t subset = 0;            // keep every subset-th query (small batches take the sorted plan)
  const char* answers_path = nullptr;  // write the int64 answers here
  for (int i = 2; i + 1 < argc; i += 2) {
    const char* k = argv[i];
    const char* v = argv[i + 1];
    if (!strcmp(k, "--S")) s.S = atoll(v);
    else if (!strcmp(k, "--U")) s.U = atoi(v);
    else if (!strcmp(k, "--Q")) s.Q = atoi(v);
    else if (!strcmp(k, "--N")) s.N = strtoull(v, 0, 10);
    else if (!strcmp(k, "--seed")) s.seed = strtoull(v, 0, 10);
    else if (!strcmp(k, "--rmax")) kn.rmax = atoi(v);
    else if (!strcmp(k, "--tile-shift")) kn.tile_shift = atoi(v);
    else if (!strcmp(k, "--chunk")) kn.chunk_items = atoll(v);
    else if (!strcmp(k, "--seg")) kn.seg_chunks = atoi(v);
    else if (!strcmp(k, "--ep-per-work")) kn.ep_per_work = atoll(v);
    else if (!strcmp(k, "--range-slots")) kn.range_slots = atoll(v);
    else if (!strcmp(k, "--threads")) nt = atoi(v);
    else if (!strcmp(k, "--all")) all = atoi(v);
    else if (!strcmp(k, "--subset")) subset = atoll(v);
    else if (!strcmp(k, "--stage-shift")) kn.stage_shift = atoi(v);
    else if (!strcmp(k, "--answers")) answers_path = v;
    else { fprintf(stderr, "unknown argument %s\n", k); return 2; }
  }
  TreeWorkload w;
  tree_generate(s, w);
  const size_t nu = size_t(s.S) * s.U;
  size_t nq = size_t(s.S) * s.Q;
  std::vector<uint32_t> kc(w.key), lc(w.qlo), hc(w.qhi);
  if (cfg == 3) {
    auto l = build_rank_lut(w.dkey);
    host_ranks(l, w.key.data(), kc.data(), nu);
    host_ranks(l, w.qlo.data(), lc.data(), nq);
    host_ranks(l, w.qhi.data(), hc.data(), nq);
  }
  TreeIndexPlan P;
  char err[256];
  if (!build_tree_index(s.S, s.U, w.coord_n(), w.parent.data(), kc.data(), kn, P, err, 256)) {
    printf("plan fail %s\n", err);
    return 1;
  }
  std::vector<int32_t> qs(nq);
  for (size_t q = 0; q < nq; ++q) qs[q] = int32_t(q / s.Q + 1);
  std::vector<size_t> keep;  // original query index of each kept query
  if (subset > 1) {
    std::vector<int32_t> qs2;
    std::vector<uint32_t> l2, h2;
    for (size_t q = 0; q < nq; q += size_t(subset)) {
      keep.push_back(q);
      qs2.push_back(qs[q]);
      l2.push_back(lc[q]);
      h2.push_back(hc[q]);
    }
    qs.swap(qs2);
    lc.swap(l2);
    hc.swap(h2);
    nq = keep.size();
  }
  BatchPlan B;
  if (!build_batch_plan(P, nq, qs.data(), lc.data(), hc.data(), all, B, err, 256)) {
    printf("batch fail %s\n", err);
    return 1;
  }
  const int U = s.U, ts = P.knobs.tile_shift;
  int rpt = 1;
  while (rpt * nt < P.knobs.rmax) rpt <<= 1;
  CHECK(rpt <= 16);
  const int span = (rpt > 1 ? nt * rpt + nt : nt * rpt) + 1;
  auto pslot = [&](int r) { return rpt > 1 ? r + r / rpt : r; };
  fprintf(stderr, "nblk=%d kitems=%zu eitems=%zu nchunk=%ld nwork=%zu ngrp=%zu rpt=%d ts=%d chunk=%ld\n", P.nblk(),
          P.kitems.size(), B.eitems.size(), P.nchunk(), B.wk_chunk.size(), B.grp_step.size(), rpt, ts,
          (long)P.knobs.chunk_items);
  std::vector<int64_t> agg(P.bkey.size()), tagg(P.tile_rank0.size()), ansr(nq, 777), g(2 * nq, 0x5a5a5a5a),
      cp(P.n_cp), ans(nq), kweight(P.kitems.size(), 0x33);
  std::vector<int> ansr_set(nq, 0), g_set(2 * nq, 0);
  // ---------------- block kernel (write_index = 1 over all blocks when all, else query blocks)
  auto block = [&](int bi, bool write_index, const BatchPlan& Bq) {
    const int b = Bq.qblk[bi];
    const int64_t k0 = P.blk_key0[b];
    const int R = int(P.blk_key0[b + 1] - k0);
    CHECK(R <= nt * rpt);
    std::vector<int64_t> wv(span, 0), Pv(span, 0), tot(size_t(nt) * rpt, 0);
    std::vector<int32_t> gstep(kGCap);
    std::vector<int64_t> gq0(kGCap + 1);
    std::vector<uint32_t> qbuf(kQCap);
    const int s0 = P.blk_step0[b];
    const int64_t gi0 = Bq.qblk_g0[bi], ge = Bq.qblk_g0[bi + 1];
    const int s1 = write_index ? P.blk_step0[b + 1] : (gi0 < ge ? Bq.grp_step[ge - 1] + 1 : s0);
    const int64_t qend = Bq.grp_q0[ge];
    auto stage = [&](int64_t g0) {
      for (int i = 0; i <= kGCap; ++i) {
        const int64_t gg = g0 + i;
        if (i < kGCap) gstep[i] = gg < ge ? Bq.grp_step[gg] : -1;
        gq0[i] = gg <= ge ? Bq.grp_q0[gg] : qend;
      }
    };
    stage(gi0);
    const int64_t qbeg = Bq.grp_q0[gi0];
    for (int i = 0; i < kQCap; ++i) qbuf[i] = (qbeg + i < qend) ? Bq.qrank[qbeg + i] : 0u;
    int64_t gi = gi0, gwin = gi0, qwin = qbeg;
    int next_q = gi < ge ? gstep[0] : -1;
    auto rebuild = [&]() {
      int64_t run_all = 0;
      for (int t = 0; t < nt; ++t) {
        for (int j = 0; j < rpt; ++j) {
          const int pi = pslot(t * rpt + j);
          CHECK(pi < span);
          tot[size_t(t) * rpt + j] += wv[pi];
          wv[pi] = 0;
          Pv[pi] = run_all;
          run_all += tot[size_t(t) * rpt + j];
        }
      }
      CHECK(pslot(nt * rpt) == span - 1);
      Pv[pslot(nt * rpt)] = run_all;
    };
    auto query = [&]() {
      const int64_t qa = gq0[gi - gwin], qb = gq0[gi - gwin + 1];
      CHECK(gi - gwin + 1 <= kGCap);
      const int64_t nqq = qb - qa;
      const bool staged = nqq <= kQCap;
      if (staged && qb > qwin + kQCap) {
        qwin = qa;
        for (int i = 0; i < kQCap; ++i) qbuf[i] = (qa + i < qend) ? Bq.qrank[qa + i] : 0u;
      }
      rebuild();
      for (int64_t sl = 0; sl < nqq; ++sl) {
        uint32_t rk;
        if (staged) {
          CHECK(qa - qwin + sl >= 0 && qa - qwin + sl < kQCap);
          rk = qbuf[qa - qwin + sl];
        } else rk = Bq.qrank[qa + sl];
        const int ylo = rk & 0xFFFF, yhi = rk >> 16;
        CHECK(ylo <= R && yhi <= R);
        ansr[qa + sl] = Pv[pslot(yhi)] - Pv[pslot(ylo)];
        ansr_set[qa + sl]++;
      }
      ++gi;
      if (gi < ge && gi - gwin >= kGCap) {
        gwin = gi;
        stage(gi);
      }
      next_q = gi < ge ? gstep[gi - gwin] : -1;
    };
    for (int st_i = s0; st_i < s1; ++st_i) {
      const int st = P.steps[st_i];
      const int64_t v = st > 0 ? st : -st;
      for (int u = 0; u < U; ++u) {
        const int r = P.evrank[size_t(st_i) * U + u];
        CHECK(r < R);
        wv[pslot(r)] += st > 0 ? w.delta[(v - 1) * U + u] : -w.delta[(v - 1) * U + u];
      }
      if (st_i == next_q) query();
    }
    CHECK(gi == ge);
    if (!write_index) return;
    rebuild();
    for (int r = 0; r < R; ++r) agg[k0 + r] = Pv[pslot(r + 1)] - Pv[pslot(r)];
    const int64_t t0 = P.blk_tile0[b];
    const int T = int(P.blk_tile0[b + 1] - t0);
    for (int t = 0; t < T; ++t) {
      const int r0 = P.tile_rank0[t0 + t], r1 = t + 1 < T ? P.tile_rank0[t0 + t + 1] : R;
      tagg[t0 + t] = Pv[pslot(r1)] - Pv[pslot(r0)];
    }
  };
  // index build: every block, no queries
  {
    BatchPlan BI;
    for (int b = 0; b < P.nblk(); ++b) BI.qblk.push_back(b);
    BI.qblk_g0.assign(size_t(P.nblk()) + 1, 0);
    BI.grp_q0.assign(1, 0);
    for (int i = 0; i < P.nblk(); ++i) block(i, true, BI);
  }
  // query side
  for (size_t i = 0; i < B.qblk.size(); ++i) block(int(i), false, B);
  for (size_t i = 0; i < nq; ++i) CHECK(ansr_set[i] == 1);
  // ---------------- hist (+kweight), seg scan
  CHECK(int64_t(P.chain_chunk0[1]) == P.nblk());
  for (int64_t c = 0; c < P.nchunk(); ++c) {
    const int ch = P.chunk_chain[c];
    const int n = P.chain_ncls[ch];
    const int64_t* src = P.chain_src[ch] ? tagg.data() : agg.data();
    std::vector<int64_t> h(n, 0);
    for (int64_t it = P.chunk_key0[c]; it < P.chunk_key0[c + 1]; ++it) {
      const uint64_t x = P.kitems[it];
      const int cl = int(BWT_KEY_CLASS(x));
      CHECK(cl < n);
      const int64_t wt = src[x & BWT_REF_MASK];
      if (c >= P.chain_chunk0[1]) kweight[it] = wt;
      h[cl] += wt;
    }
    int64_t run = 0;
    for (int y = 0; y < n; ++y) {
      cp[P.chunk_cp0[c] + y] = run;
      run += h[y];
    }
  }
  for (int ch = 0; ch < P.nchain(); ++ch) {
    const int n = P.chain_ncls[ch];
    for (int y = 0; y < n; ++y) {
      int64_t run = 0;
      for (int64_t c = P.chain_chunk0[ch]; c < P.chain_chunk0[ch + 1]; ++c) {
        const int64_t t = cp[P.chunk_cp0[c] + y];
        cp[P.chunk_cp0[c] + y] = run;
        run += t;
      }
    }
  }
  // ---------------- chunk replay: one "wavefront" per work item, lanes emulated
  const int ncls_b = P.chain_ncls.size() > 1 ? P.chain_ncls[1] : 1;
  const int stride = (ncls_b + 1) | 1;
  long groups = 0, corr_iters = 0;
  for (size_t ci = 0; ci < B.wk_chunk.size(); ++ci) {
    const int64_t c = B.wk_chunk[ci];
    CHECK(c >= P.chain_chunk0[1]);
    const int ncls = P.chain_ncls[P.chunk_chain[c]];
    CHECK(ncls == ncls_b);
    const int64_t* row = cp.data() + P.chunk_cp0[c];
    std::vector<int64_t> fen(stride, 0);
    for (int i = 0; i < stride; ++i) fen[i] = (i >= 1 && i < ncls) ? row[i] - row[i - (i & -i)] : 0;
    int64_t ep = B.wk_e0[ci];
    const int64_t e1 = B.wk_e0[ci + 1];
    CHECK(e1 > ep);
    const int64_t k0 = P.chunk_key0[c];
    const int64_t kend = std::min<int64_t>(P.chunk_key0[c + 1], k0 + int64_t(BWT_EP_POS(B.eitems[e1 - 1])));
    int64_t ki = k0;
    int64_t kcs[64];
    int64_t kws[64];
    while (true) {
      ++groups;
      int rel[64], take[64], ne = 0;
      uint64_t it[64];
      for (int l = 0; l < 64; ++l) {
        const bool ev = ep + l < e1;
        it[l] = ev ? B.eitems[ep + l] : 0;
        rel[l] = ev ? int(int64_t(BWT_EP_POS(it[l])) - (ki - k0)) : 0;
        CHECK(rel[l] >= 0);
        take[l] = ev && l + rel[l] < 64;
      }
      for (int l = 0; l < 64; ++l) {
        if (take[l]) { CHECK(l == ne); ++ne; }
      }
      const int nk = int(std::min<int64_t>(64 - ne, kend - ki));
      CHECK(nk >= 0);
      CHECK(ne > 0 || nk > 0);
      for (int l = 0; l < nk; ++l) {
        CHECK(ki + l < P.chunk_key0[c + 1]);
        const uint64_t kx = P.kitems[ki + l];
        kcs[l] = int(BWT_KEY_CLASS(kx));
        kws[l] = kweight[ki + l];
      }
      if (ne > 0) {
        const int mmax = rel[ne - 1];
        CHECK(mmax <= nk);
        corr_iters += mmax;
        int64_t vals[64];
        for (int l = 0; l < ne; ++l) {
          const int cls = int(BWT_EP_CLASS(it[l]));
          CHECK(cls < ncls);
          int64_t val = 0;
          int y = cls;
          for (int k = 0; k < 12; ++k) { CHECK(y < stride); val += fen[y]; y &= y - 1; }
          for (int m = 0; m < mmax; ++m)
            if (m < rel[l] && kcs[m] < cls) val += kws[m];
          vals[l] = val;
        }
        for (int l = 0; l < ne; ++l) {
          const uint32_t id = BWT_EP_ID(it[l]);
          CHECK(id < 2 * nq);
          CHECK(B.epos[id] == uint64_t(ep + l));  // the endpoint's place in the work order
          g[ep + l] = vals[l];  // results in work order
          g_set[ep + l]++;
        }
      }
      if (ep + ne >= e1) break;
      for (int l = 0; l < nk; ++l)
        for (int ff = int(kcs[l]) + 1; ff <= ncls; ff += ff & -ff) { CHECK(ff < stride); fen[ff] += kws[l]; }
      ep += ne;
      ki += nk;
    }
  }
  for (size_t i = 0; i < 2 * nq; ++i) CHECK(g_set[i] == 1);
  // ---------------- combine: slots then ids
  for (size_t sl = 0; sl < nq; ++sl) {
    const uint32_t id = B.qorder[sl];
    const int b = B.sblk[sl];
    const int64_t* row = cp.data() + P.chunk_cp0[b];
    CHECK((hc[id] >> ts) < uint32_t(P.chain_ncls[0]) && (lc[id] >> ts) < uint32_t(P.chain_ncls[0]));
    CHECK(B.qtile[sl] == ((lc[id] >> ts) | ((hc[id] >> ts) << 16)));
    CHECK(B.islot[id] == sl);
    const uint32_t tl = B.qtile[sl];
    const int64_t lb = g[B.epos[2 * sl + 1]] - g[B.epos[2 * sl]];
    ansr[sl] += row[tl >> 16] - row[tl & 0xFFFF] + lb;
  }
  std::vector<int64_t> stage(nq, 0);
  std::vector<int> seen(nq, 0);
  for (size_t sl = 0; sl < nq; ++sl) {
    const size_t p = (sl / BWT_COMBINE_TILE) * BWT_COMBINE_TILE + B.srank[sl];
    CHECK(p < nq && B.spos[p] < nq); seen[B.spos[p]]++; stage[B.spos[p]] = ansr[sl];
  }
  for (size_t i = 0; i < nq; ++i) CHECK(seen[i] == 1);
  for (size_t i = 0; i < nq; ++i) ans[i] = stage[B.ipos[i]];
  fprintf(stderr, "npairs=%ld nwork=%zu\n", B.npairs, B.wk_chunk.size());
  fprintf(stderr, "groups=%ld corr_iters=%ld (avg %.1f)\n", groups, corr_iters, double(corr_iters) / groups);
  if (answers_path) {
    FILE* f = fopen(answers_path, "wb");
    CHECK(f && fwrite(ans.data(), 8, nq, f) == nq && fclose(f) == 0);
  }
  unsigned char d[32];
  SHA256((unsigned char*)ans.data(), nq * 8, d);
  for (int i = 0; i < 32; ++i) printf("%02x", d[i]);
  printf("\n");
  return 0;
}
