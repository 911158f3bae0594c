// dense_emulator.cpp -- test infrastructure: a CPU emulation of the dense-carry form
// (biworlddb-artifact_amd/csrc/dense_kernels.hip) over the product's host index plan and rank
// tables.  It builds every block's net weight per distinct key, the dense carry rows, packs each
// query into the plan's 64-bit sort key and 32-bit value exactly as the key kernel does, sorts by
// the key, derives the per-rank groups, and replays every block as the persistent kernel does --
// the queries' endpoints unpacked from the sorted key and value -- then prints the answers'
// SHA-256 in (batch, query) order.
//
//   dense_emulator --S n --U n --N n [--seed n] --requests n --max-batch n --rmax n [--answers F]
#include <openssl/sha.h>

#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "dense_tables.hpp"
#include "tree_plan.hpp"
#include "tree_workload.hpp"

using namespace bwg;

#define CHECK(c)                                                         \
  do {                                                                   \
    if (!(c)) {                                                          \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      return 1;                                                          \
    }                                                                    \
  } while (0)

// the kernels' layout: key = rank << 44 | id << 17 | lo & 0x1FFFF; val = hi | (lo >> 17) << 21
constexpr int kRankShift = 44, kIdShift = 17;
constexpr uint64_t kIdMask = (uint64_t(1) << 27) - 1;
constexpr uint32_t kLoLowMask = (1u << 17) - 1, kHiMask = (1u << 21) - 1;

int main(int argc, char** argv) {
  TreeSpec spec = tree_defaults(4);
  StreamSpec ss = stream_defaults(4);
  PlanKnobs kn;
  kn.rmax = 4096;
  const char* answers_path = nullptr;
  for (int i = 1; i + 1 < argc; i += 2) {
    const char* k = argv[i];
    const char* v = argv[i + 1];
    if (!std::strcmp(k, "--S")) spec.S = std::atoll(v);
    else if (!std::strcmp(k, "--U")) spec.U = std::atoi(v);
    else if (!std::strcmp(k, "--N")) spec.N = std::strtoull(v, nullptr, 10);
    else if (!std::strcmp(k, "--seed")) spec.seed = std::strtoull(v, nullptr, 10);
    else if (!std::strcmp(k, "--requests")) ss.nreq = std::atoll(v);
    else if (!std::strcmp(k, "--max-batch")) ss.max_batch = std::atoi(v);
    else if (!std::strcmp(k, "--rmax")) kn.rmax = std::atoi(v);
    else if (!std::strcmp(k, "--answers")) answers_path = v;
    else {
      std::fprintf(stderr, "unknown argument %s\n", k);
      return 2;
    }
  }
  TreeWorkload w;
  tree_generate(spec, w);
  RequestStream rs;
  stream_generate(w, ss, rs);
  const int64_t S = spec.S, M = rs.nq();
  const int U = spec.U;
  const uint64_t NK = w.coord_n();
  CHECK(S < (int64_t(1) << 19) && M < (int64_t(1) << 27) && NK <= (uint64_t(1) << 20));

  TreeIndexPlan P;
  char err[256];
  if (!build_tree_index(S, U, NK, w.parent.data(), w.key.data(), kn, P, err, sizeof(err))) {
    std::fprintf(stderr, "plan: %s\n", err);
    return 1;
  }
  const int nblk = P.nblk();
  std::vector<int32_t> erank, estep, be0;
  CHECK(dense_rank_tables(S, P.steps, P.blk_step0, erank, estep, be0));

  // every block's net weight per distinct key: its events, enters added and exits subtracted
  std::vector<int64_t> agg(P.bkey.size(), 0);
  for (int b = 0; b < nblk; ++b)
    for (int64_t s = P.blk_step0[size_t(b)]; s < P.blk_step0[size_t(b) + 1]; ++s) {
      const int32_t v = P.steps[size_t(s)];
      const int64_t a = (v > 0 ? v : -v) - 1;
      for (int i = 0; i < U; ++i) {
        const int64_t d = w.delta[size_t(a * U + i)];
        agg[size_t(P.blk_key0[size_t(b)] + P.evrank[size_t(s * U + i)])] += v > 0 ? d : -d;
      }
    }
  // carry rows: row b at x = weight of the blocks before b with key < x
  const size_t ld = size_t(NK) + 1;
  std::vector<int64_t> carry(static_cast<size_t>(nblk) * ld), wk(ld, 0);
  for (int b = 0; b < nblk; ++b) {
    int64_t run = 0;
    for (size_t x = 0; x < ld; ++x) {
      carry[size_t(b) * ld + x] = run;
      run += wk[x];
    }
    for (int64_t r = P.blk_key0[size_t(b)]; r < P.blk_key0[size_t(b) + 1]; ++r) wk[P.bkey[size_t(r)]] += agg[size_t(r)];
  }

  // the plan: packed keys, sorted, and the first position of every rank
  std::vector<uint64_t> key(static_cast<size_t>(M));
  std::vector<uint32_t> val(static_cast<size_t>(M));
  for (int64_t q = 0; q < M; ++q) {
    const int32_t v = rs.qstate[size_t(q)];
    const uint32_t lo = rs.qlo[size_t(q)], hi = rs.qhi[size_t(q)];
    CHECK(v >= 1 && v <= S && lo <= NK && hi <= NK);
    key[size_t(q)] = (uint64_t(uint32_t(erank[size_t(v)])) << kRankShift) | (uint64_t(q) << kIdShift) |
                     uint64_t(lo & kLoLowMask);
    val[size_t(q)] = hi | ((lo >> 17) << 21);
  }
  std::vector<uint32_t> ord(static_cast<size_t>(M));
  std::iota(ord.begin(), ord.end(), 0u);
  std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
  std::vector<uint64_t> skey(static_cast<size_t>(M));
  std::vector<uint32_t> sval(static_cast<size_t>(M));
  for (int64_t j = 0; j < M; ++j) {
    skey[size_t(j)] = key[ord[size_t(j)]];
    sval[size_t(j)] = val[ord[size_t(j)]];
  }
  std::vector<int64_t> grp(static_cast<size_t>(S) + 1, M);
  for (int64_t j = M - 1; j >= 0; --j) grp[size_t(skey[size_t(j)] >> kRankShift)] = j;
  for (int64_t e = S - 1; e >= 0; --e) grp[size_t(e)] = std::min(grp[size_t(e)], grp[size_t(e) + 1]);

  // the replay: per block, events into per-rank totals; at each entered state its queries
  std::vector<int64_t> ans(static_cast<size_t>(M), 0);
  std::vector<uint8_t> done(static_cast<size_t>(M), 0);
  for (int b = 0; b < nblk; ++b) {
    const int64_t k0 = P.blk_key0[size_t(b)];
    const int R = int(P.blk_key0[size_t(b) + 1] - k0);
    const uint32_t* bk = P.bkey.data() + k0;
    std::vector<int64_t> tot(static_cast<size_t>(R), 0);
    int32_t e = be0[size_t(b)];
    for (int64_t s = P.blk_step0[size_t(b)]; s < P.blk_step0[size_t(b) + 1]; ++s) {
      const int32_t v = P.steps[size_t(s)];
      const int64_t a = (v > 0 ? v : -v) - 1;
      for (int i = 0; i < U; ++i) {
        const int64_t d = w.delta[size_t(a * U + i)];
        tot[P.evrank[size_t(s * U + i)]] += v > 0 ? d : -d;
      }
      if (v <= 0) continue;
      CHECK(estep[size_t(e)] == s && erank[size_t(v)] == e);
      for (int64_t j = grp[size_t(e)]; j < grp[size_t(e) + 1]; ++j) {
        const uint64_t kk = skey[size_t(j)];
        const uint32_t vv = sval[size_t(j)];
        const uint32_t qid = uint32_t((kk >> kIdShift) & kIdMask);
        const uint32_t lo = (uint32_t(kk) & kLoLowMask) | ((vv >> 21) << 17), hi = vv & kHiMask;
        CHECK(qid < M && !done[qid] && rs.qlo[qid] == lo && rs.qhi[qid] == hi);
        const int rl = int(std::lower_bound(bk, bk + R, lo) - bk), rh = int(std::lower_bound(bk, bk + R, hi) - bk);
        int64_t pl = 0, ph = 0;  // in-block prefixes at the endpoints' ranks
        for (int r = 0; r < rl; ++r) pl += tot[size_t(r)];
        for (int r = 0; r < rh; ++r) ph += tot[size_t(r)];
        const int64_t* row = carry.data() + size_t(b) * ld;
        ans[qid] = (row[hi] - row[lo]) + (ph - pl);
        done[qid] = 1;
      }
      ++e;
    }
  }
  for (int64_t q = 0; q < M; ++q) CHECK(done[size_t(q)]);
  unsigned char dg[32];
  SHA256(reinterpret_cast<const unsigned char*>(ans.data()), ans.size() * 8, dg);
  char hex[65];
  for (int i = 0; i < 32; ++i) std::snprintf(hex + 2 * i, 3, "%02x", dg[i]);
  if (answers_path) {
    FILE* f = std::fopen(answers_path, "wb");
    if (!f || std::fwrite(ans.data(), 8, ans.size(), f) != ans.size()) return 4;
    std::fclose(f);
  }
  std::printf("{\"answers_sha256\":\"%s\",\"n_answers\":%" PRId64 ",\"nblk\":%d}\n", hex, M, nblk);
  return 0;
}
