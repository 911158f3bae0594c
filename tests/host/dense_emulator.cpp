// dense_emulator.cpp -- test infrastructure: a CPU emulation of the dense-carry form
// (biworlddb-artifact_amd/csrc/dense_kernels.hip) over the product's host index plan and rank
// tables.  It builds every block's net weight per distinct key, the dense carry rows, and follows
// the plan's kernels tile by tile: the high-digit counts, the bucket layout, the bucket pass (each
// query packed into the 64-bit key and 32-bit value, runs appended at the bucket's cursor), the
// per-tile low-digit counts and offsets, the state pass, the per-block rank groups -- every
// position checked to be written once and inside [0, M) -- then replays every block as the
// persistent kernel does (the queries' endpoints unpacked from the grouped key and value, answers
// stored in grouped order), puts the answers in query-id order by the partition by id window and
// the scatter, and prints their SHA-256 in (batch, query) order.  Where a kernel's workgroup places
// items by LDS or global atomics, the emulation takes the items in order: any order gives the same
// positions' contents up to a permutation inside a state or a window, which the answers do not see.
//
// With --chunk c the replay follows the work items instead (BwdLaunch.chunk): every block's queries
// in walk order cut into items of c by the items kernels' arithmetic (item slots in the blocks'
// queue order, each written once), each item replayed on its own -- the events before its first
// state, then the walk from that state, stopping after its count -- and every grouped position
// answered exactly once.
//
// --hot m moves the stream's first m queries onto the state of the first of them (one group far
// larger than an item, so several items start inside it at offsets).
//
//   dense_emulator --S n --U n --N n [--seed n] --requests n --max-batch n --rmax n [--chunk c]
//                  [--hot m] [--answers F]
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

// the kernels' layouts: in bucket order key = state << 44 | id << 17 | lo & 0x1FFFF and
// val = hi | (lo >> 17) << 21; grouped by state, ends = lo | hi << 32 and the query ids apart
constexpr int kStateShift = 44, kIdShift = 17;
constexpr uint64_t kIdMask = (uint64_t(1) << 27) - 1;
constexpr uint32_t kLoLowMask = (1u << 17) - 1, kHiMask = (1u << 21) - 1;
// the plan's digits and tiles, and the unsort's windows and tiles
constexpr int kLoBits = 9, kLoBins = 1 << kLoBits;
constexpr int64_t kCountTile = 16384, kBucketTile = 12288, kStateTile = 16384;
constexpr int kWinBits = 18;
constexpr int64_t kPartTile = 8192;

int main(int argc, char** argv) {
  TreeSpec spec = tree_defaults(4);
  StreamSpec ss = stream_defaults(4);
  PlanKnobs kn;
  kn.rmax = 4096;
  const char* answers_path = nullptr;
  int64_t chunk = 0, hot = 0;
  for (int i = 1; i + 1 < argc; i += 2) {
    const char* k = argv[i];
    const char* v = argv[i + 1];
    if (!std::strcmp(k, "--S")) spec.S = std::atoll(v);
    else if (!std::strcmp(k, "--chunk")) chunk = std::atoll(v);
    else if (!std::strcmp(k, "--hot")) hot = std::atoll(v);
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
  for (int64_t i = 1; i < std::min<int64_t>(hot, rs.nq()); ++i) rs.qstate[size_t(i)] = rs.qstate[0];
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

  // the plan.  count: high-digit counts (the count kernel's tiles only split the work)
  const int nh = int((S >> kLoBits) + 1);
  std::vector<int64_t> hcount(static_cast<size_t>(nh), 0);
  for (int64_t t0 = 0; t0 < M; t0 += kCountTile)
    for (int64_t q = t0; q < std::min(M, t0 + kCountTile); ++q) ++hcount[size_t(rs.qstate[size_t(q)] >> kLoBits)];
  // layout: bucket starts (cursors) and second-level tiles per bucket
  std::vector<int64_t> hbase(static_cast<size_t>(nh) + 1, 0), tstart(static_cast<size_t>(nh) + 1, 0);
  for (int h = 0; h < nh; ++h) {
    hbase[size_t(h) + 1] = hbase[size_t(h)] + hcount[size_t(h)];
    tstart[size_t(h) + 1] = tstart[size_t(h)] + (hcount[size_t(h)] + kStateTile - 1) / kStateTile;
  }
  CHECK(hbase[size_t(nh)] == M);
  std::vector<int64_t> hcur(hbase.begin(), hbase.end() - 1);
  // bucket: per tile, keys packed, grouped by high digit, each bucket's run appended at its cursor
  std::vector<uint64_t> bkeyv(static_cast<size_t>(M));
  std::vector<uint32_t> bval(static_cast<size_t>(M));
  std::vector<uint8_t> wrote(static_cast<size_t>(M), 0);
  for (int64_t t0 = 0; t0 < M; t0 += kBucketTile) {
    const int64_t t1 = std::min(M, t0 + kBucketTile);
    std::vector<std::vector<int64_t>> run(static_cast<size_t>(nh));
    for (int64_t q = t0; q < t1; ++q) run[size_t(rs.qstate[size_t(q)] >> kLoBits)].push_back(q);
    for (int h = 0; h < nh; ++h) {
      for (int64_t q : run[size_t(h)]) {
        const int32_t v = rs.qstate[size_t(q)];
        const uint32_t lo = rs.qlo[size_t(q)], hi = rs.qhi[size_t(q)];
        CHECK(v >= 1 && v <= S && lo <= NK && hi <= NK);
        const int64_t d = hcur[size_t(h)]++;
        CHECK(d >= hbase[size_t(h)] && d < hbase[size_t(h) + 1] && !wrote[size_t(d)]);
        wrote[size_t(d)] = 1;
        bkeyv[size_t(d)] = (uint64_t(uint32_t(v)) << kStateShift) | (uint64_t(q) << kIdShift) | uint64_t(lo & kLoLowMask);
        bval[size_t(d)] = hi | ((lo >> 17) << 21);
      }
    }
  }
  // tile counts and offsets: per bucket's tile, low-digit counts; each tile's start inside every
  // state of the bucket; grp[v] = the state's first position
  const int64_t ntiles = tstart[size_t(nh)];
  std::vector<int64_t> thist(static_cast<size_t>(ntiles) * kLoBins, 0);
  auto tile_range = [&](int64_t g, int& h, int64_t& p0, int64_t& p1) {
    h = int(std::upper_bound(tstart.begin(), tstart.end() - 1, g) - tstart.begin()) - 1;
    p0 = hbase[size_t(h)] + (g - tstart[size_t(h)]) * kStateTile;
    p1 = std::min(p0 + kStateTile, hbase[size_t(h) + 1]);
  };
  for (int64_t g = 0; g < ntiles; ++g) {
    int h;
    int64_t p0, p1;
    tile_range(g, h, p0, p1);
    CHECK(p0 < p1);
    for (int64_t p = p0; p < p1; ++p) {
      const uint32_t v = uint32_t(bkeyv[size_t(p)] >> kStateShift);
      CHECK(int(v >> kLoBits) == h);
      ++thist[size_t(g) * kLoBins + (v & (kLoBins - 1))];
    }
  }
  std::vector<int64_t> grp(static_cast<size_t>(S) + 2, 0);
  for (int h = 0; h < nh; ++h) {
    std::vector<int64_t> tot(kLoBins, 0);
    for (int64_t g = tstart[size_t(h)]; g < tstart[size_t(h) + 1]; ++g)
      for (int l = 0; l < kLoBins; ++l) {
        const int64_t c = thist[size_t(g) * kLoBins + l];
        thist[size_t(g) * kLoBins + l] = tot[size_t(l)];
        tot[size_t(l)] += c;
      }
    int64_t base = 0;
    for (int l = 0; l < kLoBins; ++l) {
      const int64_t v = int64_t(h) * kLoBins + l;
      if (v <= S) grp[size_t(v)] = hbase[size_t(h)] + base;
      base += tot[size_t(l)];
    }
  }
  grp[size_t(S) + 1] = M;
  // state: per tile, grouped by low digit, each state's run at grp[v] + the tile's start
  std::vector<uint64_t> ends(static_cast<size_t>(M));
  std::vector<uint32_t> qids(static_cast<size_t>(M));
  std::fill(wrote.begin(), wrote.end(), 0);
  for (int64_t g = 0; g < ntiles; ++g) {
    int h;
    int64_t p0, p1;
    tile_range(g, h, p0, p1);
    std::vector<int64_t> local(kLoBins, 0);
    for (int64_t p = p0; p < p1; ++p) {
      const uint32_t v = uint32_t(bkeyv[size_t(p)] >> kStateShift);
      const int l = int(v & (kLoBins - 1));
      const int64_t d = grp[size_t(v)] + thist[size_t(g) * kLoBins + l] + local[size_t(l)]++;
      CHECK(d >= grp[size_t(v)] && d < grp[size_t(v) + 1] && d < M && !wrote[size_t(d)]);
      wrote[size_t(d)] = 1;
      const uint64_t kk = bkeyv[size_t(p)];
      const uint32_t vv = bval[size_t(p)];
      qids[size_t(d)] = uint32_t((kk >> kIdShift) & kIdMask);
      ends[size_t(d)] = uint64_t((uint32_t(kk) & kLoLowMask) | ((vv >> 21) << 17)) | (uint64_t(vv & kHiMask) << 32);
    }
  }
  for (int64_t j = 0; j < M; ++j) CHECK(wrote[size_t(j)]);
  // rank groups: per block, the group of the state of each of its ranks, and its last queried rank
  std::vector<int64_t> rga(static_cast<size_t>(S)), rgb(static_cast<size_t>(S));
  std::vector<int32_t> blast(static_cast<size_t>(nblk), -1);
  for (int b = 0; b < nblk; ++b)
    for (int32_t e = be0[size_t(b)]; e < be0[size_t(b) + 1]; ++e) {
      const int32_t v = P.steps[size_t(estep[size_t(e)])];
      CHECK(v >= 1 && v <= S);
      rga[size_t(e)] = grp[size_t(v)];
      rgb[size_t(e)] = grp[size_t(v) + 1];
      if (rgb[size_t(e)] > rga[size_t(e)]) blast[size_t(b)] = e;
    }

  // one block's replay from rank ea (the block's first rank: from its first step; otherwise from
  // ea's enter step, the steps before it applied first) answering `left` queries, the first at
  // offset `off` into ea's group; answers stored at their grouped positions
  std::vector<int64_t> sans(static_cast<size_t>(M), 0);
  std::vector<uint8_t> done(static_cast<size_t>(M), 0);
  auto replay = [&](int b, int32_t ea, int64_t off, int64_t left) -> bool {
    const int64_t k0 = P.blk_key0[size_t(b)];
    const int R = int(P.blk_key0[size_t(b) + 1] - k0);
    const uint32_t* bk = P.bkey.data() + k0;
    std::vector<int64_t> tot(static_cast<size_t>(R), 0);
    const int32_t e0 = be0[size_t(b)];
    const int64_t s0 = P.blk_step0[size_t(b)];
    const int64_t ss = ea == e0 ? s0 : int64_t(estep[size_t(ea)]);
    const int64_t s1 = int64_t(estep[size_t(blast[size_t(b)])]) + 1;
    auto apply = [&](int64_t s) {
      const int32_t v = P.steps[size_t(s)];
      const int64_t a = (v > 0 ? v : -v) - 1;
      for (int i = 0; i < U; ++i) {
        const int64_t d = w.delta[size_t(a * U + i)];
        tot[P.evrank[size_t(s * U + i)]] += v > 0 ? d : -d;
      }
    };
    for (int64_t s = s0; s < ss; ++s) apply(s);
    int32_t e = ea;
    for (int64_t s = ss; s < s1 && left > 0; ++s) {
      apply(s);
      const int32_t v = P.steps[size_t(s)];
      if (v <= 0) continue;
      if (estep[size_t(e)] != s || erank[size_t(v)] != e) return false;
      const int64_t qa = rga[size_t(e)] + (e == ea ? off : 0);
      const int64_t nq = std::min(rgb[size_t(e)] - qa, left);
      for (int64_t j = qa; j < qa + nq; ++j) {
        const uint32_t qid = qids[size_t(j)];
        const uint32_t lo = uint32_t(ends[size_t(j)]), hi = uint32_t(ends[size_t(j)] >> 32);
        if (!(qid < M && !done[qid] && rs.qstate[qid] == v && rs.qlo[qid] == lo && rs.qhi[qid] == hi)) return false;
        const int rl = int(std::lower_bound(bk, bk + R, lo) - bk), rh = int(std::lower_bound(bk, bk + R, hi) - bk);
        int64_t pl = 0, ph = 0;  // in-block prefixes at the endpoints' ranks
        for (int r = 0; r < rl; ++r) pl += tot[size_t(r)];
        for (int r = 0; r < rh; ++r) ph += tot[size_t(r)];
        const int64_t* row = carry.data() + size_t(b) * ld;
        sans[size_t(j)] = (row[hi] - row[lo]) + (ph - pl);
        done[qid] = 1;
      }
      left -= std::max<int64_t>(nq, 0);
      ++e;
    }
    return left == 0 || ea == e0;
  };
  if (chunk <= 0) {  // whole blocks, each up to its last queried rank
    for (int b = 0; b < nblk; ++b)
      if (blast[size_t(b)] >= 0) CHECK(replay(b, be0[size_t(b)], 0, int64_t(1) << 40));
  } else {
    // the items kernels: the blocks' queue order (heaviest first; any order gives the same items),
    // ceil(bq / chunk) items per block from ibase of its place; per block its ranks' running query
    // count, item j starting at the rank that holds ordinal j * chunk
    std::vector<int64_t> bq(static_cast<size_t>(nblk), 0);
    for (int b = 0; b < nblk; ++b)
      for (int32_t e = be0[size_t(b)]; e < be0[size_t(b) + 1]; ++e) bq[size_t(b)] += rgb[size_t(e)] - rga[size_t(e)];
    std::vector<int> order(static_cast<size_t>(nblk));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return bq[size_t(x)] > bq[size_t(y)]; });
    std::vector<int64_t> ibase(static_cast<size_t>(nblk) + 1, 0), ipos(static_cast<size_t>(nblk));
    for (int p = 0; p < nblk; ++p) {
      ipos[size_t(order[size_t(p)])] = p;
      ibase[size_t(p) + 1] = ibase[size_t(p)] + (bq[size_t(order[size_t(p)])] + chunk - 1) / chunk;
    }
    const int64_t nitems = ibase[size_t(nblk)];
    CHECK(nitems <= int64_t(nblk) + M / 256 + 1 || chunk < 256);  // the workspace's item capacity
    struct Item {
      int b;
      int32_t ea;
      int64_t off, cnt;
    };
    std::vector<Item> items(static_cast<size_t>(nitems), Item{-1, 0, 0, 0});
    for (int b = 0; b < nblk; ++b) {
      const int64_t base = ibase[size_t(ipos[size_t(b)])];
      int64_t cs = 0;
      for (int32_t e = be0[size_t(b)]; e < be0[size_t(b) + 1]; ++e) {
        const int64_t c = rgb[size_t(e)] - rga[size_t(e)], q = cs;
        for (int64_t j = (q + chunk - 1) / chunk; j * chunk < q + c; ++j) {
          CHECK(base + j < nitems && items[size_t(base + j)].b < 0);
          items[size_t(base + j)] = Item{b, e, j * chunk - q, std::min(chunk, bq[size_t(b)] - j * chunk)};
        }
        cs += c;
      }
      CHECK(cs == bq[size_t(b)]);
    }
    for (const Item& it : items) {
      CHECK(it.b >= 0 && it.cnt > 0 && it.off >= 0 && it.off < rgb[size_t(it.ea)] - rga[size_t(it.ea)]);
      CHECK(replay(it.b, it.ea, it.off, it.cnt));
    }
  }
  for (int64_t q = 0; q < M; ++q) CHECK(done[size_t(q)]);
  // the unsort: per tile of grouped positions, runs per id window appended at the window's cursor
  // (window w's region starts at w << kWinBits); then every staged answer stored at its id
  const int64_t nwin = (M + (int64_t(1) << kWinBits) - 1) >> kWinBits;
  std::vector<int64_t> cursor(static_cast<size_t>(nwin), 0), stage_ans(static_cast<size_t>(M));
  std::vector<uint32_t> stage_id(static_cast<size_t>(M));
  std::fill(wrote.begin(), wrote.end(), 0);
  for (int64_t t0 = 0; t0 < M; t0 += kPartTile) {
    const int64_t t1 = std::min(M, t0 + kPartTile);
    std::vector<std::vector<int64_t>> run(static_cast<size_t>(nwin));
    for (int64_t j = t0; j < t1; ++j) run[size_t(qids[size_t(j)] >> kWinBits)].push_back(j);
    for (int64_t wv = 0; wv < nwin; ++wv)
      for (int64_t j : run[size_t(wv)]) {
        const int64_t gpos = (wv << kWinBits) + cursor[size_t(wv)]++;
        CHECK(gpos < M && !wrote[size_t(gpos)]);
        wrote[size_t(gpos)] = 1;
        stage_ans[size_t(gpos)] = sans[size_t(j)];
        stage_id[size_t(gpos)] = qids[size_t(j)];
      }
  }
  std::vector<int64_t> ans(static_cast<size_t>(M), 0);
  std::vector<uint8_t> stored(static_cast<size_t>(M), 0);
  for (int64_t i = 0; i < M; ++i) {
    const uint32_t id = stage_id[size_t(i)];
    CHECK(wrote[size_t(i)] && id < M && (int64_t(id) >> kWinBits) == (i >> kWinBits) && !stored[id]);
    stored[id] = 1;
    ans[id] = stage_ans[size_t(i)];
  }
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
