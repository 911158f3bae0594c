// cut_emulator.cpp -- test infrastructure: the device block cut's algorithm (index_kernels.hip,
// stage 0) restated on the CPU -- e(s), the end of a block starting at step s (distinct keys of its
// steps past rmax, or max_block_steps), for every step; the path 0, e(0), e(e(0)), ... marked by
// pointer doubling rounds as the kernels run them -- against the host's sequential greedy cut;
// and the device Euler walk's algorithm (edge successors, list ranking by pointer jumping) against
// the host's DFS walk.
//   cut_emulator CFG S U N SEED RMAX [MAX_BLOCK_STEPS]
//   prints {"blocks": n, "equal": true|false, "walk_equal": true|false}
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <unordered_set>
#include <vector>

#include "tree_plan.hpp"
#include "tree_workload.hpp"

using namespace bwg;

int main(int argc, char** argv) {
  if (argc < 7) return 2;
  TreeSpec spec = tree_defaults(std::atoi(argv[1]));
  spec.S = std::atoll(argv[2]);
  spec.U = std::atoi(argv[3]);
  spec.N = std::strtoull(argv[4], nullptr, 10);
  spec.seed = std::strtoull(argv[5], nullptr, 10);
  spec.Q = 1;
  TreeWorkload w;
  tree_generate(spec, w);
  PlanKnobs kn;
  kn.rmax = std::atoi(argv[6]);
  kn.tile_shift = 0;  // balanced
  if (argc > 7) kn.max_block_steps = std::atoi(argv[7]);
  TreeIndexPlan P;
  char err[256];
  if (!plan_walk_blocks(spec.S, spec.U, w.coord_n(), w.parent.data(), w.key.data(), kn, P, err, sizeof(err))) {
    std::fprintf(stderr, "%s\n", err);
    return 3;
  }
  const int64_t n = int64_t(P.steps.size());
  std::vector<int32_t> J(size_t(n) + 1), J2(size_t(n) + 1), mark(size_t(n) + 1, 0);
  J[size_t(n)] = int32_t(n);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t s0 = 0; s0 < n; ++s0) {
    std::unordered_set<uint32_t> set;
    int64_t e = n;
    for (int64_t t = s0; t < n; ++t) {
      if (t > s0 && t - s0 >= kn.max_block_steps) { e = t; break; }
      const int64_t v = P.steps[size_t(t)] > 0 ? P.steps[size_t(t)] : -P.steps[size_t(t)];
      for (int i = 0; i < spec.U; ++i) set.insert(w.key[size_t(v - 1) * size_t(spec.U) + size_t(i)]);
      if (t > s0 && int64_t(set.size()) > kn.rmax) { e = t; break; }
    }
    J[size_t(s0)] = int32_t(e);
  }
  mark[0] = 1;
  for (int64_t reach = 1; reach <= n; reach <<= 1) {
    for (int64_t s = 0; s <= n; ++s)
      if (mark[size_t(s)]) mark[size_t(J[size_t(s)])] = 1;
    for (int64_t s = 0; s <= n; ++s) J2[size_t(s)] = J[size_t(J[size_t(s)])];
    J.swap(J2);
  }
  std::vector<int32_t> b0;
  for (int64_t s = 0; s < n; ++s)
    if (mark[size_t(s)]) b0.push_back(int32_t(s));
  b0.push_back(int32_t(n));
  // the device walk's algorithm: children by a stable sort on the parent, every edge's successor,
  // list ranking by pointer-jumping rounds, edges placed at 2S - distance, cut after the last entry
  const int64_t S = spec.S;
  std::vector<int64_t> order(static_cast<size_t>(S));
  for (int64_t i = 0; i < S; ++i) order[size_t(i)] = i + 1;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return w.parent[size_t(a)] < w.parent[size_t(b)]; });
  std::vector<int64_t> koff(size_t(S) + 2, 0), pos_of(size_t(S) + 1, 0);
  for (int64_t p = 0; p <= S + 1; ++p)
    koff[size_t(p)] = std::lower_bound(order.begin(), order.end(), p, [&](int64_t c, int64_t x) { return w.parent[size_t(c)] < x; }) - order.begin();
  for (int64_t i = 0; i < S; ++i) pos_of[size_t(order[size_t(i)])] = i;
  const int64_t END = 2 * S;
  std::vector<int64_t> nxt(size_t(END) + 1), d(size_t(END) + 1), nxt2(nxt.size()), d2(d.size());
  nxt[size_t(END)] = END;
  d[size_t(END)] = 0;
  for (int64_t c = 1; c <= S; ++c) {
    const int64_t dn = 2 * (c - 1), up = dn + 1, p = w.parent[size_t(c)], i = pos_of[size_t(c)];
    nxt[size_t(dn)] = koff[size_t(c)] < koff[size_t(c) + 1] ? 2 * (order[size_t(koff[size_t(c)])] - 1) : up;
    nxt[size_t(up)] = i + 1 < koff[size_t(p) + 1] ? 2 * (order[size_t(i) + 1] - 1) : (p == 0 ? END : 2 * (p - 1) + 1);
    d[size_t(dn)] = d[size_t(up)] = 1;
  }
  for (int64_t reach = 1; reach <= END; reach <<= 1) {
    for (int64_t e = 0; e <= END; ++e) {
      d2[size_t(e)] = d[size_t(e)] + d[size_t(nxt[size_t(e)])];
      nxt2[size_t(e)] = nxt[size_t(nxt[size_t(e)])];
    }
    d.swap(d2);
    nxt.swap(nxt2);
  }
  std::vector<int32_t> walk(static_cast<size_t>(END));
  int64_t last = 0;
  for (int64_t e = 0; e < END; ++e) {
    const int64_t pos = END - d[size_t(e)];
    walk[size_t(pos)] = (e & 1) ? -int32_t(e / 2 + 1) : int32_t(e / 2 + 1);
    if (!(e & 1)) last = std::max(last, pos);
  }
  walk.resize(size_t(last) + 1);
  std::printf("{\"blocks\": %zu, \"equal\": %s, \"walk_equal\": %s}\n", b0.size() - 1,
              b0 == P.blk_step0 ? "true" : "false", walk == P.steps ? "true" : "false");
  return 0;
}
