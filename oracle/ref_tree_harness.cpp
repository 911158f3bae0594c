// ref_tree_harness.cpp -- runs the reference's own exact-answer Solver (src/oracle.cpp:225-451)
// over a branching multi-update point-update tree.  TEST INFRASTRUCTURE ONLY: built by
// oracle/Makefile into oracle/_ref/ from the reference source where it lies, never shipped.
//
// Input: a dump written by `bw_oracle tree ... --dump FILE` (S, U, Q, parents, keys, deltas,
// query endpoints).  Mapping onto the reference model (one edit per reference state):
//   benchmark state v with updates u_0..u_{U-1}  ->  U chained reference states
//   r(v,i) = 1 + (v-1)*U + i, each an OP_INSERT of a fresh row (t = key, group 0,
//   amount = delta) on a one-row, one-group fixture whose factor is 1, so every row weighs
//   exactly its delta;  parent(r(v,0)) = r(parent(v), U-1) (or the root), parent(r(v,i)) =
//   r(v,i-1).  The answers of state v are the Solver's answers at r(v, U-1).
// BWH_STATES must equal 1 + S*U; the harness refuses any other size.
#define main bw_reference_oracle_main
#include "oracle.cpp"
#undef main

#include <cstdio>
#include <vector>

int main(int argc, char** argv) {
  if (argc < 3) { std::fprintf(stderr, "usage: tree_harness DUMP ANSWERS\n"); return 2; }
  FILE* fp = std::fopen(argv[1], "rb");
  if (!fp) return 3;
  i64 hdr[3];
  if (std::fread(hdr, sizeof(hdr), 1, fp) != 1) return 3;
  const i64 S = hdr[0], U = hdr[1], Q = hdr[2];
  if (1 + S * U != kStates || Q > kMaxQ || U < 1) {
    std::fprintf(stderr, "harness built for kStates=%d kMaxQ=%d, dump needs %lld/%lld\n", kStates, kMaxQ,
                 (long long)(1 + S * U), (long long)Q);
    return 2;
  }
  std::vector<i32> par(size_t(S) + 1, -1);
  std::vector<u32> key(size_t(S * U)), qlo(size_t(S * Q)), qhi(size_t(S * Q));
  std::vector<i64> dl(size_t(S * U));
  bool ok = std::fread(par.data() + 1, sizeof(i32), size_t(S), fp) == size_t(S);
  ok = ok && std::fread(key.data(), sizeof(u32), key.size(), fp) == key.size();
  ok = ok && std::fread(dl.data(), sizeof(i64), dl.size(), fp) == dl.size();
  ok = ok && std::fread(qlo.data(), sizeof(u32), qlo.size(), fp) == qlo.size();
  ok = ok && std::fread(qhi.data(), sizeof(u32), qhi.size(), fp) == qhi.size();
  std::fclose(fp);
  if (!ok) return 3;

  // one base row weighing nothing, one group with factor 1
  static BaseRec base_row{0, 0, 0, 0};
  Fixture f;
  f.h.n = 1;
  f.h.groups = 1;
  f.h.t_min = 0;
  f.h.t_max = 0xFFFFFFFFll;
  f.rows.p = &base_row;
  f.rows.n = 1;
  f.base_factor.assign(1, 1);
  f.group_labels.assign(1, 0);

  Workload w;
  w.per_state = true;
  w.q = i32(Q);
  w.n_base = 1;
  w.n_inserted = i32(S * U);
  w.groups = 1;
  w.parent.assign(size_t(kStates), -1);
  w.edit.assign(size_t(kStates), Edit{});
  w.queries.assign(size_t(kStates) * kMaxQ, Query{0, 1});
  w.ins_t.resize(size_t(S * U));
  w.ins_g.assign(size_t(S * U), 0);
  auto r_of = [&](i64 v, i64 i) -> i32 { return i32(1 + (v - 1) * U + i); };
  for (i64 v = 1; v <= S; ++v) {
    for (i64 i = 0; i < U; ++i) {
      const i32 r = r_of(v, i);
      w.parent[size_t(r)] = (i == 0) ? (par[size_t(v)] > 0 ? r_of(par[size_t(v)], U - 1) : 0) : r - 1;
      const size_t u = size_t((v - 1) * U + i);
      Edit e{};
      e.op = OP_INSERT;
      e.g = 0;
      e.id = 1 + i64(u);
      e.val = dl[u];
      e.t = i64(key[u]);
      w.edit[size_t(r)] = e;
      w.ins_t[u] = i64(key[u]);
      for (i64 k = 0; k < Q; ++k)
        w.queries[size_t(r) * kMaxQ + size_t(k)] =
            Query{i64(qlo[size_t((v - 1) * Q + k)]), i64(qhi[size_t((v - 1) * Q + k)])};
    }
  }
  Solver s(f, w);
  s.run();
  std::vector<i64> out(size_t(S * Q));
  for (i64 v = 1; v <= S; ++v)
    for (i64 k = 0; k < Q; ++k) out[size_t((v - 1) * Q + k)] = s.ans[size_t(r_of(v, U - 1) - 1) * size_t(Q) + size_t(k)];
  FILE* o = std::fopen(argv[2], "wb");
  if (!o) return 4;
  const bool wok = std::fwrite(out.data(), sizeof(i64), out.size(), o) == out.size();
  std::fclose(o);
  if (!wok) return 4;
  std::printf("{\"n_answers\":%zu,\"answers_sha256\":\"%s\"}\n", out.size(),
              sha::of(out.data(), out.size() * sizeof(i64)).c_str());
  return 0;
}
