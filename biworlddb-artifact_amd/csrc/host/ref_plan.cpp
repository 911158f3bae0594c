// ref_plan.cpp -- fixture, reference-native declared workload and v7 block plan (ref_plan.hpp).
#include "ref_plan.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>

namespace bwg {

bool load_ref_fixture(const std::string& path, RefFixture& f) {
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) return false;
  unsigned char h[64];
  bool ok = std::fread(h, 1, 64, fp) == 64 && std::memcmp(h, "BWBASE02", 8) == 0;
  if (ok) {
    std::memcpy(&f.n, h + 8, 8);
    std::memcpy(&f.G, h + 16, 4);
    std::memcpy(&f.t_min, h + 40, 8);
    std::memcpy(&f.t_max, h + 48, 8);
    ok = f.n > 0 && f.G > 0 && f.n < (int64_t(1) << 40) && f.G <= (1 << 20);
  }
  if (ok) {
    f.rows.resize(size_t(f.n));
    f.base_factor.resize(size_t(f.G));
    ok = std::fread(f.rows.data(), sizeof(BaseRow), f.rows.size(), fp) == f.rows.size() &&
         std::fread(f.base_factor.data(), sizeof(int64_t), f.base_factor.size(), fp) == f.base_factor.size();
  }
  std::fclose(fp);
  if (!ok) return false;
  for (int64_t x : f.base_factor)
    if (x <= 0) return false;
  return true;
}

namespace {
// splitmix64 stream of the reference (src/common.hpp:40-54)
struct RefRng {
  uint64_t s;
  explicit RefRng(uint64_t seed) : s(seed ? seed : 0x9E3779B97F4A7C15ULL) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ULL;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  uint32_t bounded(uint32_t n) { return n ? uint32_t(next() % n) : 0; }
  int64_t range(int64_t lo, int64_t hi) { return hi <= lo ? lo : lo + int64_t(next() % uint64_t(hi - lo)); }
};

// q nested ranges of one state (src/common.hpp:229-246)
void nested(RefRng& r, int64_t tmin, int64_t tmax, int q, int64_t* lo_out, int64_t* hi_out) {
  int64_t lo = tmin, hi = tmax + 1;
  const int64_t span0 = hi - lo;
  for (int k = 0; k < q; ++k) {
    int64_t span = hi - lo;
    if (span <= 4096) {
      lo = tmin + span0 / 2;
      hi = lo + 2048;
      span = hi - lo;
    }
    const int64_t da = span * int64_t(3 + r.bounded(16)) / 100;
    const int64_t db = span * int64_t(3 + r.bounded(16)) / 100;
    lo += da;
    hi -= db;
    if (hi - lo < 1024) hi = lo + 1024;
    lo_out[k] = lo;
    hi_out[k] = hi;
  }
}
}  // namespace

void make_ref_workload(const RefFixture& f, int K, int topology, int seed, int factor_period, int q, bool per_state,
                       int factor_pct, RefWorkload& w) {
  w = RefWorkload();
  w.K = K; w.topology = topology; w.seed = seed; w.factor_period = factor_period; w.q = q; w.per_state = per_state;
  const uint64_t us = uint64_t(int64_t(seed)), ut = uint64_t(int64_t(topology)), uf = uint64_t(int64_t(factor_period));
  const uint64_t qseed = us * 1000000007ULL + ut * 65537ULL + uf * 131ULL + 17ULL;
  if (per_state) {
    w.qlo.assign(size_t(K) * q, 0);
    w.qhi.assign(size_t(K) * q, 0);
    for (int v = 1; v < K; ++v) {
      RefRng r(qseed + uint64_t(v) * 2654435761ULL);
      nested(r, f.t_min, f.t_max, q, &w.qlo[size_t(v) * q], &w.qhi[size_t(v) * q]);
    }
  } else {
    w.qlo.assign(size_t(q), 0);
    w.qhi.assign(size_t(q), 0);
    RefRng r(qseed);
    nested(r, f.t_min, f.t_max, q, w.qlo.data(), w.qhi.data());
  }
  w.parent.assign(size_t(K), -1);
  for (int v = 1; v < K; ++v) w.parent[size_t(v)] = topology == 0 ? v - 1 : (topology == 1 ? (v - 1) / 2 : 0);
  w.edit.assign(size_t(K), RefEdit{});
  RefRng rng(us * 1000003ULL + ut * 7919ULL + uf * 104729ULL + 20261025ULL);
  std::vector<int64_t> gfac(size_t(f.G), 1000);
  int32_t next_group = 0;
  const int64_t N = f.n;
  for (int v = 1; v < K; ++v) {
    RefEdit e;
    const bool fstep = factor_pct >= 0 ? rng.bounded(100) < uint32_t(factor_pct)
                                       : (factor_period > 0 && v % factor_period == 0);
    if (fstep) {
      const int32_t g = next_group++ % f.G;
      const int64_t cur = gfac[size_t(g)];
      const int64_t nf = cur == 950 ? 1000 : (cur == 1000 ? 1050 : 950);
      gfac[size_t(g)] = nf;
      e.op = kFactor; e.g = g; e.val = nf;
    } else {
      const uint32_t roll = rng.bounded(100);
      if (roll < 55) {
        e.op = kReplace;
        e.id = rng.range(0, N);
        e.val = f.rows[size_t(e.id)].amount + rng.range(-100000, 100000);
      } else if (roll < 75) {
        e.op = kInsert;
        e.id = N + int64_t(w.ins_t.size());
        e.t = rng.range(f.t_min, f.t_max + 1);
        e.g = int32_t(rng.bounded(uint32_t(f.G)));
        const int64_t base_amt = f.rows[rng.bounded(uint32_t(N))].amount;
        e.val = base_amt + rng.range(-100000, 100000);
        w.ins_t.push_back(e.t);
        w.ins_g.push_back(e.g);
      } else if (roll < 90) {
        e.op = kDelete;
        e.id = rng.range(0, N);
      } else {
        e.op = kMove;
        e.id = rng.range(0, N);
        e.t = rng.range(f.t_min, f.t_max + 1);
        e.g = int32_t(rng.bounded(uint32_t(f.G)));
        e.val = f.rows[size_t(e.id)].amount + rng.range(-100000, 100000);
      }
    }
    w.edit[size_t(v)] = e;
  }
}

bool HistorySpec::chain() const {
  for (int v = 1; v < K; ++v)
    if (parent[size_t(v)] != v - 1) return false;
  return true;
}

bool load_history_spec(const std::string& path, const RefFixture& f, HistorySpec& s, std::string& err) {
  s = HistorySpec();
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) { err = "cannot open spec"; return false; }
  std::fseek(fp, 0, SEEK_END);
  const long size = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  unsigned char h[124];
  if (size < 124 || std::fread(h, 1, 124, fp) != 124 || std::memcmp(h, "BWHHIST1", 8) != 0) {
    std::fclose(fp);
    err = "not a BWHHIST1 spec";
    return false;
  }
  for (int nf = 4; nf <= 5 && !s.header_fields; ++nf) {
    int32_t i4[5] = {0, 0, 0, 0, 0};
    int64_t i8[12];
    std::memcpy(i4, h + 8, 4 * size_t(nf));
    std::memcpy(i8, h + 8 + 4 * nf, sizeof(i8));
    const int64_t K = i4[0], q = i4[1], nins = i8[1], nep = i8[4];
    if (K < 2 || K > (1 << 28) || q < 1 || q > 128 || i4[3] < 1 || nins < 0 || nep < 0 || nins > (int64_t(1) << 32) ||
        nep > (int64_t(1) << 32))
      continue;
    if (8 + 4 * nf + 96 + K * 4 + K * 40 + K * q * 16 + nins * 12 + nep * 8 != int64_t(size)) continue;
    s.header_fields = nf;
    s.K = i4[0]; s.q = i4[1]; s.fp = i4[2]; s.G = i4[3]; s.topology = nf == 5 ? i4[4] : 0;
    s.n_base = i8[0]; s.n_ins = i8[1]; s.t_min = i8[2]; s.t_max = i8[3]; s.n_ep = i8[4]; s.period_lo = i8[5];
    s.period_hi = i8[6]; s.bucket = i8[7]; s.n_inserts = i8[8]; s.n_deletes = i8[9]; s.n_credits = i8[10];
    s.n_factor_edits = i8[11];
  }
  if (!s.header_fields) {
    std::fclose(fp);
    err = "spec size matches neither header layout";
    return false;
  }
  std::fseek(fp, 8 + 4 * s.header_fields + 96, SEEK_SET);
  const size_t K = size_t(s.K), nq = K * size_t(s.q);
  struct RawEdit {  // the reference's Edit (src/common.hpp:166-175)
    uint8_t op, pad[3];
    int32_t g, pad2;
    int64_t id, val, t;
  };
  static_assert(sizeof(RawEdit) == 40, "edit record");
  std::vector<RawEdit> raw(K);
  std::vector<int64_t> qr(2 * nq);
  s.parent.resize(K);
  s.ins_t.resize(size_t(s.n_ins));
  s.ins_g.resize(size_t(s.n_ins));
  s.endpoints.resize(size_t(s.n_ep));
  const bool ok = std::fread(s.parent.data(), 4, K, fp) == K && std::fread(raw.data(), 40, K, fp) == K &&
                  std::fread(qr.data(), 16, nq, fp) == nq &&
                  std::fread(s.ins_t.data(), 8, s.ins_t.size(), fp) == s.ins_t.size() &&
                  std::fread(s.ins_g.data(), 4, s.ins_g.size(), fp) == s.ins_g.size() &&
                  std::fread(s.endpoints.data(), 8, s.endpoints.size(), fp) == s.endpoints.size();
  std::fclose(fp);
  if (!ok) { err = "short spec"; return false; }
  s.edit.resize(K);
  s.qlo.resize(nq);
  s.qhi.resize(nq);
  for (size_t i = 0; i < nq; ++i) { s.qlo[i] = qr[2 * i]; s.qhi[i] = qr[2 * i + 1]; }
  if (s.n_base != f.n || s.G != f.G) { err = "spec and base fixture disagree (rows or groups)"; return false; }
  const int64_t M = f.n + s.n_ins;
  for (int64_t k = 0; k < s.n_ins; ++k)
    if (s.ins_g[size_t(k)] < 0 || s.ins_g[size_t(k)] >= f.G) { err = "inserted row group out of range"; return false; }
  for (size_t v = 1; v < K; ++v) {
    const RawEdit& r = raw[v];
    RefEdit& e = s.edit[v];
    e.op = r.op; e.g = r.g; e.id = r.id; e.val = r.val; e.t = r.t;
    if (s.parent[v] < 0 || s.parent[v] >= int32_t(v)) { err = "spec parents must satisfy 0 <= parent[v] < v"; return false; }
    if (r.op > kFactor) { err = "unknown edit op"; return false; }
    if (r.op == kFactor ? (r.g < 0 || r.g >= f.G) : (r.id < 0 || r.id >= M)) { err = "edit target out of range"; return false; }
    if ((r.op == kInsert || r.op == kMove) && (r.g < 0 || r.g >= f.G)) { err = "edit group out of range"; return false; }
    for (int k = 0; k < s.q; ++k)
      if (s.qlo[v * size_t(s.q) + size_t(k)] > s.qhi[v * size_t(s.q) + size_t(k)]) { err = "range lo > hi"; return false; }
  }
  s.parent[0] = -1;
  return true;
}

void spec_workload(const RefFixture& f, const HistorySpec& s, RefWorkload& w, std::vector<int32_t>& entry) {
  const int K = s.K, q = s.q;
  w = RefWorkload();
  w.topology = 0;
  w.factor_period = s.fp;
  w.q = q;
  w.per_state = true;
  w.ins_t = s.ins_t;
  w.ins_g = s.ins_g;
  // children in ascending order (CSR)
  std::vector<int32_t> off(static_cast<size_t>(K) + 1, 0), kid(static_cast<size_t>(K));
  for (int v = 1; v < K; ++v) ++off[size_t(s.parent[size_t(v)]) + 1];
  for (int v = 0; v < K; ++v) off[size_t(v) + 1] += off[size_t(v)];
  {
    std::vector<int32_t> cur(off.begin(), off.end() - 1);
    for (int v = 1; v < K; ++v) kid[size_t(cur[size_t(s.parent[size_t(v)])]++)] = v;
  }
  // current row / factor state, to record what every edit overwrites
  const int64_t N = f.n;
  struct Row { uint8_t ex; int32_t g; int64_t t, amt; };
  std::vector<Row> row(size_t(N + s.n_ins));
  for (int64_t i = 0; i < N; ++i) row[size_t(i)] = {1, f.rows[size_t(i)].g, f.rows[size_t(i)].t, f.rows[size_t(i)].amount};
  for (int64_t k = 0; k < s.n_ins; ++k) row[size_t(N + k)] = {0, s.ins_g[size_t(k)], s.ins_t[size_t(k)], 0};
  std::vector<int64_t> fac(f.base_factor);
  std::vector<RefEdit> undo(static_cast<size_t>(K));
  std::vector<int32_t> state_of;  // chain state -> version whose ranges it carries
  w.edit.assign(1, RefEdit{});
  state_of.push_back(0);
  entry.assign(size_t(K), 0);
  size_t last_entry = 0;
  std::vector<std::pair<int32_t, int32_t>> stk{{0, off[0]}};
  while (!stk.empty()) {
    auto& [v, next] = stk.back();
    if (next < off[size_t(v) + 1]) {
      const int32_t c = kid[size_t(next++)];
      const RefEdit& e = s.edit[size_t(c)];
      RefEdit& u = undo[size_t(c)];
      if (e.op == kFactor) {
        u.op = kFactor; u.g = e.g; u.val = fac[size_t(e.g)];
        fac[size_t(e.g)] = e.val;
      } else {
        Row& r = row[size_t(e.id)];
        u.op = kRestore; u.id = e.id; u.ex = r.ex; u.g = r.g; u.t = r.t; u.val = r.amt;
        switch (e.op) {
          case kReplace: r.ex = 1; r.amt = e.val; break;
          case kDelete: r.ex = 0; break;
          default: r.ex = 1; r.g = e.g; r.t = e.t; r.amt = e.val; break;  // insert / move
        }
      }
      w.edit.push_back(e);
      state_of.push_back(c);
      entry[size_t(c)] = int32_t(w.edit.size() - 1);
      last_entry = w.edit.size();
      stk.push_back({c, off[size_t(c)]});
    } else {
      if (v > 0) {
        const RefEdit& u = undo[size_t(v)];
        if (u.op == kFactor) {
          fac[size_t(u.g)] = u.val;
        } else {
          Row& r = row[size_t(u.id)];
          r.ex = u.ex; r.g = u.g; r.t = u.t; r.amt = u.val;
        }
        w.edit.push_back(u);
        state_of.push_back(v);
      }
      stk.pop_back();
    }
  }
  w.edit.resize(last_entry);
  state_of.resize(last_entry);
  w.K = int(last_entry);
  w.parent.resize(size_t(w.K));
  for (int v = 0; v < w.K; ++v) w.parent[size_t(v)] = v - 1;
  w.qlo.resize(size_t(w.K) * q);
  w.qhi.resize(size_t(w.K) * q);
  for (int v = 0; v < w.K; ++v)
    for (int k = 0; k < q; ++k) {
      w.qlo[size_t(v) * q + k] = s.qlo[size_t(state_of[size_t(v)]) * q + k];
      w.qhi[size_t(v) * q + k] = s.qhi[size_t(state_of[size_t(v)]) * q + k];
    }
}

bool build_ref_plan(const RefFixture& f, const RefWorkload& w, int B, bool cpu_snapshots, RefPlan& p,
                    std::string& err) {
  if (w.topology != 0) { err = "the v7 block path needs a chain version tree (topology 0)"; return false; }
  const int K = w.K, q = w.q, G = f.G;
  const int VS = K - 1;
  if (VS < 1 || q < 1) { err = "need at least one edited state and one query"; return false; }
  if (B < 1) B = 1;
  if (B > VS) B = VS;
  p = RefPlan();
  p.K = K; p.q = q; p.B = B; p.G = G;
  // ---- endpoint classes: class_of(t) = #declared endpoints <= t (src/batch_engine_v7.hpp:559-572)
  for (int v = 1; v < K; ++v)
    for (int k = 0; k < q; ++k) {
      p.endpoints.push_back(w.qlo_of(v, k));
      p.endpoints.push_back(w.qhi_of(v, k));
      if (!w.per_state) break;
    }
  if (!w.per_state) {
    p.endpoints.clear();
    for (int k = 0; k < q; ++k) { p.endpoints.push_back(w.qlo[size_t(k)]); p.endpoints.push_back(w.qhi[size_t(k)]); }
  }
  std::sort(p.endpoints.begin(), p.endpoints.end());
  p.endpoints.erase(std::unique(p.endpoints.begin(), p.endpoints.end()), p.endpoints.end());
  const auto& E = p.endpoints;
  auto class_of = [&](int64_t t) { return int32_t(std::upper_bound(E.begin(), E.end(), t) - E.begin()); };
  p.ncls = int(E.size()) + 1;
  p.Cpfx = p.ncls + 1;
  const int Cpfx = p.Cpfx;
  p.qlb.assign(size_t(K) * q, 0);
  p.qhb.assign(size_t(K) * q, 0);
  for (int v = 1; v < K; ++v)
    for (int k = 0; k < q; ++k) {
      p.qlb[size_t(v) * q + k] = class_of(w.qlo_of(v, k));
      p.qhb[size_t(v) * q + k] = class_of(w.qhi_of(v, k));
    }
  // ---- resolved edits along the chain: previous row / factor state of every edit
  const int64_t N = f.n, M = N + int64_t(w.ins_t.size());
  struct RowState { uint8_t ex; int32_t g, c; int64_t amt; };
  std::vector<RowState> cur(static_cast<size_t>(M));
  for (int64_t i = 0; i < N; ++i) cur[size_t(i)] = {1, f.rows[size_t(i)].g, class_of(f.rows[size_t(i)].t), f.rows[size_t(i)].amount};
  for (int64_t k = 0; k < M - N; ++k) cur[size_t(N + k)] = {0, w.ins_g[size_t(k)], class_of(w.ins_t[size_t(k)]), 0};
  std::vector<int64_t> curfac(f.base_factor);
  p.node.assign(size_t(K), RefNode{});
  std::vector<uint8_t> fac_touched(size_t(G), 0);
  for (int v = 1; v < K; ++v) {
    const RefEdit& e = w.edit[size_t(v)];
    RefNode& n = p.node[size_t(v)];
    if (e.op == kFactor) {
      n.kind = 2;
      n.g_old = e.g;
      n.fac_prev = curfac[size_t(e.g)];
      n.fac_val = e.val;
      curfac[size_t(e.g)] = e.val;
      fac_touched[size_t(e.g)] = 1;
      continue;
    }
    RowState& r = cur[size_t(e.id)];
    n.kind = 1;
    n.ex_old = r.ex; n.g_old = r.g; n.c_old = r.c; n.amt_old = r.amt;
    switch (e.op) {
      case kReplace: r.ex = 1; r.amt = e.val; break;
      case kDelete: r.ex = 0; break;
      case kRestore: r.ex = e.ex; r.g = e.g; r.c = class_of(e.t); r.amt = e.val; break;
      default: r.ex = 1; r.g = e.g; r.c = class_of(e.t); r.amt = e.val; break;  // insert / move
    }
    n.ex_new = r.ex; n.g_new = r.g; n.c_new = r.c; n.amt_new = r.amt;
  }
  // ---- per-group deviation support (classes touched by row terms) and support indices
  std::vector<std::vector<int32_t>> tc(static_cast<size_t>(G));
  for (int v = 1; v < K; ++v) {
    const RefNode& n = p.node[size_t(v)];
    if (n.kind != 1) continue;
    if (n.ex_old && n.amt_old != 0) tc[size_t(n.g_old)].push_back(n.c_old);
    if (n.ex_new && n.amt_new != 0) tc[size_t(n.g_new)].push_back(n.c_new);
  }
  for (auto& x : tc) {
    std::sort(x.begin(), x.end());
    x.erase(std::unique(x.begin(), x.end()), x.end());
  }
  auto ti_of = [&](int32_t g, int32_t c) {
    const auto& x = tc[size_t(g)];
    return int32_t(std::lower_bound(x.begin(), x.end(), c) - x.begin());
  };
  for (int v = 1; v < K; ++v) {
    RefNode& n = p.node[size_t(v)];
    if (n.kind != 1) continue;
    n.ti_old = (n.ex_old && n.amt_old != 0) ? ti_of(n.g_old, n.c_old) : 0;
    n.ti_new = (n.ex_new && n.amt_new != 0) ? ti_of(n.g_new, n.c_new) : 0;
  }
  // ---- factor-touched groups, base leaf V and per-group base mass (src/batch_engine_v7.hpp:1350-1415)
  p.g2row.assign(size_t(G), -1);
  std::vector<int32_t> row2g;
  for (int32_t g = 0; g < G; ++g)
    if (fac_touched[size_t(g)]) { p.g2row[size_t(g)] = int32_t(row2g.size()); row2g.push_back(g); }
  p.n_active = int(row2g.size());
  p.vbase.assign(size_t(Cpfx), 0);
  std::vector<std::vector<std::pair<int32_t, int64_t>>> gb(row2g.size());
  for (int64_t i = 0; i < N; ++i) {
    const BaseRow& r = f.rows[size_t(i)];
    const int32_t c = class_of(r.t);
    p.vbase[size_t(c) + 1] += r.amount * f.base_factor[size_t(r.g)];
    if (p.g2row[size_t(r.g)] >= 0) gb[size_t(p.g2row[size_t(r.g)])].push_back({c, r.amount});
  }
  // ---- merged (class, base mass, deviation index) support per active group (:1578-1629)
  p.ms_off.assign(size_t(p.n_active) + 1, 0);
  for (int r = 0; r < p.n_active; ++r) {
    auto& lst = gb[size_t(r)];
    std::sort(lst.begin(), lst.end());
    std::vector<std::pair<int32_t, int64_t>> sums;  // nonzero base mass per class
    for (size_t i = 0; i < lst.size();) {
      size_t j = i;
      int64_t s = 0;
      while (j < lst.size() && lst[j].first == lst[i].first) s += lst[j++].second;
      if (s != 0) sums.push_back({lst[i].first, s});
      i = j;
    }
    const auto& dev = tc[size_t(row2g[size_t(r)])];
    size_t a = 0, d = 0;
    while (a < sums.size() || d < dev.size()) {
      if (d >= dev.size() || (a < sums.size() && sums[a].first < dev[d])) {
        p.ms_cls.push_back(sums[a].first); p.ms_base.push_back(sums[a].second); p.ms_dev.push_back(-1); ++a;
      } else if (a >= sums.size() || dev[d] < sums[a].first) {
        p.ms_cls.push_back(dev[d]); p.ms_base.push_back(0); p.ms_dev.push_back(int32_t(d)); ++d;
      } else {
        p.ms_cls.push_back(sums[a].first); p.ms_base.push_back(sums[a].second); p.ms_dev.push_back(int32_t(d));
        ++a; ++d;
      }
    }
    p.ms_off[size_t(r) + 1] = int32_t(p.ms_cls.size());
  }
  // ---- blocks, per-block factor start vectors and snapshot layout (:707-830, v7b :773-806)
  p.bv0.resize(size_t(B));
  p.bv1.resize(size_t(B));
  for (int b = 0; b < B; ++b) {
    p.bv0[size_t(b)] = 1 + int32_t(int64_t(VS) * b / B);
    p.bv1[size_t(b)] = 1 + int32_t(int64_t(VS) * (b + 1) / B);
  }
  p.bidx.assign(size_t(K), 0);
  std::vector<std::vector<int32_t>> bgroups(static_cast<size_t>(B));
  for (int b = 0; b < B; ++b)
    for (int v = p.bv0[size_t(b)]; v < p.bv1[size_t(b)]; ++v) {
      p.bidx[size_t(v)] = b;
      if (p.node[size_t(v)].kind == 2) bgroups[size_t(b)].push_back(p.node[size_t(v)].g_old);
    }
  for (auto& x : bgroups) {
    std::sort(x.begin(), x.end());
    x.erase(std::unique(x.begin(), x.end()), x.end());
  }
  p.fac_start.assign(size_t(B) * G, 0);
  {
    std::vector<int64_t> fcur(f.base_factor);
    for (int b = 0; b < B; ++b) {
      std::copy(fcur.begin(), fcur.end(), p.fac_start.begin() + int64_t(b) * G);
      for (int v = p.bv0[size_t(b)]; v < p.bv1[size_t(b)]; ++v)
        if (p.node[size_t(v)].kind == 2) fcur[size_t(p.node[size_t(v)].g_old)] = p.node[size_t(v)].fac_val;
    }
  }
  p.need_off.assign(size_t(B) * G, -1);
  std::vector<std::vector<std::pair<int32_t, int64_t>>> slots(static_cast<size_t>(G));
  int64_t pos = 0;
  for (int b = 0; b < B; ++b)
    for (int32_t g : bgroups[size_t(b)]) {
      p.need_off[size_t(b) * G + g] = pos;
      slots[size_t(g)].push_back({b, pos});
      pos += int64_t(tc[size_t(g)].size());
    }
  p.devbuf_n = pos;
  p.gb_off.assign(size_t(G) + 1, 0);
  for (int32_t g = 0; g < G; ++g) {
    p.gb_off[size_t(g) + 1] = p.gb_off[size_t(g)] + int32_t(slots[size_t(g)].size());
    for (const auto& s : slots[size_t(g)]) { p.gb_block.push_back(s.first); p.gb_reg.push_back(s.second); }
  }
  if (cpu_snapshots) {
    // deviation of each needed group at its block's start: every row term of an earlier state
    p.snap.assign(size_t(pos), 0);
    std::vector<std::vector<int64_t>> dev(static_cast<size_t>(G));
    for (int32_t g = 0; g < G; ++g) dev[size_t(g)].assign(tc[size_t(g)].size(), 0);
    for (int b = 0; b < B; ++b) {
      for (int32_t g : bgroups[size_t(b)]) {
        const int64_t off = p.need_off[size_t(b) * G + g];
        std::copy(dev[size_t(g)].begin(), dev[size_t(g)].end(), p.snap.begin() + off);
      }
      for (int v = p.bv0[size_t(b)]; v < p.bv1[size_t(b)]; ++v) {
        const RefNode& n = p.node[size_t(v)];
        if (n.kind != 1) continue;
        if (n.ex_old && n.amt_old != 0) dev[size_t(n.g_old)][size_t(n.ti_old)] -= n.amt_old;
        if (n.ex_new && n.amt_new != 0) dev[size_t(n.g_new)][size_t(n.ti_new)] += n.amt_new;
      }
    }
  }
  return true;
}

}  // namespace bwg
