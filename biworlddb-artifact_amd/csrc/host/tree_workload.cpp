// tree_workload.cpp -- generator of the benchmark tree workloads (see tree_workload.hpp).
#include "tree_workload.hpp"

#include <openssl/sha.h>
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>

#include "gen_draws.hpp"

namespace bwg {
namespace {

// Zipf(s) over M ranks by inverse CDF over zipf_cumulative's table.
struct Zipf {
  std::vector<double> cum;
  Zipf(uint64_t m, double s) : cum(zipf_cumulative(m, s)) {}
  uint64_t pick(double u) const {
    const double t = u * cum.back();
    const size_t i = size_t(std::upper_bound(cum.begin(), cum.end(), t) - cum.begin());
    return i < cum.size() ? i : cum.size() - 1;
  }
};

void log_range(uint64_t dlen, uint64_t dpos, uint64_t n, uint32_t& lo, uint32_t& hi) {
  range_from_exp(std::exp(unit(dlen) * std::log(double(n))), dpos, n, lo, hi);
}

}  // namespace

std::vector<double> zipf_cumulative(uint64_t m, double s) {
  std::vector<double> cum(m);
  double acc = 0.0;
  for (uint64_t r = 0; r < m; ++r) {
    acc += std::pow(double(r + 1), -s);
    cum[r] = acc;
  }
  return cum;
}

TreeSpec tree_defaults(int cfg) {
  TreeSpec s;
  s.cfg = cfg;
  s.seed = 20261203ULL + uint64_t(cfg);
  switch (cfg) {
    case 1: s.S = 1000; s.U = 64; s.Q = 16; s.N = 65536; break;
    case 3: s.S = 300000; s.U = 256; s.Q = 32; s.N = 4194304; break;
    default: s.S = 300000; s.U = 128; s.Q = 32; s.N = 1u << 20; break;
  }
  return s;
}

void tree_generate(const TreeSpec& spec, TreeWorkload& w) {
  w.spec = spec;
  const int64_t S = spec.S;
  const int U = spec.U, Q = spec.Q;
  const uint64_t seed = spec.seed;
  w.parent.assign(size_t(S) + 1, -1);
  // parents: sequential for the run-length model, otherwise per state
  const Stream p1(seed, 1), p2(seed, 2);
  int64_t run = 1;
  for (int64_t v = 1; v <= S; ++v) {
    int64_t p;
    if (spec.cfg == 1) {
      p = v - 1 - int64_t(p1.at(uint64_t(v)) % uint64_t(std::min<int64_t>(v, 32)));
    } else if (spec.cfg == 4) {
      if (run < 4096 && unit(p1.at(uint64_t(v))) < 63.0 / 64.0) {
        p = v - 1;
        ++run;
      } else {
        p = v - 1 - int64_t(p2.at(uint64_t(v)) % uint64_t(std::min<int64_t>(v, 1024)));
        run = 1;
      }
    } else {
      p = unit(p1.at(uint64_t(v))) < 0.9 ? v - 1
                                         : v - 1 - int64_t(p2.at(uint64_t(v)) % uint64_t(std::min<int64_t>(v, 1024)));
    }
    w.parent[size_t(v)] = int32_t(p);
  }
  const size_t nu = size_t(S) * size_t(U), nq = size_t(S) * size_t(Q);
  w.key.resize(nu);
  w.delta.resize(nu);
  w.qlo.resize(nq);
  w.qhi.resize(nq);
  w.dkey.clear();
  const Stream sk(seed, 3), sd(seed, 4), sq5(seed, 5), sq6(seed, 6);
  if (spec.cfg == 1) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)nu; ++i) w.key[size_t(i)] = uint32_t(sk.at(uint64_t(i)) % spec.N);
  } else if (spec.cfg == 3) {
    // first D distinct values of stream 7 (open addressing over key+1), then sorted
    const uint64_t D = spec.N;
    uint64_t cap = 1;
    while (cap < 2 * D) cap <<= 1;
    std::vector<uint32_t> tab(cap, 0);
    bool top_seen = false;
    const Stream s7(seed, 7);
    w.dkey.reserve(D);
    for (uint64_t j = 0; w.dkey.size() < D; ++j) {
      const uint32_t c = uint32_t(s7.at(j) >> 32);
      if (c == 0xFFFFFFFFu) {
        if (!top_seen) { top_seen = true; w.dkey.push_back(c); }
        continue;
      }
      uint64_t h = (mix64(uint64_t(c)) >> 7) & (cap - 1);
      bool fresh = true;
      while (tab[h] != 0) {
        if (tab[h] == c + 1u) { fresh = false; break; }
        h = (h + 1) & (cap - 1);
      }
      if (fresh) { tab[h] = c + 1u; w.dkey.push_back(c); }
    }
    std::sort(w.dkey.begin(), w.dkey.end());
    const Zipf z(D, 0.9);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)nu; ++i) w.key[size_t(i)] = w.dkey[z.pick(unit(sk.at(uint64_t(i))))];
  } else {
    const Zipf z(spec.N, 1.1);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)nu; ++i) w.key[size_t(i)] = uint32_t(z.pick(unit(sk.at(uint64_t(i)))));
  }
  const int64_t R = delta_bound(spec.cfg);
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)nu; ++i)
    w.delta[size_t(i)] = int64_t(sd.at(uint64_t(i)) % uint64_t(2 * R + 1)) - R;
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)nq; ++i) {
    const uint64_t a5 = sq5.at(uint64_t(i)), a6 = sq6.at(uint64_t(i));
    if (spec.cfg == 1) {
      range_uniform(a5, a6, spec.N, w.qlo[size_t(i)], w.qhi[size_t(i)]);
    } else {
      log_range(a5, a6, spec.cfg == 3 ? 0xFFFFFFFFull : spec.N, w.qlo[size_t(i)], w.qhi[size_t(i)]);
    }
  }
}

void tree_workload_sha(const TreeWorkload& w, char out_hex[65]) {
  SHA256_CTX c;
  SHA256_Init(&c);
  SHA256_Update(&c, w.parent.data() + 1, sizeof(int32_t) * size_t(w.spec.S));
  SHA256_Update(&c, w.key.data(), sizeof(uint32_t) * w.key.size());
  SHA256_Update(&c, w.delta.data(), sizeof(int64_t) * w.delta.size());
  SHA256_Update(&c, w.qlo.data(), sizeof(uint32_t) * w.qlo.size());
  SHA256_Update(&c, w.qhi.data(), sizeof(uint32_t) * w.qhi.size());
  if (!w.dkey.empty()) SHA256_Update(&c, w.dkey.data(), sizeof(uint32_t) * w.dkey.size());
  unsigned char d[32];
  SHA256_Final(d, &c);
  static const char* hx = "0123456789abcdef";
  for (int i = 0; i < 32; ++i) {
    out_hex[2 * i] = hx[d[i] >> 4];
    out_hex[2 * i + 1] = hx[d[i] & 15];
  }
  out_hex[64] = 0;
}

StreamSpec stream_defaults(int form) {
  StreamSpec s;
  s.form = form;
  if (form == 4) {
    s.nreq = 100000;
    s.max_batch = 2048;
  } else {
    s.nreq = 20000;
    s.max_batch = 4096;
    s.append_frac = 0.05;
    s.append_U = 128;
    s.rate = 20000.0;
  }
  return s;
}

void stream_scan(const TreeSpec& tree, const StreamSpec& spec, RequestStream& rs) {
  rs = RequestStream();
  rs.spec = spec;
  const uint64_t seed = tree.seed;
  const int64_t R = spec.nreq;
  rs.req_q0.assign(size_t(R) + 1, 0);
  rs.req_kind.assign(size_t(R), 0);
  rs.req_app.assign(size_t(R), 0);
  rs.arrival.assign(size_t(R), 0.0);
  const Stream d8(seed, 8), d13(seed, 13), d17(seed, 17);
  int64_t nq = 0, napp = 0;
  double t = 0.0;
  for (int64_t r = 0; r < R; ++r) {
    int64_t sz;
    int kind = 0;
    if (spec.form == 4) {
      sz = 1 + int64_t(d8.at(uint64_t(r)) % uint64_t(spec.max_batch));
    } else {
      kind = unit(d13.at(uint64_t(r))) < spec.append_frac;
      const double l = std::floor(std::exp(unit(d8.at(uint64_t(r))) * std::log(double(spec.max_batch))));
      sz = l < 1.0 ? 1 : int64_t(l);
      if (sz > spec.max_batch) sz = spec.max_batch;
      if (kind) sz = 0;
      t += -std::log1p(-unit(d17.at(uint64_t(r)))) / spec.rate;
    }
    rs.req_kind[size_t(r)] = int8_t(kind);
    rs.req_app[size_t(r)] = napp;
    rs.arrival[size_t(r)] = t;
    napp += kind;
    nq += sz;
    rs.req_q0[size_t(r) + 1] = nq;
  }
}

void stream_generate(const TreeWorkload& tree, const StreamSpec& spec, RequestStream& rs) {
  stream_scan(tree.spec, spec, rs);
  const uint64_t seed = tree.spec.seed;
  const int64_t S = tree.spec.S;
  const uint64_t N = tree.spec.N;
  const int64_t R = spec.nreq;
  const int64_t nq = rs.req_q0[size_t(R)], napp = R ? rs.req_app[size_t(R) - 1] + rs.req_kind[size_t(R) - 1] : 0;
  rs.qstate.resize(size_t(nq));
  rs.qlo.resize(size_t(nq));
  rs.qhi.resize(size_t(nq));
  const size_t na = size_t(napp) * size_t(spec.append_U);
  rs.app_parent.resize(size_t(napp));
  rs.app_key.resize(na);
  rs.app_delta.resize(na);
  // over a cfg-3 tree (ranked keys): appended keys and point queries are distinct keys drawn by
  // the tree's own Zipf(0.9) law, ranges span the 32-bit domain; else keys are their own ranks
  const bool ranked = tree.spec.cfg == 3;
  const uint64_t ND = ranked ? uint64_t(tree.dkey.size()) : N;
  auto key_of = [&](uint64_t r) { return ranked ? tree.dkey[size_t(r)] : uint32_t(r); };
  std::unique_ptr<Zipf> zipf;
  if (spec.form == 5) zipf.reset(new Zipf(ND, ranked ? 0.9 : 1.1));
  const Stream d9(seed, 9), d10(seed, 10), d11(seed, 11), d12(seed, 12), d14(seed, 14), d15(seed, 15),
      d16(seed, 16);
  for (int64_t a = 0; a < napp; ++a) {
    rs.app_parent[size_t(a)] = int32_t(1 + int64_t(d14.at(uint64_t(a)) % uint64_t(S + a)));
    for (int i = 0; i < spec.append_U; ++i) {
      const uint64_t idx = uint64_t(a) * uint64_t(spec.append_U) + uint64_t(i);
      rs.app_key[idx] = key_of(zipf->pick(unit(d15.at(idx))));
      rs.app_delta[idx] = int64_t(d16.at(idx) % 2000001ULL) - 1000000;
    }
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t r = 0; r < R; ++r) {
    if (rs.req_kind[size_t(r)]) continue;
    const uint64_t nstates = uint64_t(S + rs.req_app[size_t(r)]);
    for (int64_t g = rs.req_q0[size_t(r)]; g < rs.req_q0[size_t(r) + 1]; ++g) {
      rs.qstate[size_t(g)] = int32_t(1 + int64_t(d9.at(uint64_t(g)) % nstates));
      if (spec.form == 5 && unit(d12.at(uint64_t(g))) < 0.5) {
        const uint32_t k = key_of(zipf->pick(unit(d10.at(uint64_t(g)))));
        rs.qlo[size_t(g)] = k;
        rs.qhi[size_t(g)] = k == 0xFFFFFFFFu ? k : k + 1u;  // [k, k+1), empty at the top value
      } else {
        log_range(d10.at(uint64_t(g)), d11.at(uint64_t(g)), ranked ? 0xFFFFFFFFull : N, rs.qlo[size_t(g)],
                  rs.qhi[size_t(g)]);
      }
    }
  }
}

void stream_sha(const RequestStream& rs, char out_hex[65]) {
  SHA256_CTX c;
  SHA256_Init(&c);
  SHA256_Update(&c, rs.req_q0.data(), rs.req_q0.size() * sizeof(int64_t));
  SHA256_Update(&c, rs.req_kind.data(), rs.req_kind.size());
  SHA256_Update(&c, rs.qstate.data(), rs.qstate.size() * sizeof(int32_t));
  SHA256_Update(&c, rs.qlo.data(), rs.qlo.size() * sizeof(uint32_t));
  SHA256_Update(&c, rs.qhi.data(), rs.qhi.size() * sizeof(uint32_t));
  SHA256_Update(&c, rs.app_parent.data(), rs.app_parent.size() * sizeof(int32_t));
  SHA256_Update(&c, rs.app_key.data(), rs.app_key.size() * sizeof(uint32_t));
  SHA256_Update(&c, rs.app_delta.data(), rs.app_delta.size() * sizeof(int64_t));
  unsigned char d[32];
  SHA256_Final(d, &c);
  static const char* hx = "0123456789abcdef";
  for (int i = 0; i < 32; ++i) {
    out_hex[2 * i] = hx[d[i] >> 4];
    out_hex[2 * i + 1] = hx[d[i] & 15];
  }
  out_hex[64] = 0;
}

RankLut build_rank_lut(const std::vector<uint32_t>& keys) {
  RankLut l;
  l.row_ptr.assign(65537, 0);
  l.low16.resize(keys.size());
  for (size_t i = 0; i < keys.size(); ++i) {
    l.row_ptr[(keys[i] >> 16) + 1]++;
    l.low16[i] = uint16_t(keys[i] & 0xFFFFu);
  }
  for (size_t h = 0; h < 65536; ++h) l.row_ptr[h + 1] += l.row_ptr[h];
  return l;
}

void host_ranks(const RankLut& l, const uint32_t* in, uint32_t* out, size_t n) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    const uint32_t x = in[i];
    const uint16_t* b = l.low16.data() + l.row_ptr[x >> 16];
    const uint16_t* e = l.low16.data() + l.row_ptr[(x >> 16) + 1];
    out[i] = uint32_t(std::lower_bound(b, e, uint16_t(x & 0xFFFFu)) - l.low16.data());
  }
}

}  // namespace bwg
