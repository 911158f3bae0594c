// gen_draws.hpp -- the counter-based draws of the benchmark workloads, shared by the host
// generator (tree_workload.cpp) and the device generator (gen_kernels.hip) so both materialise
// the same bits.  Draw `idx` of stream (seed, s) is splitmix64's finaliser applied to
// base + (idx+1)*golden, with base = mix(seed ^ s*0xD1B54A32D192ED03): it depends on no other
// draw (the counter property of src/batch_engine_v7c.hpp:240-249), so any element of any array
// can be computed on its own, in any order, on any device.
#pragma once
#include <stdint.h>

#include "bwgpu_gen.h"

#ifndef BWT_HD
#ifdef __HIPCC__
#define BWT_HD __host__ __device__
#else
#define BWT_HD
#endif
#endif

namespace bwg {

BWT_HD inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Counter-based draw: element `idx` of the splitmix64 sequence seeded by (seed, stream).
struct Stream {
  uint64_t base;
  BWT_HD Stream(uint64_t seed, uint64_t stream) : base(mix64(seed ^ (stream * 0xD1B54A32D192ED03ULL))) {}
  BWT_HD uint64_t at(uint64_t idx) const { return mix64(base + (idx + 1) * 0x9E3779B97F4A7C15ULL); }
};

// Uniform double in [0, 1) from the top 53 bits (exact: a 53-bit integer times 2^-53).
BWT_HD inline double unit(uint64_t x) { return double(x >> 11) * (1.0 / 9007199254740992.0); }

// The log-uniform range [lo, hi) inside [0, n): its length is floor(e) clamped to [1, n], where
// e = exp(unit(dlen) * log(n)) evaluated by the host's libm; its start is dpos mod (n - len + 1).
BWT_HD inline void range_from_exp(double e, uint64_t dpos, uint64_t n, uint32_t& lo, uint32_t& hi) {
  const double l = __builtin_floor(e);
  uint64_t len = l < 1.0 ? 1 : uint64_t(l);
  if (len > n) len = n;
  const uint64_t a = dpos % (n - len + 1);
  lo = uint32_t(a);
  hi = uint32_t(a + len);
}

// Uniform [lo, hi) of the cfg-1 workload: the two endpoints' min and max + 1.
BWT_HD inline void range_uniform(uint64_t a5, uint64_t a6, uint64_t n, uint32_t& lo, uint32_t& hi) {
  const uint32_t a = uint32_t(a5 % n), b = uint32_t(a6 % n);
  lo = a < b ? a : b;
  hi = (a < b ? b : a) + 1u;
}

// Delta bound R of a tree config: deltas are uniform in [-R, R].
BWT_HD inline int64_t delta_bound(int cfg) { return cfg == 1 ? 1000 : (cfg == 3 ? 10000 : 1000000); }

// Zipf pick by inverse CDF over the cumulative weights cum[0..m): the first rank whose cumulative
// weight exceeds u * cum[m-1] (m - 1 if none).  The guide table narrows the search: guide[j] is
// that rank for u = j / 2^16, and since rounding of u * total is monotone in u, the rank of any
// u in [j, j+1) / 2^16 lies in [guide[j], guide[j+1]]; j = the draw's top 16 bits.
BWT_HD inline uint64_t zipf_pick(const double* cum, uint64_t m, const uint32_t* guide, uint64_t x) {
  const double t = unit(x) * cum[m - 1];
  const uint32_t j = uint32_t(x >> 48);
  uint64_t lo = guide[j], hi = guide[j + 1];  // answer in [lo, hi]
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (cum[mid] > t) hi = mid;
    else lo = mid + 1;
  }
  return lo < m ? lo : m - 1;
}

// Entries of the guide table (kZipfGuide + 1 of them).
constexpr uint32_t kZipfGuide = 65536;

// guide[j] for one j (0 <= j <= kZipfGuide): the first rank whose weight exceeds (j/2^16)*total.
BWT_HD inline uint32_t zipf_guide_at(const double* cum, uint64_t m, uint32_t j) {
  const double t = (double(j) * (1.0 / 65536.0)) * cum[m - 1];
  uint64_t lo = 0, hi = m;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (cum[mid] > t) hi = mid;
    else lo = mid + 1;
  }
  return uint32_t(lo);
}

// ---- one element of each generated array (the device kernels and the CPU emulation of them
// call these; `ex` is the exp in use, and the value it returned is handed back for the tie test)

// floor(exp(y)) of two libms can differ only when exp(y) lies within their errors (an ulp or so
// each) of an integer; every value within 2^-46 of one (64 ulps) is settled by the host's exp.
BWT_HD inline bool near_integer(double e) {
  const double k = __builtin_rint(e);
  return __builtin_fabs(e - k) <= e * 0x1p-46;
}

// Streams of the tree (3 keys, 4 deltas, 5/6 state queries) and of the request stream (9 state,
// 10/11 range, 12 point-or-range, 14 append parent, 15/16 append keys / deltas).
struct GenStreams {
  Stream k, d, q5, q6, s9, s10, s11, s12, a14, a15, a16;
  BWT_HD explicit GenStreams(uint64_t seed)
      : k(seed, 3), d(seed, 4), q5(seed, 5), q6(seed, 6), s9(seed, 9), s10(seed, 10), s11(seed, 11), s12(seed, 12),
        a14(seed, 14), a15(seed, 15), a16(seed, 16) {}
};

// Update i of the tree.
BWT_HD inline void gen_update_at(const BwtGenSpec& g, const GenStreams& s, int64_t i, uint32_t& key,
                                 int64_t& delta) {
  const uint64_t x = s.k.at(uint64_t(i));
  if (g.cfg == 1) {
    key = uint32_t(x % g.N);
  } else {
    const uint64_t r = zipf_pick(g.zipf_cum, g.zipf_n, g.zipf_guide, x);
    key = g.cfg == 3 ? g.dkey[r] : uint32_t(r);
  }
  const int64_t R = delta_bound(g.cfg);
  delta = int64_t(s.d.at(uint64_t(i)) % uint64_t(2 * R + 1)) - R;
}

// Query i of the tree's own queries (bound: the range bound; log_n = the host's log(bound)).
// Returns true when the range is a tie the host must settle.
template <class Exp>
BWT_HD inline bool gen_state_query_at(const BwtGenSpec& g, const GenStreams& s, int64_t i, uint64_t bound,
                                      double log_n, Exp ex, uint32_t& lo, uint32_t& hi) {
  const uint64_t a5 = s.q5.at(uint64_t(i)), a6 = s.q6.at(uint64_t(i));
  if (g.cfg == 1) {
    range_uniform(a5, a6, bound, lo, hi);
    return false;
  }
  const double e = ex(unit(a5) * log_n);
  range_from_exp(e, a6, bound, lo, hi);
  return near_integer(e);
}

// Query q of a request stream (req_app NULL: every query sees the tree's S states).
template <class Exp>
BWT_HD inline bool gen_stream_query_at(const BwtGenSpec& g, const GenStreams& s, int form, const int64_t* req_q0,
                                       const int64_t* req_app, int64_t nreq, int64_t q, double log_n, Exp ex,
                                       int32_t& qstate, uint32_t& lo, uint32_t& hi) {
  uint64_t nstates = uint64_t(g.S);
  if (req_app) {  // the request holding q: the last r with req_q0[r] <= q
    int64_t a = 0, b = nreq;
    while (a < b) {
      const int64_t mid = a + (b - a) / 2;
      if (req_q0[mid + 1] <= q) a = mid + 1;
      else b = mid;
    }
    nstates += uint64_t(req_app[a]);
  }
  qstate = int32_t(1 + int64_t(s.s9.at(uint64_t(q)) % nstates));
  const uint64_t a10 = s.s10.at(uint64_t(q));
  if (form == 5 && unit(s.s12.at(uint64_t(q))) < 0.5) {
    // a point query [k, k+1): over a cfg-3 tree k is a distinct key (empty at the top value)
    const uint64_t r = zipf_pick(g.zipf_cum, g.zipf_n, g.zipf_guide, a10);
    lo = g.cfg == 3 ? g.dkey[r] : uint32_t(r);
    hi = lo == 0xFFFFFFFFu ? lo : lo + 1u;
    return false;
  }
  const double e = ex(unit(a10) * log_n);
  range_from_exp(e, s.s11.at(uint64_t(q)), g.cfg == 3 ? 0xFFFFFFFFull : g.N, lo, hi);
  return near_integer(e);
}

// Append element i (i < napp*U: key and delta; i < napp also the parent of append i).
BWT_HD inline void gen_append_at(const BwtGenSpec& g, const GenStreams& s, int64_t napp, int64_t i,
                                 int32_t* app_parent, uint32_t* app_key, int64_t* app_delta) {
  if (i < napp) app_parent[i] = int32_t(1 + int64_t(s.a14.at(uint64_t(i)) % uint64_t(g.S + i)));
  const uint64_t r = zipf_pick(g.zipf_cum, g.zipf_n, g.zipf_guide, s.a15.at(uint64_t(i)));
  app_key[i] = g.cfg == 3 ? g.dkey[r] : uint32_t(r);  // over a cfg-3 tree: a distinct key
  app_delta[i] = int64_t(s.a16.at(uint64_t(i)) % 2000001ULL) - 1000000;
}

// The host's settlement of a tie: the range recomputed with the host's exp (what: 0 = state
// query i, 1 = stream query i).
inline void settle_range(const GenStreams& s, int what, uint64_t bound, double log_n,
                         double (*host_exp)(double), uint64_t i, uint32_t& lo, uint32_t& hi) {
  const uint64_t a = what == 0 ? s.q5.at(i) : s.s10.at(i), b = what == 0 ? s.q6.at(i) : s.s11.at(i);
  range_from_exp(host_exp(unit(a) * log_n), b, bound, lo, hi);
}

}  // namespace bwg
