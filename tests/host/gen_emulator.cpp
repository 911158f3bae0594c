// gen_emulator.cpp -- test infrastructure: runs the device generator's per-element functions
// (csrc/host/gen_draws.hpp, the bodies of the gen_kernels.hip kernels) on the CPU with a stand-in
// for the device's exp, settles ties as bwt_gen_settle does, and compares every array with the
// host generator (tree_generate / stream_generate).
//
//   gen_emulator tree CFG S U Q N SEED [--ulps K | --rel X | --snap X] [--no-settle]
//   gen_emulator stream FORM S N SEED NREQ MAXB APPFRAC APPU RATE [--ulps K | --rel X | --snap X] [--no-settle]
//
// The stand-in exp is the host's exp moved by K ulps (alternating sign), scaled by 1 + X, or
// snapped to the nearest integer when within a relative X of it: a libm that differs from the
// host's by a few ulps, or rounds near-integers differently, must still give identical arrays
// once ties are settled; one that differs by far more than the tie window must not.
// Prints {"mismatch": n, "ties": t, "elements": e}.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gen_draws.hpp"
#include "tree_workload.hpp"

using namespace bwg;

static int g_ulps = 0;
static double g_rel = 0, g_snap = 0;

struct StandInExp {
  double operator()(double y) const {
    double e = std::exp(y);
    if (g_rel != 0) return e * (1.0 + g_rel);
    if (g_snap != 0 && std::fabs(e - std::rint(e)) <= g_snap * e) return std::rint(e);
    // sign from a bit of y's representation, so both directions occur
    uint64_t b;
    std::memcpy(&b, &y, 8);
    const bool up = (b >> 7) & 1;
    for (int k = 0; k < g_ulps; ++k) e = std::nextafter(e, up ? INFINITY : 0.0);
    return e;
  }
};

static double host_exp(double y) { return std::exp(y); }

template <class T>
static int64_t count_diff(const std::vector<T>& a, const std::vector<T>& b) {
  if (a.size() != b.size()) return int64_t(std::max(a.size(), b.size()));
  int64_t n = 0;
  for (size_t i = 0; i < a.size(); ++i) n += a[i] != b[i];
  return n;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  bool settle = true;
  for (int i = 2; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--ulps") && i + 1 < argc) g_ulps = std::atoi(argv[i + 1]);
    if (!std::strcmp(argv[i], "--rel") && i + 1 < argc) g_rel = std::atof(argv[i + 1]);
    if (!std::strcmp(argv[i], "--snap") && i + 1 < argc) g_snap = std::atof(argv[i + 1]);
    if (!std::strcmp(argv[i], "--no-settle")) settle = false;
  }
  const bool tree_mode = !std::strcmp(argv[1], "tree");
  if ((tree_mode && argc < 8) || (!tree_mode && argc < 11)) return 2;
  int64_t mismatch = 0, ties = 0, elements = 0;
  if (tree_mode) {
    TreeSpec spec;
    spec.cfg = std::atoi(argv[2]);
    spec.S = std::atoll(argv[3]);
    spec.U = std::atoi(argv[4]);
    spec.Q = std::atoi(argv[5]);
    spec.N = std::strtoull(argv[6], nullptr, 10);
    spec.seed = std::strtoull(argv[7], nullptr, 10);
    TreeWorkload w;
    tree_generate(spec, w);
    BwtGenSpec g{};
    g.cfg = spec.cfg;
    g.U = spec.U;
    g.Q = spec.Q;
    g.S = spec.S;
    g.N = spec.N;
    g.seed = spec.seed;
    std::vector<double> cum;
    std::vector<uint32_t> guide(BWT_ZIPF_GUIDE_ENTRIES);
    if (spec.cfg != 1) {
      const uint64_t m = spec.cfg == 3 ? w.dkey.size() : spec.N;
      cum = zipf_cumulative(m, spec.cfg == 3 ? 0.9 : 1.1);
      for (uint32_t j = 0; j < BWT_ZIPF_GUIDE_ENTRIES; ++j) guide[j] = zipf_guide_at(cum.data(), m, j);
      g.zipf_cum = cum.data();
      g.zipf_n = m;
      g.zipf_guide = guide.data();
      g.dkey = w.dkey.data();
    }
    const GenStreams s(spec.seed);
    const size_t nu = size_t(spec.S) * spec.U, nq = size_t(spec.S) * spec.Q;
    std::vector<uint32_t> key(nu), qlo(nq), qhi(nq);
    std::vector<int64_t> delta(nu);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)nu; ++i) gen_update_at(g, s, i, key[size_t(i)], delta[size_t(i)]);
    const uint64_t bound = spec.cfg == 3 ? 0xFFFFFFFFull : spec.N;
    const double log_n = std::log(double(bound));
    std::vector<uint64_t> tie_idx;
    for (size_t i = 0; i < nq; ++i)
      if (gen_state_query_at(g, s, int64_t(i), bound, log_n, StandInExp{}, qlo[i], qhi[i])) tie_idx.push_back(i);
    if (settle)
      for (uint64_t i : tie_idx) settle_range(s, 0, bound, log_n, host_exp, i, qlo[i], qhi[i]);
    mismatch = count_diff(key, w.key) + count_diff(delta, w.delta) + count_diff(qlo, w.qlo) + count_diff(qhi, w.qhi);
    ties = int64_t(tie_idx.size());
    elements = int64_t(2 * nu + 2 * nq);
  } else {
    TreeSpec tree = tree_defaults(std::atoi(argv[2]) == 4 ? 4 : 2);
    tree.S = std::atoll(argv[3]);
    tree.N = std::strtoull(argv[4], nullptr, 10);
    if (std::strtoull(argv[5], nullptr, 10)) tree.seed = std::strtoull(argv[5], nullptr, 10);
    StreamSpec ss = stream_defaults(std::atoi(argv[2]));
    ss.nreq = std::atoll(argv[6]);
    ss.max_batch = std::atoi(argv[7]);
    ss.append_frac = std::atof(argv[8]);
    ss.append_U = std::atoi(argv[9]);
    ss.rate = std::atof(argv[10]);
    TreeWorkload w;
    w.spec = tree;  // the stream reads only the tree's seed, S and N
    RequestStream ref, scan;
    stream_generate(w, ss, ref);
    stream_scan(tree, ss, scan);
    BwtGenSpec g{};
    g.cfg = tree.cfg;
    g.S = tree.S;
    g.N = tree.N;
    g.seed = tree.seed;
    std::vector<double> cum = zipf_cumulative(tree.N, 1.1);
    std::vector<uint32_t> guide(BWT_ZIPF_GUIDE_ENTRIES);
    for (uint32_t j = 0; j < BWT_ZIPF_GUIDE_ENTRIES; ++j) guide[j] = zipf_guide_at(cum.data(), tree.N, j);
    g.zipf_cum = cum.data();
    g.zipf_n = tree.N;
    g.zipf_guide = guide.data();
    const GenStreams s(tree.seed);
    const int64_t R = ss.nreq, nq = scan.req_q0[size_t(R)];
    const int64_t napp = R ? scan.req_app[size_t(R) - 1] + scan.req_kind[size_t(R) - 1] : 0;
    const size_t nqs = static_cast<size_t>(nq), nas = static_cast<size_t>(napp);
    std::vector<int32_t> qstate(nqs), app_parent(nas);
    std::vector<uint32_t> qlo(nqs), qhi(nqs), app_key(nas * ss.append_U);
    std::vector<int64_t> app_delta(nas * ss.append_U);
    const double log_n = std::log(double(tree.N));
    const int64_t* rq = ss.form == 5 ? scan.req_q0.data() : nullptr;
    const int64_t* ra = ss.form == 5 ? scan.req_app.data() : nullptr;
    std::vector<uint64_t> tie_idx;
    for (int64_t q = 0; q < nq; ++q)
      if (gen_stream_query_at(g, s, ss.form, rq, ra, R, q, log_n, StandInExp{}, qstate[size_t(q)], qlo[size_t(q)],
                              qhi[size_t(q)]))
        tie_idx.push_back(uint64_t(q));
    if (settle)
      for (uint64_t i : tie_idx) settle_range(s, 1, tree.N, log_n, host_exp, i, qlo[i], qhi[i]);
    for (int64_t i = 0; i < napp * ss.append_U; ++i)
      gen_append_at(g, s, napp, i, app_parent.data(), app_key.data(), app_delta.data());
    mismatch = count_diff(qstate, ref.qstate) + count_diff(qlo, ref.qlo) + count_diff(qhi, ref.qhi) +
               count_diff(app_parent, ref.app_parent) + count_diff(app_key, ref.app_key) +
               count_diff(app_delta, ref.app_delta) + count_diff(scan.req_q0, ref.req_q0) +
               count_diff(scan.req_app, ref.req_app);
    ties = int64_t(tie_idx.size());
    elements = 3 * nq + napp * (1 + 2 * int64_t(ss.append_U));
  }
  std::printf("{\"mismatch\": %" PRId64 ", \"ties\": %" PRId64 ", \"elements\": %" PRId64 "}\n", mismatch, ties,
              elements);
  return 0;
}
