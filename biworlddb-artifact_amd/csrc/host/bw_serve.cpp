// bw_serve.cpp -- the served form: an open-loop request stream answered by continuous batching
// over a resident index, with appended branch states (one JSON line).
//
//   bw_serve [--S n --U n --N n --seed n] [--requests 20000 --max-batch 4096 --append-frac 0.05
//            --append-U 128 --rate 20000] [--tuned FILE] [--slots 2] [--budget-mb 1024]
//            [--check-alone 64] [--answers FILE] [--json FILE]
//
// Resident in HBM: the version tree's index (TreeEngine), the appended states' updates, and per
// in-flight batch one fixed buffer ("slot") whose size is the HBM budget divided among the slots.
// Requests arrive at their Poisson times (a clock started when serving starts).  Whenever a slot
// is free, every request that has arrived and fits the slot's capacity is admitted in arrival
// order: appends are copied to the device at once (a stream every batch waits on), query
// requests join the batch.  The batch is planned on the host, copied in, answered by one launch
// sequence on the slot's own stream and copied out; batches on different slots overlap on the
// device and with the host planning of the next one.  A query on an appended state is answered
// at its deepest indexed ancestor plus the overlay of the appended updates on its path.
// Latency = time the request's answers are back on the host - its arrival time.
// --check-alone K: afterwards K requests (evenly spread) are answered again, each alone in its
// own batch, and must equal their answers from the mix element for element.
// Exit codes: 2 usage, 3 unsupported input, 4 I/O, 5 GPU error, 6 alone/mix mismatch.
#include <hip/hip_runtime_api.h>
#include <openssl/sha.h>

#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "bwgpu_tree.h"
#include "tree_engine.hpp"
#include "tree_plan.hpp"
#include "tree_tune.hpp"
#include "tree_workload.hpp"

using namespace bwg;
using clk = std::chrono::steady_clock;

#define HIPX(x)                                                                                   \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess) {                                                                       \
      std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      std::exit(5);                                                                               \
    }                                                                                             \
  } while (0)
#define RUN(x)                                                                       \
  do {                                                                               \
    const int rc_ = (x);                                                             \
    if (rc_ != 0) {                                                                  \
      std::fprintf(stderr, "launch failed (%d) at %s:%d\n", rc_, __FILE__, __LINE__); \
      std::exit(5);                                                                  \
    }                                                                                \
  } while (0)

namespace {

double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

std::string sha_hex(const void* p, size_t n) {
  unsigned char d[32];
  SHA256(static_cast<const unsigned char*>(p), n, d);
  static const char* hx = "0123456789abcdef";
  std::string s(64, '0');
  for (int i = 0; i < 32; ++i) { s[size_t(2 * i)] = hx[d[i] >> 4]; s[size_t(2 * i + 1)] = hx[d[i] & 15]; }
  return s;
}

// Fixed-capacity device + pinned host buffers of one in-flight batch.  Every array has its own
// region sized for cap queries; a batch uses a prefix of each.
struct Slot {
  int64_t cap = 0;       // queries
  int64_t cap_work = 0;  // chunk work items
  int32_t nblk = 0;
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  unsigned char* dev = nullptr;
  unsigned char* host = nullptr;  // pinned: staged inputs, then the answers
  size_t bytes = 0;
  // region offsets
  size_t o_qlo, o_qhi, o_qblk, o_qblk_g0, o_grp_step, o_grp_q0, o_qorder, o_spos, o_srank, o_ipos, o_qrank, o_sblk, o_qtile, o_wk_chunk,
      o_wk_e0, o_eitems, o_ov_qid, o_ov_state, o_ansr, o_g, o_stage, o_answers;
  // the batch in flight
  bool busy = false;
  std::vector<int64_t> reqs;     // request ids
  std::vector<int64_t> req_off;  // their first query in the batch
  int64_t nq = 0;
  clk::time_point launched;
};

size_t layout(Slot& s, int64_t cap, int32_t nblk) {
  s.cap = cap;
  s.nblk = nblk;
  s.cap_work = 2 * cap + 1;
  size_t o = 0;
  auto take = [&](size_t n) {
    const size_t at = o;
    o += (n + 255) & ~size_t(255);
    return at;
  };
  s.o_qlo = take(size_t(cap) * 4);
  s.o_qhi = take(size_t(cap) * 4);
  s.o_qblk = take(size_t(nblk) * 4);
  s.o_qblk_g0 = take((size_t(nblk) + 1) * 8);
  s.o_grp_step = take(size_t(cap) * 4);
  s.o_grp_q0 = take((size_t(cap) + 1) * 8);
  s.o_qorder = take(size_t(cap) * 4);
  s.o_spos = take(size_t(cap) * 4);
  s.o_srank = take(size_t(cap) * 2);
  s.o_ipos = take(size_t(cap) * 4);
  s.o_qrank = take(size_t(cap) * 4);
  s.o_sblk = take(size_t(cap) * 4);
  s.o_qtile = take(size_t(cap) * 4);
  s.o_wk_chunk = take(size_t(s.cap_work) * 8);
  s.o_wk_e0 = take((size_t(s.cap_work) + 1) * 8);
  s.o_eitems = take(size_t(cap) * 16);
  s.o_ov_qid = take(size_t(cap) * 4);
  s.o_ov_state = take(size_t(cap) * 4);
  s.o_ansr = take(size_t(cap) * 8);
  s.o_g = take(size_t(cap) * 16);
  s.o_stage = take(size_t(cap) * 8);
  s.o_answers = take(size_t(cap) * 8);
  return o;
}

size_t bytes_per_query() {
  Slot s;
  return layout(s, 1 << 20, 0) >> 20;
}

}  // namespace

int main(int argc, char** argv) {
  TreeSpec spec = tree_defaults(2);
  StreamSpec ss = stream_defaults(5);
  PlanKnobs kn;
  LaunchKnobs lk;
  int nslots = 2, check_alone = 64;
  double budget_mb = 1024;
  std::string answers_path, json_path, tuned_path;
  for (int i = 1; i < argc; ++i) {
    if (i + 1 >= argc) { std::fprintf(stderr, "bw_serve: missing value for %s\n", argv[i]); return 2; }
    const char* k = argv[i];
    const char* v = argv[++i];
    if (!std::strcmp(k, "--S")) spec.S = std::atoll(v);
    else if (!std::strcmp(k, "--U")) spec.U = std::atoi(v);
    else if (!std::strcmp(k, "--N")) spec.N = std::strtoull(v, nullptr, 10);
    else if (!std::strcmp(k, "--seed")) spec.seed = std::strtoull(v, nullptr, 10);
    else if (!std::strcmp(k, "--requests")) ss.nreq = std::atoll(v);
    else if (!std::strcmp(k, "--max-batch")) ss.max_batch = std::atoi(v);
    else if (!std::strcmp(k, "--append-frac")) ss.append_frac = std::atof(v);
    else if (!std::strcmp(k, "--append-U")) ss.append_U = std::atoi(v);
    else if (!std::strcmp(k, "--rate")) ss.rate = std::atof(v);
    else if (!std::strcmp(k, "--tuned")) tuned_path = v;
    else if (!std::strcmp(k, "--slots")) nslots = std::atoi(v);
    else if (!std::strcmp(k, "--budget-mb")) budget_mb = std::atof(v);
    else if (!std::strcmp(k, "--check-alone")) check_alone = std::atoi(v);
    else if (!std::strcmp(k, "--answers")) answers_path = v;
    else if (!std::strcmp(k, "--json")) json_path = v;
    else { std::fprintf(stderr, "bw_serve: unknown argument %s\n", k); return 2; }
  }
  if (spec.S < 1 || spec.U < 1 || spec.N < 2 || ss.nreq < 1 || ss.max_batch < 1 || ss.append_U < 1 || ss.rate <= 0 ||
      nslots < 1 || nslots > 8 || budget_mb <= 0 || check_alone < 0) {
    std::fprintf(stderr, "bw_serve: bad sizes\n");
    return 2;
  }
  std::string tuned_from = "defaults";
  if (!tuned_path.empty() && load_tuned(tuned_path.c_str(), "served", kn, lk)) tuned_from = tuned_path;

  // ---- the tree, the stream, the resident index
  const auto g0 = clk::now();
  TreeWorkload w;
  tree_generate(spec, w);
  RequestStream rs;
  stream_generate(w, ss, rs);
  const int64_t S = spec.S, R = ss.nreq, NQ = rs.nq(), NA = rs.napp();
  const int AU = ss.append_U;
  char err[256];
  HIPX(hipSetDevice(0));
  HIPX(hipFree(nullptr));
  TreeEngine E;
  if (!E.init(S, spec.U, w.coord_n(), w.parent.data(), w.key.data(), w.delta.data(), kn,
              BwtLaunch{lk.block_threads, lk.chunk_waves, 0, 0}, err, sizeof(err))) {
    std::fprintf(stderr, "bw_serve: index: %s\n", err);
    return std::strncmp(err, "device", 6) == 0 ? 5 : 3;
  }
  hipStream_t ist;
  HIPX(hipStreamCreate(&ist));
  RUN(E.build_index(ist));
  HIPX(hipStreamSynchronize(ist));
  const int32_t nblk = E.plan.nblk();
  // appended states: resident device arrays (capacity = the stream's appends), host mirrors
  DeviceArena app;
  int32_t* d_app_parent = app.scratch<int32_t>(size_t(NA));
  uint32_t* d_app_key = app.scratch<uint32_t>(size_t(NA) * size_t(AU));
  int64_t* d_app_delta = app.scratch<int64_t>(size_t(NA) * size_t(AU));
  if (!app.ok()) return 5;
  std::vector<int32_t> app_base(static_cast<size_t>(NA));  // deepest indexed ancestor of each appended state
  for (int64_t a = 0; a < NA; ++a) {
    int64_t p = rs.app_parent[size_t(a)];
    app_base[size_t(a)] = int32_t(p <= S ? p : app_base[size_t(p - S - 1)]);
  }
  hipEvent_t app_ready;
  HIPX(hipEventCreate(&app_ready));
  HIPX(hipEventRecord(app_ready, ist));
  // slots within the HBM budget
  const size_t bpq = bytes_per_query();
  std::vector<Slot> slots(static_cast<size_t>(nslots));
  const int64_t cap = std::max<int64_t>(ss.max_batch, int64_t(budget_mb * 1048576.0 / nslots) / int64_t(bpq));
  for (Slot& s : slots) {
    s.bytes = layout(s, cap, nblk);
    HIPX(hipMalloc(reinterpret_cast<void**>(&s.dev), s.bytes));
    HIPX(hipHostMalloc(reinterpret_cast<void**>(&s.host), s.bytes, hipHostMallocDefault));
    HIPX(hipStreamCreate(&s.st));
    HIPX(hipEventCreate(&s.done));
  }
  const auto g1 = clk::now();

  // ---- one batch: plan, stage, copy in, launch, copy the answers out (all on the slot's stream)
  std::vector<int32_t> bq_state;
  std::vector<uint32_t> bq_lo, bq_hi;
  std::vector<uint32_t> ov_qid;
  std::vector<int32_t> ov_state;
  double plan_s = 0;
  auto launch_batch = [&](Slot& s) {
    const auto p0 = clk::now();
    bq_state.clear();
    bq_lo.clear();
    bq_hi.clear();
    ov_qid.clear();
    ov_state.clear();
    for (int64_t r : s.reqs)
      for (int64_t q = rs.req_q0[size_t(r)]; q < rs.req_q0[size_t(r) + 1]; ++q) {
        const int32_t v = rs.qstate[size_t(q)];
        if (v > S) {
          ov_qid.push_back(uint32_t(bq_state.size()));
          ov_state.push_back(v);
          bq_state.push_back(app_base[size_t(v - S - 1)]);
        } else {
          bq_state.push_back(v);
        }
        bq_lo.push_back(rs.qlo[size_t(q)]);
        bq_hi.push_back(rs.qhi[size_t(q)]);
      }
    const int64_t M = int64_t(bq_state.size());
    s.nq = M;
    BatchPlan bp;
    if (!build_batch_plan(E.plan, M, bq_state.data(), bq_lo.data(), bq_hi.data(), false, bp, err, sizeof(err))) {
      std::fprintf(stderr, "bw_serve: batch plan: %s\n", err);
      std::exit(3);
    }
    if (int64_t(bp.wk_chunk.size()) > s.cap_work) {
      std::fprintf(stderr, "bw_serve: work items exceed the slot\n");
      std::exit(3);
    }
    plan_s += secs(p0, clk::now());
    // stage into pinned memory, copy the used prefixes
    auto put = [&](size_t off, const void* src, size_t n) {
      if (!n) return;
      std::memcpy(s.host + off, src, n);
      HIPX(hipMemcpyAsync(s.dev + off, s.host + off, n, hipMemcpyHostToDevice, s.st));
    };
    put(s.o_qlo, bq_lo.data(), size_t(M) * 4);
    put(s.o_qhi, bq_hi.data(), size_t(M) * 4);
    put(s.o_qblk, bp.qblk.data(), bp.qblk.size() * 4);
    put(s.o_qblk_g0, bp.qblk_g0.data(), bp.qblk_g0.size() * 8);
    put(s.o_grp_step, bp.grp_step.data(), bp.grp_step.size() * 4);
    put(s.o_grp_q0, bp.grp_q0.data(), bp.grp_q0.size() * 8);
    put(s.o_qorder, bp.qorder.data(), size_t(M) * 4);
    put(s.o_spos, bp.spos.data(), size_t(M) * 4);
    put(s.o_srank, bp.srank.data(), size_t(M) * 2);
    put(s.o_ipos, bp.ipos.data(), size_t(M) * 4);
    put(s.o_qrank, bp.qrank.data(), size_t(M) * 4);
    put(s.o_sblk, bp.sblk.data(), size_t(M) * 4);
    put(s.o_qtile, bp.qtile.data(), size_t(M) * 4);
    put(s.o_wk_chunk, bp.wk_chunk.data(), bp.wk_chunk.size() * 8);
    put(s.o_wk_e0, bp.wk_e0.data(), bp.wk_e0.size() * 8);
    put(s.o_eitems, bp.eitems.data(), bp.eitems.size() * 8);
    put(s.o_ov_qid, ov_qid.data(), ov_qid.size() * 4);
    put(s.o_ov_state, ov_state.data(), ov_state.size() * 4);
    BwtBatch b{};
    b.M = M;
    b.qlo = reinterpret_cast<const uint32_t*>(s.dev + s.o_qlo);
    b.qhi = reinterpret_cast<const uint32_t*>(s.dev + s.o_qhi);
    b.nqblk = int32_t(bp.qblk.size());
    b.qblk = reinterpret_cast<const int32_t*>(s.dev + s.o_qblk);
    b.qblk_g0 = reinterpret_cast<const int64_t*>(s.dev + s.o_qblk_g0);
    b.ngrp = int64_t(bp.grp_step.size());
    b.grp_step = reinterpret_cast<const int32_t*>(s.dev + s.o_grp_step);
    b.grp_q0 = reinterpret_cast<const int64_t*>(s.dev + s.o_grp_q0);
    b.qorder = reinterpret_cast<const uint32_t*>(s.dev + s.o_qorder);
    b.spos = reinterpret_cast<const uint32_t*>(s.dev + s.o_spos);
    b.srank = reinterpret_cast<const uint16_t*>(s.dev + s.o_srank);
    b.ipos = reinterpret_cast<const uint32_t*>(s.dev + s.o_ipos);
    b.qrank = reinterpret_cast<const uint32_t*>(s.dev + s.o_qrank);
    b.sblk = reinterpret_cast<const int32_t*>(s.dev + s.o_sblk);
    b.qtile = reinterpret_cast<const uint32_t*>(s.dev + s.o_qtile);
    b.nwork = int64_t(bp.wk_chunk.size());
    b.wk_chunk = reinterpret_cast<const int64_t*>(s.dev + s.o_wk_chunk);
    b.wk_e0 = reinterpret_cast<const int64_t*>(s.dev + s.o_wk_e0);
    b.eitems = reinterpret_cast<const uint64_t*>(s.dev + s.o_eitems);
    b.ansr = reinterpret_cast<int64_t*>(s.dev + s.o_ansr);
    b.g = reinterpret_cast<int64_t*>(s.dev + s.o_g);
    b.stage = reinterpret_cast<int64_t*>(s.dev + s.o_stage);
    b.answers = reinterpret_cast<int64_t*>(s.dev + s.o_answers);
    HIPX(hipStreamWaitEvent(s.st, app_ready, 0));  // appended states admitted so far are on the device
    BwtLaunch L = E.launch;
    RUN(bwt_query_launch(&E.x, &b, &L, s.st));
    if (!ov_qid.empty()) {
      BwtOverlay o{};
      o.S = S;
      o.U = AU;
      o.n = int64_t(ov_qid.size());
      o.qid = reinterpret_cast<const uint32_t*>(s.dev + s.o_ov_qid);
      o.qstate = reinterpret_cast<const int32_t*>(s.dev + s.o_ov_state);
      o.qlo = b.qlo;
      o.qhi = b.qhi;
      o.app_parent = d_app_parent;
      o.app_key = d_app_key;
      o.app_delta = d_app_delta;
      o.answers = b.answers;
      RUN(bwt_overlay_launch(&o, s.st));
    }
    HIPX(hipMemcpyAsync(s.host + s.o_answers, s.dev + s.o_answers, size_t(M) * 8, hipMemcpyDeviceToHost, s.st));
    HIPX(hipEventRecord(s.done, s.st));
    s.busy = true;
    s.launched = clk::now();
  };
  auto admit_append = [&](int64_t r) {
    const int64_t a = rs.req_app[size_t(r)];
    HIPX(hipMemcpyAsync(d_app_parent + a, &rs.app_parent[size_t(a)], 4, hipMemcpyHostToDevice, ist));
    HIPX(hipMemcpyAsync(d_app_key + a * AU, &rs.app_key[size_t(a * AU)], size_t(AU) * 4, hipMemcpyHostToDevice, ist));
    HIPX(hipMemcpyAsync(d_app_delta + a * AU, &rs.app_delta[size_t(a * AU)], size_t(AU) * 8, hipMemcpyHostToDevice,
                        ist));
    HIPX(hipEventRecord(app_ready, ist));
  };

  // ---- serving loop (open loop: arrivals follow the stream's clock from now on)
  std::vector<int64_t> ans(static_cast<size_t>(NQ), 0);
  std::vector<double> lat;  // seconds, query requests
  lat.reserve(size_t(R));
  std::vector<double> done_at(static_cast<size_t>(R), -1.0);
  int64_t next = 0, nbatches = 0, max_batch_q = 0, nappended = 0;
  double sum_batch_q = 0;
  const auto t0 = clk::now();
  while (next < R || std::any_of(slots.begin(), slots.end(), [](const Slot& s) { return s.busy; })) {
    bool progressed = false;
    // harvest finished batches
    for (Slot& s : slots) {
      if (!s.busy || hipEventQuery(s.done) != hipSuccess) continue;
      const double now = secs(t0, clk::now());
      const int64_t* out = reinterpret_cast<const int64_t*>(s.host + s.o_answers);
      for (size_t i = 0; i < s.reqs.size(); ++i) {
        const int64_t r = s.reqs[i];
        const int64_t q0 = rs.req_q0[size_t(r)], n = rs.req_q0[size_t(r) + 1] - q0;
        std::memcpy(&ans[size_t(q0)], out + s.req_off[i], size_t(n) * 8);
        done_at[size_t(r)] = now;
        lat.push_back(now - rs.arrival[size_t(r)]);
      }
      s.busy = false;
      progressed = true;
    }
    // admit arrived requests into a free slot
    const double now = secs(t0, clk::now());
    for (Slot& s : slots) {
      if (s.busy || next >= R || rs.arrival[size_t(next)] > now) continue;
      s.reqs.clear();
      s.req_off.clear();
      int64_t nq = 0;
      while (next < R && rs.arrival[size_t(next)] <= now) {
        if (rs.req_kind[size_t(next)]) {
          admit_append(next);
          done_at[size_t(next)] = secs(t0, clk::now());
          ++nappended;
          ++next;
          continue;
        }
        const int64_t n = rs.req_q0[size_t(next) + 1] - rs.req_q0[size_t(next)];
        if (nq + n > s.cap) break;  // the HBM budget of this slot is full
        s.reqs.push_back(next);
        s.req_off.push_back(nq);
        nq += n;
        ++next;
      }
      if (s.reqs.empty()) continue;
      launch_batch(s);
      ++nbatches;
      sum_batch_q += double(nq);
      max_batch_q = std::max(max_batch_q, nq);
      progressed = true;
    }
    if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
  const double wall = secs(t0, clk::now());

  // ---- alone versus in the mix
  int64_t alone_mismatch = 0, alone_checked = 0;
  const int64_t nquery_req = R - nappended;
  std::vector<int64_t> qreqs;
  for (int64_t r = 0; r < R; ++r)
    if (!rs.req_kind[size_t(r)]) qreqs.push_back(r);
  for (int k = 0; k < check_alone && !qreqs.empty(); ++k) {
    const int64_t r = qreqs[size_t(int64_t(k) * int64_t(qreqs.size()) / std::max(1, check_alone))];
    Slot& s = slots[0];
    s.reqs.assign(1, r);
    s.req_off.assign(1, 0);
    launch_batch(s);
    HIPX(hipEventSynchronize(s.done));
    s.busy = false;
    const int64_t q0 = rs.req_q0[size_t(r)], n = rs.req_q0[size_t(r) + 1] - q0;
    const int64_t* out = reinterpret_cast<const int64_t*>(s.host + s.o_answers);
    for (int64_t i = 0; i < n; ++i) alone_mismatch += out[i] != ans[size_t(q0 + i)];
    ++alone_checked;
  }

  std::sort(lat.begin(), lat.end());
  auto pct = [&](double p) { return lat.empty() ? 0.0 : lat[std::min(lat.size() - 1, size_t(p * double(lat.size())))]; };
  const std::string sha = sha_hex(ans.data(), size_t(NQ) * 8);
  if (!answers_path.empty()) {
    FILE* f = std::fopen(answers_path.c_str(), "wb");
    if (!f) return 4;
    const bool ok = std::fwrite(ans.data(), 8, size_t(NQ), f) == size_t(NQ);
    if (std::fclose(f) != 0 || !ok) return 4;
  }
  hipDeviceProp_t prop{};
  HIPX(hipGetDeviceProperties(&prop, 0));
  const char* dev = prop.name[0] ? prop.name : prop.gcnArchName;
  const double offered = rs.arrival.empty() ? 0 : double(NQ) / std::max(1e-9, rs.arrival.back());
  char buf[4096];
  const int n = std::snprintf(
      buf, sizeof(buf),
      "{\"arm\":\"bwgpu_served\",\"form\":5,\"S\":%" PRId64 ",\"U\":%d,\"N\":%" PRIu64 ",\"seed\":%" PRIu64
      ",\"requests\":%" PRId64 ",\"query_requests\":%" PRId64 ",\"appends\":%" PRId64 ",\"rate\":%.1f,"
      "\"max_batch\":%d,\"n_answers\":%" PRId64 ",\"answers_sha256\":\"%s\",\"slots\":%d,\"budget_mb\":%.0f,"
      "\"slot_capacity_queries\":%" PRId64 ",\"batches\":%" PRId64 ",\"mean_batch_queries\":%.1f,"
      "\"max_batch_queries\":%" PRId64 ",\"wall_s\":%.4f,\"offered_queries_per_s\":%.1f,\"queries_per_s\":%.1f,"
      "\"latency_ms\":{\"p50\":%.3f,\"p90\":%.3f,\"p99\":%.3f,\"max\":%.3f},\"host_plan_s\":%.4f,"
      "\"alone\":{\"requests\":%" PRId64 ",\"mismatch\":%" PRId64 "},\"knobs_from\":\"%s\",\"setup_s\":%.3f,"
      "\"device\":\"%s\",\"cus\":%d}\n",
      S, spec.U, spec.N, spec.seed, R, nquery_req, nappended, ss.rate, ss.max_batch, NQ, sha.c_str(), nslots, budget_mb,
      cap, nbatches, nbatches ? sum_batch_q / double(nbatches) : 0.0, max_batch_q, wall, offered, double(NQ) / wall,
      pct(0.5) * 1e3, pct(0.9) * 1e3, pct(0.99) * 1e3, lat.empty() ? 0.0 : lat.back() * 1e3, plan_s, alone_checked,
      alone_mismatch, tuned_from.c_str(), secs(g0, g1), dev, prop.multiProcessorCount);
  std::fwrite(buf, 1, size_t(n), stdout);
  if (!json_path.empty()) {
    FILE* f = std::fopen(json_path.c_str(), "w");
    if (!f) return 4;
    std::fwrite(buf, 1, size_t(n), f);
    std::fclose(f);
  }
  for (Slot& s : slots) {
    (void)hipFree(s.dev);
    (void)hipHostFree(s.host);
  }
  return alone_mismatch == 0 ? 0 : 6;
}
