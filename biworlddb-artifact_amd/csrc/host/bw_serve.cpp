// bw_serve.cpp -- the served form: an open-loop request stream answered by continuous batching
// over a resident index, with appended branch states (one JSON line).
//
//   bw_serve [--cfg 2|3] [--S n --U n --N n --seed n] [--requests 20000 --max-batch 4096 --append-frac 0.05
//            --append-U 128 --rate 20000] [--tuned FILE] [--slots 2] [--budget-mb 8192] [--slot-queries 16777216]
//            [--check-alone 64] [--answers FILE] [--json FILE] [--timeline] [--tile-shift T]
//            [--form dense|two-level|hybrid] [--dense-from 1048576]
//
// Resident in HBM: the version tree's index (TreeEngine), the appended states' updates, and per
// in-flight batch one fixed buffer ("slot") whose size is the HBM budget divided among the slots
// (at most --slot-queries queries).
// Requests arrive at their Poisson times (a clock started when serving starts).  Whenever a slot
// is free, every request that has arrived and fits the slot's capacity is admitted in arrival
// order: appends join the round's contiguous range of appended states, copied to the device
// (from pinned receive buffers, one copy per array) before the round's batch launches, on a
// stream every batch waits on; query requests join the batch.  Requests land in pinned receive buffers (state, lo, hi per query, in
// arrival order), so a batch's queries are one contiguous slice of them: the host only replaces
// appended states by their deepest indexed ancestor (listing those queries for the overlay) and
// copies the slice in; the batch is planned and answered on the device and its answers copied
// straight into the pinned answer array, all on the slot's own stream, so batches on different
// slots overlap on the device and on the copy engines.  --form picks how a batch is answered:
// `two-level` (the default: include/bwgpu_tree.h, the device batch planner and the query launch
// sequence) or, over a config-2 tree, `dense` (include/bwgpu_dense.h: the dense carry of blocks of
// up to 4,096 keys resident beside the index; bucket plan by state, persistent block replay,
// unsort).  At small batches the dense replay's walks (each queried block replayed up to its state)
// cost more than the two-level form's whole batch (measured at 20,000 requests/s, ~1,850 queries
// per batch: p50 0.63 / p99 1.16 ms dense against 0.56 / 0.95 ms two-level; even at 320,000
// requests/s, ~240K per batch); over the stream offered at once (9.1M-query batches) it answers
// 1.23-1.26 G queries/s against 0.73-1.03 G.  `hybrid` keeps both indexes resident and answers a
// batch of --dense-from queries or more (2^20 by default) with the dense form, smaller ones with the
// two-level form.  A config-3 tree's 2^22 ranks need the two-level form (a dense carry cannot hold
// them).  A query on an appended state is answered at its deepest indexed ancestor plus the overlay
// of the appended updates on its path.
// Latency = time the request's answers are back on the host - its arrival time.
// Before the clock starts, one small batch per slot loads the kernels, and copies in both
// directions on every stream warm up the copy engines (their first use costs milliseconds).
// --timeline prints each batch's host launch steps and completion times on stderr.  --tile-shift
// overrides the tuned key-tile width (0: the balanced rule).
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

#include <omp.h>

#include "bwgpu_dense.h"
#include "bwgpu_tree.h"
#include "dense_engine.hpp"
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

// Fixed-capacity device + pinned host buffers of one in-flight batch: the raw queries (pinned
// staging and device copies), the overlay's query lists, and the device planner's slot (plan
// outputs, result buffers, workspace).  A batch uses a prefix of each.
struct Slot {
  int64_t cap = 0;  // queries
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  hipEvent_t start = nullptr;
  hipEvent_t mark[3] = {};  // --timeline: after the copies in, after the plan, before the copy out
  unsigned char* host = nullptr;  // pinned: overlay ids | overlay states | answers of a batch run alone
  size_t h_ovq = 0, h_ovs = 0, h_ans = 0, host_bytes = 0;
  DeviceArena in;
  int32_t* d_state = nullptr;
  uint32_t* d_lo = nullptr;
  uint32_t* d_hi = nullptr;
  uint32_t* d_ovq = nullptr;
  int32_t* d_ovs = nullptr;
  DevicePlanSlot plan;  // two-level form
  DenseBatch db;        // dense form: plan and replay workspace, flags
  int64_t* d_dans = nullptr;
  // the batch in flight
  bool busy = false;
  std::vector<int64_t> reqs;     // request ids
  std::vector<int64_t> req_off;  // their first query in the batch
  int64_t nq = 0;
  int64_t q0 = 0;                // the batch's first query in the stream
  clk::time_point launched;
};

size_t host_layout(Slot& s, int64_t cap) {
  size_t o = 0;
  auto take = [&](size_t n) {
    const size_t at = o;
    o += (n + 255) & ~size_t(255);
    return at;
  };
  const size_t n = size_t(cap);
  s.h_ovq = take(n * 4);
  s.h_ovs = take(n * 4);
  s.h_ans = take(n * 8);
  return o;
}

// Device bytes of one slot of cap queries (inputs, overlay lists, plan outputs and results,
// planner workspace).
size_t slot_device_bytes(const TreeEngine& E, int64_t cap) {
  size_t work = 0;
  if (bwt_plan_workspace(&E.x, &E.px, cap, &work) != 0) return 0;
  const size_t n = size_t(cap);
  const size_t outputs = n * (4 + 8 + 4 + 8 + 4 + 4 + 4 + 4 + 16 + 16 + 16 + 8 + 4 + 2 + 4 + 8 + 16 + 8 + 8);
  return work + outputs + n * 20 + 4096;
}
// ... and for the dense form: its workspace, the answers, the inputs and overlay lists
size_t dense_slot_device_bytes(const DenseEngine& D, int64_t cap) {
  size_t work = 0;
  if (bwd_answer_workspace(&D.x, cap, &work) != 0) return 0;
  return work + size_t(cap) * (8 + 20) + 4096;
}

}  // namespace

int main(int argc, char** argv) {
  // the tree: config 2 (keys are their own ranks), or --cfg 3 (distinct 32-bit keys with the
  // rank-LUT resident: every batch's raw endpoints and every appended key are ranked)
  int tcfg = 2;
  for (int i = 1; i + 1 < argc; ++i)
    if (!std::strcmp(argv[i], "--cfg")) tcfg = std::atoi(argv[i + 1]);
  if (tcfg != 2 && tcfg != 3) { std::fprintf(stderr, "bw_serve: --cfg 2|3\n"); return 2; }
  const bool ranked = tcfg == 3;
  TreeSpec spec = tree_defaults(tcfg);
  StreamSpec ss = stream_defaults(5);
  PlanKnobs kn;
  LaunchKnobs lk;
  int nslots = 2, check_alone = 64;
  int64_t slot_queries = int64_t(1) << 24;  // most queries one batch takes
  int tile_shift_arg = -1;  // --tile-shift: key-tile width override (0: the balanced rule)
  std::string form_arg;     // --form dense|two-level|hybrid (default: two-level)
  int64_t dense_from = int64_t(1) << 20;  // --dense-from: hybrid batches of at least this many queries go dense
  bool timeline = false;  // --timeline: per-batch host times on stderr  // most queries one batch takes (batches pipeline across slots)
  double budget_mb = 8192;
  std::string answers_path, json_path, tuned_path;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--timeline")) { timeline = true; continue; }
    if (i + 1 >= argc) { std::fprintf(stderr, "bw_serve: missing value for %s\n", argv[i]); return 2; }
    const char* k = argv[i];
    const char* v = argv[++i];
    if (!std::strcmp(k, "--cfg")) continue;
    else if (!std::strcmp(k, "--S")) spec.S = std::atoll(v);
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
    else if (!std::strcmp(k, "--slot-queries")) slot_queries = std::atoll(v);
    else if (!std::strcmp(k, "--check-alone")) check_alone = std::atoi(v);
    else if (!std::strcmp(k, "--tile-shift")) tile_shift_arg = std::atoi(v);
    else if (!std::strcmp(k, "--form")) form_arg = v;
    else if (!std::strcmp(k, "--dense-from")) dense_from = std::atoll(v);
    else if (!std::strcmp(k, "--answers")) answers_path = v;
    else if (!std::strcmp(k, "--json")) json_path = v;
    else { std::fprintf(stderr, "bw_serve: unknown argument %s\n", k); return 2; }
  }
  if (spec.S < 1 || spec.U < 1 || spec.N < 2 || ss.nreq < 1 || ss.max_batch < 1 || ss.append_U < 1 || ss.rate <= 0 ||
      nslots < 1 || nslots > 8 || budget_mb <= 0 || check_alone < 0) {
    std::fprintf(stderr, "bw_serve: bad sizes\n");
    return 2;
  }
  if (!form_arg.empty() && form_arg != "dense" && form_arg != "two-level" && form_arg != "hybrid") {
    std::fprintf(stderr, "bw_serve: --form dense|two-level|hybrid\n");
    return 2;
  }
  const bool dense = form_arg == "dense";    // every batch dense
  const bool hybrid = form_arg == "hybrid";  // batches of dense_from or more queries dense, the rest two-level
  const bool with_dense = dense || hybrid;
  if (with_dense && ranked) {
    std::fprintf(stderr, "bw_serve: the dense form needs keys that are their own ranks (config 2)\n");
    return 3;
  }
  std::string tuned_from = "defaults";
  if (!tuned_path.empty() && load_tuned(tuned_path.c_str(), "served", kn, lk)) tuned_from = tuned_path;
  if (dense) {  // blocks of up to 4,096 keys (the dense replay's LDS, a 47 GB carry), indexed by
    kn.rmax = 4096;  // 1,024-thread workgroups (at most 16 ranks per thread)
    lk.block_threads = 1024;
  }
  // the served knobs were tuned on config 2's 2^20 keys; over config 3's 2^22 ranks their key
  // tiles (2^9 keys: 8,193 tiles) still fit the 14-bit class field and measured best (p99 0.93 ms
  // against 1.27 ms with the balanced rule's 2^11 tiles); a rank space the field cannot hold in
  // tiles that narrow falls back to the balanced rule
  if (ranked && kn.tile_shift > 0 &&
      (uint64_t(spec.N) >> kn.tile_shift) + 1 > uint64_t(BWT_MAX_CLASSES - 1))
    kn.tile_shift = 0;
  if (tile_shift_arg >= 0) kn.tile_shift = tile_shift_arg;

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
  // ranked: the index is built over the keys' ranks; the LUT (65,537 row pointers, the low 16
  // bits of every distinct key) stays on the device for the batches
  RankLut lut;
  std::vector<uint32_t> kranks;
  const uint32_t* kcoord = w.key.data();
  DeviceArena lutmem;
  const uint32_t* d_row = nullptr;
  const uint16_t* d_low = nullptr;
  if (ranked) {
    lut = build_rank_lut(w.dkey);
    kranks.resize(w.key.size());
    host_ranks(lut, w.key.data(), kranks.data(), w.key.size());
    kcoord = kranks.data();
    d_row = lutmem.upload(lut.row_ptr);
    d_low = lutmem.upload(lut.low16);
    if (!lutmem.ok()) return 5;
  }
  TreeEngine E;
  if (!E.init(S, spec.U, w.coord_n(), w.parent.data(), kcoord, w.delta.data(), kn,
              BwtLaunch{lk.block_threads, lk.chunk_waves, 0, 0}, err, sizeof(err))) {
    std::fprintf(stderr, "bw_serve: index: %s\n", err);
    return std::strncmp(err, "device", 6) == 0 ? 5 : 3;
  }
  hipStream_t ist;
  HIPX(hipStreamCreate(&ist));
  RUN(E.build_index(ist));
  HIPX(hipStreamSynchronize(ist));
  // the dense form's index: E itself (--form dense: built with blocks of 4,096 keys), or beside the
  // two-level one for the hybrid form
  TreeEngine E4;
  PlanKnobs kn4 = kn;
  kn4.rmax = 4096;
  if (hybrid) {
    if (!E4.init(S, spec.U, w.coord_n(), w.parent.data(), kcoord, w.delta.data(), kn4, BwtLaunch{1024, lk.chunk_waves, 0, 0},
                 err, sizeof(err))) {
      std::fprintf(stderr, "bw_serve: dense index: %s\n", err);
      return std::strncmp(err, "device", 6) == 0 ? 5 : 3;
    }
    RUN(E4.build_index(ist));
    HIPX(hipStreamSynchronize(ist));
  }
  DenseEngine D;
  if (with_dense && !D.init(hybrid ? E4 : E, ist, err, sizeof(err))) {
    std::fprintf(stderr, "bw_serve: %s\n", err);
    return 5;
  }
  // the dense replay's shape from the tuned file's dense entry (its block size is this index's)
  DenseKnobs dk;
  if (with_dense && !tuned_path.empty() && load_dense_tuned(tuned_path.c_str(), dk) && dk.rmax == 4096)
    D.launch = BwdLaunch{dk.threads, dk.wg_per_cu, 0, BWD_STORE_AUTO, dk.chunk};
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
  // pinned receive buffers: every request's queries land here in arrival order, and the answers
  // come back into p_ans at the same offsets
  int32_t* p_state = nullptr;
  uint32_t *p_lo = nullptr, *p_hi = nullptr;
  int64_t* p_ans = nullptr;
  HIPX(hipHostMalloc(reinterpret_cast<void**>(&p_state), size_t(std::max<int64_t>(NQ, 1)) * 4, hipHostMallocDefault));
  HIPX(hipHostMalloc(reinterpret_cast<void**>(&p_lo), size_t(std::max<int64_t>(NQ, 1)) * 4, hipHostMallocDefault));
  HIPX(hipHostMalloc(reinterpret_cast<void**>(&p_hi), size_t(std::max<int64_t>(NQ, 1)) * 4, hipHostMallocDefault));
  HIPX(hipHostMalloc(reinterpret_cast<void**>(&p_ans), size_t(std::max<int64_t>(NQ, 1)) * 8, hipHostMallocDefault));
  std::memcpy(p_state, rs.qstate.data(), size_t(NQ) * 4);
  std::memcpy(p_lo, rs.qlo.data(), size_t(NQ) * 4);
  std::memcpy(p_hi, rs.qhi.data(), size_t(NQ) * 4);
  // appended states land in pinned receive buffers too (keys ranked at admission over a ranked tree)
  int32_t* p_app_parent = nullptr;
  uint32_t* p_app_key = nullptr;
  int64_t* p_app_delta = nullptr;
  const size_t nap = size_t(std::max<int64_t>(NA, 1));
  HIPX(hipHostMalloc(reinterpret_cast<void**>(&p_app_parent), nap * 4, hipHostMallocDefault));
  HIPX(hipHostMalloc(reinterpret_cast<void**>(&p_app_key), nap * size_t(AU) * 4, hipHostMallocDefault));
  HIPX(hipHostMalloc(reinterpret_cast<void**>(&p_app_delta), nap * size_t(AU) * 8, hipHostMallocDefault));
  std::memcpy(p_app_parent, rs.app_parent.data(), size_t(NA) * 4);
  std::memcpy(p_app_key, rs.app_key.data(), size_t(NA) * size_t(AU) * 4);
  std::memcpy(p_app_delta, rs.app_delta.data(), size_t(NA) * size_t(AU) * 8);
  hipEvent_t app_ready;
  HIPX(hipEventCreate(&app_ready));
  HIPX(hipEventRecord(app_ready, ist));
  hipStream_t ast;  // the appended states' copies
  HIPX(hipStreamCreateWithFlags(&ast, hipStreamNonBlocking));
  // slots within the HBM budget: the largest capacity whose device buffers fit the slot's share
  std::vector<Slot> slots(static_cast<size_t>(nslots));
  const size_t share = size_t(budget_mb * 1048576.0 / nslots);
  int64_t cap = ss.max_batch;
  {
    auto bytes = [&](int64_t c) {
      return (with_dense ? dense_slot_device_bytes(D, c) : 0) + (dense ? 0 : slot_device_bytes(E, c));
    };
    const int64_t probe = int64_t(1) << 20;
    const size_t pb = bytes(probe);
    if (pb == 0) { std::fprintf(stderr, "bw_serve: planner workspace\n"); return 5; }
    const double per_q = double(pb) / double(probe);
    cap = std::max<int64_t>(cap, int64_t(double(share) / per_q));
    while (cap > ss.max_batch && bytes(cap) > share) cap = cap * 15 / 16;
    cap = std::max<int64_t>(std::min(cap, std::max<int64_t>(slot_queries, 1)), ss.max_batch);
  }
  for (Slot& s : slots) {
    s.cap = cap;
    s.host_bytes = host_layout(s, cap);
    HIPX(hipHostMalloc(reinterpret_cast<void**>(&s.host), s.host_bytes, hipHostMallocDefault));
    s.d_state = s.in.scratch<int32_t>(size_t(cap));
    s.d_lo = s.in.scratch<uint32_t>(size_t(cap));
    s.d_hi = s.in.scratch<uint32_t>(size_t(cap));
    s.d_ovq = s.in.scratch<uint32_t>(size_t(cap));
    s.d_ovs = s.in.scratch<int32_t>(size_t(cap));
    const bool planned = (!with_dense || (alloc_dense_batch(D, cap, s.db, err, sizeof(err)) &&
                                          (s.d_dans = s.db.arena.scratch<int64_t>(size_t(cap))) != nullptr)) &&
                         (dense || alloc_plan_slot(E, cap, s.plan, err, sizeof(err)));
    if (!s.in.ok() || !planned) {
      std::fprintf(stderr, "bw_serve: slot allocation: %s\n", s.in.ok() ? err : "inputs");
      return 5;
    }
    HIPX(hipStreamCreate(&s.st));
    HIPX(hipEventCreate(&s.done));
    HIPX(hipEventCreate(&s.start));
    for (auto& m : s.mark) HIPX(hipEventCreate(&m));
  }
  const auto g1 = clk::now();

  // ---- one batch: stage the raw queries, copy in, plan and answer on the device, copy the
  //      answers out (all on the slot's stream)
  double plan_s = 0;  // host time spent staging batches
  const BwtLaunch L = E.launch;
  // dst: pinned host array receiving the batch's answers (the stream's answer array at the
  // batch's offset, or the slot's own buffer for a request re-answered alone)
  int64_t dense_batches = 0;
  // force: -1 the form's choice, 0 two-level, 1 dense (the warm-up runs each path once per slot)
  auto launch_batch = [&](Slot& s, int64_t* dst, int force = -1) {
    const auto p0 = clk::now();
    uint32_t* hoq = reinterpret_cast<uint32_t*>(s.host + s.h_ovq);
    int32_t* hos = reinterpret_cast<int32_t*>(s.host + s.h_ovs);
    // the batch's requests are consecutive query requests (appends carry no queries), so their
    // queries are one contiguous slice of the receive buffers
    const int64_t q0 = rs.req_q0[size_t(s.reqs.front())];
    const int64_t M = rs.req_q0[size_t(s.reqs.back()) + 1] - q0;
    // queries on appended states: answered at their deepest indexed ancestor plus the overlay;
    // the slice is scanned in parallel pieces (a count pass, then each piece fills its part)
    int64_t nov = 0;
    const int npc = M >= (int64_t(1) << 16) ? std::max(1, omp_get_max_threads()) : 1;
    std::vector<int64_t> pcnt(size_t(npc) + 1, 0);
#pragma omp parallel for num_threads(npc) schedule(static, 1) if (npc > 1)
    for (int t = 0; t < npc; ++t) {
      int64_t c = 0;
      for (int64_t q = q0 + M * t / npc; q < q0 + M * (t + 1) / npc; ++q) c += rs.qstate[size_t(q)] > S;
      pcnt[size_t(t) + 1] = c;
    }
    for (int t = 0; t < npc; ++t) pcnt[size_t(t) + 1] += pcnt[size_t(t)];
    nov = pcnt[size_t(npc)];
    if (nov) {
#pragma omp parallel for num_threads(npc) schedule(static, 1) if (npc > 1)
      for (int t = 0; t < npc; ++t) {
        int64_t o = pcnt[size_t(t)];
        for (int64_t q = q0 + M * t / npc; q < q0 + M * (t + 1) / npc; ++q) {
          const int32_t v = rs.qstate[size_t(q)];
          if (v > S) {
            hoq[o] = uint32_t(q - q0);
            hos[o++] = v;
            p_state[q] = app_base[size_t(v - S - 1)];
          }
        }
      }
    }
    const auto p1 = clk::now();
    int32_t* hs = p_state + q0;
    uint32_t* hl = p_lo + q0;
    uint32_t* hh = p_hi + q0;
    s.q0 = q0;
    s.nq = M;
    HIPX(hipEventRecord(s.start, s.st));
    HIPX(hipMemcpyAsync(s.d_state, hs, size_t(M) * 4, hipMemcpyHostToDevice, s.st));
    HIPX(hipMemcpyAsync(s.d_lo, hl, size_t(M) * 4, hipMemcpyHostToDevice, s.st));
    HIPX(hipMemcpyAsync(s.d_hi, hh, size_t(M) * 4, hipMemcpyHostToDevice, s.st));
    if (nov) {
      HIPX(hipMemcpyAsync(s.d_ovq, hoq, size_t(nov) * 4, hipMemcpyHostToDevice, s.st));
      HIPX(hipMemcpyAsync(s.d_ovs, hos, size_t(nov) * 4, hipMemcpyHostToDevice, s.st));
    }
    if (ranked) {  // raw endpoints -> ranks (#keys < x), in place
      BwtRankArgs rl{M, s.d_lo, s.d_lo, d_row, d_low}, rh{M, s.d_hi, s.d_hi, d_row, d_low};
      RUN(bwt_rank_remap_launch(&rl, s.st));
      RUN(bwt_rank_remap_launch(&rh, s.st));
    }
    HIPX(hipStreamWaitEvent(s.st, app_ready, 0));  // appended states admitted so far are on the device
    if (timeline) HIPX(hipEventRecord(s.mark[0], s.st));
    const auto p2 = clk::now();
    const bool dn = force >= 0 ? force == 1 : (dense || (hybrid && M >= dense_from));
    if (dn && force < 0) ++dense_batches;
    if (dn)
      RUN(bwd_plan_launch(&D.x, M, s.d_state, s.d_lo, s.d_hi, s.db.flags, s.db.work, s.db.work_bytes, s.st));
    else
      RUN(device_plan(E, s.plan, M, s.d_state, s.d_lo, s.d_hi, s.st));
    if (timeline) HIPX(hipEventRecord(s.mark[1], s.st));
    const auto p3 = clk::now();
    int64_t* answers = dn ? s.d_dans : s.plan.b.answers;
    if (dn) {
      RUN(bwd_replay_launch(&D.x, M, answers, &D.launch, s.db.work, s.db.work_bytes, s.st));
    } else {
      BwtLaunch l = L;
      RUN(bwt_query_launch(&E.x, &s.plan.b, &l, s.st));
    }
    if (nov) {
      BwtOverlay o{};
      o.S = S;
      o.U = AU;
      o.n = nov;
      o.qid = s.d_ovq;
      o.qstate = s.d_ovs;
      o.qlo = s.d_lo;
      o.qhi = s.d_hi;
      o.app_parent = d_app_parent;
      o.app_key = d_app_key;
      o.app_delta = d_app_delta;
      o.answers = answers;
      RUN(bwt_overlay_launch(&o, s.st));
    }
    if (timeline) HIPX(hipEventRecord(s.mark[2], s.st));
    HIPX(hipMemcpyAsync(dst, answers, size_t(M) * 8, hipMemcpyDeviceToHost, s.st));
    HIPX(hipEventRecord(s.done, s.st));
    s.busy = true;
    s.launched = clk::now();
    if (timeline)
      std::fprintf(stderr, "launch %" PRId64 " queries: scan %.3f ms, copies-in %.3f ms, plan launches %.3f ms, rest %.3f ms\n",
                   M, secs(p0, p1) * 1e3, secs(p1, p2) * 1e3, secs(p2, p3) * 1e3, secs(p3, s.launched) * 1e3);
    plan_s += secs(p0, s.launched);
  };
  // appends admitted but not yet copied: the contiguous range [app_lo, app_hi) of appended states
  int64_t app_lo = 0, app_hi = 0;
  auto admit_append = [&](int64_t r) {
    const int64_t a = rs.req_app[size_t(r)];
    if (ranked) {  // an appended key is a distinct key: its rank, from the LUT row of its high bits
      uint32_t* keys = p_app_key + size_t(a * AU);
      for (int i = 0; i < AU; ++i) {
        const uint32_t x = rs.app_key[size_t(a * AU) + size_t(i)];
        const uint16_t* b = lut.low16.data() + lut.row_ptr[x >> 16];
        const uint16_t* e = lut.low16.data() + lut.row_ptr[(x >> 16) + 1];
        keys[i] = uint32_t(std::lower_bound(b, e, uint16_t(x & 0xFFFFu)) - lut.low16.data());
      }
    }
    if (app_hi == app_lo) app_lo = app_hi = a;
    app_hi = a + 1;  // appends are admitted in order
  };
  auto flush_appends = [&]() {
    if (app_hi == app_lo) return;
    const size_t n = size_t(app_hi - app_lo);
    const auto f0 = clk::now();
    HIPX(hipMemcpyAsync(d_app_parent + app_lo, p_app_parent + app_lo, n * 4, hipMemcpyHostToDevice, ast));
    const auto f1 = clk::now();
    HIPX(hipMemcpyAsync(d_app_key + app_lo * AU, p_app_key + app_lo * AU, n * size_t(AU) * 4, hipMemcpyHostToDevice, ast));
    const auto f2 = clk::now();
    HIPX(hipMemcpyAsync(d_app_delta + app_lo * AU, p_app_delta + app_lo * AU, n * size_t(AU) * 8, hipMemcpyHostToDevice,
                        ast));
    const auto f3 = clk::now();
    HIPX(hipEventRecord(app_ready, ast));
    if (timeline)
      std::fprintf(stderr, "appends %zu: copies %.3f %.3f %.3f ms, record %.3f ms\n", n, secs(f0, f1) * 1e3,
                   secs(f1, f2) * 1e3, secs(f2, f3) * 1e3, secs(f3, clk::now()) * 1e3);
    app_lo = app_hi;
  };

  // ---- warm-up: one small batch per slot (first launches load the kernels), not timed
  {
    int64_t r0 = 0;
    while (r0 < R && rs.req_kind[size_t(r0)]) ++r0;
    if (r0 < R)
      for (Slot& s : slots) {
        s.reqs.assign(1, r0);
        s.req_off.assign(1, 0);
        for (int f = hybrid ? 0 : -1; f <= (hybrid ? 1 : -1); ++f) {  // hybrid: each path once
          launch_batch(s, reinterpret_cast<int64_t*>(s.host + s.h_ans), f);
          HIPX(hipEventSynchronize(s.done));
          s.busy = false;
        }
      }
    // copies: the runtime's first copies through each copy engine pay a one-time set-up
    // (measured ~5-6 ms, once per engine), so before the clock starts copies go in both
    // directions on every stream at once, at several sizes
    for (size_t n = size_t(4) << 10; n <= (size_t(16) << 20); n *= 16) {
      for (Slot& s : slots) {
        const size_t m = std::min({n, size_t(NQ) * 4, size_t(s.cap) * 4});
        HIPX(hipMemcpyAsync(s.d_lo, p_lo, m, hipMemcpyHostToDevice, s.st));
        HIPX(hipMemcpyAsync(s.host + s.h_ans, with_dense ? s.d_dans : s.plan.b.answers, m, hipMemcpyDeviceToHost, s.st));
      }
      if (NA > 0) {
        const size_t m = std::min(n, size_t(NA) * size_t(AU) * 8);
        HIPX(hipMemcpyAsync(d_app_delta, p_app_delta, m, hipMemcpyHostToDevice, ast));
      }
      for (Slot& s : slots) HIPX(hipStreamSynchronize(s.st));
      HIPX(hipStreamSynchronize(ast));
    }
    // and every append array at one append's size, 64 KiB and whole (the device arrays are
    // overwritten at admission)
    if (NA > 0) {
      const size_t whole[3] = {size_t(NA) * 4, size_t(NA) * size_t(AU) * 4, size_t(NA) * size_t(AU) * 8};
      const size_t one[3] = {4, size_t(AU) * 4, size_t(AU) * 8};
      void* dst[3] = {d_app_parent, d_app_key, d_app_delta};
      const void* src[3] = {p_app_parent, p_app_key, p_app_delta};
      for (int k = 0; k < 3; ++k)
        for (const size_t n : {one[k], std::min<size_t>(whole[k], size_t(64) << 10), whole[k]})
          HIPX(hipMemcpyAsync(dst[k], src[k], n, hipMemcpyHostToDevice, ast));
      HIPX(hipStreamSynchronize(ast));
    }
    plan_s = 0;
  }
  // ---- serving loop (open loop: arrivals follow the stream's clock from now on)
  std::vector<double> gpu_ms;  // per batch: stream time from its first copy in to its answers out
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
      for (size_t i = 0; i < s.reqs.size(); ++i) {  // the answers are in p_ans already
        const int64_t r = s.reqs[i];
        done_at[size_t(r)] = now;
        lat.push_back(now - rs.arrival[size_t(r)]);
      }
      float t = 0;
      HIPX(hipEventElapsedTime(&t, s.start, s.done));
      gpu_ms.push_back(double(t));
      if (timeline) {
        float c = 0, p = 0, q = 0, o = 0;
        HIPX(hipEventElapsedTime(&c, s.start, s.mark[0]));
        HIPX(hipEventElapsedTime(&p, s.mark[0], s.mark[1]));
        HIPX(hipEventElapsedTime(&q, s.mark[1], s.mark[2]));
        HIPX(hipEventElapsedTime(&o, s.mark[2], s.done));
        std::fprintf(stderr,
                     "batch %" PRId64 " queries: launched %.3f ms, harvested %.3f ms, stream %.3f ms (copies in %.3f, "
                     "plan %.3f, answers %.3f, copy out %.3f)\n",
                     s.nq, secs(t0, s.launched) * 1e3, now * 1e3, double(t), double(c), double(p), double(q), double(o));
      }
      s.busy = false;
      progressed = true;
    }
    // admit arrived requests into a free slot
    const double now = secs(t0, clk::now());
    for (Slot& s : slots) {
      if (s.busy || next >= R || rs.arrival[size_t(next)] > now) continue;
      const auto a0 = clk::now();
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
      const auto a1 = clk::now();
      flush_appends();  // the round's appended states reach the device before its batch
      if (timeline)
        std::fprintf(stderr, "admit at %.3f ms: %zu requests %.3f ms, appends copy %.3f ms\n", secs(t0, a0) * 1e3,
                     s.reqs.size(), secs(a0, a1) * 1e3, secs(a1, clk::now()) * 1e3);
      if (s.reqs.empty()) continue;
      launch_batch(s, p_ans + rs.req_q0[size_t(s.reqs.front())]);
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
    launch_batch(s, reinterpret_cast<int64_t*>(s.host + s.h_ans));
    HIPX(hipEventSynchronize(s.done));
    s.busy = false;
    const int64_t q0 = rs.req_q0[size_t(r)], n = rs.req_q0[size_t(r) + 1] - q0;
    const int64_t* out = reinterpret_cast<const int64_t*>(s.host + s.h_ans);
    for (int64_t i = 0; i < n; ++i) alone_mismatch += out[i] != p_ans[size_t(q0 + i)];
    ++alone_checked;
  }

  // the dense plan flags an out-of-range query (each slot's flag holds its last dense batch's)
  if (with_dense)
    for (Slot& s : slots) {
      int64_t f = 0;
      HIPX(hipMemcpy(&f, s.db.flags, sizeof(f), hipMemcpyDeviceToHost));
      if (f) {
        std::fprintf(stderr, "bw_serve: the dense plan flagged a query (%" PRId64 ")\n", f);
        return 3;
      }
    }
  // latency of the second half of the requests (after any start-up transient)
  std::vector<double> lat2;
  for (int64_t r = R / 2; r < R; ++r)
    if (!rs.req_kind[size_t(r)] && done_at[size_t(r)] >= 0) lat2.push_back(done_at[size_t(r)] - rs.arrival[size_t(r)]);
  std::sort(lat2.begin(), lat2.end());
  const double p99_late = lat2.empty() ? 0.0 : lat2[std::min(lat2.size() - 1, size_t(0.99 * double(lat2.size())))];
  std::sort(gpu_ms.begin(), gpu_ms.end());
  double gpu_mean = 0;
  for (double x : gpu_ms) gpu_mean += x;
  if (!gpu_ms.empty()) gpu_mean /= double(gpu_ms.size());
  const double gpu_p99 = gpu_ms.empty() ? 0.0 : gpu_ms[std::min(gpu_ms.size() - 1, size_t(0.99 * double(gpu_ms.size())))];
  std::sort(lat.begin(), lat.end());
  auto pct = [&](double p) { return lat.empty() ? 0.0 : lat[std::min(lat.size() - 1, size_t(p * double(lat.size())))]; };
  const std::string sha = sha_hex(p_ans, size_t(NQ) * 8);
  if (!answers_path.empty()) {
    FILE* f = std::fopen(answers_path.c_str(), "wb");
    if (!f) return 4;
    const bool ok = std::fwrite(p_ans, 8, size_t(NQ), f) == size_t(NQ);
    if (std::fclose(f) != 0 || !ok) return 4;
  }
  hipDeviceProp_t prop{};
  HIPX(hipGetDeviceProperties(&prop, 0));
  const char* dev = prop.name[0] ? prop.name : prop.gcnArchName;
  const double offered = rs.arrival.empty() ? 0 : double(NQ) / std::max(1e-9, rs.arrival.back());
  char buf[4096];
  const int n = std::snprintf(
      buf, sizeof(buf),
      "{\"arm\":\"bwgpu_served\",\"form\":5,\"tree_cfg\":%d,\"rank_lut\":%s,\"S\":%" PRId64 ",\"U\":%d,\"N\":%" PRIu64 ",\"seed\":%" PRIu64
      ",\"requests\":%" PRId64 ",\"query_requests\":%" PRId64 ",\"appends\":%" PRId64 ",\"rate\":%.1f,"
      "\"max_batch\":%d,\"n_answers\":%" PRId64 ",\"answers_sha256\":\"%s\",\"slots\":%d,\"budget_mb\":%.0f,"
      "\"slot_capacity_queries\":%" PRId64 ",\"batches\":%" PRId64 ",\"mean_batch_queries\":%.1f,"
      "\"max_batch_queries\":%" PRId64 ",\"wall_s\":%.4f,\"offered_queries_per_s\":%.1f,\"queries_per_s\":%.1f,"
      "\"latency_ms\":{\"p50\":%.3f,\"p90\":%.3f,\"p99\":%.3f,\"max\":%.3f},\"planner\":\"device\",\"answer_form\":\"%s\","
      "\"dense_from\":%" PRId64 ",\"dense_batches\":%" PRId64 ","
      "\"host_stage_s\":%.4f,\"slot_device_bytes\":%zu,\"p99_second_half_ms\":%.3f,"
      "\"batch_gpu_ms\":{\"mean\":%.4f,\"p99\":%.4f},"
      "\"alone\":{\"requests\":%" PRId64 ",\"mismatch\":%" PRId64 "},\"knobs_from\":\"%s\",\"setup_s\":%.3f,"
      "\"device\":\"%s\",\"cus\":%d}\n",
      tcfg, ranked ? "true" : "false", S, spec.U, spec.N, spec.seed, R, nquery_req, nappended, ss.rate, ss.max_batch, NQ, sha.c_str(), nslots, budget_mb,
      cap, nbatches, nbatches ? sum_batch_q / double(nbatches) : 0.0, max_batch_q, wall, offered, double(NQ) / wall,
      pct(0.5) * 1e3, pct(0.9) * 1e3, pct(0.99) * 1e3, lat.empty() ? 0.0 : lat.back() * 1e3,
      dense ? "dense" : hybrid ? "hybrid" : "two-level", hybrid ? dense_from : (dense ? 0 : -1), dense_batches, plan_s, slots[0].in.bytes + slots[0].plan.arena.bytes + slots[0].db.arena.bytes,
      p99_late * 1e3, gpu_mean, gpu_p99, alone_checked,
      alone_mismatch, tuned_from.c_str(), secs(g0, g1), dev, prop.multiProcessorCount);
  std::fwrite(buf, 1, size_t(n), stdout);
  if (!json_path.empty()) {
    FILE* f = std::fopen(json_path.c_str(), "w");
    if (!f) return 4;
    std::fwrite(buf, 1, size_t(n), f);
    std::fclose(f);
  }
  for (Slot& s : slots) (void)hipHostFree(s.host);
  for (void* p : {static_cast<void*>(p_state), static_cast<void*>(p_lo), static_cast<void*>(p_hi), static_cast<void*>(p_ans),
                  static_cast<void*>(p_app_parent), static_cast<void*>(p_app_key), static_cast<void*>(p_app_delta)})
    (void)hipHostFree(p);
  return alone_mismatch == 0 ? 0 : 6;
}
