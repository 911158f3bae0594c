// bw_batch.cpp -- the batched form: a stream of query batches answered over a resident index
// (one JSON line, the phase/JSON style of the reference GPU workers, src/panel_batch_gpu.cpp:
// 41-67,258-293).
//
//   bw_batch [--S n --U n --N n --seed n] [--requests 100000 --max-batch 2048]
//            [--rmax --tile-shift --chunk --ep-per-work --range-slots --seg --threads --waves] [--tuned FILE]
//            [--warmup 1 --repeat 3] [--plain-sample 8] [--stages] [--answers FILE] [--json FILE]
//            [--stream-sha] [--device-plan] [--device-gen] [--device-index] [--device-all]
//
// The index (block weights, carry rows) is built once on the device.  Every batch of the stream
// is then answered by one launch sequence (block replay, chunk replay, combine) captured in a
// hipGraph: all batches are planned together (their queries are independent), so one replay of
// the graph answers the whole stream.  Answers are in (batch, query) order.  For comparison the
// first --plain-sample batches are also run the plain way, each on its own: one full pass
// (index rebuild + answers, bwt_pass_launch) per batch, and one query-only launch per batch.
// --stages also times the launch sequence's kernels one by one (HIP events on the stream they run
// on, outside the graph) and reports each one's algorithmic bytes.
// --device-plan also plans the stream on the device (include/bwgpu_plan.h), compares that plan
// array for array with the host plan, and replays the device plan.  --device-gen materialises the
// stream's queries with the device generator (include/bwgpu_gen.h), compares them element for
// element with the host generator's, and feeds the device arrays to the planner and the replays.
// --device-index (implies --device-plan) also builds the index on the device (include/bwgpu_index.h:
// the Euler walk, the block cut and every table; the host only checks the inputs), compares it array for array
// with the host-built index, and answers the stream once more over it (device plan + query
// sequence); the answers must be the same.  --device-all then runs the whole pipeline on the device
// from the seed -- updates and queries generated, index built, stream planned and answered,
// answers hashed -- timed phase by phase; its answers must be the same too.
// --dense [--dense-rmax 4096 --dense-threads 1024 --dense-wg 1 --dense-max-wg 0; with --tuned, the
// file's "dense" entry (bw_tune --form dense) gives the defaults] also answers the
// stream with the dense-carry form (include/bwgpu_dense.h): blocks of up to --dense-rmax keys, the
// persistent replay at --dense-threads threads, --dense-wg workgroups per CU and at most
// --dense-max-wg workgroups in all (0: no cap), answers stored by --dense-store (0 auto, 1 direct
// random stores, 2 staged: sorted order, then partition and die-local window scatter; the
// BWD_STORE_* modes); its answers must be the same.  --dense-overlap
// also times the steps pipelined (the next plan on a second stream beside a replay capped at
// --dense-max-wg workgroups).
// --noalloc-check also compares its allocation-free entry point with the workspace entry point.
// Exit codes: 2 usage, 3 unsupported input, 4 I/O, 5 GPU error, 6 answers differ.
#include <hip/hip_runtime_api.h>
#include <openssl/sha.h>

#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "bwgpu_dense.h"
#include "bwgpu_extract.h"
#include "bwgpu_tree.h"
#include "dense_engine.hpp"
#include "fnv_split.hpp"
#include "tree_engine.hpp"
#include "tree_gen.hpp"
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

static double ms_between(clk::time_point a, clk::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

static std::string sha_hex(const void* p, size_t n) {
  unsigned char d[32];
  SHA256(static_cast<const unsigned char*>(p), n, d);
  static const char* hx = "0123456789abcdef";
  std::string s(64, '0');
  for (int i = 0; i < 32; ++i) { s[size_t(2 * i)] = hx[d[i] >> 4]; s[size_t(2 * i + 1)] = hx[d[i] & 15]; }
  return s;
}

int main(int argc, char** argv) {
  // --form 5 only serves --stream-sha (the served form's stream over its cfg-2 tree, or with
  // --cfg 3 over a cfg-3 tree)
  int form = 4, tcfg = 0;
  for (int i = 1; i + 1 < argc; ++i) {
    if (!std::strcmp(argv[i], "--form")) form = std::atoi(argv[i + 1]);
    if (!std::strcmp(argv[i], "--cfg")) tcfg = std::atoi(argv[i + 1]);
  }
  if (form != 4 && form != 5) { std::fprintf(stderr, "bw_batch: --form 4|5\n"); return 2; }
  if (tcfg == 0) tcfg = form == 4 ? 4 : 2;
  if ((form == 4 && tcfg != 4) || (form == 5 && tcfg != 2 && tcfg != 3)) {
    std::fprintf(stderr, "bw_batch: --cfg 4 (batched), 2 or 3 (served stream)\n");
    return 2;
  }
  TreeSpec spec = tree_defaults(tcfg);
  StreamSpec ss = stream_defaults(form);
  PlanKnobs kn;
  LaunchKnobs lk;
  int warmup = 1, repeat = 3, plain_sample = 8;
  bool want_sha = false, stages = false, dev_plan = false, dev_gen = false, dev_index = false, dev_all = false;
  bool dense = false, noalloc_check = false;
  int dense_rmax = 4096, dense_threads = 1024, dense_wg = 1, dense_maxwg = 0, dense_store = BWD_STORE_AUTO, dense_chunk = 0;
  bool dense_overlap = false;
  int fork = 1;  // replay the block and level-B stages on two streams
  std::string answers_path, json_path, tuned_path;
  bool have_knob[8] = {false, false, false, false, false, false, false, false};
  bool have_dense[5] = {false, false, false, false, false};  // --dense-rmax / -threads / -wg / -store / -chunk given
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--stream-sha")) { want_sha = true; continue; }
    if (!std::strcmp(argv[i], "--stages")) { stages = true; continue; }
    if (!std::strcmp(argv[i], "--device-plan")) { dev_plan = true; continue; }
    if (!std::strcmp(argv[i], "--device-gen")) { dev_gen = true; continue; }
    if (!std::strcmp(argv[i], "--device-index")) { dev_index = dev_plan = true; continue; }
    if (!std::strcmp(argv[i], "--device-all")) { dev_all = true; continue; }
    if (!std::strcmp(argv[i], "--dense")) { dense = dev_plan = true; continue; }
    if (!std::strcmp(argv[i], "--noalloc-check")) { noalloc_check = dense = dev_plan = true; continue; }
    if (!std::strcmp(argv[i], "--dense-overlap")) { dense_overlap = dense = dev_plan = true; continue; }
    if (i + 1 >= argc) { std::fprintf(stderr, "bw_batch: missing value for %s\n", argv[i]); return 2; }
    const char* k = argv[i];
    const char* v = argv[++i];
    if (!std::strcmp(k, "--form") || !std::strcmp(k, "--cfg")) continue;
    else if (!std::strcmp(k, "--S")) spec.S = std::atoll(v);
    else if (!std::strcmp(k, "--U")) spec.U = std::atoi(v);
    else if (!std::strcmp(k, "--N")) spec.N = std::strtoull(v, nullptr, 10);
    else if (!std::strcmp(k, "--seed")) spec.seed = std::strtoull(v, nullptr, 10);
    else if (!std::strcmp(k, "--requests")) ss.nreq = std::atoll(v);
    else if (!std::strcmp(k, "--max-batch")) ss.max_batch = std::atoi(v);
    else if (!std::strcmp(k, "--append-frac")) ss.append_frac = std::atof(v);
    else if (!std::strcmp(k, "--append-U")) ss.append_U = std::atoi(v);
    else if (!std::strcmp(k, "--rate")) ss.rate = std::atof(v);
    else if (!std::strcmp(k, "--rmax")) { kn.rmax = std::atoi(v); have_knob[0] = true; }
    else if (!std::strcmp(k, "--tile-shift")) { kn.tile_shift = std::atoi(v); have_knob[1] = true; }
    else if (!std::strcmp(k, "--chunk")) { kn.chunk_items = std::atoll(v); have_knob[2] = true; }
    else if (!std::strcmp(k, "--ep-per-work")) { kn.ep_per_work = std::atoll(v); have_knob[3] = true; }
    else if (!std::strcmp(k, "--range-slots")) { kn.range_slots = std::atoll(v); have_knob[6] = true; }
    else if (!std::strcmp(k, "--stage-shift")) kn.stage_shift = std::atoi(v);
    else if (!std::strcmp(k, "--bt-table-mb")) kn.bt_table_mb = std::atoll(v);
    else if (!std::strcmp(k, "--fork")) fork = std::atoi(v);
    else if (!std::strcmp(k, "--dense-rmax")) { dense_rmax = std::atoi(v); have_dense[0] = true; }
    else if (!std::strcmp(k, "--dense-threads")) { dense_threads = std::atoi(v); have_dense[1] = true; }
    else if (!std::strcmp(k, "--dense-wg")) { dense_wg = std::atoi(v); have_dense[2] = true; }
    else if (!std::strcmp(k, "--dense-max-wg")) dense_maxwg = std::atoi(v);
    else if (!std::strcmp(k, "--dense-store")) { dense_store = std::atoi(v); have_dense[3] = true; }
    else if (!std::strcmp(k, "--dense-chunk")) { dense_chunk = std::atoi(v); have_dense[4] = true; }
    else if (!std::strcmp(k, "--seg")) kn.seg_chunks = std::atoi(v);
    else if (!std::strcmp(k, "--threads")) { lk.block_threads = std::atoi(v); have_knob[4] = true; }
    else if (!std::strcmp(k, "--waves")) { lk.chunk_waves = std::atoi(v); have_knob[5] = true; }
    else if (!std::strcmp(k, "--tuned")) tuned_path = v;
    else if (!std::strcmp(k, "--warmup")) warmup = std::atoi(v);
    else if (!std::strcmp(k, "--repeat")) repeat = std::atoi(v);
    else if (!std::strcmp(k, "--plain-sample")) plain_sample = std::atoi(v);
    else if (!std::strcmp(k, "--answers")) answers_path = v;
    else if (!std::strcmp(k, "--json")) json_path = v;
    else { std::fprintf(stderr, "bw_batch: unknown argument %s\n", k); return 2; }
  }
  if (spec.S < 1 || spec.U < 1 || spec.N < 2 || ss.nreq < 1 || ss.max_batch < 1 || repeat < 1 || warmup < 0 ||
      plain_sample < 0) {
    std::fprintf(stderr, "bw_batch: bad sizes\n");
    return 2;
  }
  // knobs chosen by the autotuner (tree_tune) for this device and workload shape, unless given
  std::string tuned_from = "defaults", dense_from = "defaults";
  if (!tuned_path.empty()) {
    PlanKnobs tk = kn;
    LaunchKnobs tl = lk;
    if (load_tuned(tuned_path.c_str(), "batched", tk, tl)) {
      if (!have_knob[0]) kn.rmax = tk.rmax;
      if (!have_knob[1]) kn.tile_shift = tk.tile_shift;
      if (!have_knob[2]) kn.chunk_items = tk.chunk_items;
      if (!have_knob[3]) kn.ep_per_work = tk.ep_per_work;
      if (!have_knob[6]) kn.range_slots = tk.range_slots;
      if (!have_knob[4]) lk.block_threads = tl.block_threads;
      if (!have_knob[5]) lk.chunk_waves = tl.chunk_waves;
      tuned_from = tuned_path;
    }
    DenseKnobs dk{dense_rmax, dense_threads, dense_wg, dense_store, dense_chunk};
    if (load_dense_tuned(tuned_path.c_str(), dk)) {  // the dense form's own entry (bw_tune --form dense)
      if (!have_dense[0]) dense_rmax = dk.rmax;
      if (!have_dense[1]) dense_threads = dk.threads;
      if (!have_dense[2]) dense_wg = dk.wg_per_cu;
      if (!have_dense[3]) dense_store = dk.store;
      if (!have_dense[4]) dense_chunk = dk.chunk;
      dense_from = tuned_path;
    }
  }

  // ---- generate: the tree and the request stream
  const auto t0 = clk::now();
  TreeWorkload w;
  tree_generate(spec, w);
  const auto ts0 = clk::now();
  RequestStream rs;
  stream_generate(w, ss, rs);
  const double stream_host_ms = ms_between(ts0, clk::now());
  if (want_sha) {
    // with --device-gen the per-query and per-append arrays hashed are the device generator's
    uint32_t ties = 0;
    if (dev_gen) {
      char gerr[256];
      HIPX(hipSetDevice(0));
      RequestStream scan;
      stream_scan(w.spec, ss, scan);
      DeviceStream G;
      if (!device_generate_stream(w.spec, scan, G, nullptr, gerr, sizeof(gerr), w.dkey)) {
        std::fprintf(stderr, "bw_batch: device generation: %s\n", gerr);
        return 5;
      }
      auto down = [](auto& v, const auto* d) {
        if (!v.empty()) HIPX(hipMemcpy(v.data(), d, v.size() * sizeof(v[0]), hipMemcpyDeviceToHost));
        return 0;
      };
      for (auto* v : {&rs.qlo, &rs.qhi, &rs.app_key}) std::fill(v->begin(), v->end(), 0u);
      if (down(rs.qstate, G.qstate) || down(rs.qlo, G.qlo) || down(rs.qhi, G.qhi) ||
          down(rs.app_parent, G.app_parent) || down(rs.app_key, G.app_key) || down(rs.app_delta, G.app_delta))
        return 5;
      rs.req_q0 = scan.req_q0;
      rs.req_kind = scan.req_kind;
      ties = G.ties;
    }
    char hex[65];
    stream_sha(rs, hex);
    std::printf("{\"stream_sha256\":\"%s\",\"nq\":%" PRId64 ",\"napp\":%" PRId64 ",\"device_gen\":%s,\"ties\":%u}\n",
                hex, rs.nq(), rs.napp(), dev_gen ? "true" : "false", ties);
    return 0;
  }
  if (form != 4) { std::fprintf(stderr, "bw_batch: the served form runs in bw_serve\n"); return 2; }
  const int64_t M = rs.nq();
  const auto t1 = clk::now();

  // ---- index on the device, then one plan for every query of the stream
  char err[256];
  HIPX(hipSetDevice(0));
  HIPX(hipFree(nullptr));
  TreeEngine E;
  const auto ti0 = clk::now();
  if (!E.init(spec.S, spec.U, w.coord_n(), w.parent.data(), w.key.data(), w.delta.data(), kn,
              BwtLaunch{lk.block_threads, lk.chunk_waves, 0, 0}, err, sizeof(err))) {
    std::fprintf(stderr, "bw_batch: index: %s\n", err);
    return std::strncmp(err, "device", 6) == 0 ? 5 : 3;
  }
  const double host_index_ms = ms_between(ti0, clk::now());  // host index plan + its upload
  // the device generator: the stream's queries materialised on the GPU (timed on the second run:
  // the first loads code), compared element for element with the host generator's
  DeviceStream DS;
  double dgen_gpu_ms = -1, dgen_wall_ms = -1;
  int64_t dgen_mismatch = -1;
  if (dev_gen) {
    for (int rep = 0; rep < 2; ++rep) {
      const auto g0 = clk::now();
      RequestStream scan;
      stream_scan(w.spec, ss, scan);
      if (!device_generate_stream(w.spec, scan, DS, nullptr, err, sizeof(err), w.dkey)) {
        std::fprintf(stderr, "bw_batch: device generation: %s\n", err);
        return 5;
      }
      dgen_wall_ms = ms_between(g0, clk::now());
    }
    dgen_gpu_ms = DS.gpu_ms;
    dgen_mismatch = count_device_mismatch(DS.qstate, rs.qstate) + count_device_mismatch(DS.qlo, rs.qlo) +
                    count_device_mismatch(DS.qhi, rs.qhi);
  }
  const auto t2 = clk::now();
  BatchPlan plan;
  if (!build_batch_plan(E.plan, M, rs.qstate.data(), rs.qlo.data(), rs.qhi.data(), false, plan, err, sizeof(err))) {
    std::fprintf(stderr, "bw_batch: batch plan: %s\n", err);
    return 3;
  }
  const auto t3 = clk::now();
  DeviceArena io;
  uint32_t* d_qlo = dev_gen ? DS.qlo : io.upload(rs.qlo);
  uint32_t* d_qhi = dev_gen ? DS.qhi : io.upload(rs.qhi);
  DeviceBatch DB;
  if (!io.ok() || !upload_batch(E, std::move(plan), d_qlo, d_qhi, DB, err, sizeof(err))) {
    std::fprintf(stderr, "bw_batch: upload failed\n");
    return 5;
  }
  HIPX(hipDeviceSynchronize());
  const auto t4 = clk::now();

  hipStream_t st;
  HIPX(hipStreamCreate(&st));
  hipEvent_t e0, e1;
  HIPX(hipEventCreate(&e0));
  HIPX(hipEventCreate(&e1));
  // index build (once; timed on the device)
  RUN(E.build_index(st));
  HIPX(hipEventRecord(e0, st));
  RUN(E.build_index(st));
  HIPX(hipEventRecord(e1, st));
  HIPX(hipEventSynchronize(e1));
  float index_ms = 0;
  HIPX(hipEventElapsedTime(&index_ms, e0, e1));
  // the device planner: the same plan built on the GPU from the device copies of the queries,
  // compared array for array with the host plan; its batch is then the one replayed
  DevicePlanSlot dslot;
  DeviceArena qa;  // the device copy of the query states (device planner)
  const int32_t* d_qstate = nullptr;
  BwtBatch* QB = &DB.b;
  double dplan_ms = -1;
  int64_t plan_mismatch = -1;
  if (dev_plan) {
    d_qstate = dev_gen ? DS.qstate : qa.upload(rs.qstate);
    if (!qa.ok() || !alloc_plan_slot(E, M, dslot, err, sizeof(err))) {
      std::fprintf(stderr, "bw_batch: device plan: %s\n", qa.ok() ? err : "upload");
      return 5;
    }
    RUN(device_plan(E, dslot, M, d_qstate, d_qlo, d_qhi, st));  // warm
    HIPX(hipStreamSynchronize(st));
    std::vector<double> pm;
    for (int i = 0; i < std::max(1, repeat); ++i) {
      HIPX(hipEventRecord(e0, st));
      RUN(device_plan(E, dslot, M, d_qstate, d_qlo, d_qhi, st));
      HIPX(hipEventRecord(e1, st));
      HIPX(hipEventSynchronize(e1));
      float t = 0;
      HIPX(hipEventElapsedTime(&t, e0, e1));
      pm.push_back(t);
    }
    dplan_ms = 0;
    for (double t : pm) dplan_ms += t / double(pm.size());  // average, like every other timing here
    int64_t cnt[5];
    HIPX(hipMemcpy(cnt, dslot.dcounts, sizeof(cnt), hipMemcpyDeviceToHost));
    if (cnt[4]) { std::fprintf(stderr, "bw_batch: device plan flagged out-of-range queries\n"); return 3; }
    const BatchPlan& H = DB.plan;
    plan_mismatch = 0;
    auto cmp = [&](const char* name, const void* dev, const void* host, size_t n, size_t elem) {
      std::vector<unsigned char> tmp(n * elem);
      if (n) HIPX(hipMemcpy(tmp.data(), dev, n * elem, hipMemcpyDeviceToHost));
      if (n && std::memcmp(tmp.data(), host, n * elem) != 0) {
        std::fprintf(stderr, "bw_batch: device plan differs in %s\n", name);
        ++plan_mismatch;
      }
      return 0;
    };
    if (cnt[0] != int64_t(H.qblk.size()) || cnt[1] != int64_t(H.grp_step.size()) ||
        cnt[2] != int64_t(H.wk_chunk.size()) || cnt[3] != H.npairs) {
      std::fprintf(stderr, "bw_batch: device plan counts %ld %ld %ld %ld vs host %zu %zu %zu %ld\n", (long)cnt[0],
                   (long)cnt[1], (long)cnt[2], (long)cnt[3], H.qblk.size(), H.grp_step.size(), H.wk_chunk.size(),
                   (long)H.npairs);
      ++plan_mismatch;
    } else {
      const BwtBatch& d = dslot.b;
      cmp("qblk", d.qblk, H.qblk.data(), H.qblk.size(), 4);
      cmp("qblk_g0", d.qblk_g0, H.qblk_g0.data(), H.qblk_g0.size(), 8);
      cmp("grp_step", d.grp_step, H.grp_step.data(), H.grp_step.size(), 4);
      cmp("grp_q0", d.grp_q0, H.grp_q0.data(), H.grp_q0.size(), 8);
      cmp("qorder", d.qorder, H.qorder.data(), H.qorder.size(), 4);
      cmp("qrank", d.qrank, H.qrank.data(), H.qrank.size(), 4);
      cmp("sblk", d.sblk, H.sblk.data(), H.sblk.size(), 4);
      cmp("qtile", d.qtile, H.qtile.data(), H.qtile.size(), 4);
      cmp("wk_chunk", d.wk_chunk, H.wk_chunk.data(), H.wk_chunk.size(), 8);
      cmp("wk_e0", d.wk_e0, H.wk_e0.data(), H.wk_e0.size(), 8);
      cmp("eitems", d.eitems, H.eitems.data(), H.eitems.size(), 8);
      cmp("epos", d.epos, H.epos.data(), H.epos.size(), 4);
      cmp("spos", d.spos, H.spos.data(), H.spos.size(), 4);
      cmp("srank", d.srank, H.srank.data(), H.srank.size(), 2);
      cmp("ipos", d.ipos, H.ipos.data(), H.ipos.size(), 4);
    }
    QB = &dslot.b;
  }
  // the whole stream: one query launch sequence, captured once, replayed
  BwtLaunch L = E.launch;
  hipStream_t side;
  hipEvent_t fk, jn;
  HIPX(hipStreamCreate(&side));
  HIPX(hipEventCreateWithFlags(&fk, hipEventDisableTiming));
  HIPX(hipEventCreateWithFlags(&jn, hipEventDisableTiming));
  auto query = [&]() { return fork ? query_forked(E, QB, st, side, fk, jn) : bwt_query_launch(&E.x, QB, &L, st); };
  RUN(query());  // warm (also sets the kernels' LDS limits)
  HIPX(hipStreamSynchronize(st));
  hipGraph_t graph;
  hipGraphExec_t gexec;
  HIPX(hipStreamBeginCapture(st, hipStreamCaptureModeGlobal));
  RUN(query());
  HIPX(hipStreamEndCapture(st, &graph));
  HIPX(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
  for (int i = 0; i < warmup; ++i) HIPX(hipGraphLaunch(gexec, st));
  HIPX(hipStreamSynchronize(st));
  // every timed replay's answers are checked afterwards by their device FNV-1a (outside the clock)
  size_t gfnv_wb = 0;
  if (bw_fnv1a_workspace(size_t(M) * 8, &gfnv_wb)) return 5;
  DeviceArena gfa;
  void* gfnv_work = gfa.alloc(gfnv_wb);
  uint64_t* d_gfnv = gfa.scratch<uint64_t>(size_t(2 * repeat));
  if (!gfa.ok()) {
    std::fprintf(stderr, "bw_batch: device: FNV workspace\n");
    return 5;
  }
  int nfnv = 0;
  std::vector<double> run_ms;
  for (int i = 0; i < repeat; ++i) {
    HIPX(hipEventRecord(e0, st));
    HIPX(hipGraphLaunch(gexec, st));
    HIPX(hipEventRecord(e1, st));
    HIPX(hipEventSynchronize(e1));
    float t = 0;
    HIPX(hipEventElapsedTime(&t, e0, e1));
    run_ms.push_back(t);
    RUN(bw_fnv1a_device(QB->answers, size_t(M) * 8, kFnvStart, gfnv_work, gfnv_wb, d_gfnv + nfnv++, st));
  }
  // the two-level form's whole step with its device plan: plan + graph replay in one event pair
  std::vector<double> step2_ms;
  if (dev_plan) {
    for (int i = 0; i < repeat; ++i) {
      HIPX(hipEventRecord(e0, st));
      RUN(device_plan(E, dslot, M, d_qstate, d_qlo, d_qhi, st));
      HIPX(hipGraphLaunch(gexec, st));
      HIPX(hipEventRecord(e1, st));
      HIPX(hipEventSynchronize(e1));
      float t = 0;
      HIPX(hipEventElapsedTime(&t, e0, e1));
      step2_ms.push_back(t);
      RUN(bw_fnv1a_device(QB->answers, size_t(M) * 8, kFnvStart, gfnv_work, gfnv_wb, d_gfnv + nfnv++, st));
    }
  }
  std::vector<uint64_t> replay_fnv(static_cast<size_t>(nfnv));
  HIPX(hipMemcpy(replay_fnv.data(), d_gfnv, sizeof(uint64_t) * size_t(nfnv), hipMemcpyDeviceToHost));
  double step2_avg = 0;
  for (double t : step2_ms) step2_avg += t / double(step2_ms.size());
  // per-kernel timing of the same sequence (outside the graph), averaged over `repeat` runs
  double st_ms[3] = {0, 0, 0};
  if (stages) {
    hipEvent_t ev[4];
    BwtLaunch Lq = L;  // the query sequence's launches, as bwt_query_launch makes them
    Lq.write_index = 0;
    Lq.carry_in_block = 1;
    for (auto& e : ev) HIPX(hipEventCreate(&e));
    for (int i = 0; i < repeat; ++i) {
      HIPX(hipEventRecord(ev[0], st));
      RUN(bwt_block_replay_launch(&E.x, QB, &Lq, st));
      HIPX(hipEventRecord(ev[1], st));
      RUN(bwt_chunk_replay_launch(&E.x, QB, &Lq, st));
      HIPX(hipEventRecord(ev[2], st));
      RUN(bwt_combine_launch(&E.x, QB, &Lq, st));
      HIPX(hipEventRecord(ev[3], st));
      HIPX(hipEventSynchronize(ev[3]));
      for (int k = 0; k < 3; ++k) {
        float t = 0;
        HIPX(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
        st_ms[k] += double(t) / repeat;
      }
    }
    for (auto& e : ev) HIPX(hipEventDestroy(e));
  }
  // algorithmic bytes per kernel (what each must read and write once):
  //  block replay: per replayed event its u16 rank + i64 delta, per step its id, per query its
  //    packed ranks and key tiles (2 x u32) + result (i64), per listed block its level-C row
  //  chunk replay: per replayed key item its class word (u32) + weight (i64), per work item its
  //    chunk's carry row, per endpoint item (u64) its read and its result (i64)
  //  combine (query side: the block replay added the level-C carry): per slot its partial
  //    result, the 2 level-B places (u32) and results, its staging index (u16) + position (u32)
  //    and staged value; per query id its staging position (u32), the staged value, the answer
  int64_t kit_run = 0;
  {
    const BatchPlan& P2 = DB.plan;
    for (size_t i = 0; i < P2.wk_chunk.size(); ++i) {
      const int64_t c = P2.wk_chunk[i];
      kit_run += std::min<int64_t>(E.plan.chunk_key0[size_t(c) + 1] - E.plan.chunk_key0[size_t(c)],
                                   int64_t(BWT_EP_POS(P2.eitems[size_t(P2.wk_e0[i + 1] - 1)])));
    }
  }
  const double sb_block = double(DB.plan.steps_run) * spec.U * 10.0 + double(DB.plan.steps_run) * 4.0 +
                          double(M) * 16.0 + double(DB.plan.qblk.size()) * double(E.x.ncls_c) * 8.0;
  const double sb_chunk = double(kit_run) * 12.0 + double(DB.plan.wk_chunk.size()) * double(E.x.ncls_b) * 8.0 +
                          double(DB.plan.eitems.size()) * 16.0;
  const double sb_combine = double(M) * (8.0 + 8.0 + 16.0 + 2.0 + 4.0 + 8.0 + 4.0 + 8.0 + 8.0);
  // answers off the device while three host threads write and hash them (bw_extract_answers)
  std::vector<int64_t> ans(static_cast<size_t>(M));
  BwExtractResult xr{};
  const int xrc = bw_extract_answers(QB->answers, size_t(M), nullptr,
                                     answers_path.empty() ? nullptr : answers_path.c_str(), ans.data(),
                                     BW_EXTRACT_FNV_ON_DEVICE, &xr);
  int64_t replay_mismatch = 0;
  for (uint64_t f : replay_fnv) replay_mismatch += f != xr.fnv1a;
  if (replay_mismatch) std::fprintf(stderr, "bw_batch: %" PRId64 " timed graph replays' answers differ\n", replay_mismatch);
  if (xrc) {
    std::fprintf(stderr, "bw_batch: extract failed (%d)\n", xrc);
    return xrc < 0 ? 4 : 5;
  }
  const std::string sha = xr.sha256_hex;

  // ---- the dense-carry form (--dense, include/bwgpu_dense.h): an engine cut into blocks of up
  // to dense_rmax keys, one dense carry row per block (built once), then per step the whole
  // stream planned (one radix sort) and answered (one persistent block-replay launch, captured
  // in a hipGraph); its answers must equal the two-level form's
  char dense_json[4096] = "null", noalloc_json[512] = "null";
  int64_t dense_mismatch = -1, noalloc_mismatch = -1;
  if (dense) {
    PlanKnobs dk = kn;
    dk.rmax = dense_rmax;
    TreeEngine E4;
    const auto d0 = clk::now();
    if (!E4.init(spec.S, spec.U, w.coord_n(), w.parent.data(), w.key.data(), w.delta.data(), dk,
                 BwtLaunch{1024, lk.chunk_waves, 0, 0}, err, sizeof(err))) {
      std::fprintf(stderr, "bw_batch: dense index: %s\n", err);
      return std::strncmp(err, "device", 6) == 0 ? 5 : 3;
    }
    RUN(E4.build_index(st));
    DenseEngine D;
    D.launch = BwdLaunch{dense_threads, dense_wg, dense_overlap ? 0 : dense_maxwg, dense_store, dense_chunk};
    if (!D.init(E4, st, err, sizeof(err))) {
      std::fprintf(stderr, "bw_batch: %s\n", err);
      return 5;
    }
    const double dense_index_wall = ms_between(d0, clk::now());
    int32_t rep_wg = 0, rep_res = 0;
    size_t rep_lds = 0;
    if (bwd_replay_shape(&D.x, M, &D.launch, &rep_wg, &rep_res, &rep_lds)) {
      std::fprintf(stderr, "bw_batch: dense replay shape\n");
      return 5;
    }
    DenseBatch DBn;
    int64_t* d_ans = nullptr;
    if (!alloc_dense_batch(D, M, DBn, err, sizeof(err)) || !(d_ans = DBn.arena.scratch<int64_t>(size_t(M)))) {
      std::fprintf(stderr, "bw_batch: %s\n", err);
      return 5;
    }
    auto plan_d = [&]() { return bwd_plan_launch(&D.x, M, d_qstate, d_qlo, d_qhi, DBn.flags, DBn.work, DBn.work_bytes, st); };
    RUN(plan_d());
    RUN(bwd_replay_launch(&D.x, M, d_ans, &D.launch, DBn.work, DBn.work_bytes, st));  // warm
    HIPX(hipStreamSynchronize(st));
    int64_t flag = 0;
    HIPX(hipMemcpy(&flag, DBn.flags, sizeof(flag), hipMemcpyDeviceToHost));
    if (flag) { std::fprintf(stderr, "bw_batch: dense plan flagged out-of-range queries\n"); return 3; }
    // a step is two graph launches: the plan's kernels, then the replay's (block replay and, for
    // staged stores, the answers' partition and scatter); both captured once
    hipGraph_t pg, dg;
    hipGraphExec_t pgx, dgx;
    HIPX(hipStreamBeginCapture(st, hipStreamCaptureModeGlobal));
    RUN(plan_d());
    HIPX(hipStreamEndCapture(st, &pg));
    HIPX(hipGraphInstantiate(&pgx, pg, nullptr, nullptr, 0));
    HIPX(hipStreamBeginCapture(st, hipStreamCaptureModeGlobal));
    RUN(bwd_replay_launch(&D.x, M, d_ans, &D.launch, DBn.work, DBn.work_bytes, st));
    HIPX(hipStreamEndCapture(st, &dg));
    HIPX(hipGraphInstantiate(&dgx, dg, nullptr, nullptr, 0));
    hipEvent_t em;
    HIPX(hipEventCreate(&em));
    for (int i = 0; i < warmup; ++i) {
      HIPX(hipGraphLaunch(pgx, st));
      HIPX(hipGraphLaunch(dgx, st));
    }
    HIPX(hipStreamSynchronize(st));
    // every timed step's answers are checked afterwards by their device FNV-1a (outside the clock)
    size_t fnv_wb = 0;
    if (bw_fnv1a_workspace(size_t(M) * 8, &fnv_wb)) return 5;
    void* fnv_work = DBn.arena.alloc(fnv_wb);
    uint64_t* d_fnv = DBn.arena.scratch<uint64_t>(size_t(repeat));
    if (!fnv_work || !d_fnv) {
      std::fprintf(stderr, "bw_batch: device: FNV workspace\n");
      return 5;
    }
    std::vector<double> step_ms, dplan_v, drep_v;
    for (int i = 0; i < repeat; ++i) {
      HIPX(hipEventRecord(e0, st));
      HIPX(hipGraphLaunch(pgx, st));
      HIPX(hipEventRecord(em, st));
      HIPX(hipGraphLaunch(dgx, st));
      HIPX(hipEventRecord(e1, st));
      HIPX(hipEventSynchronize(e1));
      float t = 0, tp = 0;
      HIPX(hipEventElapsedTime(&t, e0, e1));
      HIPX(hipEventElapsedTime(&tp, e0, em));
      step_ms.push_back(t);
      dplan_v.push_back(tp);
      drep_v.push_back(double(t) - double(tp));
      RUN(bw_fnv1a_device(d_ans, size_t(M) * 8, kFnvStart, fnv_work, fnv_wb, d_fnv + i, st));
    }
    std::vector<uint64_t> step_fnv(static_cast<size_t>(repeat));
    HIPX(hipMemcpy(step_fnv.data(), d_fnv, sizeof(uint64_t) * size_t(repeat), hipMemcpyDeviceToHost));
    int64_t flag_after = 0;
    HIPX(hipMemcpy(&flag_after, DBn.flags, sizeof(flag_after), hipMemcpyDeviceToHost));
    BwExtractResult dxr{};
    const int dxrc = bw_extract_answers(d_ans, size_t(M), st, nullptr, nullptr, BW_EXTRACT_FNV_ON_DEVICE, &dxr);
    if (dxrc) return dxrc < 0 ? 4 : 5;
    dense_mismatch = sha == dxr.sha256_hex ? 0 : 1;
    if (dense_mismatch) std::fprintf(stderr, "bw_batch: dense-carry answers differ from the two-level form's\n");
    int64_t step_mismatch = flag_after != 0;
    for (uint64_t f : step_fnv) step_mismatch += f != dxr.fnv1a;
    if (step_mismatch) std::fprintf(stderr, "bw_batch: %" PRId64 " timed dense steps' answers differ\n", step_mismatch);
    dense_mismatch += step_mismatch;
    uint32_t work_items = 0;  // with chunk > 0: the items the timed replays took
    if (dense_chunk > 0 && bwd_replay_work_items(&D.x, M, DBn.work, DBn.work_bytes, &work_items)) return 5;
    // the replay's phases (block replay, answer partition, answer scatter), timed apart: the mean
    // over `repeat` replays of the planned stream
    double phase_ms[3] = {0, 0, 0};
    for (int i = 0; i < repeat; ++i) {
      float ph[3] = {0, 0, 0};
      RUN(bwd_replay_timed(&D.x, M, d_ans, &D.launch, DBn.work, DBn.work_bytes, st, ph));
      for (int k = 0; k < 3; ++k) phase_ms[k] += ph[k] / repeat;
    }
    // the stream shifted by one query (every query array 4 bytes past a 16-byte boundary, so the
    // plan packs keys with its unaligned form): its answers are the whole stream's, shifted by one
    int64_t offset_mismatch = 0;
    {
      int64_t* d_ans2 = DBn.arena.scratch<int64_t>(size_t(M));
      if (!d_ans2) {
        std::fprintf(stderr, "bw_batch: device: offset check allocation\n");
        return 5;
      }
      RUN(bwd_answer_launch(&D.x, M - 1, d_qstate + 1, d_qlo + 1, d_qhi + 1, d_ans2, DBn.flags, &D.launch, DBn.work,
                            DBn.work_bytes, st));
      HIPX(hipStreamSynchronize(st));
      std::vector<int64_t> full(static_cast<size_t>(M)), shifted(static_cast<size_t>(M - 1));
      HIPX(hipMemcpy(full.data(), d_ans, sizeof(int64_t) * size_t(M), hipMemcpyDeviceToHost));
      HIPX(hipMemcpy(shifted.data(), d_ans2, sizeof(int64_t) * size_t(M - 1), hipMemcpyDeviceToHost));
      HIPX(hipMemcpy(&flag, DBn.flags, sizeof(flag), hipMemcpyDeviceToHost));
      for (int64_t i = 0; i + 1 < M; ++i) offset_mismatch += shifted[size_t(i)] != full[size_t(i + 1)];
      offset_mismatch += flag != 0;
      if (offset_mismatch) std::fprintf(stderr, "bw_batch: the shifted stream's answers differ (%" PRId64 ")\n", offset_mismatch);
      dense_mismatch += offset_mismatch;
    }
    // edge cases of the plan, checked against root-path sums computed here on the host (independent
    // of the kernels): the first m queries of the stream (a) all moved to one state (one bucket and
    // one state group: the plan's counters at their worst contention), (b) moved to four states
    // round robin, (c) as (b) with every 101st query out of range (state 0 or S+1, lo or hi above
    // NK: flags[0] = 1 and answer 0), (d) batches of 1 and 7 queries of (b)
    int64_t edge_mismatch = 0;
    const int64_t edge_m = std::min<int64_t>(M, int64_t(1) << 20);
    {
      const int64_t S = spec.S;
      const int U = spec.U;
      const uint64_t NK = w.coord_n();
      const int32_t four[4] = {1, int32_t(S / 3 + 1), int32_t(S / 2 + 1), int32_t(S)};
      // the updates on the root->v path by key, with their prefix sums: F_v(x) = pre[#keys < x]
      struct PathSums {
        std::vector<uint32_t> k;
        std::vector<int64_t> pre;
        int64_t F(uint32_t x) const { return pre[size_t(std::lower_bound(k.begin(), k.end(), x) - k.begin())]; }
      };
      auto path_sums = [&](int32_t v) {
        std::vector<std::pair<uint32_t, int64_t>> u;
        for (int32_t a = v; a >= 1; a = w.parent[size_t(a)])
          for (int i = 0; i < U; ++i) u.emplace_back(w.key[size_t(a - 1) * U + i], w.delta[size_t(a - 1) * U + i]);
        std::sort(u.begin(), u.end());
        PathSums p;
        p.pre.push_back(0);
        for (const auto& kv : u) {
          p.k.push_back(kv.first);
          p.pre.push_back(p.pre.back() + kv.second);
        }
        return p;
      };
      std::vector<PathSums> ps;
      for (int32_t v : four) ps.push_back(path_sums(v));
      std::vector<int32_t> qs(static_cast<size_t>(edge_m));
      std::vector<uint32_t> ql(rs.qlo.begin(), rs.qlo.begin() + edge_m), qh(rs.qhi.begin(), rs.qhi.begin() + edge_m);
      DeviceArena ea;
      int32_t* e_qs = ea.scratch<int32_t>(size_t(edge_m));
      uint32_t* e_ql = ea.scratch<uint32_t>(size_t(edge_m));
      uint32_t* e_qh = ea.scratch<uint32_t>(size_t(edge_m));
      int64_t* e_ans = ea.scratch<int64_t>(size_t(edge_m));
      if (!ea.ok()) {
        std::fprintf(stderr, "bw_batch: device: edge-case allocation\n");
        return 5;
      }
      std::vector<int64_t> got(static_cast<size_t>(edge_m));
      // answers the first m queries of (qs, ql, qh); expect(i) is query i's answer, or -1 for "out of range"
      auto check = [&](int64_t m, bool bad_expected, auto expect_state) -> int64_t {
        HIPX(hipMemcpy(e_qs, qs.data(), sizeof(int32_t) * size_t(m), hipMemcpyHostToDevice));
        HIPX(hipMemcpy(e_ql, ql.data(), sizeof(uint32_t) * size_t(m), hipMemcpyHostToDevice));
        HIPX(hipMemcpy(e_qh, qh.data(), sizeof(uint32_t) * size_t(m), hipMemcpyHostToDevice));
        RUN(bwd_answer_launch(&D.x, m, e_qs, e_ql, e_qh, e_ans, DBn.flags, &D.launch, DBn.work, DBn.work_bytes, st));
        HIPX(hipStreamSynchronize(st));
        int64_t f = 0;
        HIPX(hipMemcpy(&f, DBn.flags, sizeof(f), hipMemcpyDeviceToHost));
        HIPX(hipMemcpy(got.data(), e_ans, sizeof(int64_t) * size_t(m), hipMemcpyDeviceToHost));
        int64_t bad = (f != (bad_expected ? 1 : 0));
        for (int64_t i = 0; i < m; ++i) {
          const int j = expect_state(i);  // index into `four`, or -1: out of range
          const int64_t want = j < 0 ? 0 : ps[size_t(j)].F(qh[size_t(i)]) - ps[size_t(j)].F(ql[size_t(i)]);
          bad += got[size_t(i)] != want;
        }
        return bad;
      };
      for (int64_t i = 0; i < edge_m; ++i) qs[size_t(i)] = four[2];
      const int64_t ea1 = check(edge_m, false, [](int64_t) { return 2; });
      for (int64_t i = 0; i < edge_m; ++i) qs[size_t(i)] = four[i % 4];
      const int64_t eb = check(edge_m, false, [](int64_t i) { return int(i % 4); });
      const int64_t ed = check(1, false, [](int64_t i) { return int(i % 4); }) + check(7, false, [](int64_t i) { return int(i % 4); });
      for (int64_t i = 0; i < edge_m; i += 101) {
        switch ((i / 101) % 4) {
          case 0: qs[size_t(i)] = 0; break;
          case 1: qs[size_t(i)] = int32_t(S + 1); break;
          case 2: ql[size_t(i)] = uint32_t(NK + 1); break;
          default: qh[size_t(i)] = uint32_t(NK + 1); break;
        }
      }
      const int64_t ec = check(edge_m, true, [](int64_t i) { return i % 101 == 0 ? -1 : int(i % 4); });
      edge_mismatch = ea1 + eb + ec + ed;
      if (edge_mismatch)
        std::fprintf(stderr, "bw_batch: dense edge cases differ: one state %" PRId64 ", four %" PRId64 ", out of range %" PRId64
                     ", tiny %" PRId64 "\n", ea1, eb, ec, ed);
      dense_mismatch += edge_mismatch;
    }
    // --dense-overlap: consecutive steps pipelined.  Step i+1's plan runs on a second stream, into
    // a second workspace, beside step i's replay, which is capped at --dense-max-wg workgroups so
    // that compute units stay free for the plan; a plan waits until the replay two steps back has
    // left its workspace, a replay waits for its plan.  The clock brackets all `repeat` steps,
    // the first plan (not overlapped) included.
    char overlap_json[384] = "null";
    if (dense_overlap) {
      DenseBatch DB2;
      if (!alloc_dense_batch(D, M, DB2, err, sizeof(err))) {
        std::fprintf(stderr, "bw_batch: %s\n", err);
        return 5;
      }
      DenseBatch* wsp[2] = {&DBn, &DB2};
      BwdLaunch lo = D.launch;
      lo.max_workgroups = dense_maxwg;
      hipStream_t sp;
      HIPX(hipStreamCreate(&sp));
      hipEvent_t ep[2], er[2], o0, o1;
      for (int k = 0; k < 2; ++k) {
        HIPX(hipEventCreateWithFlags(&ep[k], hipEventDisableTiming));
        HIPX(hipEventCreateWithFlags(&er[k], hipEventDisableTiming));
      }
      HIPX(hipEventCreate(&o0));
      HIPX(hipEventCreate(&o1));
      auto plan_on = [&](int w) {
        return bwd_plan_launch(&D.x, M, d_qstate, d_qlo, d_qhi, wsp[w]->flags, wsp[w]->work, wsp[w]->work_bytes, sp);
      };
      auto run_steps = [&](int n) {
        RUN(plan_on(0));
        HIPX(hipEventRecord(ep[0], sp));
        for (int i = 0; i < n; ++i) {
          const int w = i & 1;
          HIPX(hipStreamWaitEvent(st, ep[w], 0));
          RUN(bwd_replay_launch(&D.x, M, d_ans, &lo, wsp[w]->work, wsp[w]->work_bytes, st));
          HIPX(hipEventRecord(er[w], st));
          if (i + 1 < n) {
            const int w2 = (i + 1) & 1;
            if (i >= 1) HIPX(hipStreamWaitEvent(sp, er[w2], 0));  // replay i-1 has left workspace w2
            RUN(plan_on(w2));
            HIPX(hipEventRecord(ep[w2], sp));
          }
        }
        return 0;
      };
      HIPX(hipDeviceSynchronize());
      if (run_steps(std::max(2, warmup))) return 5;
      HIPX(hipDeviceSynchronize());
      HIPX(hipEventRecord(o0, st));
      HIPX(hipStreamWaitEvent(sp, o0, 0));
      if (run_steps(repeat)) return 5;
      HIPX(hipEventRecord(o1, st));
      HIPX(hipEventSynchronize(o1));
      HIPX(hipDeviceSynchronize());
      float tov = 0;
      HIPX(hipEventElapsedTime(&tov, o0, o1));
      int64_t fo = 0, f2 = 0;
      HIPX(hipMemcpy(&fo, DBn.flags, sizeof(fo), hipMemcpyDeviceToHost));
      HIPX(hipMemcpy(&f2, DB2.flags, sizeof(f2), hipMemcpyDeviceToHost));
      BwExtractResult oxr{};
      const int oxrc = bw_extract_answers(d_ans, size_t(M), st, nullptr, nullptr, BW_EXTRACT_FNV_ON_DEVICE, &oxr);
      if (oxrc) return oxrc < 0 ? 4 : 5;
      const int64_t omis = (sha == oxr.sha256_hex ? 0 : 1) + (fo != 0) + (f2 != 0);
      if (omis) std::fprintf(stderr, "bw_batch: the overlapped steps' answers differ\n");
      dense_mismatch += omis;
      const double ostep = double(tov) / repeat;
      std::snprintf(overlap_json, sizeof(overlap_json),
                    "{\"max_workgroups\":%d,\"step_ms\":%.4f,\"queries_per_s\":%.1f,\"repeat\":%d,\"mismatch\":%" PRId64
                    "}",
                    dense_maxwg, ostep, double(M) / (ostep * 1e-3), repeat, omis);
      for (int k = 0; k < 2; ++k) {
        HIPX(hipEventDestroy(ep[k]));
        HIPX(hipEventDestroy(er[k]));
      }
      HIPX(hipEventDestroy(o0));
      HIPX(hipEventDestroy(o1));
      HIPX(hipStreamDestroy(sp));
    }
    auto mean = [](const std::vector<double>& v) {
      double a = 0;
      for (double x : v) a += x;
      return a / double(v.size());
    };
    const double sm = mean(step_ms);
    std::snprintf(dense_json, sizeof(dense_json),
                  "{\"rmax\":%d,\"nblk\":%d,\"threads\":%d,\"wg_per_cu\":%d,\"wg_resident\":%d,\"workgroups\":%d,"
                  "\"lds_bytes\":%zu,\"store\":%d,\"chunk\":%d,\"work_items\":%u,\"knobs_from\":\"%s\","
                  "\"carry_bytes\":%zu,\"carry_build_ms\":%.3f,"
                  "\"index_wall_ms\":%.3f,\"step_ms\":{\"avg\":%.4f,\"min\":%.4f,\"max\":%.4f,\"repeat\":%d},"
                  "\"plan_ms\":%.4f,\"replay_ms\":%.4f,\"replay_phases_ms\":{\"block\":%.4f,\"partition\":%.4f,"
                  "\"scatter\":%.4f},\"queries_per_s\":%.1f,\"answers_sha256\":\"%s\","
                  "\"offset_mismatch\":%" PRId64 ",\"step_mismatch\":%" PRId64 ",\"edge\":{\"queries\":%" PRId64
                  ",\"mismatch\":%" PRId64 "},\"overlap\":%s,\"mismatch\":%" PRId64 "}",
                  dense_rmax, E4.x.nblk, dense_threads, dense_wg, rep_res, rep_wg, rep_lds, dense_store, dense_chunk, work_items,
                  dense_from.c_str(), D.carry_bytes, D.build_ms, dense_index_wall, sm,
                  *std::min_element(step_ms.begin(), step_ms.end()), *std::max_element(step_ms.begin(), step_ms.end()),
                  repeat, mean(dplan_v), mean(drep_v), phase_ms[0], phase_ms[1], phase_ms[2], double(M) / (sm * 1e-3),
                  dxr.sha256_hex, offset_mismatch, step_mismatch, edge_m, edge_mismatch, overlap_json, dense_mismatch);
    HIPX(hipGraphExecDestroy(dgx));
    HIPX(hipGraphDestroy(dg));
    HIPX(hipGraphExecDestroy(pgx));
    HIPX(hipGraphDestroy(pg));
    HIPX(hipEventDestroy(em));
    // --noalloc-check: the allocation-free entry point (library-held scratch) against the
    // workspace entry point on two prefixes of the stream in a row, the second smaller; both must
    // also equal the whole stream's answers.  A batch beyond its scratch must be refused.
    if (noalloc_check) {
      size_t need = 0, held = 0;
      int64_t m1 = std::min<int64_t>(M, int64_t(1) << 21);
      while (m1 > 1 && (bwd_answer_noalloc_capacity(&D.x, m1, &need, &held) != 0 || need > held)) m1 = m1 * 3 / 4;
      const int64_t m2 = std::max<int64_t>(1, m1 / 3 + 7 < m1 ? m1 / 3 + 7 : m1 - 1);
      // the largest batch the scratch holds over this index
      int64_t cap_lo = 0, cap_hi = int64_t(1) << 24;
      while (cap_lo < cap_hi) {
        const int64_t mid = (cap_lo + cap_hi + 1) / 2;
        size_t cn = 0, ch = 0;
        if (bwd_answer_noalloc_capacity(&D.x, mid, &cn, &ch) == 0 && cn <= ch) cap_lo = mid; else cap_hi = mid - 1;
      }
      DeviceArena na;
      int64_t* a_ref = na.scratch<int64_t>(size_t(m1));
      int64_t* a_na = na.scratch<int64_t>(size_t(m1));
      int64_t* f_na = na.scratch<int64_t>(1);
      if (!na.ok()) return 5;
      noalloc_mismatch = 0;
      std::vector<int64_t> hr(static_cast<size_t>(m1)), hn(static_cast<size_t>(m1));
      int64_t diff[2][3] = {{0, 0, 0}, {0, 0, 0}};  // per call: ref vs noalloc, ref vs stream, noalloc vs stream
      int call = 0;
      for (const int64_t m : {m1, m2}) {
        RUN(bwd_answer_launch(&D.x, m, d_qstate, d_qlo, d_qhi, a_ref, DBn.flags, &D.launch, DBn.work, DBn.work_bytes,
                              st));
        RUN(bwd_answer_noalloc_launch(&D.x, m, d_qstate, d_qlo, d_qhi, a_na, f_na, &D.launch, st));
        HIPX(hipStreamSynchronize(st));
        HIPX(hipMemcpy(hr.data(), a_ref, size_t(m) * 8, hipMemcpyDeviceToHost));
        HIPX(hipMemcpy(hn.data(), a_na, size_t(m) * 8, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < m; ++i) {
          diff[call][0] += hr[size_t(i)] != hn[size_t(i)];
          diff[call][1] += hr[size_t(i)] != ans[size_t(i)];
          diff[call][2] += hn[size_t(i)] != ans[size_t(i)];
        }
        noalloc_mismatch += diff[call][0] + diff[call][2];
        ++call;
      }
      int64_t fl = 0;
      HIPX(hipMemcpy(&fl, f_na, sizeof(fl), hipMemcpyDeviceToHost));
      noalloc_mismatch += fl != 0;
      // the whole stream needs far more than the scratch: refused before any device work
      const bool refused = M > m1 && bwd_answer_noalloc_launch(&D.x, M, d_qstate, d_qlo, d_qhi, a_na, f_na, &D.launch,
                                                               st) != 0;
      if (M > m1 && !refused) ++noalloc_mismatch;
      std::snprintf(noalloc_json, sizeof(noalloc_json),
                    "{\"m1\":%" PRId64 ",\"m2\":%" PRId64 ",\"scratch_bytes\":%zu,\"m1_needs_bytes\":%zu,"
                    "\"capacity_queries\":%" PRId64 ",\"refuses_oversize\":%s,\"diff\":[[%" PRId64 ",%" PRId64
                    ",%" PRId64 "],[%" PRId64 ",%" PRId64 ",%" PRId64 "]],\"mismatch\":%" PRId64 "}",
                    m1, m2, held, need, cap_lo, refused ? "true" : "false", diff[0][0], diff[0][1], diff[0][2],
                    diff[1][0], diff[1][1], diff[1][2], noalloc_mismatch);
    }
  }

  // ---- the index tables built on the device (--device-index): compared with the host-built
  // tables array for array, then the stream planned and answered over that index
  char didx_json[768] = "null";
  int64_t didx_mismatch = -1;
  if (dev_index) {
    TreeEngine ED;
    if (!ED.init_device_index(spec.S, spec.U, w.coord_n(), w.parent.data(), w.key.data(), nullptr, w.delta.data(),
                              kn, L, st, err, sizeof(err), true, true)) {
      std::fprintf(stderr, "bw_batch: %s\n", err);
      return 5;
    }
    didx_mismatch = 0;
    auto same = [&](const char* name, const void* a, const void* b, size_t bytes) {
      std::vector<unsigned char> ha(bytes), hb(bytes);
      if (bytes && (hipMemcpy(ha.data(), a, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
                    hipMemcpy(hb.data(), b, bytes, hipMemcpyDeviceToHost) != hipSuccess || ha != hb)) {
        std::fprintf(stderr, "bw_batch: device index differs in %s\n", name);
        ++didx_mismatch;
      }
    };
    const TreeIndexPlan& H = E.plan;
    const size_t nst = H.steps.size(), nbk = size_t(H.nblk()), KT = size_t(H.KT);
    if (ED.plan.steps != H.steps || ED.plan.blk_step0 != H.blk_step0 || ED.device_index.nkeys != int64_t(H.bkey.size()) ||
        ED.device_index.nslots != int64_t(H.tile_rank0.size()) || ED.plan.chunk_key0 != H.chunk_key0 ||
        ED.plan.chunk_cp0 != H.chunk_cp0 || ED.plan.seg_chunk0 != H.seg_chunk0) {
      std::fprintf(stderr, "bw_batch: device index shape differs\n");
      ++didx_mismatch;
    } else {
      same("evrank", ED.x.evrank, E.x.evrank, nst * size_t(spec.U) * 2);
      same("blk_key0", ED.x.blk_key0, E.x.blk_key0, (nbk + 1) * 8);
      same("blk_tile0", ED.x.blk_tile0, E.x.blk_tile0, (nbk + 1) * 8);
      same("bkey", ED.x.bkey, E.x.bkey, H.bkey.size() * 4);
      same("tile_rank0", ED.x.tile_rank0, E.x.tile_rank0, H.tile_rank0.size() * 2);
      same("kitems", ED.x.kitems, E.x.kitems, H.kitems.size() * 8);
      same("state_step", ED.px.state_step, E.px.state_step, (size_t(spec.S) + 1) * 4);
      same("step_block", ED.px.step_block, E.px.step_block, nst * 4);
      same("chain_key0", ED.px.chain_key0, E.px.chain_key0, (KT + 2) * 8);
      same("tile_tb0", ED.px.tile_tb0, E.px.tile_tb0, (KT + 1) * 8);
      same("tb_block", ED.px.tb_block, E.px.tb_block, H.tb_block.size() * 4);
      same("tb_pos", ED.px.tb_pos, E.px.tb_pos, H.tb_pos.size() * 8);
      if (ED.px.bt_pos && E.px.bt_pos) same("bt_pos", ED.px.bt_pos, E.px.bt_pos, nbk * KT * 4);
    }
    // the stream over the device-built index: index weights, device plan, query sequence
    RUN(ED.build_index(st));
    RUN(device_plan(ED, dslot, M, d_qstate, d_qlo, d_qhi, st));
    RUN(bwt_query_launch(&ED.x, &dslot.b, &L, st));
    BwExtractResult dx{};
    const int drc = bw_extract_answers(dslot.b.answers, size_t(M), st, nullptr, nullptr, BW_EXTRACT_FNV_ON_DEVICE, &dx);
    if (drc) return drc < 0 ? 4 : 5;
    if (sha != dx.sha256_hex) {
      std::fprintf(stderr, "bw_batch: answers over the device index differ\n");
      ++didx_mismatch;
    }
    std::snprintf(didx_json, sizeof(didx_json),
                  "{\"host_ms\":%.3f,\"gpu_ms\":%.3f,\"total_ms\":%.3f,\"host_plan_ms\":%.3f,"
                  "\"nkeys\":%" PRId64 ",\"nslots\":%" PRId64 ",\"answers_sha256\":\"%s\",\"mismatch\":%" PRId64 "}",
                  ED.device_index.host_ms, ED.device_index.gpu_ms, ED.device_index.total_ms, host_index_ms,
                  ED.device_index.nkeys, ED.device_index.nslots, dx.sha256_hex, didx_mismatch);
  }

  // ---- everything on the device (--device-all): the tree's updates and the stream's queries
  // generated, the index built, the stream planned and answered, the answers hashed -- from the
  // seed and the host's parent array (a sequential recurrence) to the answers' SHA-256
  char dall_json[512] = "null";
  int64_t dall_mismatch = -1;
  if (dev_all) {
    HIPX(hipStreamSynchronize(st));
    const auto a0 = clk::now();
    DeviceTree DT;
    RequestStream scan;
    stream_scan(w.spec, ss, scan);
    DeviceStream DS2;
    if (!device_generate_tree(spec, w.dkey, DT, st, err, sizeof(err)) ||
        !device_generate_stream(w.spec, scan, DS2, st, err, sizeof(err), w.dkey)) {
      std::fprintf(stderr, "bw_batch: device generation: %s\n", err);
      return 5;
    }
    const auto a1 = clk::now();
    TreeEngine EA;
    if (!EA.init_device_index(spec.S, spec.U, w.coord_n(), w.parent.data(), nullptr, DT.key, nullptr, kn, L, st, err,
                              sizeof(err), true, true, DT.delta)) {
      std::fprintf(stderr, "bw_batch: %s\n", err);
      return 5;
    }
    RUN(EA.build_index(st));
    HIPX(hipStreamSynchronize(st));
    const auto a2 = clk::now();
    DevicePlanSlot slot2;
    if (!alloc_plan_slot(EA, DS2.nq, slot2, err, sizeof(err))) {
      std::fprintf(stderr, "bw_batch: %s\n", err);
      return 5;
    }
    HIPX(hipDeviceSynchronize());
    const auto a3 = clk::now();
    RUN(device_plan(EA, slot2, DS2.nq, DS2.qstate, DS2.qlo, DS2.qhi, st));
    RUN(bwt_query_launch(&EA.x, &slot2.b, &L, st));
    HIPX(hipStreamSynchronize(st));
    const auto a4 = clk::now();
    BwExtractResult ax{};
    const int arc = bw_extract_answers(slot2.b.answers, size_t(DS2.nq), st, nullptr, nullptr, BW_EXTRACT_FNV_ON_DEVICE,
                                       &ax);
    if (arc) return arc < 0 ? 4 : 5;
    const auto a5 = clk::now();
    dall_mismatch = (DS2.nq == M && sha == ax.sha256_hex) ? 0 : 1;
    if (dall_mismatch) std::fprintf(stderr, "bw_batch: the all-device run's answers differ\n");
    std::snprintf(dall_json, sizeof(dall_json),
                  "{\"generate_ms\":%.3f,\"index_ms\":%.3f,\"alloc_ms\":%.3f,\"plan_answer_ms\":%.3f,"
                  "\"extract_ms\":%.3f,\"total_ms\":%.3f,\"answers_sha256\":\"%s\",\"mismatch\":%" PRId64 "}",
                  ms_between(a0, a1), ms_between(a1, a2), ms_between(a2, a3), ms_between(a3, a4), ms_between(a4, a5),
                  ms_between(a0, a5), ax.sha256_hex, dall_mismatch);
  }

  // ---- plain comparison on the first batches: each batch on its own
  const int64_t K = std::min<int64_t>(plain_sample, ss.nreq);
  double plain_pass_ms = 0, plain_query_ms = 0;
  int64_t plain_q = 0, plain_mismatch = 0;
  for (int64_t r = 0; r < K; ++r) {
    const int64_t a = rs.req_q0[size_t(r)], b = rs.req_q0[size_t(r) + 1];
    BatchPlan bp;
    if (!build_batch_plan(E.plan, b - a, rs.qstate.data() + a, rs.qlo.data() + a, rs.qhi.data() + a, true, bp, err,
                          sizeof(err))) {
      std::fprintf(stderr, "bw_batch: plain plan: %s\n", err);
      return 3;
    }
    DeviceBatch one;
    if (!upload_batch(E, std::move(bp), d_qlo + a, d_qhi + a, one, err, sizeof(err))) return 5;
    RUN(bwt_pass_launch(&E.x, &one.b, &L, st));  // warm
    HIPX(hipEventRecord(e0, st));
    RUN(bwt_pass_launch(&E.x, &one.b, &L, st));
    HIPX(hipEventRecord(e1, st));
    HIPX(hipEventSynchronize(e1));
    float t = 0;
    HIPX(hipEventElapsedTime(&t, e0, e1));
    plain_pass_ms += t;
    HIPX(hipEventRecord(e0, st));
    RUN(bwt_query_launch(&E.x, &one.b, &L, st));
    HIPX(hipEventRecord(e1, st));
    HIPX(hipEventSynchronize(e1));
    HIPX(hipEventElapsedTime(&t, e0, e1));
    plain_query_ms += t;
    std::vector<int64_t> one_ans(static_cast<size_t>(b - a));
    HIPX(hipMemcpy(one_ans.data(), one.b.answers, size_t(b - a) * sizeof(int64_t), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < b - a; ++i) plain_mismatch += one_ans[size_t(i)] != ans[size_t(a + i)];
    plain_q += b - a;
  }
  HIPX(hipGraphExecDestroy(gexec));
  HIPX(hipGraphDestroy(graph));

  std::sort(run_ms.begin(), run_ms.end());
  double avg = 0;
  for (double x : run_ms) avg += x;
  avg /= double(run_ms.size());
  hipDeviceProp_t prop{};
  HIPX(hipGetDeviceProperties(&prop, 0));
  const char* dev = prop.name[0] ? prop.name : prop.gcnArchName;
  const BatchPlan& BP = DB.plan;
  char buf[8192];
  const int n = std::snprintf(
      buf, sizeof(buf),
      "{\"arm\":\"bwgpu_batched\",\"form\":4,\"S\":%" PRId64 ",\"U\":%d,\"N\":%" PRIu64 ",\"seed\":%" PRIu64
      ",\"batches\":%" PRId64 ",\"max_batch\":%d,\"n_answers\":%" PRId64 ",\"answers_sha256\":\"%s\","
      "\"plan\":{\"nblk\":%d,\"nqblk\":%zu,\"ngrp\":%zu,\"neitems\":%zu,\"nwork\":%zu,\"rmax\":%d,\"tile_shift\":%d,"
      "\"chunk_items\":%" PRId64 ",\"ep_per_work\":%" PRId64 ",\"range_slots\":%" PRId64 ",\"npairs\":%" PRId64
      ",\"threads\":%d,\"waves\":%d,\"knobs_from\":\"%s\"},"
      "\"phases_ms\":{\"generate\":%.3f,\"index_plan_upload\":%.3f,\"batch_plan\":%.3f,\"batch_upload\":%.3f,"
      "\"extract\":%.3f,\"index_build\":%.4f},\"answers_fnv1a\":\"%016" PRIx64 "\","
      "\"graph_ms\":{\"avg\":%.4f,\"min\":%.4f,\"max\":%.4f,\"repeat\":%d,\"warmup\":%d,\"checked\":%d,"
      "\"replay_mismatch\":%" PRId64 "},"
      "\"queries_per_s\":%.1f,"
      "\"stage_ms\":{\"block_replay\":%.4f,\"chunk_replay\":%.4f,\"combine\":%.4f},"
      "\"stage_bytes\":{\"block_replay\":%.0f,\"chunk_replay\":%.0f,\"combine\":%.0f},"
      "\"plain\":{\"batches\":%" PRId64 ",\"queries\":%" PRId64 ",\"pass_ms\":%.4f,\"query_ms\":%.4f,"
      "\"pass_queries_per_s\":%.1f,\"query_queries_per_s\":%.1f,\"mismatch\":%" PRId64 "},"
      "\"device_plan\":{\"used\":%s,\"ms\":%.4f,\"mismatch\":%" PRId64 ",\"step_ms\":%.4f,\"step_queries_per_s\":%.1f},"
      "\"device_index\":%s,\"device_all\":%s,\"dense\":%s,\"noalloc\":%s,\"device_gen\":{\"used\":%s,\"gpu_ms\":%.4f,\"wall_ms\":%.3f,\"host_ms\":%.3f,\"ties\":%u,"
      "\"mismatch\":%" PRId64 "},\"forked\":%d,"
      "\"device_bytes\":%zu,\"device\":\"%s\",\"cus\":%d}\n",
      spec.S, spec.U, spec.N, spec.seed, ss.nreq, ss.max_batch, M, sha.c_str(), E.plan.nblk(), BP.qblk.size(),
      BP.grp_step.size(), BP.eitems.size(), BP.wk_chunk.size(), E.plan.knobs.rmax, E.plan.knobs.tile_shift,
      E.plan.knobs.chunk_items, E.plan.knobs.ep_per_work, E.plan.knobs.range_slots, BP.npairs, L.block_threads,
      L.chunk_waves, tuned_from.c_str(),
      ms_between(t0, t1), ms_between(t1, t2), ms_between(t2, t3), ms_between(t3, t4), xr.total_ms, index_ms, xr.fnv1a,
      avg, run_ms.front(), run_ms.back(), repeat, warmup, nfnv, replay_mismatch, double(M) / (avg * 1e-3), st_ms[0], st_ms[1], st_ms[2],
      sb_block, sb_chunk, sb_combine, K, plain_q, plain_pass_ms,
      plain_query_ms, plain_pass_ms > 0 ? double(plain_q) / (plain_pass_ms * 1e-3) : 0.0,
      plain_query_ms > 0 ? double(plain_q) / (plain_query_ms * 1e-3) : 0.0, plain_mismatch,
      dev_plan ? "true" : "false", dplan_ms, plan_mismatch, step2_avg,
      step2_avg > 0 ? double(M) / (step2_avg * 1e-3) : 0.0, didx_json, dall_json, dense_json, noalloc_json, dev_gen ? "true" : "false", dgen_gpu_ms, dgen_wall_ms,
      stream_host_ms, DS.ties, dgen_mismatch, fork, E.arena.bytes + io.bytes + DB.arena.bytes + dslot.arena.bytes,
      dev, prop.multiProcessorCount);
  std::fwrite(buf, 1, size_t(n), stdout);
  if (!json_path.empty()) {
    FILE* f = std::fopen(json_path.c_str(), "w");
    if (!f) return 4;
    std::fwrite(buf, 1, size_t(n), f);
    std::fclose(f);
  }
  return plain_mismatch == 0 && plan_mismatch <= 0 && dgen_mismatch <= 0 && didx_mismatch <= 0 && dall_mismatch <= 0 &&
                     dense_mismatch <= 0 && noalloc_mismatch <= 0 && replay_mismatch == 0
             ? 0
             : 6;
}
