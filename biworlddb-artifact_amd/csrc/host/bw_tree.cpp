// bw_tree.cpp -- timed worker for the generalised GPU batch path (one process per run), with the
// phase/JSON contract of the reference GPU workers (src/panel_batch_gpu.cpp:41-67,258-293):
// generate (load), setup (host plans), upload, execute (device passes), extract (D2H + SHA-256).
//
//   bw_tree --cfg 2 [--S n --U n --Q n --N n --seed n] [--mode pass|query]
//           [--rmax 2048 --tile-shift 10 --chunk 4096 --seg 32 --threads 256 --waves 4]
//           [--warmup 1 --repeat 5] [--answers FILE] [--json FILE] [--extract overlap] [--device-gen]
// --mode pass  : the plain form -- every timed pass rebuilds the index (block replay with weights,
//                chunk histograms, carry scan) and answers the declared queries.
// --mode query : the index is built once (timed separately), every timed pass answers the
//                declared queries over the resident index.
// --extract overlap|serial|both : how the answers leave the device and are hashed -- overlapped
//                (bw_extract_answers with the FNV-1a on the device, the default), one step after
//                another on the host as the reference workers do, or both in turn (timed side by
//                side, checksums compared).
// --device-gen : also materialise the workload's update and query arrays with the device generator
//                (include/bwgpu_gen.h) and compare them element for element with the host's.
// --device-plan : every timed pass also plans its batch on the device (include/bwgpu_plan.h; in pass
//                mode with every block listed, bwt_plan_all_blocks_launch) instead of using the host
//                plan built before the clock starts.
// Exit codes follow the reference: 2 usage, 3 unsupported input, 4 I/O, 5 GPU error; 6 when the
// serial and overlapped extracts disagree.
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

#include "bwgpu_extract.h"
#include "bwgpu_tree.h"
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
static uint64_t fnv1a(const void* data, size_t n) {
  const uint8_t* p = static_cast<const uint8_t*>(data);
  uint64_t h = 1469598103934665603ULL;
  for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 1099511628211ULL;
  return h;
}

int main(int argc, char** argv) {
  int cfg = 2;
  for (int i = 1; i + 1 < argc; ++i)
    if (!std::strcmp(argv[i], "--cfg")) cfg = std::atoi(argv[i + 1]);
  if (cfg < 1 || cfg > 5) { std::fprintf(stderr, "bw_tree: --cfg must be 1..5\n"); return 2; }
  TreeSpec spec = tree_defaults(cfg);
  PlanKnobs kn;
  int threads = 256, waves = 4, warmup = 1, repeat = 5;
  std::string answers_path, json_path, mode = "pass", extract_mode = "overlap";
  bool wsha = false, verify_remap = false, device_gen = false, dplan = false;
  int64_t copy_probe = 0;  // bytes of a device-to-device copy run first (HBM counter scale)
  // --tuned FILE: the plain form's knobs from the autotuning file ("plain3" for config 3's 2^22
  // ranks, "plain" otherwise), as defaults the flags override
  std::string tuned_from = "defaults";
  for (int i = 1; i + 1 < argc; ++i)
    if (!std::strcmp(argv[i], "--tuned")) {
      LaunchKnobs tl;
      tl.block_threads = threads;
      tl.chunk_waves = waves;
      if (load_tuned(argv[i + 1], cfg == 3 ? "plain3" : "plain", kn, tl)) {
        threads = tl.block_threads;
        waves = tl.chunk_waves;
        tuned_from = argv[i + 1];
      }
    }
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--workload-sha")) { wsha = true; continue; }
    if (!std::strcmp(argv[i], "--verify-remap")) { verify_remap = true; continue; }
    if (!std::strcmp(argv[i], "--device-gen")) { device_gen = true; continue; }
    if (!std::strcmp(argv[i], "--device-plan")) { dplan = true; continue; }
    if (i + 1 >= argc) { std::fprintf(stderr, "bw_tree: missing value for %s\n", argv[i]); return 2; }
    const char* k = argv[i];
    const char* v = argv[++i];
    if (!std::strcmp(k, "--cfg")) continue;
    else if (!std::strcmp(k, "--S")) spec.S = std::atoll(v);
    else if (!std::strcmp(k, "--U")) spec.U = std::atoi(v);
    else if (!std::strcmp(k, "--Q")) spec.Q = std::atoi(v);
    else if (!std::strcmp(k, "--N")) spec.N = std::strtoull(v, nullptr, 10);
    else if (!std::strcmp(k, "--seed")) spec.seed = std::strtoull(v, nullptr, 10);
    else if (!std::strcmp(k, "--mode")) mode = v;
    else if (!std::strcmp(k, "--rmax")) kn.rmax = std::atoi(v);
    else if (!std::strcmp(k, "--tile-shift")) kn.tile_shift = std::atoi(v);
    else if (!std::strcmp(k, "--chunk")) kn.chunk_items = std::atoll(v);
    else if (!std::strcmp(k, "--ep-per-work")) kn.ep_per_work = std::atoll(v);
    else if (!std::strcmp(k, "--range-slots")) kn.range_slots = std::atoll(v);
    else if (!std::strcmp(k, "--stage-shift")) kn.stage_shift = std::atoi(v);
    else if (!std::strcmp(k, "--seg")) kn.seg_chunks = std::atoi(v);
    else if (!std::strcmp(k, "--max-block-steps")) kn.max_block_steps = std::atoi(v);
    else if (!std::strcmp(k, "--threads")) threads = std::atoi(v);
    else if (!std::strcmp(k, "--waves")) waves = std::atoi(v);
    else if (!std::strcmp(k, "--warmup")) warmup = std::atoi(v);
    else if (!std::strcmp(k, "--repeat")) repeat = std::atoi(v);
    else if (!std::strcmp(k, "--answers")) answers_path = v;
    else if (!std::strcmp(k, "--extract")) extract_mode = v;
    else if (!std::strcmp(k, "--json")) json_path = v;
    else if (!std::strcmp(k, "--copy-probe")) copy_probe = std::atoll(v);
    else if (!std::strcmp(k, "--tuned")) continue;
    else { std::fprintf(stderr, "bw_tree: unknown argument %s\n", k); return 2; }
  }
  if (spec.S < 1 || spec.U < 1 || spec.Q < 1 || spec.N < 1 || repeat < 1 || warmup < 0 ||
      (mode != "pass" && mode != "query") ||
      (extract_mode != "overlap" && extract_mode != "serial" && extract_mode != "both")) {
    std::fprintf(stderr, "bw_tree: bad sizes or mode\n");
    return 2;
  }

  // ---- load: generate the declared workload
  const auto t0 = clk::now();
  TreeWorkload w;
  tree_generate(spec, w);
  if (wsha) {
    // with --device-gen the update and query arrays hashed are the device generator's
    uint32_t ties = 0;
    if (device_gen) {
      char gerr[256];
      HIPX(hipSetDevice(0));
      DeviceTree G;
      if (!device_generate_tree(spec, w.dkey, G, nullptr, gerr, sizeof(gerr))) {
        std::fprintf(stderr, "bw_tree: device generation: %s\n", gerr);
        return 5;
      }
      std::fill(w.key.begin(), w.key.end(), 0u);
      std::fill(w.qlo.begin(), w.qlo.end(), 0u);
      std::fill(w.qhi.begin(), w.qhi.end(), 0u);
      std::fill(w.delta.begin(), w.delta.end(), int64_t(0));
      HIPX(hipMemcpy(w.key.data(), G.key, w.key.size() * 4, hipMemcpyDeviceToHost));
      HIPX(hipMemcpy(w.delta.data(), G.delta, w.delta.size() * 8, hipMemcpyDeviceToHost));
      HIPX(hipMemcpy(w.qlo.data(), G.qlo, w.qlo.size() * 4, hipMemcpyDeviceToHost));
      HIPX(hipMemcpy(w.qhi.data(), G.qhi, w.qhi.size() * 4, hipMemcpyDeviceToHost));
      ties = G.ties;
    }
    char hex[65];
    tree_workload_sha(w, hex);
    std::printf("{\"workload_sha256\":\"%s\",\"device_gen\":%s,\"ties\":%u}\n", hex,
                device_gen ? "true" : "false", ties);
    return 0;
  }
  const auto t1 = clk::now();

  // ---- setup: coordinates (ranks for the sparse key space), index plan, declared batch plan
  const int64_t S = spec.S;
  const size_t nu = size_t(S) * size_t(spec.U), nq = size_t(S) * size_t(spec.Q);
  RankLut lut;
  std::vector<uint32_t> kc, lc, hc;
  const uint32_t *kcoord = w.key.data(), *lcoord = w.qlo.data(), *hcoord = w.qhi.data();
  if (cfg == 3) {
    lut = build_rank_lut(w.dkey);
    kc.resize(nu); lc.resize(nq); hc.resize(nq);
    host_ranks(lut, w.key.data(), kc.data(), nu);
    host_ranks(lut, w.qlo.data(), lc.data(), nq);
    host_ranks(lut, w.qhi.data(), hc.data(), nq);
    kcoord = kc.data(); lcoord = lc.data(); hcoord = hc.data();
  }
  char err[256];
  HIPX(hipSetDevice(0));
  HIPX(hipFree(nullptr));
  TreeEngine E;
  if (!E.init(S, spec.U, w.coord_n(), w.parent.data(), kcoord, w.delta.data(), kn, BwtLaunch{threads, waves, 0, 0},
              err, sizeof(err))) {
    std::fprintf(stderr, "bw_tree: index: %s\n", err);
    return std::strncmp(err, "device", 6) == 0 ? 5 : 3;
  }
  const TreeIndexPlan& P = E.plan;
  std::vector<int32_t> qstate(nq);
  for (size_t q = 0; q < nq; ++q) qstate[q] = int32_t(q / size_t(spec.Q) + 1);
  BatchPlan BPlan;
  if (!build_batch_plan(P, int64_t(nq), qstate.data(), lcoord, hcoord, mode == "pass", BPlan, err, sizeof(err))) {
    std::fprintf(stderr, "bw_tree: batch plan: %s\n", err);
    return 3;
  }
  const auto t2 = clk::now();

  // ---- upload the batch inputs (the index arrays went up in init)
  const auto tu0 = clk::now();
  DeviceArena io;
  uint32_t *d_qlo_raw = nullptr, *d_qhi_raw = nullptr, *d_row = nullptr;
  uint16_t* d_low = nullptr;
  uint32_t* d_qlo;
  uint32_t* d_qhi;
  if (cfg == 3) {  // raw 32-bit endpoints are resolved by rank on the device, inside the pass
    d_qlo_raw = io.upload(w.qlo);
    d_qhi_raw = io.upload(w.qhi);
    d_row = io.upload(lut.row_ptr);
    d_low = io.upload(lut.low16);
    d_qlo = io.scratch<uint32_t>(nq);
    d_qhi = io.scratch<uint32_t>(nq);
  } else {
    d_qlo = io.upload(w.qlo);
    d_qhi = io.upload(w.qhi);
  }
  DeviceBatch DB;
  if (!io.ok() || !upload_batch(E, std::move(BPlan), d_qlo, d_qhi, DB, err, sizeof(err))) {
    std::fprintf(stderr, "bw_tree: upload failed\n");
    return 5;
  }
  const BatchPlan& BP = DB.plan;
  if (cfg == 3) DB.b.qtile = nullptr;  // combine reads the tiles of the ranks remapped on the device
  // --device-plan: the pass plans its batch on the device (include/bwgpu_plan.h) inside the
  // timed region, from the device copies of the query states and (remapped) endpoints
  DevicePlanSlot dslot;
  DeviceArena qsa;
  int32_t* d_qs = nullptr;
  if (dplan) {
    d_qs = qsa.upload(qstate);
    if (!qsa.ok() || !alloc_plan_slot(E, int64_t(nq), dslot, err, sizeof(err))) {
      std::fprintf(stderr, "bw_tree: device plan: %s\n", qsa.ok() ? err : "upload");
      return 5;
    }
  }
  const BwtBatch& Bt = dplan ? dslot.b : DB.b;
  BwtIndex& X = E.x;
  const size_t g_dev_bytes = E.arena.bytes + io.bytes + DB.arena.bytes;
  HIPX(hipDeviceSynchronize());
  const auto tu1 = clk::now();

  hipStream_t stream;
  HIPX(hipStreamCreate(&stream));
  if (copy_probe > 0) {
    void *ca = nullptr, *cb = nullptr;
    HIPX(hipMalloc(&ca, size_t(copy_probe)));
    HIPX(hipMalloc(&cb, size_t(copy_probe)));
    HIPX(hipMemsetAsync(ca, 1, size_t(copy_probe), stream));
    HIPX(hipMemcpyAsync(cb, ca, size_t(copy_probe), hipMemcpyDeviceToDevice, stream));
    HIPX(hipStreamSynchronize(stream));
    HIPX(hipFree(ca));
    HIPX(hipFree(cb));
  }
  BwtLaunch L{threads, waves, 0, 0};
  BwtRankArgs rk[2];
  if (cfg == 3) {
    rk[0] = BwtRankArgs{int64_t(nq), d_qlo_raw, d_qlo, d_row, d_low};
    rk[1] = BwtRankArgs{int64_t(nq), d_qhi_raw, d_qhi, d_row, d_low};
  }
  double index_ms = 0;
  // stage events: remap | block | hist | scan | replay | combine
  const int NST = 6;
  hipEvent_t ev[NST + 1];
  for (auto& e : ev) HIPX(hipEventCreate(&e));
  hipEvent_t evp;
  HIPX(hipEventCreate(&evp));
  double plan_stage_ms = 0;
  auto build_index = [&]() { RUN(E.build_index(stream)); };
  if (mode == "query") {
    build_index();  // warm
    HIPX(hipEventRecord(ev[0], stream));
    build_index();
    HIPX(hipEventRecord(ev[1], stream));
    HIPX(hipEventSynchronize(ev[1]));
    float t;
    HIPX(hipEventElapsedTime(&t, ev[0], ev[1]));
    index_ms = t;
  }
  auto run_pass = [&](bool timed) {
    BwtLaunch wl = L;
    wl.write_index = mode == "pass";
    wl.carry_in_block = mode != "pass";  // as bwt_query_launch / bwt_pass_launch do
    if (timed) HIPX(hipEventRecord(ev[0], stream));
    if (cfg == 3)
      for (auto& r : rk) RUN(bwt_rank_remap_launch(&r, stream));
    if (timed) HIPX(hipEventRecord(ev[1], stream));
    if (dplan) {
      RUN(device_plan(E, dslot, int64_t(nq), d_qs, d_qlo, d_qhi, stream));
      if (mode == "pass") RUN(bwt_plan_all_blocks_launch(&X, &dslot.b, stream));
    }
    if (timed) HIPX(hipEventRecord(evp, stream));
    RUN(bwt_block_replay_launch(&X, &Bt, &wl, stream));
    if (timed) HIPX(hipEventRecord(ev[2], stream));
    if (mode == "pass") RUN(bwt_chunk_hist_launch(&X, &wl, stream));
    if (timed) HIPX(hipEventRecord(ev[3], stream));
    if (mode == "pass") RUN(bwt_carry_scan_launch(&X, &wl, stream));
    if (timed) HIPX(hipEventRecord(ev[4], stream));
    RUN(bwt_chunk_replay_launch(&X, &Bt, &wl, stream));
    if (timed) HIPX(hipEventRecord(ev[5], stream));
    RUN(bwt_combine_launch(&X, &Bt, &wl, stream));
    if (timed) HIPX(hipEventRecord(ev[6], stream));
  };
  for (int i = 0; i < warmup; ++i) run_pass(false);
  HIPX(hipStreamSynchronize(stream));
  std::vector<double> pass_ms;
  double stage_ms[NST] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < repeat; ++i) {
    run_pass(true);
    HIPX(hipEventSynchronize(ev[NST]));
    float t = 0;
    HIPX(hipEventElapsedTime(&t, ev[0], ev[NST]));
    pass_ms.push_back(t);
    for (int s = 0; s < NST; ++s) {
      HIPX(hipEventElapsedTime(&t, ev[s], ev[s + 1]));
      stage_ms[s] += t / repeat;
    }
    HIPX(hipEventElapsedTime(&t, ev[1], evp));  // the device plan, inside the block stage's events
    plan_stage_ms += t / repeat;
    stage_ms[1] -= t / repeat;
  }
  int64_t remap_mismatch = -1;
  if (cfg == 3 && verify_remap) {  // device CSR rank-LUT output against the host ranks
    std::vector<uint32_t> dl(nq), dh(nq);
    HIPX(hipMemcpy(dl.data(), d_qlo, nq * 4, hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(dh.data(), d_qhi, nq * 4, hipMemcpyDeviceToHost));
    remap_mismatch = 0;
    for (size_t i = 0; i < nq; ++i) remap_mismatch += (dl[i] != lc[i]) + (dh[i] != hc[i]);
  }
  // ---- the device generator: the same update and query arrays materialised on the GPU, compared
  // element for element with the host generator's (timed on its second run: the first loads code)
  double dgen_gpu_ms = -1, dgen_wall_ms = -1;
  int64_t dgen_mismatch = -1;
  uint32_t dgen_ties = 0;
  if (device_gen) {
    DeviceTree DT;
    for (int rep = 0; rep < 2; ++rep) {
      const auto g0 = clk::now();
      if (!device_generate_tree(spec, w.dkey, DT, nullptr, err, sizeof(err))) {
        std::fprintf(stderr, "bw_tree: device generation: %s\n", err);
        return 5;
      }
      dgen_wall_ms = ms_between(g0, clk::now());
    }
    dgen_gpu_ms = DT.gpu_ms;
    dgen_ties = DT.ties;
    dgen_mismatch = count_device_mismatch(DT.key, w.key) + count_device_mismatch(DT.delta, w.delta) + count_device_mismatch(DT.qlo, w.qlo) +
                    count_device_mismatch(DT.qhi, w.qhi);
  }
  char dgen_json[256] = "null";
  if (device_gen)
    std::snprintf(dgen_json, sizeof(dgen_json),
                  "{\"gpu_ms\":%.4f,\"wall_ms\":%.3f,\"host_ms\":%.3f,\"ties\":%u,\"mismatch\":%" PRId64 "}",
                  dgen_gpu_ms, dgen_wall_ms, ms_between(t0, t1), dgen_ties, dgen_mismatch);

  // ---- download + extract: the answers off the device, written, hashed (SHA-256, FNV-1a 64).
  // serial: one step after another, as the reference workers do; overlap: bw_extract_answers
  // (chunks copied while three threads write and hash them); both: serial, then overlap, and the
  // two must agree.
  std::vector<int64_t> ans(nq);
  std::string sha;
  uint64_t fn = 0;
  double dl_ms = 0, ex_ms = 0, ser[5] = {-1, -1, -1, -1, -1};
  BwExtractResult xr{};
  bool extract_same = true;
  if (extract_mode != "overlap") {
    const auto a = clk::now();
    HIPX(hipMemcpy(ans.data(), Bt.answers, nq * sizeof(int64_t), hipMemcpyDeviceToHost));
    const auto b = clk::now();
    if (!answers_path.empty()) {
      FILE* f = std::fopen(answers_path.c_str(), "wb");
      if (!f) return 4;
      const bool ok = std::fwrite(ans.data(), sizeof(int64_t), nq, f) == nq;
      if (std::fclose(f) != 0 || !ok) return 4;
    }
    const auto c = clk::now();
    sha = sha_hex(ans.data(), nq * sizeof(int64_t));
    const auto d = clk::now();
    fn = fnv1a(ans.data(), nq * sizeof(int64_t));
    const auto e = clk::now();
    const double v[5] = {ms_between(a, b), ms_between(b, c), ms_between(c, d), ms_between(d, e), ms_between(a, e)};
    std::copy(v, v + 5, ser);
    dl_ms = ser[0];
    ex_ms = ser[4] - ser[0];
  }
  if (extract_mode != "serial") {
    const int xrc = bw_extract_answers(Bt.answers, nq, nullptr, answers_path.empty() ? nullptr : answers_path.c_str(),
                                       ans.data(), BW_EXTRACT_FNV_ON_DEVICE, &xr);
    if (xrc) {
      std::fprintf(stderr, "bw_tree: extract failed (%d)\n", xrc);
      return xrc < 0 ? 4 : 5;
    }
    if (extract_mode == "both") extract_same = sha == xr.sha256_hex && fn == xr.fnv1a;
    sha = xr.sha256_hex;
    fn = xr.fnv1a;
    dl_ms = xr.copy_ms;
    ex_ms = xr.total_ms - xr.copy_ms;
  }
  char ext_json[512];
  std::snprintf(ext_json, sizeof(ext_json),
                "{\"mode\":\"%s\",\"serial\":{\"copy\":%.3f,\"write\":%.3f,\"sha256\":%.3f,\"fnv1a\":%.3f,"
                "\"total\":%.3f},\"overlap\":{\"copy\":%.3f,\"write\":%.3f,\"sha256\":%.3f,\"fnv1a\":%.3f,"
                "\"total\":%.3f},\"same\":%s}",
                extract_mode.c_str(), ser[0], ser[1], ser[2], ser[3], ser[4], xr.copy_ms, xr.write_ms, xr.sha_ms,
                xr.fnv_ms, xr.total_ms, extract_same ? "true" : "false");

  std::sort(pass_ms.begin(), pass_ms.end());
  double avg = 0;
  for (double x : pass_ms) avg += x;
  avg /= double(pass_ms.size());
  hipDeviceProp_t prop{};
  HIPX(hipGetDeviceProperties(&prop, 0));
  const char* dev = prop.name[0] ? prop.name : prop.gcnArchName;
  // algorithmic bytes of one pass: every update (u32 key + i64 delta) and every query
  // (2 x u32 endpoint + i64 answer) once
  const double bytes_alg = double(nu) * 12.0 + double(nq) * 16.0;
  // algorithmic bytes of each stage (what the stage must read and write once):
  //  block replay: per walked event u16 rank + i64 delta; per step its id; per query its u32
  //    packed ranks + i64 result; with the index write the block's keys (u32 in, i64 weight
  //    out) and tiles (u16 in, i64 out)
  //  chunk histogram: every key item (u64) + its i64 weight, level-B weights out in chain
  //    order, the dense i64 rows out
  //  carry scan: the rows in and out; chunk replay: per walked key item its class word (u32) +
  //    weight (i64), per work item its chunk's carry row, per endpoint item (u64) its read and
  //    its result (i64); combine per slot: block and tiles (2 x u32), 2 level-C row reads, the
  //    in-block part, the 2 level-B places (u32) and results, its staging index (u16) + position
  //    (u32) and staged value; per query id: its staging position (u32), the staged value, the
  //    answer
  int64_t steps_run = mode == "pass" ? int64_t(P.steps.size()) : BP.steps_run, kit_run = 0;
  for (size_t i = 0; i < BP.wk_chunk.size(); ++i) {
    const int64_t c = BP.wk_chunk[i];
    kit_run += std::min<int64_t>(P.chunk_key0[size_t(c) + 1] - P.chunk_key0[size_t(c)],
                                 int64_t(BWT_EP_POS(BP.eitems[size_t(BP.wk_e0[i + 1] - 1)])));
  }
  const double nkeys_d = double(P.bkey.size()), ntile_d = double(P.tile_rank0.size());
  const int x_ncls_b = P.chain_ncls.size() > 1 ? P.chain_ncls[1] : 1;
  const double sb_remap = cfg == 3 ? double(nq) * 16.0 : 0.0;
  // query mode: the block replay also adds the level-C carry (query tiles + the block's row)
  const bool carry_in_block = mode != "pass";
  const double sb_block = double(steps_run) * spec.U * 10.0 + double(steps_run) * 4.0 + double(nq) * 12.0 +
                          (mode == "pass" ? nkeys_d * 12.0 + ntile_d * 10.0 : 0.0) +
                          (carry_in_block ? double(nq) * 4.0 + double(BP.qblk.size()) * double(P.KT) * 8.0 : 0.0);
  const double nkit_b = double(P.kitems.size()) - ntile_d;
  const double sb_hist = mode == "pass" ? double(P.kitems.size()) * 16.0 + nkit_b * 8.0 + double(P.n_cp) * 8.0 : 0.0;
  const double sb_scan = mode == "pass" ? double(P.n_cp) * 16.0 : 0.0;
  const double sb_replay = double(kit_run) * 12.0 + double(BP.wk_chunk.size()) * double(x_ncls_b) * 8.0 +
                           double(BP.eitems.size()) * 16.0;
  const double sb_combine = double(nq) * (carry_in_block ? 66.0 : 90.0);
  char buf[8192];
  int n = std::snprintf(
      buf, sizeof(buf),
      "{\"arm\":\"bwgpu_tree\",\"mode\":\"%s\",\"cfg\":%d,\"S\":%" PRId64 ",\"U\":%d,\"Q\":%d,\"N\":%" PRIu64
      ",\"NK\":%" PRIu64 ",\"seed\":%" PRIu64 ",\"n_answers\":%zu,\"answers_sha256\":\"%s\",\"answers_fn\":\"%016" PRIx64
      "\",\"plan\":{\"nsteps\":%zu,\"nblk\":%d,\"nkeys\":%zu,\"ntiles\":%zu,\"nkitems\":%zu,\"neitems\":%zu,"
      "\"nchunk\":%" PRId64 ",\"nwork\":%zu,\"nchain\":%d,\"nseg\":%" PRId64 ",\"max_cls\":%d,\"rmax\":%d,"
      "\"tile_shift\":%d,\"chunk_items\":%" PRId64 ",\"ngrp\":%zu,\"ep_per_work\":%" PRId64 ",\"seg_chunks\":%d,\"threads\":%d,\"waves\":%d,\"knobs_from\":\"%s\"},"
      "\"phases_ms\":{\"generate\":%.3f,\"setup\":%.3f,\"upload\":%.3f,\"download\":%.3f,\"extract\":%.3f,"
      "\"index_build\":%.4f},"
      "\"pass_ms\":{\"avg\":%.4f,\"min\":%.4f,\"max\":%.4f,\"repeat\":%d,\"warmup\":%d},"
      "\"stage_ms\":{\"rank_remap\":%.4f,\"device_plan\":%.4f,\"block_replay\":%.4f,\"chunk_hist\":%.4f,\"carry_scan\":%.4f,"
      "\"chunk_replay\":%.4f,\"combine\":%.4f},"
      "\"stage_bytes\":{\"rank_remap\":%.0f,\"block_replay\":%.0f,\"chunk_hist\":%.0f,\"carry_scan\":%.0f,"
      "\"chunk_replay\":%.0f,\"combine\":%.0f},"
      "\"device_plan\":%s,\"remap_mismatch\":%" PRId64 ",\"extract\":%s,\"device_gen\":%s,\"queries_per_s\":%.1f,\"bytes_alg\":%.0f,\"device_bytes\":%zu,\"device\":\"%s\",\"cus\":%d}\n",
      mode.c_str(), cfg, S, spec.U, spec.Q, spec.N, w.coord_n(), spec.seed, nq, sha.c_str(), fn, P.steps.size(),
      P.nblk(), P.bkey.size(), P.tile_rank0.size(), P.kitems.size(), BP.eitems.size(), P.nchunk(), BP.wk_chunk.size(),
      P.nchain(), P.nseg(), P.max_cls, kn.rmax, P.knobs.tile_shift, P.knobs.chunk_items, BP.grp_step.size(), kn.ep_per_work,
      kn.seg_chunks, threads, waves, tuned_from.c_str(),
      ms_between(t0, t1), ms_between(t1, t2), ms_between(tu0, tu1), dl_ms, ex_ms,
      index_ms, avg, pass_ms.front(), pass_ms.back(), repeat, warmup, stage_ms[0], plan_stage_ms, stage_ms[1], stage_ms[2],
      stage_ms[3], stage_ms[4], stage_ms[5], sb_remap, sb_block, sb_hist, sb_scan, sb_replay, sb_combine,
      dplan ? "true" : "false", remap_mismatch, ext_json, dgen_json, double(nq) / (avg * 1e-3), bytes_alg, g_dev_bytes, dev, prop.multiProcessorCount);
  std::fwrite(buf, 1, size_t(n), stdout);
  if (!json_path.empty()) {
    FILE* f = std::fopen((json_path + ".tmp").c_str(), "w");
    if (!f) return 4;
    std::fwrite(buf, 1, size_t(n), f);
    std::fclose(f);
    if (std::rename((json_path + ".tmp").c_str(), json_path.c_str()) != 0) return 4;
  }
  return extract_same ? 0 : 6;
}
