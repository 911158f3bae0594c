// bw_panel.cpp -- timed worker of the reference drop-in path, with the CLI, phases and JSON of the
// reference GPU workers (src/panel_batch_gpu.cpp:41-67,258-293, src/panel_batch_gpu2.cpp,
// src/panel_gpu_session.cpp:21-166).  It drives only the drop-in C-ABI (gpu_v7.h, gpu_v7b.h,
// gpu_session.h) plus the device carry (bwgpu_carry.h).
//
//   bw_panel --fixture F --topology 0 --seed S --factor-period P --q Q --answers A --json J
//            [--base-id ID] [--repeat R] [--arm v7|v7b|sess|sess_global] [--states K]
//            [--blocks B] [--threads T] [--carry gpu|cpu] [--per-state]
//   bw_panel --fixture BASE --spec SPEC --answers A --json J [--arm ...] [--plan-only]
//       a BWHHIST1 declared history (src/real_history_spec.hpp, src/real_branch_spec.hpp) over its
//       version-0 base fixture instead of the generated workload: chain histories as they are,
//       branching ones as the chain of their depth-first walk (entries apply an edit, exits undo
//       it); the answers are every version's, in (version, range) order.  --plan-only stops after
//       the host plan and prints its shape (no GPU call).
// BWH_RANGE_SCOPE=per_state selects per-state ranges as in the reference.
// Exit codes: 2 usage, 3 unsupported input, 4 I/O, 5 GPU error (src/panel_batch_gpu.cpp:32-39).
#include <hip/hip_runtime_api.h>
#include <openssl/sha.h>

#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "bwgpu_carry.h"
#include "gpu_session.h"
#include "gpu_v7.h"
#include "gpu_v7b.h"
#include "ref_plan.hpp"

using namespace bwg;
using clk = std::chrono::steady_clock;

extern "C" int bw_sess_run_global_launch(const GpuSessRunArgs* a);

#define HIPX(x)                                                                                   \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess) {                                                                       \
      std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return 5;                                                                                   \
    }                                                                                             \
  } while (0)
#define LAUNCH(x)                                                                 \
  do {                                                                            \
    const int rc_ = (x);                                                          \
    if (rc_ != 0) {                                                               \
      std::fprintf(stderr, "launch failed (%d) at %s:%d\n", rc_, __FILE__, __LINE__); \
      return 5;                                                                   \
    }                                                                             \
  } while (0)

static int64_t ns_between(clk::time_point a, clk::time_point b) {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(b - a).count();
}

// Every device array of a run lives in one arena: the inputs first (one contiguous range, sent by
// one copy from a pinned staging buffer), then the work arrays.  Allocation happens with the host
// plan (setup); the upload phase packs the inputs into the staging buffer on the host threads and
// copies them in one transfer.
struct RunArena {
  struct Input {
    void** dst;
    const void* src;
    size_t bytes, off;
  };
  std::vector<Input> inputs;
  std::vector<std::pair<void**, size_t>> places;  // every array's offset
  size_t in_bytes = 0, total = 0;
  unsigned char* dev = nullptr;
  unsigned char* stage = nullptr;
  static size_t align(size_t v) { return (v + 255) & ~size_t(255); }
  template <typename T>
  void input(T** d, const std::vector<T>& v) {
    const size_t b = std::max<size_t>(v.size(), 1) * sizeof(T);
    inputs.push_back({reinterpret_cast<void**>(d), v.data(), v.size() * sizeof(T), in_bytes});
    places.push_back({reinterpret_cast<void**>(d), in_bytes});
    in_bytes = align(in_bytes + b);
  }
  template <typename T>
  void work(T** d, size_t n) {
    if (total < in_bytes) total = in_bytes;
    places.push_back({reinterpret_cast<void**>(d), total});
    total = align(total + std::max<size_t>(n, 1) * sizeof(T));
  }
  size_t stage_bytes = 0;
  // out_bytes: the largest result read back through the staging buffer
  hipError_t allocate(size_t out_bytes) {
    if (total < in_bytes) total = in_bytes;
    stage_bytes = std::max<size_t>(std::max(in_bytes, out_bytes), 256);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&dev), std::max<size_t>(total, 256));
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&stage), stage_bytes);
    if (e == hipSuccess)
      for (auto& pl : places) *pl.first = dev + pl.second;
    // the process's first large host-to-device transfer starts the DMA engine (a one-time cost of
    // milliseconds to a hundred; small copies take a kernel path and do not): paid here, with the
    // allocation, as a resident library pays it once at initialisation.  The staging buffer holds
    // no input yet, so nothing of the run is transferred early.
    const size_t warm = std::min<size_t>(std::min(stage_bytes, std::max<size_t>(total, 256)), size_t(4) << 20);
    if (e == hipSuccess) e = hipMemcpy(dev, stage, warm, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(stage, dev, warm, hipMemcpyDeviceToHost);  // and the other direction's
    // the host threads that pack the inputs are started here too
    int nth = 0;
#pragma omp parallel
    {
#pragma omp atomic
      ++nth;
    }
    (void)nth;
    return e;
  }
  // n bytes of device memory at src into dst, through the pinned staging buffer
  hipError_t download(void* dst, const void* src, size_t n) {
    for (size_t o = 0; o < n; o += stage_bytes) {
      const size_t k = std::min(stage_bytes, n - o);
      const hipError_t e = hipMemcpy(stage, static_cast<const unsigned char*>(src) + o, k, hipMemcpyDeviceToHost);
      if (e != hipSuccess) return e;
      std::memcpy(static_cast<unsigned char*>(dst) + o, stage, k);
    }
    return hipSuccess;
  }
  int64_t pack_ns = 0;
  hipError_t upload() {
    const auto p0 = clk::now();
    // 256 KiB pieces over the host threads, then one pinned-to-device copy
    constexpr size_t kPiece = size_t(256) << 10;
    std::vector<std::pair<size_t, size_t>> piece;  // (input, offset within it)
    for (size_t i = 0; i < inputs.size(); ++i)
      for (size_t o = 0; o < inputs[i].bytes; o += kPiece) piece.push_back({i, o});
#pragma omp parallel for schedule(dynamic)
    for (int64_t k = 0; k < int64_t(piece.size()); ++k) {
      const Input& in = inputs[piece[size_t(k)].first];
      const size_t o = piece[size_t(k)].second, n = std::min(kPiece, in.bytes - o);
      std::memcpy(stage + in.off + o, static_cast<const unsigned char*>(in.src) + o, n);
    }
    pack_ns = ns_between(p0, clk::now());
    return hipMemcpy(dev, stage, in_bytes, hipMemcpyHostToDevice);
  }
  void release() {
    if (dev) (void)hipFree(dev);
    if (stage) (void)hipHostFree(stage);
    dev = stage = nullptr;
  }
  ~RunArena() { release(); }
};

int main(int argc, char** argv) {
  std::string fixture, base_id, answers_path, json_path, arm = "v7b", carry = "gpu", spec_path;
  int topology = 0, seed = 7, fp = 4, q = 2, repeat = 0, K = 2049, B = 256, threads = 256;
  bool per_state = false, plan_only = false;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--per-state")) { per_state = true; continue; }
    if (!std::strcmp(argv[i], "--plan-only")) { plan_only = true; continue; }
    if (i + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", argv[i]); return 2; }
    const char* k = argv[i];
    const char* v = argv[++i];
    if (!std::strcmp(k, "--fixture")) fixture = v;
    else if (!std::strcmp(k, "--base-id")) base_id = v;
    else if (!std::strcmp(k, "--topology")) topology = std::atoi(v);
    else if (!std::strcmp(k, "--seed")) seed = std::atoi(v);
    else if (!std::strcmp(k, "--factor-period")) fp = std::atoi(v);
    else if (!std::strcmp(k, "--q")) q = std::atoi(v);
    else if (!std::strcmp(k, "--repeat")) repeat = std::atoi(v);
    else if (!std::strcmp(k, "--answers")) answers_path = v;
    else if (!std::strcmp(k, "--json")) json_path = v;
    else if (!std::strcmp(k, "--arm")) arm = v;
    else if (!std::strcmp(k, "--states")) K = std::atoi(v);
    else if (!std::strcmp(k, "--blocks")) B = std::atoi(v);
    else if (!std::strcmp(k, "--threads")) threads = std::atoi(v);
    else if (!std::strcmp(k, "--carry")) carry = v;
    else if (!std::strcmp(k, "--spec")) spec_path = v;
    else { std::fprintf(stderr, "unknown argument %s\n", k); return 2; }
  }
  if (fixture.empty() || answers_path.empty() || json_path.empty()) {
    std::fprintf(stderr, "usage: bw_panel --fixture PATH --topology T --seed S --factor-period P --q Q "
                         "--answers A --json J [--arm v7|v7b|sess|sess_global] [--states K] [--blocks B]\n");
    return 2;
  }
  if (q < 1 || q > 128 || K < 2 || (arm != "v7" && arm != "v7b" && arm != "sess" && arm != "sess_global") ||
      (carry != "gpu" && carry != "cpu")) {
    std::fprintf(stderr, "unsupported q / states / arm / carry\n");
    return 2;
  }
  if (const char* sc = std::getenv("BWH_RANGE_SCOPE")) per_state = per_state || std::strcmp(sc, "per_state") == 0;
  const bool sess = arm == "sess" || arm == "sess_global";

  const auto t0 = clk::now();
  RefFixture f;
  if (!load_ref_fixture(fixture, f)) { std::fprintf(stderr, "cannot load fixture\n"); return 3; }
  const auto t1 = clk::now();
  RefWorkload w;
  std::string err;
  HistorySpec spec;
  std::vector<int32_t> entry;  // spec: chain state of every version
  const bool from_spec = !spec_path.empty();
  if (from_spec) {
    if (!load_history_spec(spec_path, f, spec, err)) { std::fprintf(stderr, "cannot use spec: %s\n", err.c_str()); return 3; }
    spec_workload(f, spec, w, entry);
    K = w.K;
    q = spec.q;
    fp = spec.fp;
    per_state = true;
    topology = spec.chain() ? 0 : 1;
  } else {
    make_ref_workload(f, K, topology, seed, fp, q, per_state, -1, w);
  }
  RefPlan p;
  if (!build_ref_plan(f, w, B, arm == "v7", p, err)) { std::fprintf(stderr, "unsupported case: %s\n", err.c_str()); return 3; }
  char spec_json[512] = "";
  if (from_spec)
    std::snprintf(spec_json, sizeof(spec_json),
                  ",\"spec\":{\"versions\":%d,\"q\":%d,\"branching\":%s,\"header_fields\":%d,\"walk_states\":%d,"
                  "\"inserts\":%" PRId64 ",\"deletes\":%" PRId64 ",\"credits\":%" PRId64 ",\"factor_edits\":%" PRId64
                  ",\"n_ins\":%" PRId64 "}",
                  spec.K, spec.q, spec.chain() ? "false" : "true", spec.header_fields, w.K, spec.n_inserts,
                  spec.n_deletes, spec.n_credits, spec.n_factor_edits, spec.n_ins);
  if (plan_only) {
    std::printf("{\"plan_only\":true,\"n_states\":%d,\"q\":%d,\"blocks\":%d,\"ncls\":%d,\"G\":%d,\"factor_groups\":%d%s}\n",
                K, q, p.B, p.ncls, p.G, p.n_active, spec_json);
    return 0;
  }
  B = p.B;
  const int G = p.G, Cpfx = p.Cpfx, V = K - 1;
  std::vector<GpuV7Node> n72;
  std::vector<GpuV7bNode> n40;
  if (arm == "v7") {
    n72.assign(size_t(K), GpuV7Node{});
    for (int v = 1; v < K; ++v) {
      const RefNode& r = p.node[size_t(v)];
      GpuV7Node& n = n72[size_t(v)];
      n.kind = r.kind;
      n.g_old = r.g_old; n.c_old = r.c_old; n.ti_old = r.ti_old;
      n.g_new = r.g_new; n.c_new = r.c_new; n.ti_new = r.ti_new;
      n.amt_old = r.amt_old; n.amt_new = r.amt_new; n.fac_val = r.fac_val; n.fac_prev = r.fac_prev;
      n.ex_old = r.ex_old; n.ex_new = r.ex_new;
    }
  } else {
    if (G > 32767) { std::fprintf(stderr, "unsupported: the 40-byte node holds int16 group ids\n"); return 3; }
    n40.assign(size_t(K), GpuV7bNode{});
    for (int v = 1; v < K; ++v) {
      const RefNode& r = p.node[size_t(v)];
      GpuV7bNode& n = n40[size_t(v)];
      n.kind = uint8_t((r.kind == 1 ? 1 : 2) | (r.ex_old ? 4 : 0) | (r.ex_new ? 8 : 0));
      n.g_old = int16_t(r.g_old); n.g_new = int16_t(r.g_new);
      n.c_old = r.c_old; n.c_new = r.c_new; n.ti_old = r.ti_old; n.ti_new = r.ti_new;
      if (r.kind == 1) { n.a = r.amt_old; n.b = r.amt_new; } else { n.a = r.fac_val; n.b = r.fac_prev; }
    }
  }
  const auto t2 = clk::now();

  // ---- GPU arena (allocated with the plan), then the upload
  HIPX(hipSetDevice(0));
  HIPX(hipFree(nullptr));
  const auto ta0 = clk::now();
  RunArena ar;
  void* d_nodes = nullptr;
  GpuV7Node* d_n72 = nullptr;
  GpuV7bNode* d_n40 = nullptr;
  if (arm == "v7") ar.input(&d_n72, n72);
  else ar.input(&d_n40, n40);
  int32_t *d_v0, *d_v1, *d_g2row, *d_msoff, *d_msdev, *d_qlb, *d_qhb, *d_bidx, *d_gboff, *d_gbblk;
  int64_t *d_fac, *d_mscls, *d_msbase, *d_need, *d_gbreg, *d_vbase, *d_devbuf, *d_devbak = nullptr;
  int64_t *d_lvals, *d_dV, *d_carry, *d_ans, *d_fen = nullptr, *d_leaf = nullptr;
  ar.input(&d_v0, p.bv0); ar.input(&d_v1, p.bv1); ar.input(&d_g2row, p.g2row); ar.input(&d_msoff, p.ms_off);
  ar.input(&d_msdev, p.ms_dev); ar.input(&d_qlb, p.qlb); ar.input(&d_qhb, p.qhb); ar.input(&d_bidx, p.bidx);
  ar.input(&d_gboff, p.gb_off); ar.input(&d_gbblk, p.gb_block); ar.input(&d_fac, p.fac_start);
  ar.input(&d_mscls, p.ms_cls); ar.input(&d_msbase, p.ms_base); ar.input(&d_need, p.need_off);
  ar.input(&d_gbreg, p.gb_reg); ar.input(&d_vbase, p.vbase);
  if (arm == "v7") ar.input(&d_devbuf, p.snap);
  else ar.work(&d_devbuf, size_t(p.devbuf_n));
  const size_t lv_n = size_t(V) * 2 * q, dv_n = size_t(B) * Cpfx, ans_n = size_t(V) * q;
  ar.work(&d_lvals, lv_n); ar.work(&d_dV, dv_n); ar.work(&d_carry, dv_n); ar.work(&d_ans, ans_n);
  if (sess) { ar.work(&d_fen, dv_n); ar.work(&d_leaf, dv_n); ar.work(&d_devbak, size_t(p.devbuf_n)); }
  HIPX(ar.allocate(ans_n * sizeof(int64_t)));
  d_nodes = arm == "v7" ? static_cast<void*>(d_n72) : static_cast<void*>(d_n40);
  const auto tu0 = clk::now();
  HIPX(ar.upload());
  if (arm != "v7") HIPX(hipMemset(d_devbuf, 0, std::max<int64_t>(p.devbuf_n, 1) * 8));
  HIPX(hipDeviceSynchronize());
  const auto tu1 = clk::now();

  hipEvent_t e0, e1, e2, e3, e4, e5;
  HIPX(hipEventCreate(&e0)); HIPX(hipEventCreate(&e1)); HIPX(hipEventCreate(&e2));
  HIPX(hipEventCreate(&e3)); HIPX(hipEventCreate(&e4)); HIPX(hipEventCreate(&e5));
  std::vector<int64_t> dV_h, cp_h;
  float snap_ms = 0, kernel_ms = 0, carry_ms = 0, combine_ms = 0;
  // ---- snapshots (v7b / session: on device)
  HIPX(hipEventRecord(e0, 0));
  if (arm == "v7b") {
    GpuV7bSnapArgs sa{K, static_cast<const GpuV7bNode*>(d_nodes), d_bidx, d_gboff, d_gbblk, d_gbreg, d_devbuf};
    LAUNCH(gpu_v7b_snap_launch(&sa));
  } else if (sess) {
    GpuSessSnapArgs sa{K, static_cast<const GpuSessNode*>(d_nodes), d_bidx, d_gboff, d_gbblk, d_gbreg, d_devbuf};
    LAUNCH(gpu_sess_snap_launch(&sa));
    HIPX(hipMemcpy(d_devbak, d_devbuf, std::max<int64_t>(p.devbuf_n, 1) * 8, hipMemcpyDeviceToDevice));
  }
  HIPX(hipEventRecord(e1, 0));
  HIPX(hipEventSynchronize(e1));
  HIPX(hipEventElapsedTime(&snap_ms, e0, e1));

  // one batch: replay, carry, combine (the session restores its snapshots first)
  auto batch = [&](bool timed) -> int {
    if (sess) {
      HIPX(hipMemcpy(d_devbuf, d_devbak, std::max<int64_t>(p.devbuf_n, 1) * 8, hipMemcpyDeviceToDevice));
      GpuSessResetArgs ra{B, Cpfx, d_fen, d_leaf};
      LAUNCH(gpu_sess_reset_launch(&ra));
    }
    HIPX(hipEventRecord(e2, 0));
    if (arm == "v7") {
      GpuV7Args a{B, G, Cpfx, p.ncls, q, static_cast<const GpuV7Node*>(d_nodes), d_v0, d_v1, d_fac, d_g2row,
                  d_msoff, d_mscls, d_msbase, d_msdev, d_devbuf, d_need, d_qlb, d_qhb, d_lvals, d_dV,
                  threads, int((2 * Cpfx + G) * 8)};
      LAUNCH(gpu_v7_launch(&a));
    } else if (arm == "v7b") {
      GpuV7bArgs a{B, G, Cpfx, p.ncls, q, static_cast<const GpuV7bNode*>(d_nodes), d_v0, d_v1, d_fac, d_g2row,
                   d_msoff, d_mscls, d_msbase, d_msdev, d_devbuf, d_need, d_qlb, d_qhb, d_lvals, d_dV,
                   threads, int((2 * Cpfx + G) * 8)};
      LAUNCH(gpu_v7b_launch(&a));
    } else {
      GpuSessRunArgs a{B, G, Cpfx, p.ncls, q, static_cast<const GpuSessNode*>(d_nodes), d_v0, d_v1, d_fac, d_g2row,
                       d_msoff, d_mscls, d_msbase, d_msdev, d_devbuf, d_need, d_qlb, d_qhb, d_lvals, d_dV,
                       d_fen, d_leaf, threads};
      LAUNCH(arm == "sess" ? gpu_sess_run_launch(&a) : bw_sess_run_global_launch(&a));
    }
    HIPX(hipEventRecord(e3, 0));
    if (carry == "gpu") {
      BwCarryArgs ca{B, Cpfx, d_vbase, d_dV, d_carry};
      LAUNCH(bw_carry_launch(&ca, nullptr));
    } else {
      // the reference's serial host carry (src/panel_batch_gpu.cpp:191-204)
      dV_h.resize(dv_n);
      cp_h.assign(dv_n, 0);
      HIPX(hipMemcpy(dV_h.data(), d_dV, dv_n * 8, hipMemcpyDeviceToHost));
      std::vector<int64_t> vstart(p.vbase);
      for (int b = 0; b < B; ++b) {
        int64_t* cp = &cp_h[size_t(b) * Cpfx];
        for (int k = 1; k < Cpfx; ++k) cp[k] = cp[k - 1] + vstart[size_t(k)];
        const int64_t* dv = &dV_h[size_t(b) * Cpfx];
        for (int k = 1; k < Cpfx; ++k) vstart[size_t(k)] += dv[k];
      }
      HIPX(hipMemcpy(d_carry, cp_h.data(), dv_n * 8, hipMemcpyHostToDevice));
    }
    HIPX(hipEventRecord(e4, 0));
    if (arm == "v7") {
      GpuV7CombineArgs c{V, q, Cpfx, d_qlb, d_qhb, d_carry, d_bidx, d_lvals, d_ans};
      LAUNCH(gpu_combine_launch(&c));
    } else if (arm == "v7b") {
      GpuV7bCombineArgs c{V, q, Cpfx, d_qlb, d_qhb, d_carry, d_bidx, d_lvals, d_ans};
      LAUNCH(gpu_v7b_combine_launch(&c));
    } else {
      GpuSessCombineArgs c{V, q, Cpfx, d_qlb, d_qhb, d_carry, d_bidx, d_lvals, d_ans};
      LAUNCH(gpu_sess_combine_launch(&c));
    }
    HIPX(hipEventRecord(e5, 0));
    HIPX(hipEventSynchronize(e5));
    if (timed) {
      float t;
      HIPX(hipEventElapsedTime(&t, e2, e3)); kernel_ms = t;
      HIPX(hipEventElapsedTime(&t, e3, e4)); carry_ms = t;
      HIPX(hipEventElapsedTime(&t, e4, e5)); combine_ms = t;
    }
    return 0;
  };
  const auto tb0 = clk::now();
  if (batch(true)) return 5;
  const auto tb1 = clk::now();
  int64_t steady_ns = 0;
  if (sess && repeat > 0) {  // session: further batches over the resident plan
    const auto ts0 = clk::now();
    for (int r = 0; r < repeat; ++r)
      if (batch(false)) return 5;
    steady_ns = ns_between(ts0, clk::now()) / repeat;
  }
  std::vector<int64_t> answers(ans_n);
  HIPX(ar.download(answers.data(), d_ans, ans_n * 8));
  if (from_spec) {  // every version's answers, from the chain state of its entry
    std::vector<int64_t> by_version(size_t(spec.K - 1) * q);
    for (int v = 1; v < spec.K; ++v)
      for (int k = 0; k < q; ++k)
        by_version[size_t(v - 1) * q + k] = answers[size_t(entry[size_t(v)] - 1) * q + k];
    answers.swap(by_version);
  }
  const size_t out_n = answers.size();
  const auto t3 = clk::now();

  // ---- extract
  FILE* out = std::fopen(answers_path.c_str(), "wb");
  if (!out) return 4;
  const bool wok = std::fwrite(answers.data(), 8, answers.size(), out) == answers.size();
  std::fclose(out);
  if (!wok) return 4;
  uint64_t fnv = 1469598103934665603ULL;
  const uint8_t* bp = reinterpret_cast<const uint8_t*>(answers.data());
  for (size_t i = 0; i < out_n * 8; ++i) fnv = (fnv ^ bp[i]) * 1099511628211ULL;
  unsigned char dg[32];
  SHA256(bp, out_n * 8, dg);
  char sha[65];
  for (int i = 0; i < 32; ++i) std::snprintf(sha + 2 * i, 3, "%02x", dg[i]);
  const auto t4 = clk::now();
  ar.release();
  const auto t5 = clk::now();

  // setup: the host plan and the device arena (with its pinned staging buffer)
  const int64_t alloc_ns = ns_between(ta0, tu0);
  const int64_t load_ns = ns_between(t0, t1), setup_ns = ns_between(t1, t2) + alloc_ns, upload_ns = ns_between(tu0, tu1);
  const int64_t batch_ns = ns_between(tb0, tb1), download_ns = ns_between(tb1, t3) - steady_ns * repeat;
  const int64_t execute_ns = upload_ns + int64_t(snap_ms * 1e6) + batch_ns + download_ns;
  const int64_t extract_ns = ns_between(t3, t4), cleanup_ns = ns_between(t4, t5);
  hipDeviceProp_t prop{};
  HIPX(hipGetDeviceProperties(&prop, 0));
  char buf[4096];
  const int n = std::snprintf(
      buf, sizeof(buf),
      "{\"arm\":\"bwgpu_%s\",\"base_id\":\"%s\",\"fixture\":\"%s\",\"topology\":%d,\"seed\":%d,\"factor_period\":%d,"
      "\"q\":%d,\"repeat\":%d,\"n_states\":%d,\"n_answers\":%zu,\"range_scope\":\"%s\",\"answers_fn\":\"%016" PRIx64
      "\",\"answers_sha256\":\"%s\",\"phases_ns\":{\"load\":%" PRId64 ",\"setup\":%" PRId64 ",\"execute\":%" PRId64
      ",\"extract\":%" PRId64 ",\"cleanup\":%" PRId64 "},\"gpu\":{\"device\":\"%s\",\"blocks\":%d,\"threads\":%d,"
      "\"Cpfx\":%d,\"G\":%d,\"carry\":\"%s\",\"alloc_ns\":%" PRId64 ",\"upload_ns\":%" PRId64 ",\"upload_bytes\":%zu,\"pack_ns\":%" PRId64 ",\"snap_ns\":%" PRId64
      ",\"kernel_ns\":%" PRId64
      ",\"carry_ns\":%" PRId64 ",\"combine_ns\":%" PRId64 ",\"download_ns\":%" PRId64 ",\"steady_batch_ns\":%" PRId64
      "},\"full_ns\":%" PRId64 "%s}\n",
      arm.c_str(), base_id.c_str(), fixture.c_str(), topology, seed, fp, q, repeat, from_spec ? spec.K : K, out_n,
      per_state ? "per_state" : "shared", fnv, sha, load_ns, setup_ns, execute_ns, extract_ns, cleanup_ns, prop.name, B,
      threads, Cpfx, G, carry.c_str(), alloc_ns, upload_ns, ar.in_bytes, ar.pack_ns, int64_t(snap_ms * 1e6), int64_t(kernel_ms * 1e6),
      int64_t(carry_ms * 1e6), int64_t(combine_ms * 1e6), download_ns, steady_ns,
      load_ns + setup_ns + execute_ns + extract_ns + cleanup_ns, spec_json);
  if (n < 0 || size_t(n) >= sizeof(buf)) return 4;
  std::fwrite(buf, 1, size_t(n), stdout);
  FILE* jf = std::fopen((json_path + ".tmp").c_str(), "w");
  if (!jf) return 4;
  std::fwrite(buf, 1, size_t(n), jf);
  std::fclose(jf);
  if (std::rename((json_path + ".tmp").c_str(), json_path.c_str()) != 0) return 4;
  return 0;
}
