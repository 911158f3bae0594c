// bw_tree.cpp -- timed worker for the generalised GPU batch path (one process per run), with the
// phase/JSON contract of the reference GPU workers (src/panel_batch_gpu.cpp:41-67,258-293):
// load/generate, setup (host plan), upload, execute (device passes), extract (D2H + SHA-256).
//
//   bw_tree --cfg 2 [--S n --U n --Q n --N n --seed n] [--rmax 2048 --tile-shift 11
//           --chunk 16384 --seg 32 --threads 256 --waves 4] [--warmup 1 --repeat 5]
//           [--answers FILE] [--json FILE]
//
// Exit codes follow the reference: 2 usage, 3 unsupported input, 4 I/O, 5 GPU error.
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

#include "bwgpu_tree.h"
#include "tree_plan.hpp"
#include "tree_workload.hpp"

using namespace bwg;
using clk = std::chrono::steady_clock;

#define HIPX(x)                                                                                 \
  do {                                                                                          \
    hipError_t e_ = (x);                                                                        \
    if (e_ != hipSuccess) {                                                                     \
      std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      std::exit(5);                                                                             \
    }                                                                                           \
  } while (0)

static double ms_between(clk::time_point a, clk::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

template <typename T>
static T* upload(const std::vector<T>& v, size_t* total) {
  T* d = nullptr;
  const size_t n = std::max<size_t>(v.size(), 1) * sizeof(T);
  HIPX(hipMalloc(&d, n));
  if (!v.empty()) HIPX(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  *total += n;
  return d;
}
template <typename T>
static T* scratch(size_t n, size_t* total) {
  T* d = nullptr;
  HIPX(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)));
  *total += std::max<size_t>(n, 1) * sizeof(T);
  return d;
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
  std::string answers_path, json_path;
  bool wsha = false;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--workload-sha")) { wsha = true; continue; }
    if (i + 1 >= argc) { std::fprintf(stderr, "bw_tree: missing value for %s\n", argv[i]); return 2; }
    const char* k = argv[i];
    const char* v = argv[++i];
    if (!std::strcmp(k, "--cfg")) continue;
    else if (!std::strcmp(k, "--S")) spec.S = std::atoll(v);
    else if (!std::strcmp(k, "--U")) spec.U = std::atoi(v);
    else if (!std::strcmp(k, "--Q")) spec.Q = std::atoi(v);
    else if (!std::strcmp(k, "--N")) spec.N = std::strtoull(v, nullptr, 10);
    else if (!std::strcmp(k, "--seed")) spec.seed = std::strtoull(v, nullptr, 10);
    else if (!std::strcmp(k, "--rmax")) kn.rmax = std::atoi(v);
    else if (!std::strcmp(k, "--tile-shift")) kn.tile_shift = std::atoi(v);
    else if (!std::strcmp(k, "--chunk")) kn.chunk_items = std::atoll(v);
    else if (!std::strcmp(k, "--seg")) kn.seg_chunks = std::atoi(v);
    else if (!std::strcmp(k, "--max-block-steps")) kn.max_block_steps = std::atoi(v);
    else if (!std::strcmp(k, "--threads")) threads = std::atoi(v);
    else if (!std::strcmp(k, "--waves")) waves = std::atoi(v);
    else if (!std::strcmp(k, "--warmup")) warmup = std::atoi(v);
    else if (!std::strcmp(k, "--repeat")) repeat = std::atoi(v);
    else if (!std::strcmp(k, "--answers")) answers_path = v;
    else if (!std::strcmp(k, "--json")) json_path = v;
    else { std::fprintf(stderr, "bw_tree: unknown argument %s\n", k); return 2; }
  }
  if (spec.S < 1 || spec.U < 1 || spec.Q < 1 || spec.N < 1 || repeat < 1 || warmup < 0) {
    std::fprintf(stderr, "bw_tree: bad sizes\n");
    return 2;
  }

  // ---- load: generate the declared workload
  const auto t0 = clk::now();
  TreeWorkload w;
  tree_generate(spec, w);
  if (wsha) {
    char hex[65];
    tree_workload_sha(w, hex);
    std::printf("{\"workload_sha256\":\"%s\"}\n", hex);
    return 0;
  }
  const auto t1 = clk::now();

  // ---- setup: coordinates (ranks for the sparse key space) and the host plan
  const size_t nu = size_t(spec.S) * size_t(spec.U), nq = size_t(spec.S) * size_t(spec.Q);
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
  TreePlan P;
  char err[256];
  if (!build_tree_plan(spec.S, spec.U, spec.Q, w.coord_n(), w.parent.data(), kcoord, lcoord, hcoord, kn, P, err,
                       sizeof(err))) {
    std::fprintf(stderr, "bw_tree: plan: %s\n", err);
    return 3;
  }
  const auto t2 = clk::now();

  // ---- upload (inputs + plan), allocate scratch
  HIPX(hipSetDevice(0));
  HIPX(hipFree(nullptr));
  const auto tu0 = clk::now();
  size_t dev_bytes = 0;
  BwtPassArgs a{};
  a.S = spec.S; a.U = spec.U; a.Q = spec.Q;
  a.nblk = P.nblk();
  a.block_threads = threads;
  a.rmax = kn.rmax;
  a.max_cls = P.max_cls;
  a.nchunk = P.nchunk();
  a.nchain = P.nchain();
  a.chunk_waves = waves;
  uint32_t *d_key, *d_qlo, *d_qhi;
  uint32_t *d_key_raw = nullptr, *d_qlo_raw = nullptr, *d_qhi_raw = nullptr, *d_row = nullptr;
  uint16_t* d_low = nullptr;
  if (cfg == 3) {
    d_key_raw = upload(w.key, &dev_bytes);
    d_qlo_raw = upload(w.qlo, &dev_bytes);
    d_qhi_raw = upload(w.qhi, &dev_bytes);
    d_row = upload(lut.row_ptr, &dev_bytes);
    d_low = upload(lut.low16, &dev_bytes);
    d_key = scratch<uint32_t>(nu, &dev_bytes);
    d_qlo = scratch<uint32_t>(nq, &dev_bytes);
    d_qhi = scratch<uint32_t>(nq, &dev_bytes);
  } else {
    d_key = upload(w.key, &dev_bytes);
    d_qlo = upload(w.qlo, &dev_bytes);
    d_qhi = upload(w.qhi, &dev_bytes);
  }
  a.key = d_key; a.qlo = d_qlo; a.qhi = d_qhi;
  a.delta = upload(w.delta, &dev_bytes);
  a.steps = upload(P.steps, &dev_bytes);
  a.blk_step0 = upload(P.blk_step0, &dev_bytes);
  a.blk_key0 = upload(P.blk_key0, &dev_bytes);
  a.bkey = upload(P.bkey, &dev_bytes);
  a.blk_tile0 = upload(P.blk_tile0, &dev_bytes);
  a.tile_rank0 = upload(P.tile_rank0, &dev_bytes);
  a.items = upload(P.items, &dev_bytes);
  a.chunk_item0 = upload(P.chunk_item0, &dev_bytes);
  a.chunk_chain = upload(P.chunk_chain, &dev_bytes);
  a.chunk_cp0 = upload(P.chunk_cp0, &dev_bytes);
  a.chain_ncls = upload(P.chain_ncls, &dev_bytes);
  a.chain_src = upload(P.chain_src, &dev_bytes);
  a.chain_out = upload(P.chain_out, &dev_bytes);
  a.chain_chunk0 = upload(P.chain_chunk0, &dev_bytes);
  a.nseg = P.nseg();
  a.seg_chunk0 = upload(P.seg_chunk0, &dev_bytes);
  a.seg_cp0 = upload(P.seg_cp0, &dev_bytes);
  a.chain_seg0 = upload(P.chain_seg0, &dev_bytes);
  a.segsum = scratch<int64_t>(size_t(P.n_segsum), &dev_bytes);
  a.agg = scratch<int64_t>(P.bkey.size(), &dev_bytes);
  a.tagg = scratch<int64_t>(P.tile_rank0.size(), &dev_bytes);
  a.cp = scratch<int64_t>(size_t(P.n_cp), &dev_bytes);
  a.g = scratch<int64_t>(4 * nq, &dev_bytes);
  a.ansr = scratch<int64_t>(nq, &dev_bytes);
  a.answers = scratch<int64_t>(nq, &dev_bytes);
  HIPX(hipDeviceSynchronize());
  const auto tu1 = clk::now();

  hipStream_t stream;
  HIPX(hipStreamCreate(&stream));
  BwtRankArgs rk[3];
  if (cfg == 3) {
    rk[0] = BwtRankArgs{int64_t(nu), d_key_raw, d_key, d_row, d_low};
    rk[1] = BwtRankArgs{int64_t(nq), d_qlo_raw, d_qlo, d_row, d_low};
    rk[2] = BwtRankArgs{int64_t(nq), d_qhi_raw, d_qhi, d_row, d_low};
  }
  // stage events: remap | block | hist | scan | replay | combine
  const int NST = 6;
  hipEvent_t ev[NST + 1];
  for (auto& e : ev) HIPX(hipEventCreate(&e));
  auto run_pass = [&](bool timed) {
    int rc = 0;
    if (timed) HIPX(hipEventRecord(ev[0], stream));
    if (cfg == 3)
      for (auto& r : rk)
        if ((rc = bwt_rank_remap_launch(&r, stream))) return rc;
    if (timed) HIPX(hipEventRecord(ev[1], stream));
    if ((rc = bwt_block_replay_launch(&a, stream))) return rc;
    if (timed) HIPX(hipEventRecord(ev[2], stream));
    if ((rc = bwt_chunk_hist_launch(&a, stream))) return rc;
    if (timed) HIPX(hipEventRecord(ev[3], stream));
    if ((rc = bwt_carry_scan_launch(&a, stream))) return rc;
    if (timed) HIPX(hipEventRecord(ev[4], stream));
    if ((rc = bwt_chunk_replay_launch(&a, stream))) return rc;
    if (timed) HIPX(hipEventRecord(ev[5], stream));
    if ((rc = bwt_combine_launch(&a, stream))) return rc;
    if (timed) HIPX(hipEventRecord(ev[6], stream));
    return 0;
  };
  for (int i = 0; i < warmup; ++i)
    if (run_pass(false)) return 5;
  HIPX(hipStreamSynchronize(stream));
  std::vector<double> pass_ms;
  double stage_ms[NST] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < repeat; ++i) {
    if (run_pass(true)) return 5;
    HIPX(hipEventSynchronize(ev[NST]));
    float t = 0;
    HIPX(hipEventElapsedTime(&t, ev[0], ev[NST]));
    pass_ms.push_back(t);
    for (int s = 0; s < NST; ++s) {
      HIPX(hipEventElapsedTime(&t, ev[s], ev[s + 1]));
      stage_ms[s] += t / repeat;
    }
  }
  const auto tx0 = clk::now();
  std::vector<int64_t> ans(nq);
  HIPX(hipMemcpy(ans.data(), a.answers, nq * sizeof(int64_t), hipMemcpyDeviceToHost));
  const auto tx1 = clk::now();
  if (!answers_path.empty()) {
    FILE* f = std::fopen(answers_path.c_str(), "wb");
    if (!f) return 4;
    const bool ok = std::fwrite(ans.data(), sizeof(int64_t), nq, f) == nq;
    if (std::fclose(f) != 0 || !ok) return 4;
  }
  const std::string sha = sha_hex(ans.data(), nq * sizeof(int64_t));
  const uint64_t fn = fnv1a(ans.data(), nq * sizeof(int64_t));
  const auto tx2 = clk::now();

  std::sort(pass_ms.begin(), pass_ms.end());
  double avg = 0;
  for (double x : pass_ms) avg += x;
  avg /= double(pass_ms.size());
  hipDeviceProp_t prop{};
  HIPX(hipGetDeviceProperties(&prop, 0));
  // algorithmic bytes of one pass: every update (u32 key + i64 delta) and every query
  // (2 x u32 endpoint + i64 answer) once
  const double bytes_alg = double(nu) * 12.0 + double(nq) * 16.0;
  // algorithmic bytes of each stage: what the stage must read and write once
  int64_t n_enter = 0, n_keyitems = 0;
  for (int32_t st : P.steps) n_enter += st > 0;
  for (uint64_t x : P.items) n_keyitems += !(x & BWT_ITEM_EP);
  const double nsteps_d = double(P.steps.size()), nkeys_d = double(P.bkey.size()), ntile_d = double(P.tile_rank0.size());
  const double nitems_d = double(P.items.size()), nep_d = double(2 * nq);
  const double sb_remap = cfg == 3 ? double(nu + 2 * nq) * 8.0 : 0.0;
  const double sb_block = nsteps_d * spec.U * 12.0 + nsteps_d * 4.0 + double(n_enter) * spec.Q * 16.0 +
                          nkeys_d * 12.0 + ntile_d * 10.0;
  const double sb_hist = nitems_d * 8.0 + double(n_keyitems) * 8.0 + double(P.n_cp) * 8.0;
  const double sb_scan = double(P.n_cp) * 16.0;
  const double sb_replay = nitems_d * 8.0 + double(n_keyitems) * 8.0 + nep_d * 2.0 * 8.0;
  const double sb_combine = double(nq) * 48.0;
  char buf[8192];
  int n = std::snprintf(
      buf, sizeof(buf),
      "{\"arm\":\"bwgpu_tree\",\"cfg\":%d,\"S\":%" PRId64 ",\"U\":%d,\"Q\":%d,\"N\":%" PRIu64 ",\"NK\":%" PRIu64
      ",\"seed\":%" PRIu64 ",\"n_answers\":%zu,\"answers_sha256\":\"%s\",\"answers_fn\":\"%016" PRIx64 "\","
      "\"plan\":{\"nsteps\":%zu,\"nblk\":%d,\"nkeys\":%zu,\"ntiles\":%zu,\"nitems\":%zu,\"nchunk\":%" PRId64
      ",\"nchain\":%d,\"nseg\":%" PRId64 ",\"max_cls\":%d,\"rmax\":%d,\"tile_shift\":%d,\"chunk_items\":%" PRId64
      ",\"seg_chunks\":%d,\"threads\":%d,\"waves\":%d},"
      "\"phases_ms\":{\"generate\":%.3f,\"setup\":%.3f,\"upload\":%.3f,\"download\":%.3f,\"extract\":%.3f},"
      "\"pass_ms\":{\"avg\":%.4f,\"min\":%.4f,\"max\":%.4f,\"repeat\":%d,\"warmup\":%d},"
      "\"stage_ms\":{\"rank_remap\":%.4f,\"block_replay\":%.4f,\"chunk_hist\":%.4f,\"carry_scan\":%.4f,"
      "\"chunk_replay\":%.4f,\"combine\":%.4f},"
      "\"stage_bytes\":{\"rank_remap\":%.0f,\"block_replay\":%.0f,\"chunk_hist\":%.0f,\"carry_scan\":%.0f,"
      "\"chunk_replay\":%.0f,\"combine\":%.0f},"
      "\"queries_per_s\":%.1f,\"bytes_alg\":%.0f,\"device_bytes\":%zu,\"device\":\"%s\",\"cus\":%d}\n",
      cfg, spec.S, spec.U, spec.Q, spec.N, w.coord_n(), spec.seed, nq, sha.c_str(), fn, P.steps.size(), P.nblk(),
      P.bkey.size(), P.tile_rank0.size(), P.items.size(), P.nchunk(), P.nchain(), P.nseg(), P.max_cls, kn.rmax,
      kn.tile_shift, kn.chunk_items, kn.seg_chunks, threads, waves, ms_between(t0, t1), ms_between(t1, t2),
      ms_between(tu0, tu1), ms_between(tx0, tx1), ms_between(tx1, tx2), avg, pass_ms.front(), pass_ms.back(), repeat,
      warmup, stage_ms[0], stage_ms[1], stage_ms[2], stage_ms[3], stage_ms[4], stage_ms[5], sb_remap, sb_block,
      sb_hist, sb_scan, sb_replay, sb_combine, double(nq) / (avg * 1e-3),
      bytes_alg, dev_bytes, prop.name, prop.multiProcessorCount);
  std::fwrite(buf, 1, size_t(n), stdout);
  if (!json_path.empty()) {
    FILE* f = std::fopen((json_path + ".tmp").c_str(), "w");
    if (!f) return 4;
    std::fwrite(buf, 1, size_t(n), f);
    std::fclose(f);
    if (std::rename((json_path + ".tmp").c_str(), json_path.c_str()) != 0) return 4;
  }
  return 0;
}
