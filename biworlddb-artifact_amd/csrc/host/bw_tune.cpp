// bw_tune.cpp -- runs the autotuner (tree_tune) for one workload form on this device and stores
// the chosen knobs in a tuned-knobs file read by bw_batch / bw_serve (--tuned FILE).
//
//   bw_tune --form batched|served [--S n --U n --N n --seed n] [--sample-requests R]
//           [--repeat 3] [--start FILE] --out FILE
// --start: begin the descent from FILE's knobs for this form (else from the defaults).
// The sample is the stream's first R requests (default: the whole batched stream; 400 served
// requests, about the size of one served batch): knobs are chosen at the density of queries per
// block the form actually runs at.
// One JSON line per trial on stderr, the chosen knobs as one JSON line on stdout.
#include <hip/hip_runtime_api.h>

#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tree_tune.hpp"
#include "tree_workload.hpp"

using namespace bwg;

int main(int argc, char** argv) {
  std::string form = "batched", out, start;
  int64_t sample = -1;
  int repeat = 3;
  int cfg_form = 4;
  for (int i = 1; i + 1 < argc; ++i)
    if (!std::strcmp(argv[i], "--form")) form = argv[i + 1];
  if (form == "served") cfg_form = 5;
  else if (form != "batched") { std::fprintf(stderr, "bw_tune: --form batched|served\n"); return 2; }
  TreeSpec spec = tree_defaults(cfg_form == 4 ? 4 : 2);
  for (int i = 1; i < argc; ++i) {
    if (i + 1 >= argc) { std::fprintf(stderr, "bw_tune: missing value for %s\n", argv[i]); return 2; }
    const char* k = argv[i];
    const char* v = argv[++i];
    if (!std::strcmp(k, "--form")) continue;
    else if (!std::strcmp(k, "--S")) spec.S = std::atoll(v);
    else if (!std::strcmp(k, "--U")) spec.U = std::atoi(v);
    else if (!std::strcmp(k, "--N")) spec.N = std::strtoull(v, nullptr, 10);
    else if (!std::strcmp(k, "--seed")) spec.seed = std::strtoull(v, nullptr, 10);
    else if (!std::strcmp(k, "--sample-requests")) sample = std::atoll(v);
    else if (!std::strcmp(k, "--repeat")) repeat = std::atoi(v);
    else if (!std::strcmp(k, "--start")) start = v;
    else if (!std::strcmp(k, "--out")) out = v;
    else { std::fprintf(stderr, "bw_tune: unknown argument %s\n", k); return 2; }
  }
  if (sample < 0) sample = cfg_form == 4 ? stream_defaults(4).nreq : 400;
  if (out.empty() || sample < 1 || repeat < 1) { std::fprintf(stderr, "bw_tune: --out FILE required\n"); return 2; }
  TreeWorkload w;
  tree_generate(spec, w);
  StreamSpec ss = stream_defaults(cfg_form);
  ss.nreq = sample;
  RequestStream rs;
  stream_generate(w, ss, rs);
  // the sample's queries over states of the generated tree (appended states are not part of the
  // resident index: the served form answers them through its overlay)
  std::vector<int32_t> qs;
  std::vector<uint32_t> lo, hi;
  for (int64_t q = 0; q < rs.nq(); ++q)
    if (rs.qstate[size_t(q)] <= spec.S) {
      qs.push_back(rs.qstate[size_t(q)]);
      lo.push_back(rs.qlo[size_t(q)]);
      hi.push_back(rs.qhi[size_t(q)]);
    }
  if (hipSetDevice(0) != hipSuccess) return 5;
  std::vector<TuneTrial> log;
  PlanKnobs k0;
  LaunchKnobs l0;
  if (!start.empty() && !load_tuned(start.c_str(), form.c_str(), k0, l0)) {
    std::fprintf(stderr, "bw_tune: no %s entry in %s\n", form.c_str(), start.c_str());
    return 2;
  }
  const TuneTrial best = autotune(w, w.key.data(), int64_t(qs.size()), qs.data(), lo.data(), hi.data(), k0, l0, repeat,
                                  log);
  for (const TuneTrial& t : log)
    std::fprintf(stderr,
                 "{\"rmax\":%d,\"tile_shift\":%d,\"chunk_items\":%" PRId64 ",\"ep_per_work\":%" PRId64
                 ",\"range_slots\":%" PRId64 ",\"block_threads\":%d,\"chunk_waves\":%d,\"ms\":%.4f}\n",
                 t.kn.rmax, t.kn.tile_shift, t.kn.chunk_items, t.kn.ep_per_work, t.kn.range_slots, t.lk.block_threads,
                 t.lk.chunk_waves, t.ms);
  if (best.ms <= 0) { std::fprintf(stderr, "bw_tune: no configuration ran\n"); return 5; }
  hipDeviceProp_t prop{};
  (void)hipGetDeviceProperties(&prop, 0);
  if (!save_tuned(out.c_str(), form.c_str(), best.kn, best.lk, prop.gcnArchName, best.ms)) return 4;
  std::printf("{\"form\":\"%s\",\"sample_queries\":%zu,\"trials\":%zu,\"rmax\":%d,\"tile_shift\":%d,\"chunk_items\":%" PRId64
              ",\"ep_per_work\":%" PRId64 ",\"range_slots\":%" PRId64 ",\"block_threads\":%d,\"chunk_waves\":%d,"
              "\"ms\":%.4f}\n",
              form.c_str(), qs.size(), log.size(), best.kn.rmax, best.kn.tile_shift, best.kn.chunk_items,
              best.kn.ep_per_work, best.kn.range_slots, best.lk.block_threads, best.lk.chunk_waves, best.ms);
  return 0;
}
