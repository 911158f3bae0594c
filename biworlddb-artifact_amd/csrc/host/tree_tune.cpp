// tree_tune.cpp -- see tree_tune.hpp.
#include "tree_tune.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "bwgpu_tree.h"
#include "tree_engine.hpp"

namespace bwg {
namespace {

std::string read_file(const char* path) {
  std::string s;
  FILE* f = std::fopen(path, "rb");
  if (!f) return s;
  char buf[4096];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) s.append(buf, n);
  std::fclose(f);
  return s;
}

// The object text of key `form` at the top level of a flat two-level JSON file ("" if absent).
std::string entry_of(const std::string& text, const char* form) {
  const std::string k = std::string("\"") + form + "\"";
  const size_t at = text.find(k);
  if (at == std::string::npos) return "";
  const size_t b = text.find('{', at);
  const size_t e = text.find('}', b);
  if (b == std::string::npos || e == std::string::npos) return "";
  return text.substr(b, e - b + 1);
}

bool number_of(const std::string& obj, const char* key, double& out) {
  const std::string k = std::string("\"") + key + "\"";
  const size_t at = obj.find(k);
  if (at == std::string::npos) return false;
  const size_t c = obj.find(':', at);
  if (c == std::string::npos) return false;
  char* end = nullptr;
  out = std::strtod(obj.c_str() + c + 1, &end);
  return end != obj.c_str() + c + 1;
}

std::string entry_text(const PlanKnobs& kn, const LaunchKnobs& lk, const char* device, double ms) {
  char b[512];
  std::snprintf(b, sizeof(b),
                "{\"rmax\": %d, \"tile_shift\": %d, \"chunk_items\": %lld, \"ep_per_work\": %lld, "
                "\"range_slots\": %lld, \"block_threads\": %d, \"chunk_waves\": %d, \"ms\": %.4f, "
                "\"device\": \"%s\"}",
                kn.rmax, kn.tile_shift, (long long)kn.chunk_items, (long long)kn.ep_per_work,
                (long long)kn.range_slots, lk.block_threads, lk.chunk_waves, ms, device ? device : "");
  return b;
}

struct Timer {
  hipStream_t st = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  Timer() {
    (void)hipStreamCreate(&st);
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
  }
  ~Timer() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipStreamDestroy(st);
  }
};

}  // namespace

bool load_tuned(const char* path, const char* form, PlanKnobs& kn, LaunchKnobs& lk) {
  const std::string obj = entry_of(read_file(path), form);
  if (obj.empty()) return false;
  double v;
  PlanKnobs k = kn;
  LaunchKnobs l = lk;
  if (number_of(obj, "rmax", v)) k.rmax = int(v);
  if (number_of(obj, "tile_shift", v)) k.tile_shift = int(v);
  if (number_of(obj, "chunk_items", v)) k.chunk_items = int64_t(v);
  if (number_of(obj, "ep_per_work", v)) k.ep_per_work = int64_t(v);
  if (number_of(obj, "range_slots", v)) k.range_slots = int64_t(v);
  if (number_of(obj, "block_threads", v)) l.block_threads = int(v);
  if (number_of(obj, "chunk_waves", v)) l.chunk_waves = int(v);
  kn = k;
  lk = l;
  return true;
}

namespace {

// Rewrites the file with `form`'s entry replaced by `entry`, keeping the other forms' entries.
bool save_entry(const char* path, const char* form, const std::string& entry) {
  const std::string text = read_file(path);
  const char* forms[] = {"batched", "served", "plain", "plain3", "dense"};
  std::string out = "{\n";
  bool first = true;
  for (const char* f : forms) {
    std::string e = std::strcmp(f, form) == 0 ? entry : entry_of(text, f);
    if (e.empty()) continue;
    out += (first ? "  \"" : ",\n  \"") + std::string(f) + "\": " + e;
    first = false;
  }
  out += "\n}\n";
  FILE* fp = std::fopen(path, "wb");
  if (!fp) return false;
  const bool ok = std::fwrite(out.data(), 1, out.size(), fp) == out.size();
  return std::fclose(fp) == 0 && ok;
}

}  // namespace

bool save_tuned(const char* path, const char* form, const PlanKnobs& kn, const LaunchKnobs& lk, const char* device,
                double ms) {
  return save_entry(path, form, entry_text(kn, lk, device, ms));
}

bool load_dense_tuned(const char* path, DenseKnobs& dk) {
  const std::string obj = entry_of(read_file(path), "dense");
  if (obj.empty()) return false;
  double v;
  DenseKnobs k = dk;
  if (number_of(obj, "rmax", v)) k.rmax = int(v);
  if (number_of(obj, "threads", v)) k.threads = int(v);
  if (number_of(obj, "wg_per_cu", v)) k.wg_per_cu = int(v);
  if (number_of(obj, "store", v)) k.store = int(v);
  if (number_of(obj, "chunk", v)) k.chunk = int(v);
  dk = k;
  return true;
}

bool save_dense_tuned(const char* path, const DenseKnobs& dk, const char* device, double ms) {
  char b[256];
  std::snprintf(b, sizeof(b),
                "{\"rmax\": %d, \"threads\": %d, \"wg_per_cu\": %d, \"store\": %d, \"chunk\": %d, \"ms\": %.4f, "
                "\"device\": \"%s\"}",
                dk.rmax, dk.threads, dk.wg_per_cu, dk.store, dk.chunk, ms, device ? device : "");
  return save_entry(path, "dense", b);
}

TuneTrial autotune(const TreeWorkload& w, const uint32_t* kcoord, int64_t M, const int32_t* qstate,
                   const uint32_t* qlo, const uint32_t* qhi, PlanKnobs kn, LaunchKnobs lk, int repeat,
                   std::vector<TuneTrial>& log) {
  Timer tm;
  DeviceArena io;
  uint32_t* d_qlo = io.upload(qlo, size_t(M));
  uint32_t* d_qhi = io.upload(qhi, size_t(M));
  std::unique_ptr<TreeEngine> eng;
  std::unique_ptr<DeviceBatch> db;
  PlanKnobs eng_kn{}, db_kn{};
  bool have_eng = false, have_db = false;
  char err[256];
  auto trial = [&](const PlanKnobs& k, const LaunchKnobs& l) -> double {
    const bool same_index = have_eng && k.rmax == eng_kn.rmax && k.tile_shift == eng_kn.tile_shift &&
                            k.chunk_items == eng_kn.chunk_items;
    if (!same_index) {
      have_db = false;
      db.reset();
      eng.reset(new TreeEngine());
      have_eng = eng->init(w.spec.S, w.spec.U, w.coord_n(), w.parent.data(), kcoord, w.delta.data(), k,
                           BwtLaunch{l.block_threads, l.chunk_waves, 0, 0}, err, sizeof(err));
      if (!have_eng) return -1;
      eng_kn = k;
      if (eng->build_index(tm.st) != 0) return -1;
    }
    if (!have_db || k.ep_per_work != db_kn.ep_per_work || k.range_slots != db_kn.range_slots) {
      TreeIndexPlan& P = eng->plan;
      const PlanKnobs saved = P.knobs;
      P.knobs.ep_per_work = k.ep_per_work;
      P.knobs.range_slots = k.range_slots;
      BatchPlan bp;
      const bool ok = build_batch_plan(P, M, qstate, qlo, qhi, false, bp, err, sizeof(err));
      P.knobs = saved;
      if (!ok) return -1;
      db.reset(new DeviceBatch());
      if (!upload_batch(*eng, std::move(bp), d_qlo, d_qhi, *db, err, sizeof(err))) return -1;
      have_db = true;
      db_kn = k;
    }
    BwtLaunch L{l.block_threads, l.chunk_waves, 0, 0};
    if (bwt_query_launch(&eng->x, &db->b, &L, tm.st) != 0) return -1;  // warm; rejects bad shapes
    double best = 1e30;
    for (int r = 0; r < repeat; ++r) {
      (void)hipEventRecord(tm.a, tm.st);
      if (bwt_query_launch(&eng->x, &db->b, &L, tm.st) != 0) return -1;
      (void)hipEventRecord(tm.b, tm.st);
      if (hipEventSynchronize(tm.b) != hipSuccess) return -1;
      float t = 0;
      (void)hipEventElapsedTime(&t, tm.a, tm.b);
      best = std::min(best, double(t));
    }
    return best;
  };
  TuneTrial cur{kn, lk, trial(kn, lk)};
  if (have_eng) {  // automatic knobs (0) resolved to the values the planner chose
    cur.kn.tile_shift = eng->plan.knobs.tile_shift;
    cur.kn.chunk_items = eng->plan.knobs.chunk_items;
    eng_kn = cur.kn;
    db_kn = cur.kn;
  }
  log.push_back(cur);
  // candidate values per knob (the current value is always among the tried ones)
  const int rmax_c[] = {1024, 2048, 4096, 0};
  const int ts_c[] = {9, 10, 11, 0};
  const int64_t chunk_c[] = {512, 1024, 2048, 4096};
  const int64_t epw_c[] = {2048, 8192, 32768, 0};
  const int64_t rs_c[] = {int64_t(1) << 20, int64_t(1) << 21, int64_t(1) << 22, int64_t(1) << 24};
  const int thr_c[] = {128, 256, 512, 0};
  const int wav_c[] = {1, 2, 4, 0};
  for (int knob = 0; knob < 7; ++knob) {
    for (int i = 0; i < 4; ++i) {
      TuneTrial t = cur;
      bool none = false;
      switch (knob) {
        case 0: none = !rmax_c[i]; t.kn.rmax = rmax_c[i]; break;
        case 1: none = !ts_c[i]; t.kn.tile_shift = ts_c[i]; break;
        case 2: t.kn.chunk_items = chunk_c[i]; break;
        case 3: none = !epw_c[i]; t.kn.ep_per_work = epw_c[i]; break;
        case 4: t.kn.range_slots = rs_c[i]; break;
        case 5: none = !thr_c[i]; t.lk.block_threads = thr_c[i]; break;
        default: none = !wav_c[i]; t.lk.chunk_waves = wav_c[i]; break;
      }
      if (none) continue;
      if (std::memcmp(&t.kn, &cur.kn, sizeof(PlanKnobs)) == 0 && t.lk.block_threads == cur.lk.block_threads &&
          t.lk.chunk_waves == cur.lk.chunk_waves)
        continue;
      t.ms = trial(t.kn, t.lk);
      log.push_back(t);
      if (t.ms > 0 && (cur.ms < 0 || t.ms < cur.ms)) cur = t;
    }
    // the engine must reflect the current best before the next knob is varied
    (void)trial(cur.kn, cur.lk);
  }
  return cur;
}

}  // namespace bwg
