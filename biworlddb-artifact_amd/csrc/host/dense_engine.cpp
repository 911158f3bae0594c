// dense_engine.cpp -- see dense_engine.hpp.
#include "dense_engine.hpp"

#include <cstdio>

namespace bwg {

bool DenseEngine::init(const TreeEngine& e, hipStream_t st, char* err, size_t errlen) {
  const TreeIndexPlan& P = e.plan;
  if (e.x.S > BWD_MAX_STATES || P.NK > BWD_MAX_NK || e.x.rmax > BWD_MAX_RMAX) {
    std::snprintf(err, errlen, "dense form limits: S <= %d, NK <= %u, rmax <= %d", BWD_MAX_STATES, BWD_MAX_NK,
                  BWD_MAX_RMAX);
    return false;
  }
  std::vector<int32_t> erank, estep, be0;
  if (!dense_rank_tables(e.x.S, P.steps, P.blk_step0, erank, estep, be0)) {
    std::snprintf(err, errlen, "dense form: the walk does not enter every state exactly once");
    return false;
  }
  x = BwdIndex{};
  x.S = e.x.S;
  x.U = e.x.U;
  x.nblk = e.x.nblk;
  x.rmax = e.x.rmax;
  x.NK = P.NK;
  x.nsteps = int64_t(P.steps.size());
  x.carry_ld = ((int64_t(P.NK) + 1 + 63) / 64) * 64;
  x.delta = e.x.delta;
  x.steps = e.x.steps;
  x.blk_step0 = e.x.blk_step0;
  x.evrank = e.x.evrank;
  x.blk_key0 = e.x.blk_key0;
  x.bkey = e.x.bkey;
  x.blk_e0 = arena.upload(be0);
  x.state_erank = arena.upload(erank);
  x.erank_step = arena.upload(estep);
  carry_bytes = size_t(x.nblk) * size_t(x.carry_ld) * sizeof(int64_t);
  x.carry = static_cast<int64_t*>(arena.alloc(carry_bytes));
  if (!arena.ok()) {
    std::snprintf(err, errlen, "device: dense carry allocation (%zu bytes): %s", carry_bytes,
                  hipGetErrorString(arena.err));
    return false;
  }
  size_t wb = 0;
  int rc = bwd_carry_workspace(&x, &wb);
  DeviceArena work;
  void* w = rc ? nullptr : work.alloc(wb);
  if (rc || !work.ok()) {
    std::snprintf(err, errlen, "device: dense carry workspace (%d)", rc);
    return false;
  }
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    std::snprintf(err, errlen, "device: events");
    return false;
  }
  (void)hipEventRecord(a, st);
  rc = bwd_carry_build_launch(&x, e.x.agg, w, wb, st);
  (void)hipEventRecord(b, st);
  const hipError_t se = hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc || se != hipSuccess) {
    std::snprintf(err, errlen, "device: dense carry build (%d, %s)", rc, hipGetErrorString(se));
    return false;
  }
  build_ms = ms;
  return true;
}

bool alloc_dense_batch(const DenseEngine& d, int64_t cap, DenseBatch& b, char* err, size_t errlen) {
  size_t wb = 0;
  const int rc = bwd_answer_workspace(&d.x, cap, &wb);
  if (rc) {
    std::snprintf(err, errlen, "dense batch workspace (%d)", rc);
    return false;
  }
  b.cap = cap;
  b.work_bytes = wb;
  b.work = b.arena.alloc(wb);
  b.flags = b.arena.scratch<int64_t>(1);
  if (!b.arena.ok()) {
    std::snprintf(err, errlen, "device: dense batch allocation: %s", hipGetErrorString(b.arena.err));
    return false;
  }
  return true;
}

}  // namespace bwg
