// tree_engine.cpp -- see tree_engine.hpp.
#include "tree_engine.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>

namespace bwg {

void DeviceArena::release() {
  for (void* p : ptrs_) (void)hipFree(p);
  ptrs_.clear();
  bytes = 0;
}

void* DeviceArena::alloc(size_t n) {
  if (err != hipSuccess) return nullptr;
  void* d = nullptr;
  const hipError_t e = hipMalloc(&d, n);
  if (e != hipSuccess) {
    err = e;
    return nullptr;
  }
  ptrs_.push_back(d);
  bytes += n;
  return d;
}

void DeviceArena::copy_in(void* d, const void* s, size_t n) {
  if (err != hipSuccess) return;
  const hipError_t e = hipMemcpy(d, s, n, hipMemcpyHostToDevice);
  if (e != hipSuccess) err = e;
}

bool TreeEngine::init(int64_t S, int U, uint64_t NK, const int32_t* parent, const uint32_t* key,
                      const int64_t* delta, const PlanKnobs& knobs, const BwtLaunch& l, char* err, size_t errlen) {
  if (!build_tree_index(S, U, NK, parent, key, knobs, plan, err, errlen)) return false;
  launch = l;
  const TreeIndexPlan& P = plan;
  x = BwtIndex{};
  x.S = S;
  x.U = U;
  x.nblk = P.nblk();
  x.rmax = P.knobs.rmax;
  x.max_cls = P.max_cls;
  x.nchain = P.nchain();
  x.ncls_c = P.chain_ncls[0];
  x.ncls_b = P.chain_ncls.size() > 1 ? P.chain_ncls[1] : 1;
  x.tile_shift = P.knobs.tile_shift;
  x.nchunk = P.nchunk();
  x.nchunk_c = P.chain_chunk0[1];
  x.nseg = P.nseg();
  x.delta = arena.upload(delta, size_t(S) * size_t(U));
  x.steps = arena.upload(P.steps);
  x.blk_step0 = arena.upload(P.blk_step0);
  x.evrank = arena.upload(P.evrank);
  x.blk_key0 = arena.upload(P.blk_key0);
  x.bkey = arena.upload(P.bkey);
  x.blk_tile0 = arena.upload(P.blk_tile0);
  x.tile_rank0 = arena.upload(P.tile_rank0);
  x.kitems = arena.upload(P.kitems);
  x.chunk_key0 = arena.upload(P.chunk_key0);
  x.chunk_chain = arena.upload(P.chunk_chain);
  x.chunk_cp0 = arena.upload(P.chunk_cp0);
  x.chain_ncls = arena.upload(P.chain_ncls);
  x.chain_src = arena.upload(P.chain_src);
  x.chain_out = arena.upload(P.chain_out);
  x.chain_chunk0 = arena.upload(P.chain_chunk0);
  x.seg_chunk0 = arena.upload(P.seg_chunk0);
  x.seg_cp0 = arena.upload(P.seg_cp0);
  x.chain_seg0 = arena.upload(P.chain_seg0);
  x.agg = arena.scratch<int64_t>(P.bkey.size());
  x.tagg = arena.scratch<int64_t>(P.tile_rank0.size());
  x.cp = arena.scratch<int64_t>(size_t(P.n_cp));
  x.segsum = arena.scratch<int64_t>(size_t(P.n_segsum));
  x.kweight = arena.scratch<int64_t>(P.kitems.size());
  // the device planner's tables
  px = BwtPlanIndex{};
  px.nsteps = int64_t(P.steps.size());
  px.NK = P.NK;
  px.KT = P.KT;
  px.stage_shift = P.knobs.stage_shift;
  px.chunk_items = P.knobs.chunk_items;
  px.ep_per_work = P.knobs.ep_per_work;
  px.range_slots = P.knobs.range_slots;
  px.state_step = arena.upload(P.state_step);
  {
    std::vector<int32_t> sb(P.steps.size());
    for (int32_t b = 0; b < P.nblk(); ++b)
      for (int32_t s = P.blk_step0[size_t(b)]; s < P.blk_step0[size_t(b) + 1]; ++s) sb[size_t(s)] = b;
    px.step_block = arena.upload(sb);
  }
  px.tile_tb0 = arena.upload(P.tile_tb0);
  px.tb_block = arena.upload(P.tb_block);
  px.tb_pos = arena.upload(P.tb_pos);
  px.chain_key0 = arena.upload(P.chain_key0);
  // the endpoint positions tabulated per (block, tile), when the table fits its budget
  px.bt_pos = nullptr;
  const size_t nbt = size_t(P.nblk()) * size_t(P.KT);
  if (nbt * sizeof(int32_t) <= size_t(std::max<int64_t>(0, P.knobs.bt_table_mb)) << 20) {
    std::vector<int32_t> bt(nbt);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < int64_t(P.KT); ++t) {
      int64_t i = P.tile_tb0[size_t(t)];
      const int64_t e = P.tile_tb0[size_t(t) + 1];
      const int32_t end = int32_t(P.chain_key0[size_t(t) + 2] - P.chain_key0[size_t(t) + 1]);
      for (int32_t b = 0; b < P.nblk(); ++b) {
        while (i < e && P.tb_block[size_t(i)] < b) ++i;
        bt[size_t(b) * size_t(P.KT) + size_t(t)] = i < e ? int32_t(P.tb_pos[size_t(i)]) : end;
      }
    }
    px.bt_pos = arena.upload(bt);
  }
  // the index-build batch: every block, no queries
  std::vector<int32_t> allb(size_t(P.nblk()));
  for (int32_t b = 0; b < P.nblk(); ++b) allb[size_t(b)] = b;
  std::vector<int64_t> zeros(size_t(P.nblk()) + 1, 0);
  all_blocks = BwtBatch{};
  all_blocks.nqblk = P.nblk();
  all_blocks.qblk = arena.upload(allb);
  all_blocks.qblk_g0 = arena.upload(zeros);
  all_blocks.grp_q0 = all_blocks.qblk_g0;  // one entry (0) is read
  if (!arena.ok()) {
    if (err && errlen) std::snprintf(err, errlen, "device upload: %s", hipGetErrorString(arena.err));
    return false;
  }
  return true;
}

bool TreeEngine::init_device_index(int64_t S, int U, uint64_t NK, const int32_t* parent, const uint32_t* key,
                                   const uint32_t* d_key, const int64_t* delta, const PlanKnobs& knobs,
                                   const BwtLaunch& l, hipStream_t st, char* err, size_t errlen, bool device_cut,
                                   bool device_walk, const int64_t* d_delta) {
  auto fail = [&](const char* what, int rc) {
    if (err && errlen) std::snprintf(err, errlen, "device index: %s (%d)", what, rc);
    return false;
  };
  const auto t0 = std::chrono::steady_clock::now();
  device_cut = device_cut || device_walk || !key;  // without host keys nothing can be cut on the host
  if (!key && !d_key) return fail("no keys", 0);
  if (!plan_walk_blocks(S, U, NK, parent, key, knobs, plan, err, errlen, !device_cut, !device_walk)) return false;
  const auto t1 = std::chrono::steady_clock::now();
  launch = l;
  TreeIndexPlan& P = plan;
  DeviceArena tmp;  // the build's workspace and counts, freed at the end
  struct Events {  // destroyed on every return
    hipEvent_t a = nullptr, b = nullptr;
    ~Events() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
  } ev;
  if (hipEventCreate(&ev.a) != hipSuccess || hipEventCreate(&ev.b) != hipSuccess) return fail("events", 0);
  hipEvent_t e0 = ev.a, e1 = ev.b;
  (void)hipEventRecord(e0, st);
  int rc;
  x = BwtIndex{};
  if (device_walk) {  // the Euler walk on the device, then the walk to the host
    size_t wwb = 0;
    if ((rc = bwt_index_walk_workspace(S, &wwb))) return fail("walk workspace", rc);
    void* wwork = tmp.alloc(wwb);
    const int32_t* d_parent = tmp.upload(parent, size_t(S) + 1);
    int32_t* d_steps = arena.scratch<int32_t>(2 * size_t(S));
    int64_t* d_ns = tmp.scratch<int64_t>(1);
    if (!arena.ok() || !tmp.ok()) return fail("allocation", int(arena.ok() ? tmp.err : arena.err));
    if ((rc = bwt_index_walk_launch(S, d_parent, d_steps, d_ns, wwork, wwb, st))) return fail("walk", rc);
    int64_t ns = 0;
    if (hipMemcpyAsync(&ns, d_ns, 8, hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
      return fail("walk length", 0);
    if (ns < 1 || ns > 2 * S || ns >= (int64_t(1) << 31) / U) return fail("walk length", int(ns));
    P.steps.assign(size_t(ns), 0);
    if (hipMemcpyAsync(P.steps.data(), d_steps, size_t(ns) * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail("walk download", 0);
    x.steps = d_steps;
  } else {
    x.steps = arena.upload(P.steps);
  }
  const int64_t nsteps = int64_t(P.steps.size());
  BwtIndexBuild in{};
  in.S = S;
  in.U = U;
  in.rmax = P.knobs.rmax;
  in.tile_shift = P.knobs.tile_shift;
  in.KT = P.KT;
  in.nsteps = nsteps;
  in.steps = x.steps;
  in.key = d_key ? d_key : tmp.upload(key, size_t(S) * size_t(U));
  if (device_cut) {  // the block cut on the device, then its block starts to the host
    size_t cb = 0;
    if ((rc = bwt_index_cut_workspace(&in, &cb))) return fail("cut workspace", rc);
    void* cwork = tmp.alloc(cb);
    int32_t* d_b0 = arena.scratch<int32_t>(size_t(nsteps) + 1);
    int32_t* d_nb = tmp.scratch<int32_t>(1);
    if (!arena.ok() || !tmp.ok()) return fail("allocation", int(arena.ok() ? tmp.err : arena.err));
    const int32_t mbs = int32_t(std::min<int64_t>(P.knobs.max_block_steps, int64_t(1) << 30));
    if ((rc = bwt_index_cut_launch(&in, mbs, d_b0, d_nb, cwork, cb, st))) return fail("cut", rc);
    int32_t nb = 0;
    if (hipMemcpyAsync(&nb, d_nb, 4, hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
      return fail("cut count", 0);
    if (nb < 1 || nb > nsteps) return fail("cut count", nb);
    P.blk_step0.assign(size_t(nb) + 1, 0);
    if (hipMemcpyAsync(P.blk_step0.data(), d_b0, P.blk_step0.size() * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail("cut download", 0);
    x.blk_step0 = d_b0;
  } else {
    x.blk_step0 = arena.upload(P.blk_step0);
  }
  const int32_t nblk = P.nblk();
  in.nblk = nblk;
  in.blk_step0 = x.blk_step0;
  size_t wb = 0;
  rc = bwt_index_workspace(&in, &wb);
  if (rc) return fail("workspace", rc);
  void* work = tmp.alloc(wb);
  int32_t* nkeys = tmp.scratch<int32_t>(size_t(nblk));
  int32_t* ntiles = tmp.scratch<int32_t>(size_t(nblk));
  BwtIndexTables T{};
  T.evrank = arena.scratch<uint16_t>(size_t(nsteps) * size_t(U));
  T.state_step = arena.scratch<int32_t>(size_t(S) + 1);
  T.step_block = arena.scratch<int32_t>(size_t(nsteps));
  if (!arena.ok() || !tmp.ok()) return fail("allocation", int(arena.ok() ? tmp.err : arena.err));
  if ((rc = bwt_index_keys_launch(&in, &T, nkeys, ntiles, work, wb, st))) return fail("stage 1", rc);
  // the two counts' exclusive scans (on the host: nblk entries each)
  const size_t nb = static_cast<size_t>(nblk);
  std::vector<int32_t> hk(nb), ht(nb);
  if (hipMemcpyAsync(hk.data(), nkeys, hk.size() * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(ht.data(), ntiles, ht.size() * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail("counts", 0);
  P.blk_key0.assign(size_t(nblk) + 1, 0);
  P.blk_tile0.assign(size_t(nblk) + 1, 0);
  for (int32_t b = 0; b < nblk; ++b) {
    if (hk[size_t(b)] < 1 || hk[size_t(b)] > P.knobs.rmax)
      return fail("a block exceeds rmax or holds a key outside the tiles", hk[size_t(b)]);
    P.blk_key0[size_t(b) + 1] = P.blk_key0[size_t(b)] + hk[size_t(b)];
    P.blk_tile0[size_t(b) + 1] = P.blk_tile0[size_t(b)] + ht[size_t(b)];
  }
  const int64_t nkeys_total = P.blk_key0[size_t(nblk)], nslots = P.blk_tile0[size_t(nblk)];
  x.blk_key0 = arena.upload(P.blk_key0);
  x.blk_tile0 = arena.upload(P.blk_tile0);
  T.blk_key0 = x.blk_key0;
  T.blk_tile0 = x.blk_tile0;
  T.bkey = arena.scratch<uint32_t>(size_t(nkeys_total));
  T.tile_rank0 = arena.scratch<uint16_t>(size_t(nslots));
  T.kitems = arena.scratch<uint64_t>(size_t(nslots + nkeys_total));
  T.chain_key0 = arena.scratch<int64_t>(size_t(P.KT) + 2);
  T.tile_tb0 = arena.scratch<int64_t>(size_t(P.KT) + 1);
  T.tb_block = arena.scratch<int32_t>(size_t(nslots));
  T.tb_pos = arena.scratch<int64_t>(size_t(nslots));
  const size_t nbt = size_t(nblk) * size_t(P.KT);
  T.bt_pos = nbt * sizeof(int32_t) <= size_t(std::max<int64_t>(0, P.knobs.bt_table_mb)) << 20
                 ? arena.scratch<int32_t>(nbt)
                 : nullptr;
  if (!arena.ok()) return fail("allocation", int(arena.err));
  if ((rc = bwt_index_chains_launch(&in, &T, nkeys_total, nslots, work, wb, st))) return fail("stage 2", rc);
  (void)hipEventRecord(e1, st);
  P.chain_key0.assign(size_t(P.KT) + 2, 0);
  if (hipMemcpyAsync(P.chain_key0.data(), T.chain_key0, P.chain_key0.size() * 8, hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail("chain offsets", 0);
  float gpu_ms = 0;
  (void)hipEventElapsedTime(&gpu_ms, e0, e1);
  plan_chunks(P);
  // the index as init() lays it out
  x.S = S;
  x.U = U;
  x.nblk = nblk;
  x.rmax = P.knobs.rmax;
  x.max_cls = P.max_cls;
  x.nchain = P.nchain();
  x.ncls_c = P.chain_ncls[0];
  x.ncls_b = P.chain_ncls.size() > 1 ? P.chain_ncls[1] : 1;
  x.tile_shift = P.knobs.tile_shift;
  x.nchunk = P.nchunk();
  x.nchunk_c = P.chain_chunk0[1];
  x.nseg = P.nseg();
  x.delta = d_delta ? d_delta : arena.upload(delta, size_t(S) * size_t(U));
  x.evrank = T.evrank;
  x.bkey = T.bkey;
  x.tile_rank0 = T.tile_rank0;
  x.kitems = T.kitems;
  x.chunk_key0 = arena.upload(P.chunk_key0);
  x.chunk_chain = arena.upload(P.chunk_chain);
  x.chunk_cp0 = arena.upload(P.chunk_cp0);
  x.chain_ncls = arena.upload(P.chain_ncls);
  x.chain_src = arena.upload(P.chain_src);
  x.chain_out = arena.upload(P.chain_out);
  x.chain_chunk0 = arena.upload(P.chain_chunk0);
  x.seg_chunk0 = arena.upload(P.seg_chunk0);
  x.seg_cp0 = arena.upload(P.seg_cp0);
  x.chain_seg0 = arena.upload(P.chain_seg0);
  x.agg = arena.scratch<int64_t>(size_t(nkeys_total));
  x.tagg = arena.scratch<int64_t>(size_t(nslots));
  x.cp = arena.scratch<int64_t>(size_t(P.n_cp));
  x.segsum = arena.scratch<int64_t>(size_t(P.n_segsum));
  x.kweight = arena.scratch<int64_t>(size_t(nslots + nkeys_total));
  px = BwtPlanIndex{};
  px.nsteps = nsteps;
  px.NK = P.NK;
  px.KT = P.KT;
  px.stage_shift = P.knobs.stage_shift;
  px.chunk_items = P.knobs.chunk_items;
  px.ep_per_work = P.knobs.ep_per_work;
  px.range_slots = P.knobs.range_slots;
  px.state_step = T.state_step;
  px.step_block = T.step_block;
  px.tile_tb0 = T.tile_tb0;
  px.tb_block = T.tb_block;
  px.tb_pos = T.tb_pos;
  px.chain_key0 = T.chain_key0;
  px.bt_pos = T.bt_pos;
  std::vector<int32_t> allb(nb);
  for (int32_t b = 0; b < nblk; ++b) allb[size_t(b)] = b;
  std::vector<int64_t> zeros(nb + 1, 0);
  all_blocks = BwtBatch{};
  all_blocks.nqblk = nblk;
  all_blocks.qblk = arena.upload(allb);
  all_blocks.qblk_g0 = arena.upload(zeros);
  all_blocks.grp_q0 = all_blocks.qblk_g0;
  if (!arena.ok()) return fail("upload", int(arena.err));
  if (hipStreamSynchronize(st) != hipSuccess) return fail("sync", 0);
  device_index.host_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  device_index.gpu_ms = gpu_ms;
  device_index.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  device_index.nkeys = nkeys_total;
  device_index.nslots = nslots;
  return true;
}

int TreeEngine::build_index(hipStream_t st) {
  BwtLaunch w = launch;
  w.write_index = 1;
  int rc;
  if ((rc = bwt_block_replay_launch(&x, &all_blocks, &w, st))) return rc;
  if ((rc = bwt_chunk_hist_launch(&x, &w, st))) return rc;
  return bwt_carry_scan_launch(&x, &w, st);
}

bool upload_batch(const TreeEngine& e, BatchPlan&& plan, const uint32_t* d_qlo, const uint32_t* d_qhi,
                  DeviceBatch& out, char* err, size_t errlen) {
  out.arena.release();
  out.arena.err = hipSuccess;
  out.plan = std::move(plan);
  const BatchPlan& B = out.plan;
  BwtBatch& b = out.b;
  b = BwtBatch{};
  b.M = B.M;
  b.qlo = d_qlo;
  b.qhi = d_qhi;
  b.nqblk = int32_t(B.qblk.size());
  b.qblk = out.arena.upload(B.qblk);
  b.qblk_g0 = out.arena.upload(B.qblk_g0);
  b.ngrp = int64_t(B.grp_step.size());
  b.grp_step = out.arena.upload(B.grp_step);
  b.grp_q0 = out.arena.upload(B.grp_q0);
  b.qorder = out.arena.upload(B.qorder);
  b.qrank = out.arena.upload(B.qrank);
  b.sblk = out.arena.upload(B.sblk);
  b.qtile = out.arena.upload(B.qtile);
  b.nwork = int64_t(B.wk_chunk.size());
  b.wk_chunk = out.arena.upload(B.wk_chunk);
  b.wk_e0 = out.arena.upload(B.wk_e0);
  b.eitems = out.arena.upload(B.eitems);
  b.epos = out.arena.upload(B.epos);
  b.spos = out.arena.upload(B.spos);
  b.srank = out.arena.upload(B.srank);
  b.ipos = out.arena.upload(B.ipos);
  b.ansr = out.arena.scratch<int64_t>(size_t(B.M));
  b.g = out.arena.scratch<int64_t>(2 * size_t(B.M));
  b.stage = out.arena.scratch<int64_t>(size_t(B.M));
  b.answers = out.arena.scratch<int64_t>(size_t(B.M));
  (void)e;
  if (!out.arena.ok()) {
    if (err && errlen) std::snprintf(err, errlen, "batch upload: %s", hipGetErrorString(out.arena.err));
    return false;
  }
  return true;
}

int query_forked(const TreeEngine& e, const BwtBatch* b, hipStream_t st, hipStream_t side, hipEvent_t fork,
                 hipEvent_t join) {
  BwtLaunch l = e.launch;  // as bwt_query_launch sets it
  l.write_index = 0;
  l.carry_in_block = 1;
  hipError_t he;
  int rc;
  if ((he = hipEventRecord(fork, st)) != hipSuccess || (he = hipStreamWaitEvent(side, fork, 0)) != hipSuccess)
    return int(he);
  if ((rc = bwt_chunk_replay_launch(&e.x, b, &l, side))) return rc;
  if ((he = hipEventRecord(join, side)) != hipSuccess) return int(he);
  if ((rc = bwt_block_replay_launch(&e.x, b, &l, st))) return rc;
  if ((he = hipStreamWaitEvent(st, join, 0)) != hipSuccess) return int(he);
  return bwt_combine_launch(&e.x, b, &l, st);
}

bool alloc_plan_slot(const TreeEngine& e, int64_t cap, DevicePlanSlot& s, char* err, size_t errlen) {
  s.arena.release();
  s.arena.err = hipSuccess;
  s.cap = cap;
  s.b = BwtBatch{};
  const size_t n = size_t(cap);
  const size_t nq = size_t(e.x.nblk);  // every block can be listed (bwt_plan_all_blocks_launch)
  BwtBatch& b = s.b;
  b.qblk = s.arena.scratch<int32_t>(nq);
  b.qblk_g0 = s.arena.scratch<int64_t>(nq + 1);
  b.grp_step = s.arena.scratch<int32_t>(n);
  b.grp_q0 = s.arena.scratch<int64_t>(n + 1);
  b.qorder = s.arena.scratch<uint32_t>(n);
  b.qrank = s.arena.scratch<uint32_t>(n);
  b.sblk = s.arena.scratch<int32_t>(n);
  b.qtile = s.arena.scratch<uint32_t>(n);
  b.wk_chunk = s.arena.scratch<int64_t>(2 * n);
  b.wk_e0 = s.arena.scratch<int64_t>(2 * n + 1);
  b.eitems = s.arena.scratch<uint64_t>(2 * n);
  b.epos = s.arena.scratch<uint32_t>(2 * n);
  b.spos = s.arena.scratch<uint32_t>(n);
  b.srank = s.arena.scratch<uint16_t>(n);
  b.ipos = s.arena.scratch<uint32_t>(n);
  b.ansr = s.arena.scratch<int64_t>(n);
  b.g = s.arena.scratch<int64_t>(2 * n);
  b.stage = s.arena.scratch<int64_t>(n);
  b.answers = s.arena.scratch<int64_t>(n);
  s.dcounts = s.arena.scratch<int64_t>(8);
  int rc = bwt_plan_workspace(&e.x, &e.px, cap, &s.work_bytes);
  if (rc) {
    if (err && errlen) std::snprintf(err, errlen, "plan workspace: error %d", rc);
    return false;
  }
  s.work = s.arena.alloc(s.work_bytes);
  if (!s.arena.ok()) {
    if (err && errlen) std::snprintf(err, errlen, "plan slot: %s", hipGetErrorString(s.arena.err));
    return false;
  }
  return true;
}

int device_plan(const TreeEngine& e, DevicePlanSlot& s, int64_t M, const int32_t* d_qstate, const uint32_t* d_qlo,
                const uint32_t* d_qhi, hipStream_t st) {
  if (M < 1 || M > s.cap) return int(hipErrorInvalidValue);
  s.b.qlo = d_qlo;
  s.b.qhi = d_qhi;
  // the workspace was sized for cap queries; a smaller batch needs no more
  return bwt_plan_launch(&e.x, &e.px, d_qstate, M, &s.b, s.dcounts, s.work, s.work_bytes, st);
}

}  // namespace bwg
