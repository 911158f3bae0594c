/* bwgpu_plan.h -- C-ABI of the device batch planner: builds a BwtBatch (include/bwgpu_tree.h)
 * from a batch of queries (state, lo, hi) on the GPU, with no host round trip.
 *
 * Counterpart in the reference: the host-side per-query planning of BatchEngineV7::build --
 * endpoint class bounds qlb/qhb per query (src/batch_engine_v7.hpp:666-698) and the radix-sort
 * driven plan construction (src/batch_engine_v7.hpp:1069-1348) -- which the reference GPU
 * drivers run on the CPU before every launch (src/panel_batch_gpu.cpp:86-121).  Here the same
 * role (turning a query batch into the launch plan of the replay kernels) runs as device sorts,
 * scans and searches over the resident index.
 *
 * The result is the plan the host planner (build_batch_plan) makes for the same queries, array
 * for array: slots ordered by (enter step, query id); step groups and queried blocks; in-block
 * ranks and key tiles; level-B endpoint items ordered by (slot range, chunk, position, endpoint)
 * cut into work items at every ep_per_work-th item of a (range, chunk) run (moved one item on
 * where it would part a slot's two endpoints); staging positions grouped by (query-id window,
 * die part, slot).  Counts are known on the device only: they are written to dcounts, and the
 * replay kernels read them from there (BwtBatch.dcounts).  No call allocates or synchronises;
 * the caller provides every output array (capacities below) and one workspace. */
#ifndef BWGPU_PLAN_H
#define BWGPU_PLAN_H
#include <stddef.h>
#include <stdint.h>

#include "bwgpu_tree.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Index-side tables the planner reads (device pointers; built once with the index). */
typedef struct BwtPlanIndex {
  int64_t nsteps;              /* steps of the preorder walk */
  uint64_t NK;                 /* endpoint coordinates are in [0, NK] */
  int32_t KT;                  /* key tiles (level-C classes) */
  int32_t stage_shift;         /* query ids per staging window = 2^stage_shift */
  int64_t chunk_items;         /* level-B chunk length (key items) */
  int64_t ep_per_work;         /* endpoint items per work item (>= 64) */
  int64_t range_slots;         /* slots per range of the work order (>= 1) */
  const int32_t* state_step;   /* S+1: enter step of each state */
  const int32_t* step_block;   /* nsteps: block of each step */
  const int64_t* tile_tb0;     /* KT+1: offsets into tb_block / tb_pos per key tile */
  const int32_t* tb_block;     /* blocks holding keys of the tile, ascending */
  const int64_t* tb_pos;       /* where those keys start on the tile's level-B chain */
  const int64_t* chain_key0;   /* KT+2: key-item offsets of the chains (chain 1+t = tile t) */
  const int32_t* bt_pos;       /* NULL, or nblk*KT: [b*KT+t] = position on chain 1+t before which
                                  lie the tile's keys of blocks < b (the search above, tabulated) */
} BwtPlanIndex;

/* Workspace bytes for a batch of M queries (M >= 1). */
int bwt_plan_workspace(const BwtIndex* x, const BwtPlanIndex* p, int64_t M, size_t* bytes);

/* Plan M queries (1 <= M <= BWT_MAX_BATCH): qstate[M] in [1, S], b->qlo / b->qhi[M] in [0, NK].
 * Output arrays of b, with capacities: qblk min(nblk, M); qblk_g0 min(nblk, M)+1; grp_step M;
 * grp_q0 M+1; qorder, qrank, sblk, qtile, spos, srank, ipos M; wk_chunk 2M; wk_e0 2M+1;
 * eitems, epos 2M (the planner writes through these pointers).  Sets b->M, b->dcounts = dcounts and
 * b->nqblk / b->ngrp / b->nwork to those capacities (upper bounds).  dcounts[5]: nqblk, ngrp,
 * nwork, npairs, and an error flag (nonzero: some query was out of range; it was planned as a
 * query of state 1 over [0, 0), and the batch's answers are not meaningful). */
int bwt_plan_launch(const BwtIndex* x, const BwtPlanIndex* p, const int32_t* qstate, int64_t M, BwtBatch* b,
                    int64_t* dcounts, void* work, size_t work_bytes, void* stream);

/* After bwt_plan_launch on the same stream: list every block of the index in b (qblk[i] = i, with
 * its step groups), as a pass that also writes the index needs (bwt_pass_launch); sets b->nqblk =
 * nblk and dcounts[0].  Capacities: qblk nblk, qblk_g0 nblk+1.  Counterpart: the host planner's
 * all_blocks option (build_batch_plan); the reference's drivers always launch every block
 * (src/panel_batch_gpu.cpp:172-178). */
int bwt_plan_all_blocks_launch(const BwtIndex* x, BwtBatch* b, void* stream);

#ifdef __cplusplus
}
#endif
#endif
