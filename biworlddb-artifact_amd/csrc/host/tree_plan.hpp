// tree_plan.hpp -- host plans of the generalised batch path (the BatchEngineV7::build analogue,
// src/batch_engine_v7.hpp:576-704 and build_v7_blocks :707-830, for branching multi-update
// trees).
//   TreeIndexPlan  static, update-side: Euler linearisation, blocks sized to the LDS Fenwick
//                  tree, per-block distinct keys / key tiles / in-block event ranks, and the two
//                  carry levels' chains, chunks and scan segments.
//   BatchPlan      per query batch: queries in slots grouped by block and step (step groups),
//                  level-B endpoint items grouped by chunk and position.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace bwg {

struct PlanKnobs {
  int rmax = 2048;              // max distinct keys per block (LDS Fenwick size of the block replay)
  int tile_shift = 0;           // key tile = key >> tile_shift (level-B classes = 1 << tile_shift);
                                // 0 = balanced: ceil(log2(NK+1) / 2), so both levels' trees are small
  int64_t chunk_items = 0;      // key items per carry chunk on level B (0 = auto); level C has
                                // one chunk per block
  int64_t ep_per_work = 2048;   // endpoints per chunk work item (batch plans)
  int64_t range_slots = 1 << 22;  // batch plans: work items are ordered by ranges of this many
                                  // query slots, so the level-B results written at any time fall
                                  // in one window of the slot-ordered result array
  int seg_chunks = 32;          // chunks per carry-scan segment
  int stage_shift = 18;         // batch plans: query ids per staging window of the combine
                                // (2^18 answers = 2 MiB, held in one die's cache)
  int64_t bt_table_mb = 1024;   // device planner: largest (block, key tile) position table to
                                // build (0: the planner searches instead)
  int max_block_steps = 1 << 30;
};

struct TreeIndexPlan {
  int64_t S = 0;
  int U = 0;
  uint64_t NK = 0;  // coordinates are in [0, NK]
  PlanKnobs knobs;
  int KT = 0;       // key tiles (level-C classes)
  // Euler blocks
  std::vector<int32_t> steps, blk_step0;
  std::vector<int64_t> blk_key0, blk_tile0;
  std::vector<uint32_t> bkey;
  std::vector<uint16_t> tile_rank0, evrank;
  std::vector<int32_t> state_block, state_step;  // S+1 (-1 for the root)
  // carry chains: chain 0 = level C over all blocks; chain 1+t = level B of tile t
  std::vector<uint64_t> kitems;
  std::vector<int64_t> chunk_key0, chunk_cp0, chain_chunk0, chain_key0;
  std::vector<int32_t> chunk_chain, chain_ncls, chain_src, chain_out;
  std::vector<int64_t> seg_chunk0, seg_cp0, chain_seg0;
  int64_t n_cp = 0, n_segsum = 0;
  int max_cls = 0;
  // endpoint positions on level B: blocks holding keys of tile t and where they start
  std::vector<int64_t> tile_tb0;
  std::vector<int32_t> tb_block;
  std::vector<int64_t> tb_pos;
  int32_t nblk() const { return int32_t(blk_step0.size()) - 1; }
  int64_t nchunk() const { return int64_t(chunk_chain.size()); }
  int64_t nseg() const { return int64_t(seg_chunk0.size()) - 1; }
  int32_t nchain() const { return int32_t(chain_ncls.size()); }
};

struct BatchPlan {
  int64_t M = 0;
  std::vector<int32_t> qblk;      // listed blocks
  std::vector<int64_t> qblk_g0;   // their step groups
  std::vector<int32_t> grp_step;  // step of each group
  std::vector<int64_t> grp_q0;    // its slots
  std::vector<uint32_t> qorder;   // slot -> query id
  std::vector<uint32_t> islot;    // query id -> slot
  std::vector<uint32_t> spos;     // staging positions (windows of 2^stage_shift query ids, then
                                  //   die part, then slot), ascending within each slot tile
  std::vector<uint16_t> srank;    // slot -> index of its staging position in its tile's list
  std::vector<uint32_t> ipos;     // query id -> staging position
  std::vector<uint32_t> qrank;    // slot -> in-block ranks lo | hi << 16
  std::vector<int32_t> sblk;      // slot -> block
  std::vector<uint32_t> qtile;    // slot -> key tiles of the endpoints lo | hi << 16
  std::vector<int64_t> wk_chunk, wk_e0;
  std::vector<uint64_t> eitems;   // level B only, endpoint id = 2 * slot + side
  std::vector<uint32_t> epos;     // endpoint id -> its index in eitems
  int64_t npairs = 0;             // slots whose two endpoints share a key tile (one level-B item
                                  // pair, adjacent in item order and in one work item)
  int64_t steps_run = 0;          // block-replay steps the batch needs (statistics)
};

// parent: S+1 entries (parent[0] = -1, parent[v] < v); key: S*U coordinates in [0, NK).
bool build_tree_index(int64_t S, int U, uint64_t NK, const int32_t* parent, const uint32_t* key,
                      const PlanKnobs& knobs, TreeIndexPlan& plan, char* err, size_t errlen);
// Its three stages: (1) checks and tile width, then (walk = true) the Euler walk and (cut = true)
// the greedy block cut (steps, blk_step0, KT); (2) per block its sorted distinct keys, tiles and event ranks, the chains and the
// level-B position tables; (3) the chunk and segment tables, from chain_key0 and blk_tile0 (and
// the chain class counts of stage 2).  The device index build runs stage 2 on the GPU.
bool plan_walk_blocks(int64_t S, int U, uint64_t NK, const int32_t* parent, const uint32_t* key,
                      const PlanKnobs& knobs, TreeIndexPlan& plan, char* err, size_t errlen, bool cut = true,
                      bool walk = true);
bool plan_block_tables(const uint32_t* key, TreeIndexPlan& plan, char* err, size_t errlen);
void plan_chunks(TreeIndexPlan& plan);

// M queries (state in [1,S], endpoint coordinates in [0, NK]).  all_blocks lists every block
// (needed when the same launch also writes the index).
bool build_batch_plan(const TreeIndexPlan& x, int64_t M, const int32_t* qstate, const uint32_t* qlo,
                      const uint32_t* qhi, bool all_blocks, BatchPlan& b, char* err, size_t errlen);

}  // namespace bwg
