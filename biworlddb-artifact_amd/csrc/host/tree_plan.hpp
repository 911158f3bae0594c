// tree_plan.hpp -- host plan of the generalised batch path (the BatchEngineV7::build analogue,
// src/batch_engine_v7.hpp:576-704 and build_v7_blocks :707-830, for branching multi-update
// trees): Euler linearisation, blocks sized to the LDS Fenwick tree, per-block distinct keys
// and key tiles, and the two carry levels' chains / chunks / scan segments.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace bwg {

struct PlanKnobs {
  int rmax = 2048;          // max distinct keys per block (LDS Fenwick size of the block replay)
  int tile_shift = 11;      // key tile = key >> tile_shift; level-B classes = 1 << tile_shift
  int64_t chunk_items = 16384;  // items per carry chunk
  int seg_chunks = 32;      // chunks per carry-scan segment
  int max_block_steps = 1 << 30;  // optional cap on steps per block
};

struct TreePlan {
  int64_t S = 0;
  int U = 0, Q = 0;
  uint64_t NK = 0;  // coordinates are in [0, NK]
  PlanKnobs knobs;
  // Euler blocks
  std::vector<int32_t> steps, blk_step0;
  std::vector<int64_t> blk_key0, blk_tile0;
  std::vector<uint32_t> bkey;
  std::vector<uint16_t> tile_rank0;
  std::vector<int32_t> state_block;  // S+1: block of the state's enter step (-1 for the root)
  // carry chains
  std::vector<uint64_t> items;
  std::vector<int64_t> chunk_item0, chunk_cp0, chain_chunk0;
  std::vector<int32_t> chunk_chain, chain_ncls, chain_src, chain_out;
  std::vector<int64_t> seg_chunk0, seg_cp0, chain_seg0;
  int64_t n_cp = 0, n_segsum = 0;
  int max_cls = 0;
  int32_t nblk() const { return int32_t(blk_step0.size()) - 1; }
  int64_t nchunk() const { return int64_t(chunk_chain.size()); }
  int64_t nseg() const { return int64_t(seg_chunk0.size()) - 1; }
  int32_t nchain() const { return int32_t(chain_ncls.size()); }
  size_t host_bytes() const;
};

// parent: S+1 entries (parent[0] = -1, parent[v] < v); key: S*U coordinates in [0, NK);
// qlo/qhi: S*Q coordinates in [0, NK].  Returns false with a message on invalid input.
bool build_tree_plan(int64_t S, int U, int Q, uint64_t NK, const int32_t* parent, const uint32_t* key,
                     const uint32_t* qlo, const uint32_t* qhi, const PlanKnobs& knobs, TreePlan& plan,
                     char* err, size_t errlen);

}  // namespace bwg
