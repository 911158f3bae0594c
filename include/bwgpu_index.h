/* bwgpu_index.h -- C-ABI of the device index build: the Euler walk of the version tree, its cut
 * into blocks and the per-block and per-chain tables of BwtIndex / BwtPlanIndex
 * (include/bwgpu_tree.h, include/bwgpu_plan.h), on the GPU from the parent array and the keys.
 *
 * Counterpart in the reference: the host plan build BatchEngineV7::build -- per-row chains by
 * radix sort, supports, base scans (src/batch_engine_v7.hpp:576-704,1069-1415) -- which the
 * reference GPU drivers run on the CPU before the launch (src/panel_batch_gpu.cpp:86-121).  Here
 * the tables are those of the Euler-walk index; they equal the host build (plan_block_tables in
 * csrc/host/tree_plan.hpp) array for array:
 *   stage 0 (bwt_index_cut_launch)     the greedy block cut (below), or the host's cut;
 *   stage 1 (bwt_index_keys_launch)    one workgroup per block: its events' keys through an LDS
 *            hash set, the distinct keys sorted in LDS (bitonic), every event's rank by binary
 *            search; the block of every step, the enter step of every state; per block its
 *            distinct-key and key-tile counts (the caller scans them into blk_key0 / blk_tile0);
 *   stage 2 (bwt_index_chains_launch)  per block its keys and tile slots (level-C items); the
 *            slots sorted by tile (stable radix sort), their key counts scanned into level-B
 *            chain positions; chain_key0, tile_tb0, tb_block, tb_pos, the level-B items, and the
 *            dense (block, tile) position table bt_pos.
 * Launches are asynchronous on `stream`; return value: a hipError_t (0 = success). */
#ifndef BWGPU_INDEX_H
#define BWGPU_INDEX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct BwtIndexBuild {
  int64_t S;                 /* states 1..S */
  int32_t U;                 /* updates per state */
  int32_t rmax;              /* at most rmax distinct keys per block (8 <= rmax <= 8192) */
  int32_t tile_shift;        /* key tile = key >> tile_shift, 1..12 */
  int32_t KT;                /* key tiles: (NK >> tile_shift) + 1 */
  int64_t nsteps;            /* steps of the walk (+v enter, -v exit) */
  int32_t nblk;              /* blocks */
  const int32_t* steps;      /* device, nsteps */
  const int32_t* blk_step0;  /* device, nblk + 1 */
  const uint32_t* key;       /* device, S*U coordinates (update i of state v at (v-1)*U + i) */
} BwtIndexBuild;

typedef struct BwtIndexTables {  /* device arrays; capacities in brackets */
  uint16_t* evrank;          /* [nsteps*U] rank of every event's key in its block */
  int32_t* state_step;       /* [S+1] enter step of every state (-1 for state 0) */
  int32_t* step_block;       /* [nsteps] block of every step */
  const int64_t* blk_key0;   /* [nblk+1] exclusive scan of stage 1's key counts (input of stage 2) */
  const int64_t* blk_tile0;  /* [nblk+1] exclusive scan of its tile counts (input of stage 2) */
  uint32_t* bkey;            /* [nkeys] sorted distinct keys of every block */
  uint16_t* tile_rank0;      /* [nslots] first rank of each key tile present in a block */
  uint64_t* kitems;          /* [nslots + nkeys] level-C then level-B chain items */
  int64_t* chain_key0;       /* [KT+2] */
  int64_t* tile_tb0;         /* [KT+1] */
  int32_t* tb_block;         /* [nslots] */
  int64_t* tb_pos;           /* [nslots] */
  int32_t* bt_pos;           /* NULL, or [nblk*KT] */
} BwtIndexTables;

/* Workspace bytes of both stages. */
int bwt_index_workspace(const BwtIndexBuild* in, size_t* bytes);

/* Stage 1: writes t->evrank, t->state_step, t->step_block and, per block, nkeys[b] / ntiles[b]
 * (device, nblk each).  A block with more than rmax distinct keys (the cut never makes one) or a
 * key whose tile is >= KT is an error the caller finds in nkeys (> rmax); stage 2 must not run. */
int bwt_index_keys_launch(const BwtIndexBuild* in, BwtIndexTables* t, int32_t* nkeys, int32_t* ntiles, void* work,
                          size_t work_bytes, void* stream);

/* Stage 2, after stage 1 on the same workspace: t->blk_key0 / t->blk_tile0 hold the exclusive
 * scans of nkeys / ntiles, whose totals are nkeys_total / nslots. */
int bwt_index_chains_launch(const BwtIndexBuild* in, BwtIndexTables* t, int64_t nkeys_total, int64_t nslots,
                            void* work, size_t work_bytes, void* stream);

/* The Euler walk on the device: children in ascending order by a stable radix sort of (parent,
 * state); every tree edge's successor in the preorder walk (down to the first child or back up;
 * from a child up to its next sibling or further up); the walk's order by list ranking (pointer
 * jumping over the 2S edges, log2(2S) rounds).  parent: device, S+1 entries, parent[v] in [0, v)
 * (the caller checks).  Writes steps (capacity 2S; +v enters v, -v leaves it; the walk ends with
 * the last entry, as the host walk does) and *nsteps (device int64). */
int bwt_index_walk_workspace(int64_t S, size_t* bytes);
int bwt_index_walk_launch(int64_t S, const int32_t* parent, int32_t* steps, int64_t* nsteps, void* work,
                          size_t work_bytes, void* stream);

/* The greedy block cut on the device (stage 0): the cut the host makes step by step -- a block
 * takes steps while the distinct keys of its steps stay within rmax and its step count within
 * max_block_steps -- found in parallel: for every step s, one workgroup computes e(s), the end of
 * a block that would start at s (an LDS hash set of the keys, step by step); the blocks are then
 * the path 0, e(0), e(e(0)), ..., marked by pointer doubling (log2(blocks) rounds) and
 * compacted.  Uses in->steps, in->key, in->U, in->rmax, in->nsteps.  Writes blk_step0 (capacity
 * nsteps + 1; ends with nsteps) and *nblk (device int32). */
int bwt_index_cut_workspace(const BwtIndexBuild* in, size_t* bytes);
int bwt_index_cut_launch(const BwtIndexBuild* in, int32_t max_block_steps, int32_t* blk_step0, int32_t* nblk,
                         void* work, size_t work_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif
