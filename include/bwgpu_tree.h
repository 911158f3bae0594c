/* bwgpu_tree.h -- C-ABI of the generalised GPU batch path: exact int64 range sums over a
 * branching version tree whose states each carry many point updates.
 *
 * The reference GPU path (src/gpu_v7.h:49-50, src/gpu_v7b.h:50-52, src/gpu_session.h:21-24)
 * only replays single-edit *chains* with a class domain small enough for shared memory
 * (src/batch_engine_v7.hpp:707-709, src/gpu_v7.cu:102-103).  Its CPU executor for branching
 * trees is a serial preorder apply/undo pass (src/batch_engine_v7.hpp:1021-1058).  These entry
 * points generalise the v7 algorithm (block-parallel chain execution, Fenwick/prefix replay,
 * carry, combine; src/batch_engine_v7.hpp:832-968) to that case; they have no counterpart in
 * the reference ABI.  No HIP or torch types appear here: device pointers are plain pointers,
 * a stream is an opaque pointer (NULL = the default stream), every call returns 0 or a
 * hipError_t value, like the reference launchers (src/gpu_v7.cu:98-121).
 *
 * Model.  States 1..S (0 is an empty root); state v applies U point updates (key, delta) on
 * top of its parent and asks Q half-open range sums [lo, hi) over keys, answered after its own
 * updates.  Keys and endpoints are coordinates in [0, NK] (for a sparse 32-bit key space they
 * are ranks, see bwt_rank_remap_launch).  Answers are int64 in canonical order (v-1)*Q + k,
 * exactly the reference oracle's order (src/oracle.cpp:459-466).
 *
 * Algorithm (the plan is built on the host by bwt_plan_build, the passes run on the GPU):
 *  1. The tree is linearised by its preorder walk into steps (+v enter, -v exit = undo);
 *     consecutive steps form blocks whose distinct keys fit a shared-memory Fenwick tree.
 *  2. Block replay (one workgroup per block): local Fenwick over the block's distinct keys;
 *     for each query endpoint the in-block prefix; at the end the block's net weight per
 *     distinct key (agg) and per key tile (tagg).
 *  3. Carry over earlier blocks, split into two key levels so that every structure fits
 *     shared memory: level C = one chain over all blocks with class = key tile (key >> TS);
 *     level B = one chain per tile with class = key & (2^TS - 1).  Chains are cut into chunks;
 *     a chunk histogram + a segmented column scan gives each chunk its dense carry; a chunk
 *     replay adds the chunk-local Fenwick prefix.  (v7's dV / carrypref / lvals, generalised.)
 *  4. Combine: answer = in-block(hi) - in-block(lo) + carry(hi) - carry(lo).
 */
#ifndef BWGPU_TREE_H
#define BWGPU_TREE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Item encoding of the carry chains (64 bits): bit 63 = endpoint item, bits 40..51 = class,
 * bits 0..39 = reference (endpoint id 2*((v-1)*Q+k)+side for endpoints; weight index for keys). */
#define BWT_ITEM_EP (1ull << 63)
#define BWT_CLASS_SHIFT 40
#define BWT_REF_MASK ((1ull << 40) - 1)

typedef struct BwtPassArgs {
  /* sizes */
  int64_t S;             /* edited states */
  int32_t U, Q;          /* updates and queries per state */
  int32_t nblk;          /* Euler blocks */
  int32_t block_threads; /* threads of the block-replay workgroup (multiple of 64, <= 1024) */
  int32_t rmax;          /* max distinct keys of one block (shared-memory Fenwick size) */
  int32_t max_cls;       /* max classes of any chain (<= 4096) */
  int64_t nchunk;        /* carry chunks over all chains */
  int32_t nchain;
  int32_t chunk_waves;   /* waves per workgroup of the chunk kernels */
  /* inputs: coordinates (keys / ranks), deltas, query endpoints */
  const uint32_t* key;   /* S*U */
  const int64_t* delta;  /* S*U */
  const uint32_t* qlo;   /* S*Q */
  const uint32_t* qhi;   /* S*Q */
  /* plan: Euler blocks */
  const int32_t* steps;      /* +v enter, -v exit */
  const int32_t* blk_step0;  /* nblk+1 */
  const int64_t* blk_key0;   /* nblk+1, offsets into bkey / agg */
  const uint32_t* bkey;      /* sorted distinct keys of every block */
  const int64_t* blk_tile0;  /* nblk+1, offsets into tile_rank0 / tagg */
  const uint16_t* tile_rank0;/* first in-block rank of each key tile present in the block */
  /* plan: carry chains */
  const uint64_t* items;     /* all chains' items, chain-major */
  const int64_t* chunk_item0;/* nchunk+1 */
  const int32_t* chunk_chain;/* nchunk */
  const int64_t* chunk_cp0;  /* nchunk, offset of the chunk's dense carry row in cp */
  const int32_t* chain_ncls; /* nchain */
  const int32_t* chain_src;  /* nchain: 0 = weights from agg (level B), 1 = from tagg (level C) */
  const int32_t* chain_out;  /* nchain: 0 or 1, which half of g receives the endpoint values */
  const int64_t* chain_chunk0; /* nchain+1 */
  /* plan: carry-scan segments (runs of consecutive chunks of one chain) */
  int64_t nseg;
  const int64_t* seg_chunk0; /* nseg+1 */
  const int64_t* seg_cp0;    /* nseg, offset of the segment's row in segsum */
  const int64_t* chain_seg0; /* nchain+1 */
  int64_t* segsum;           /* scratch: sum over segments of ncls(chain) */
  /* scratch / outputs */
  int64_t* agg;    /* blk_key0[nblk] */
  int64_t* tagg;   /* blk_tile0[nblk] */
  int64_t* cp;     /* sum over chunks of ncls(chain) */
  int64_t* g;      /* 2 * 2*S*Q endpoint carries (level C half, level B half) */
  int64_t* ansr;   /* S*Q in-block differences */
  int64_t* answers;/* S*Q */
} BwtPassArgs;

/* One complete pass (block replay, chunk histogram, carry scan, chunk replay, combine) on
 * `stream`.  Asynchronous; safe to capture into a hipGraph (no host synchronisation, no
 * allocation, no memset). */
int bwt_pass_launch(const BwtPassArgs* a, void* stream);

/* Individual stages (same arguments), for profiling and the batched form. */
int bwt_block_replay_launch(const BwtPassArgs* a, void* stream);
int bwt_chunk_hist_launch(const BwtPassArgs* a, void* stream);
int bwt_carry_scan_launch(const BwtPassArgs* a, void* stream);
int bwt_chunk_replay_launch(const BwtPassArgs* a, void* stream);
int bwt_combine_launch(const BwtPassArgs* a, void* stream);

/* CSR rank-LUT over a sorted set of D distinct 32-bit keys (src/sparse_engine.hpp:62-97 is the
 * reference's bitmap+rank-LUT; this is its CSR form): row h = keys whose high 16 bits equal h,
 * row_ptr[h] = number of keys with high bits < h (65537 entries), low16[i] = low 16 bits of
 * key i.  rank(x) = #keys < x = row_ptr[x>>16] + #(low16 in row x>>16 that are < (x & 0xFFFF)).
 * Maps n values in place-free fashion: out[i] = rank(in[i]). */
typedef struct BwtRankArgs {
  int64_t n;
  const uint32_t* in;
  uint32_t* out;
  const uint32_t* row_ptr; /* 65537 */
  const uint16_t* low16;   /* D */
} BwtRankArgs;
int bwt_rank_remap_launch(const BwtRankArgs* a, void* stream);

#ifdef __cplusplus
}
#endif
#endif
