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
 * hipError_t value, like the reference launchers (src/gpu_v7.cu:98-121).  Nothing allocates,
 * synchronises or memsets, so every sequence of these calls can be captured into a hipGraph.
 *
 * Model.  States 1..S (0 is an empty root); state v applies U point updates (key, delta) on
 * top of its parent.  A query (v, lo, hi) asks the sum of the deltas on the root->v path with
 * lo <= key < hi.  Keys and endpoints are coordinates in [0, NK] (for a sparse 32-bit key space
 * they are ranks, see bwt_rank_remap_launch).  Query q's answer goes to answers[q]; the
 * reference's declared workload (Q queries per state, order (v-1)*Q+k, src/oracle.cpp:459-466)
 * is the batch whose query q belongs to state q/Q+1.
 *
 * Algorithm.  The tree is linearised by its preorder walk into steps (+v enter, -v exit = undo);
 * consecutive steps form blocks whose distinct keys fit LDS.  answer(q) = F(hi) - F(lo),
 * F(x) = sum over the events before q's step with key < x, split as
 *   in-block part  -- block replay over the block's own distinct keys (local ranks);
 *   carry level C  -- one chain over all blocks, class = key tile (key >> TS), cut into one
 *                     chunk per block, so its carry is a dense row per block read directly;
 *   carry level B  -- one chain per key tile, class = key & (2^TS - 1), cut into chunks.
 * Chunk histograms + a segmented column scan give every chunk its dense carry; a level-B chunk
 * replay adds the chunk-local prefix at each query endpoint.  (These are v7's lvals / dV /
 * carrypref / combine, generalised.)  The *index* (block weights, chunk carries) depends only
 * on the updates and is built once; a *batch* (queries) only replays the blocks and chunks it
 * touches.
 */
#ifndef BWGPU_TREE_H
#define BWGPU_TREE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Classes are 14-bit: a chain has at most BWT_MAX_CLASSES classes (level C: key tiles, so
 * NK >> tile_shift < BWT_MAX_CLASSES - 1; level B: offsets within a tile). */
#define BWT_CLASS_BITS 14
#define BWT_CLASS_MASK ((1u << BWT_CLASS_BITS) - 1)
#define BWT_MAX_CLASSES (1 << BWT_CLASS_BITS)
/* Key item of a carry chain: class << 40 | weight index (into agg for level B, tagg for C). */
#define BWT_CLASS_SHIFT 40
#define BWT_REF_MASK ((1ull << 40) - 1)
#define BWT_KEY_CLASS(item) ((uint32_t)((item) >> BWT_CLASS_SHIFT) & BWT_CLASS_MASK)
/* Endpoint item of a batch: local position << 44 | class << 30 | endpoint id (2 slot + side).
 * The position (19 bits) is the number of the chunk's key items that precede the endpoint; the
 * id (30 bits) bounds a batch to BWT_MAX_BATCH queries. */
#define BWT_EP_CLASS_SHIFT 30
#define BWT_EP_ID_MASK ((1ull << 30) - 1)
#define BWT_EP_POS_SHIFT 44
#define BWT_EP_POS_MASK 0x7FFFFull
#define BWT_EP_POS(item) (((item) >> BWT_EP_POS_SHIFT) & BWT_EP_POS_MASK)
#define BWT_EP_CLASS(item) ((uint32_t)((item) >> BWT_EP_CLASS_SHIFT) & BWT_CLASS_MASK)
#define BWT_EP_ID(item) ((uint32_t)((item) & BWT_EP_ID_MASK))
#define BWT_MAX_BATCH ((1ll << 29) - 1)

/* Static, update-side structure (resident across batches). */
typedef struct BwtIndex {
  int64_t S;
  int32_t U;
  int32_t nblk;
  int32_t rmax;      /* max distinct keys of one block (LDS Fenwick size of the block replay) */
  int32_t max_cls;   /* max classes of any chain (<= BWT_MAX_CLASSES) */
  int32_t nchain;
  int32_t ncls_c;    /* classes of chain 0 (level C: key tiles) */
  int32_t ncls_b;    /* classes of every other chain (level B: keys within a tile) */
  int32_t tile_shift;/* TS: level-C class of a coordinate x is x >> TS */
  int64_t nchunk;
  int64_t nchunk_c;  /* chunks of chain 0 = nblk (chunk b = block b); they come first */
  int64_t nseg;
  const int64_t* delta;       /* S*U update deltas (state-major) */
  const int32_t* steps;       /* +v enter, -v exit */
  const int32_t* blk_step0;   /* nblk+1 */
  const uint16_t* evrank;     /* nsteps*U: in-block rank of the key of update i of step s (s*U+i) */
  const int64_t* blk_key0;    /* nblk+1, offsets into bkey / agg */
  const uint32_t* bkey;       /* sorted distinct keys of every block (coordinates) */
  const int64_t* blk_tile0;   /* nblk+1, offsets into tile_rank0 / tagg */
  const uint16_t* tile_rank0; /* first in-block rank of each key tile present in the block */
  const uint64_t* kitems;     /* key items of all chains, chain-major */
  const int64_t* chunk_key0;  /* nchunk+1 */
  const int32_t* chunk_chain; /* nchunk */
  const int64_t* chunk_cp0;   /* nchunk: offset of the chunk's dense carry row in cp */
  const int32_t* chain_ncls;  /* nchain */
  const int32_t* chain_src;   /* nchain: 0 = weights from agg (level B), 1 = from tagg (level C) */
  const int32_t* chain_out;   /* nchain: which half of g receives the endpoint values */
  const int64_t* chain_chunk0;/* nchain+1 */
  const int64_t* seg_chunk0;  /* nseg+1: carry-scan segments = runs of chunks of one chain */
  const int64_t* seg_cp0;     /* nseg */
  const int64_t* chain_seg0;  /* nchain+1 */
  int64_t* agg;     /* net weight of every block per distinct key */
  int64_t* tagg;    /* net weight of every block per key tile */
  int64_t* cp;      /* chunk carries (dense rows) */
  int64_t* segsum;  /* scan scratch */
  int64_t* kweight; /* weight of every key item in chain order (written by the histogram) */
} BwtIndex;

/* One query batch (built on the host by the batch planner; all device pointers).  Queries are
 * held in *slots*, grouped by block and, inside a block, by the step of their state. */
typedef struct BwtBatch {
  int64_t M;                  /* queries */
  const uint32_t* qlo;        /* M endpoint coordinates, by query id */
  const uint32_t* qhi;
  int32_t nqblk;              /* blocks holding queried states */
  int32_t pad0;
  const int32_t* qblk;        /* nqblk block ids */
  const int64_t* qblk_g0;     /* nqblk+1 offsets into the step groups */
  int64_t ngrp;               /* step groups: the queries of one state */
  const int32_t* grp_step;    /* ngrp global enter steps */
  const int64_t* grp_q0;      /* ngrp+1 offsets into the slots */
  const uint32_t* qorder;     /* M: query id of each slot */
  const uint32_t* qrank;      /* M: the slot's endpoints' in-block ranks, lo | hi << 16 */
  const int32_t* sblk;        /* M: block of each slot */
  const uint32_t* qtile;      /* M: the slot's endpoints' key tiles (level-C classes), lo | hi << 16;
                               NULL: combine takes them from qlo/qhi (coordinates >> TS) */
  int64_t nwork;              /* level-B work items: a chunk and a run of its endpoints */
  const int64_t* wk_chunk;    /* nwork chunk ids */
  const int64_t* wk_e0;       /* nwork+1 offsets into eitems */
  const uint64_t* eitems;     /* level-B endpoint items per work item, by position; id = 2*slot+side */
  const uint32_t* epos;       /* 2M, by endpoint id: the endpoint's index in eitems (and in g) */
  int64_t* ansr;              /* M, by slot: in-block part */
  int64_t* g;                 /* 2M, in work order (g[i] belongs to eitems[i]): level-B carry */
  const uint32_t* spos;       /* M: staging positions of the slots of each tile of
                                 BWT_COMBINE_TILE slots, ascending within the tile */
  const uint16_t* srank;      /* M: index of the slot's staging position in its tile's list */
  const uint32_t* ipos;       /* M: staging position of each query id's value (= spos of its slot) */
  int64_t* stage;             /* M staging buffer: finished values grouped by windows of query ids */
  int64_t* answers;           /* M, by query id */
  const int64_t* dcounts;     /* NULL, or device counts [nqblk, ngrp, nwork, ...] written by the
                                 device planner (include/bwgpu_plan.h); nqblk / ngrp / nwork
                                 above are then upper bounds, and the replay kernels loop over
                                 the device counts */
} BwtBatch;

/* The combine kernels map workgroups to slot tiles / query-id blocks XCD-major: on a part with
 * BWT_XCDS accelerator dies that take workgroups round robin, die x handles the x-th of BWT_XCDS
 * contiguous parts of the tiles (slots, BWT_COMBINE_TILE each) or blocks (query ids,
 * BWT_COMBINE_THREADS each).  The planner groups the staging buffer by (query-id window, slot
 * part) so each die's staging stores and loads stay in its own cache (the mapping affects speed
 * only, never results). */
#define BWT_XCDS 8
#define BWT_COMBINE_THREADS 256
#define BWT_COMBINE_TILE 2048

typedef struct BwtLaunch {
  int32_t block_threads;  /* block replay workgroup (multiple of 64, <= 1024) */
  int32_t chunk_waves;    /* wavefronts per workgroup of the chunk kernels (1..4), one chunk each */
                          /* block replay: each thread owns the smallest power of two >= rmax /
                           * block_threads of the block's ranks (at most 16) */
  int32_t write_index;    /* block replay also writes agg / tagg (needs every block listed) */
  int32_t carry_in_block; /* query side (write_index 0, index built): the block replay also adds
                             the level-C carry (its block's dense row) to every slot's in-block
                             part, and the combine does not; bwt_query_launch sets it,
                             bwt_pass_launch clears it */
} BwtLaunch;

/* Index build: agg/tagg are written by a block replay with write_index set; then: */
int bwt_chunk_hist_launch(const BwtIndex* x, const BwtLaunch* l, void* stream);
int bwt_carry_scan_launch(const BwtIndex* x, const BwtLaunch* l, void* stream);
/* Query side. */
int bwt_block_replay_launch(const BwtIndex* x, const BwtBatch* b, const BwtLaunch* l, void* stream);
int bwt_chunk_replay_launch(const BwtIndex* x, const BwtBatch* b, const BwtLaunch* l, void* stream);
int bwt_combine_launch(const BwtIndex* x, const BwtBatch* b, const BwtLaunch* l, void* stream);
/* Everything for one batch over a cold index (block replay writing the index, histogram, scan,
 * chunk replay, combine): the plain "one launch sequence per query batch" form. */
int bwt_pass_launch(const BwtIndex* x, const BwtBatch* b, const BwtLaunch* l, void* stream);
/* The query side only, over a resident index (block replay, chunk replay, combine). */
int bwt_query_launch(const BwtIndex* x, const BwtBatch* b, const BwtLaunch* l, void* stream);

/* Queries on states appended after the index was built (the served form).  Appended state
 * S+1+a (a = 0, 1, ...) has parent app_parent[a] and U updates app_key/app_delta[a*U .. a*U+U).
 * A query on an appended state is answered as its deepest indexed ancestor p (a query of the
 * batch) plus this overlay: the updates of the appended states on the path from the state up
 * to p with lo <= key < hi, added to answers[qid].  One wavefront per overlay query. */
typedef struct BwtOverlay {
  int64_t S;                  /* indexed states (appended ones are numbered after them) */
  int32_t U;                  /* updates per appended state */
  int32_t pad0;
  int64_t n;                  /* overlay queries */
  const uint32_t* qid;        /* n: query ids within the batch */
  const int32_t* qstate;      /* n: their appended states */
  const uint32_t* qlo;        /* by query id */
  const uint32_t* qhi;
  const int32_t* app_parent;  /* appended states' parents (any state numbered lower) */
  const uint32_t* app_key;    /* appended updates: key coordinates */
  const int64_t* app_delta;
  int64_t* answers;           /* by query id: += overlay part */
} BwtOverlay;
int bwt_overlay_launch(const BwtOverlay* o, void* stream);

/* CSR rank-LUT over a sorted set of D distinct 32-bit keys (src/sparse_engine.hpp:62-97 is the
 * reference's bitmap+rank-LUT; this is its CSR form): row h = keys whose high 16 bits equal h,
 * row_ptr[h] = number of keys with high bits < h (65537 entries), low16[i] = low 16 bits of
 * key i.  out[i] = rank(in[i]) = #keys < in[i]. */
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
