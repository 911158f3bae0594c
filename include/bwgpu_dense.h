/* bwgpu_dense.h -- C-ABI of the dense-carry batched form: a batch of exact int64 range-sum
 * queries over a branching version tree, answered by one bucket sort of the queries by state and
 * one persistent block-replay kernel over a resident index whose carry is stored dense.
 *
 * Counterparts in the reference.  The v7 algorithm splits every answer into a block-local part
 * (lvals: block replay from a zero prefix, src/gpu_v7.cu:10-82) and the carry of the blocks
 * before (carrypref, built on the host by the serial carry pass src/panel_batch_gpu.cpp:191-204,
 * combined in gpu_combine_kernel src/gpu_v7.cu:84-96).  Its carry is one dense row per block
 * over the class domain (B x Cpfx int64).  Here the class domain is a 2^20-key space and the
 * tree branches (the Euler linearisation of include/bwgpu_tree.h), and the carry is likewise
 * one dense row per block, NK+1 entries wide: on a 288 GB device the rows of a 300,000-state
 * tree cut into ~5,600 blocks take ~47 GB and are built once.  A query then needs one read of
 * its block's row at each endpoint, no second carry level, no endpoint sort and no combine.
 * The per-query planning the reference does on the host (qlb/qhb, src/batch_engine_v7.hpp:
 * 666-698) is one device bucket sort of the queries by their state.
 *
 * No HIP or torch types appear here: device pointers are plain pointers, a stream is an opaque
 * pointer (NULL = the default stream), every call returns 0 or a hipError_t value.  Nothing
 * allocates or synchronises; the caller owns every buffer, workspaces included.
 *
 * Model (as include/bwgpu_tree.h).  States 1..S (0 is the empty root); state v applies U point
 * updates (key, delta) on top of its parent.  A query (v, lo, hi) asks the sum of the deltas on
 * the root->v path with lo <= key < hi; its answer is F_v(hi) - F_v(lo), F_v(x) = sum of the
 * deltas on the path with key < x.  Keys and endpoints are coordinates in [0, NK].
 */
#ifndef BWGPU_DENSE_H
#define BWGPU_DENSE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Limits of this form: a query's state (19 bits), id (27 bits) and endpoints (up to 2^20
 * each) are packed into one 64-bit sort key and one 32-bit value. */
#define BWD_MAX_NK (1u << 20)
#define BWD_MAX_STATES ((1 << 19) - 1)
#define BWD_MAX_QUERIES (1ll << 27)
#define BWD_MAX_RMAX 4096

/* The resident index.  The walk, blocks, keys and ranks are those of a BwtIndex (its first
 * fields have the same meaning); blk_e0 / state_erank / erank_step and carry are this form's. */
typedef struct BwdIndex {
  int64_t S;
  int32_t U;
  int32_t nblk;
  int32_t rmax;              /* max distinct keys of one block (<= BWD_MAX_RMAX) */
  int32_t pad0;
  uint64_t NK;               /* coordinates are in [0, NK], NK <= BWD_MAX_NK */
  int64_t nsteps;            /* steps of the preorder walk (2S - 1 when every state is entered) */
  int64_t carry_ld;          /* entries per carry row (>= NK + 1, a multiple of 64) */
  const int64_t* delta;      /* S*U update deltas (state-major) */
  const int32_t* steps;      /* nsteps: +v enter, -v exit */
  const int32_t* blk_step0;  /* nblk+1 */
  const uint16_t* evrank;    /* nsteps*U: in-block rank of the key of update i of step s (s*U+i) */
  const int64_t* blk_key0;   /* nblk+1, offsets into bkey */
  const uint32_t* bkey;      /* sorted distinct keys of every block */
  const int32_t* blk_e0;     /* nblk+1: enter steps before each block's first step */
  const int32_t* state_erank;/* S+1: preorder rank of each state (its enter step's index among
                                the enter steps); entry 0 unused */
  const int32_t* erank_step; /* S: the step of the e-th enter */
  int64_t* carry;            /* nblk*carry_ld: carry[b*carry_ld + x] = sum of the deltas of the
                                walk's events before block b's first step with key < x */
} BwdIndex;

/* Carry build from the blocks' net weight per distinct key (agg[blk_key0[b] + r] for the r-th
 * smallest key of block b: what bwt_block_replay_launch writes with write_index set). */
int bwd_carry_workspace(const BwdIndex* x, size_t* bytes);
int bwd_carry_build_launch(const BwdIndex* x, const int64_t* agg, void* work, size_t work_bytes, void* stream);

/* Launch shape of the persistent block replay: workgroup threads (256, 512 or 1024; each thread
 * owns rmax/threads in-block ranks, at most 16), workgroups per compute unit (at most as many as
 * the replay's registers and LDS let reside), a cap on the workgroups (0: none) that leaves
 * compute units free for work on another stream, such as the next batch's plan, the work item
 * (chunk 0: a workgroup takes a whole block; chunk c: a block's queries, in walk order, are cut
 * into items of c, each taken by its own workgroup, which replays the block's events up to its
 * first state; the items of a block are queued together, so few blocks' carry rows are read at a
 * time), and how answers reach their query ids:
 *   BWD_STORE_DIRECT  the replay stores each answer at its query id (one random 8-byte store each);
 *   BWD_STORE_STAGED  the replay stores answers in sorted order, then a partition pass groups them
 *                     by windows of 2^18 query ids and a scatter pass stores each window from one
 *                     die (two streaming passes instead of random stores);
 *   BWD_STORE_AUTO    staged when the batch's ids span more than one window. */
#define BWD_STORE_AUTO 0
#define BWD_STORE_DIRECT 1
#define BWD_STORE_STAGED 2
#define BWD_MIN_CHUNK 256
typedef struct BwdLaunch {
  int32_t threads;
  int32_t wg_per_cu;
  int32_t max_workgroups;
  int32_t store;             /* BWD_STORE_* */
  int32_t chunk;             /* queries per work item (0: whole blocks; else >= BWD_MIN_CHUNK) */
} BwdLaunch;

/* A planned batch of M queries: the queries grouped by state (any order inside a state), the group
 * of the state of every preorder rank, and every block's last rank with queries.  Lives in the
 * answer workspace. */
int bwd_answer_workspace(const BwdIndex* x, int64_t M, size_t* bytes);

/* Plan: group M queries (qstate[M] in [1, S], qlo / qhi[M] in [0, NK]) by state into the
 * workspace.  An out-of-range query sets flags[0] to 1 and is answered 0 (planned over [0, 0));
 * flags is one int64 on the device, written by this call (2 would report an inconsistent plan,
 * which the plan's own counts rule out).  Kernels only: capturable in a hipGraph. */
int bwd_plan_launch(const BwdIndex* x, int64_t M, const int32_t* qstate, const uint32_t* qlo, const uint32_t* qhi,
                    int64_t* flags, void* work, size_t work_bytes, void* stream);

/* The block replay's launch for a batch of M queries under *l (host query, no launch): its
 * workgroups, the workgroups per CU that reside (l->wg_per_cu, at most as many as its registers and
 * LDS allow) and its LDS per workgroup.  No counterpart in the reference (its replay's grid is one
 * block per chain block). */
int bwd_replay_shape(const BwdIndex* x, int64_t M, const BwdLaunch* l, int32_t* workgroups, int32_t* wg_per_cu,
                     size_t* lds_bytes);
/* The work items the last replay with chunk > 0 took from this workspace (a blocking copy of one
 * count; undefined before such a replay).  No counterpart in the reference. */
int bwd_replay_work_items(const BwdIndex* x, int64_t M, const void* work, size_t work_bytes, uint32_t* items);

/* Replay: one launch of the persistent block kernel answers the planned batch (answers[M], by
 * query id): its workgroups take blocks from a device work queue until none is left, so blocks
 * of uneven length and query count share the device.  Kernels only (capturable in a hipGraph). */
int bwd_replay_launch(const BwdIndex* x, int64_t M, int64_t* answers, const BwdLaunch* l, void* work,
                      size_t work_bytes, void* stream);
/* Measurement: bwd_replay_launch once on `stream` with HIP events at its phase boundaries; waits
 * for it and writes the phases' times in ms: ms[0] the persistent block replay, ms[1] the answer
 * partition, ms[2] the answer scatter (both 0 when the answers are stored by id directly).  Not for
 * graph capture (it synchronises). */
int bwd_replay_timed(const BwdIndex* x, int64_t M, int64_t* answers, const BwdLaunch* l, void* work, size_t work_bytes,
                     void* stream, float* ms);
/* Both. */
int bwd_answer_launch(const BwdIndex* x, int64_t M, const int32_t* qstate, const uint32_t* qlo, const uint32_t* qhi,
                      int64_t* answers, int64_t* flags, const BwdLaunch* l, void* work, size_t work_bytes,
                      void* stream);

/* Allocation-free variant of bwd_answer_launch, for callers that cannot allocate device memory
 * (tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  It
 * plans and answers in a library-held scratch array of 64 MiB per GPU (held until the process
 * ends) instead of a caller's workspace, and calls no hipMalloc, hipFree or allocating library
 * routine; the caller still passes inputs, answers and flags.
 * A batch whose plan needs more scratch than that (about 1.9M queries over 300,000 states; see
 * bwd_answer_noalloc_capacity) is rejected with hipErrorInvalidValue and a message on stderr.
 * Every call on a GPU shares that GPU's array and leaves its contents behind: calls that overlap
 * on one GPU (two streams, two threads) race and must be serialised by the caller.  Kernels only,
 * like the workspace entry points. */
int bwd_answer_noalloc_launch(const BwdIndex* x, int64_t M, const int32_t* qstate, const uint32_t* qlo,
                              const uint32_t* qhi, int64_t* answers, int64_t* flags, const BwdLaunch* l, void* stream);
/* Scratch bytes a batch of M queries needs, and the bytes the allocation-free entry point holds. */
int bwd_answer_noalloc_capacity(const BwdIndex* x, int64_t M, size_t* bytes_needed, size_t* bytes_held);

#ifdef __cplusplus
}
#endif
#endif
