/* bwgpu_gen.h -- C-ABI of the device workload generator: materialises the update and query
 * arrays of the benchmark workloads on the GPU, bit-identical to the host generator
 * (biworlddb-artifact_amd/csrc/host/tree_workload.cpp), so a batch stream can be generated,
 * planned and answered without its queries ever crossing PCIe.
 *
 * Counterpart in the reference: the declared-workload generator's split into a cheap sequential
 * draw-index scan and a parallel, bit-identical materialisation from counter-based splitmix64
 * draws (src/batch_engine_v7c.hpp:240-249 draw_at, :332-410 phase 1 scan / phase 2
 * materialisation), which the reference runs on CPU threads.  Here the sequential parts stay on
 * the host (parents, request sizes, the Zipf inverse-CDF table, the distinct-key set of cfg 3)
 * and every per-element array is materialised by the GPU.
 *
 * Every draw is a function of its own index, so any device computes the same integers.  The two
 * floating-point steps are made exact: (1) a Zipf pick compares unit(x) * total (one correctly
 * rounded multiply) against the host-built cumulative table, searched within a 65,537-entry guide
 * table (bwt_gen_guide_launch); (2) a log-uniform length floor(exp(y)) depends on the libm's exp
 * only when exp(y) lies within a few ulps of an integer -- the device flags those draws ("ties")
 * and bwt_gen_settle recomputes them with the host's exp, the one the host generator uses.
 * Launches are asynchronous on `stream` (a hipStream_t, NULL = default stream); bwt_gen_settle
 * synchronises it.  Return value: a hipError_t (0 = success). */
#ifndef BWGPU_GEN_H
#define BWGPU_GEN_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BWT_ZIPF_GUIDE_ENTRIES 65537

typedef struct BwtGenSpec {
  int32_t cfg;                 /* tree config 1..5 (key draw, delta bound, query ranges) */
  int32_t U, Q;                /* updates and queries per state */
  int64_t S;                   /* states 1..S */
  uint64_t N;                  /* key domain (cfg 3: the number of distinct keys) */
  uint64_t seed;
  const double* zipf_cum;      /* device: zipf_n cumulative Zipf weights (NULL for cfg 1) */
  uint64_t zipf_n;             /* < 2^32 */
  const uint32_t* zipf_guide;  /* device: BWT_ZIPF_GUIDE_ENTRIES, written by bwt_gen_guide_launch */
  const uint32_t* dkey;        /* cfg 3: device copy of the zipf_n sorted distinct keys */
} BwtGenSpec;

/* Draws whose floor(exp(.)) the host settles: idx[0 .. min(*count, cap)) (device arrays).  The
 * query launches reset *count.  Element indices of a launch must be < 2^32. */
typedef struct BwtGenTies {
  uint32_t* idx;
  uint32_t* count;
  uint32_t cap;
} BwtGenTies;

/* guide[j] = the first rank whose cumulative weight exceeds (j / 65536) * total, j = 0..65536. */
int bwt_gen_guide_launch(const double* zipf_cum, uint64_t zipf_n, uint32_t* guide, void* stream);

/* key[S*U], delta[S*U] of the tree (update i of state v is element (v-1)*U + i). */
int bwt_gen_updates_launch(const BwtGenSpec* g, uint32_t* key, int64_t* delta, void* stream);

/* The tree's own queries: qlo / qhi[S*Q] (cfg 1 uniform, else log-uniform over [0, N), raw
 * 32-bit endpoints for cfg 3). */
int bwt_gen_state_queries_launch(const BwtGenSpec* g, uint32_t* qlo, uint32_t* qhi, BwtGenTies ties, void* stream);

/* The queries of a request stream over the tree (form 4 batched, 5 served): req_q0[nreq+1] and,
 * for the served form, req_app[nreq] come from the host scan (stream_scan); nq = req_q0[nreq].
 * Writes qstate, qlo, qhi[nq].  For the served form g->zipf_* is the key law of its points and
 * appended keys: Zipf(1.1) over the N keys of a cfg-2 tree, or over a cfg-3 tree the tree's own
 * Zipf(0.9) over its distinct keys (g->dkey), its ranges then spanning the 32-bit domain. */
int bwt_gen_stream_launch(const BwtGenSpec* g, int32_t form, const int64_t* req_q0, const int64_t* req_app,
                          int64_t nreq, int64_t nq, int32_t* qstate, uint32_t* qlo, uint32_t* qhi, BwtGenTies ties,
                          void* stream);

/* The appended branch states of the served form: app_parent[napp], app_key / app_delta[napp*append_U]. */
int bwt_gen_appends_launch(const BwtGenSpec* g, int64_t napp, int32_t append_U, int32_t* app_parent,
                           uint32_t* app_key, int64_t* app_delta, void* stream);

/* Synchronises `stream`, then rewrites the tied ranges of the last state-query (what = 0) or
 * stream (what = 1) launch with the host's exp.  hipErrorInvalidValue if more than ties.cap
 * draws tied (nothing is rewritten then). */
int bwt_gen_settle(const BwtGenSpec* g, int32_t what, uint32_t* qlo, uint32_t* qhi, BwtGenTies ties, void* stream);

#ifdef __cplusplus
}
#endif
#endif
