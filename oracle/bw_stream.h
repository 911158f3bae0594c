/* bw_stream.h -- the oracle's restatement of the request streams of the batched and served forms
 * (test infrastructure).  The reference has no request streams (its GPU workers answer one
 * declared batch per run, src/panel_batch_gpu.cpp:41-67); these streams follow the counter-based
 * design of bw_twl.h (every draw is twl_draw(seed, stream, index) of the tree's seed) and the
 * shapes written in the benchmark configs.  The product has an independent C++ implementation
 * (biworlddb-artifact_amd/csrc/host/tree_workload.cpp); tests/test_oracle.py checks both give
 * byte-identical streams.
 *
 * Batched form (over a config-4 tree): nreq query batches; batch b holds
 *   size_b = 1 + bounded(draw(8,b), max_batch)                       (uniform in [1, max_batch])
 * queries; query g (global index, batches in order) asks
 *   state = 1 + bounded(draw(9,g), S),   range = loglen(draw(10,g), draw(11,g), N)  (as cfg 2).
 * Served form (over a config-2 tree, or a config-3 tree: see twl_stream_generate): nreq requests; request r is an append when
 *   u01(draw(13,r)) < append_frac, else a query batch of log-uniform size
 *   size_r = clamp(floor(exp(u01(draw(8,r)) * ln(max_batch))), 1, max_batch).
 * Append a (a-th append) creates state S+a+1 with parent 1 + bounded(draw(14,a), S+a) and
 * append_U updates: key Zipf(1.1) over [0,N) by u01(draw(15, a*append_U+i)), delta
 * bounded(draw(16, a*append_U+i), 2*10^6+1) - 10^6.  Query g of a batch asks a state uniform over
 * the states existing when its request arrives, 1 + bounded(draw(9,g), S + appends before it),
 * and is a point query [k, k+1) with k Zipf(1.1) by u01(draw(10,g)) when u01(draw(12,g)) < 1/2,
 * else the cfg-2 log-uniform range of draw(10,g), draw(11,g).  Arrivals are Poisson:
 * arrival_r = arrival_{r-1} + (-log1p(-u01(draw(17,r))) / rate), arrival_{-1} = 0.
 * Answers are in (request, query) order; append requests hold none.
 */
#ifndef BW_STREAM_H
#define BW_STREAM_H
#include "bw_twl.h"

typedef struct {
  int form;            /* 4 = batched, 5 = served */
  int64_t nreq;
  int max_batch;
  double append_frac;  /* served form */
  int append_U;
  double rate;         /* served form: requests per second */
  /* generated */
  int64_t* req_q0;     /* nreq+1 offsets into the queries */
  int8_t* req_kind;    /* 0 = query batch, 1 = append */
  double* arrival;     /* served form */
  int64_t nq;
  int32_t* qstate;
  uint32_t* qlo;
  uint32_t* qhi;
  int64_t napp;
  int32_t* app_parent; /* napp */
  uint32_t* app_key;   /* napp * append_U */
  int64_t* app_delta;
} twl_stream_t;

static void twl_stream_defaults(int form, twl_stream_t* s) {
  memset(s, 0, sizeof(*s));
  s->form = form;
  if (form == 4) {
    s->nreq = 100000;
    s->max_batch = 2048;
  } else {
    s->nreq = 20000;
    s->max_batch = 4096;
    s->append_frac = 0.05;
    s->append_U = 128;
    s->rate = 20000.0;
  }
}

/* w: the generated tree (cfg 4 for the batched form, cfg 2 for the served form). */
static int twl_stream_generate(const twl_t* w, twl_stream_t* s) {
  const uint64_t seed = w->seed;
  const int64_t S = w->S;
  const uint64_t N = w->N;
  s->req_q0 = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s->nreq + 1));
  s->req_kind = (int8_t*)malloc((size_t)s->nreq + 1);
  s->arrival = (double*)malloc(sizeof(double) * (size_t)(s->nreq + 1));
  if (!s->req_q0 || !s->req_kind || !s->arrival) return -1;
  /* sizes and kinds */
  int64_t nq = 0, napp = 0;
  double t = 0.0;
  s->req_q0[0] = 0;
  for (int64_t r = 0; r < s->nreq; ++r) {
    int64_t sz;
    int kind = 0;
    if (s->form == 4) {
      sz = 1 + (int64_t)(twl_draw(seed, 8, (uint64_t)r) % (uint64_t)s->max_batch);
    } else {
      kind = twl_u01(twl_draw(seed, 13, (uint64_t)r)) < s->append_frac;
      const double l = floor(exp(twl_u01(twl_draw(seed, 8, (uint64_t)r)) * log((double)s->max_batch)));
      sz = l < 1.0 ? 1 : (int64_t)l;
      if (sz > s->max_batch) sz = s->max_batch;
      if (kind) sz = 0;
      t += -log1p(-twl_u01(twl_draw(seed, 17, (uint64_t)r))) / s->rate;
    }
    s->req_kind[r] = (int8_t)kind;
    s->arrival[r] = t;
    napp += kind;
    nq += sz;
    s->req_q0[r + 1] = nq;
  }
  s->nq = nq;
  s->napp = napp;
  s->qstate = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nq ? nq : 1));
  s->qlo = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(nq ? nq : 1));
  s->qhi = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(nq ? nq : 1));
  const size_t na = (size_t)napp * (size_t)s->append_U;
  s->app_parent = (int32_t*)malloc(sizeof(int32_t) * (size_t)(napp ? napp : 1));
  s->app_key = (uint32_t*)malloc(sizeof(uint32_t) * (na ? na : 1));
  s->app_delta = (int64_t*)malloc(sizeof(int64_t) * (na ? na : 1));
  if (!s->qstate || !s->qlo || !s->qhi || !s->app_parent || !s->app_key || !s->app_delta) return -1;
  /* over a cfg-3 tree: appended keys and point queries are distinct keys by the tree's Zipf(0.9),
   * ranges over the 32-bit domain; else keys are their own ranks */
  const int ranked = w->cfg == 3;
  double* cum = NULL;
  if (s->form == 5) cum = twl_zipf_table(N, ranked ? 0.9 : 1.1);
  int64_t a = 0;
  for (int64_t r = 0; r < s->nreq; ++r) {
    if (s->req_kind[r]) {
      s->app_parent[a] = (int32_t)(1 + (int64_t)(twl_draw(seed, 14, (uint64_t)a) % (uint64_t)(S + a)));
      for (int i = 0; i < s->append_U; ++i) {
        const uint64_t idx = (uint64_t)a * (uint64_t)s->append_U + (uint64_t)i;
        const uint64_t pk = twl_zipf_pick(cum, N, twl_u01(twl_draw(seed, 15, idx)));
        s->app_key[idx] = ranked ? w->dkey[pk] : (uint32_t)pk;
        s->app_delta[idx] = (int64_t)(twl_draw(seed, 16, idx) % 2000001ULL) - 1000000;
      }
      ++a;
      continue;
    }
    const uint64_t nstates = (uint64_t)(S + a);
    for (int64_t g = s->req_q0[r]; g < s->req_q0[r + 1]; ++g) {
      s->qstate[g] = (int32_t)(1 + (int64_t)(twl_draw(seed, 9, (uint64_t)g) % nstates));
      if (s->form == 5 && twl_u01(twl_draw(seed, 12, (uint64_t)g)) < 0.5) {
        const uint64_t pk = twl_zipf_pick(cum, N, twl_u01(twl_draw(seed, 10, (uint64_t)g)));
        const uint32_t k = ranked ? w->dkey[pk] : (uint32_t)pk;
        s->qlo[g] = k;
        s->qhi[g] = k == 0xFFFFFFFFu ? k : k + 1u; /* [k, k+1), empty at the top value */
      } else {
        twl_loglen_range(twl_draw(seed, 10, (uint64_t)g), twl_draw(seed, 11, (uint64_t)g),
                         ranked ? 0xFFFFFFFFull : N, &s->qlo[g], &s->qhi[g]);
      }
    }
  }
  free(cum);
  return 0;
}

static void twl_stream_free(twl_stream_t* s) {
  free(s->req_q0); free(s->req_kind); free(s->arrival); free(s->qstate); free(s->qlo); free(s->qhi);
  free(s->app_parent); free(s->app_key); free(s->app_delta);
  memset(s, 0, sizeof(*s));
}
#endif
