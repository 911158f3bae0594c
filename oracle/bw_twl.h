/* bw_twl.h -- the oracle's own restatement of the tree-workload generator (test infrastructure).
 *
 * The reference never generated the branching, multi-update, point-update workloads of the
 * benchmark configs; its generators are src/common.hpp:279-374 (one edit per state) and the
 * counter-based splitmix64 variant src/batch_engine_v7c.hpp:240-249.  This generator follows the
 * counter-based design of the latter (every draw is a pure function of (seed, stream, index),
 * so it parallelises and reproduces on any host) and the shapes written in the benchmark
 * configs.  The product has an independent C++ implementation of the same definition
 * (biworlddb-artifact_amd/csrc/host/tree_workload.cpp); tests/test_oracle.py checks that both
 * produce byte-identical workloads (same SHA-256 over every generated array).
 *
 * Definition (state 0 is an empty root; states 1..S are edited; each has U updates and Q
 * range queries answered after its own updates):
 *   base(stream)      = mix64(seed ^ stream*0xD1B54A32D192ED03)
 *   draw(stream, i)   = mix64(base(stream) + (i+1)*0x9E3779B97F4A7C15)     (splitmix64)
 *   u01(x)            = (x >> 11) * 2^-53
 *   bounded(x, n)     = x mod n
 * parents   cfg1: v-1-bounded(draw(1,v), min(v,32))
 *           cfg2/5: u01(draw(1,v)) < 0.9 ? v-1 : v-1-bounded(draw(2,v), min(v,1024))
 *           cfg4: chain continues (parent v-1) while the run is < 4096 and u01(draw(1,v)) < 63/64,
 *                 else v-1-bounded(draw(2,v), min(v,1024)) and a new run starts
 * keys      idx=(v-1)*U+i; cfg1 bounded(draw(3,idx),N); cfg2/4/5 Zipf(1.1) over [0,N) by inverse
 *           CDF of u01(draw(3,idx)); cfg3 dkey[Zipf(0.9) index over D distinct keys]
 * deltas    bounded(draw(4,idx), 2R+1) - R  (R = 1000, 10^6, 10^4 for cfg1, cfg2/4/5, cfg3)
 * queries   idx=(v-1)*Q+k; cfg1: a,b = bounded(draw(5|6,idx),N), [min, max+1)
 *           cfg2/4/5: len = clamp(floor(exp(u01(draw(5,idx))*log(N))),1,N),
 *                     lo = bounded(draw(6,idx), N-len+1), hi = lo+len
 *           cfg3: the same with N := 2^32-1 over raw 32-bit endpoints (resolved by rank)
 * cfg3 keys: draw(7,j)>>32 for j=0,1,... keeping first occurrences until D distinct, sorted.
 * Zipf(s) over M ranks: cum[r] = sum_{i<=r} pow(i+1,-s) (sequential double sum),
 *           index = first r with cum[r] > u*cum[M-1] (M-1 if none).
 */
#ifndef BW_TWL_H
#define BW_TWL_H
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t twl_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static inline uint64_t twl_draw(uint64_t seed, uint64_t stream, uint64_t idx) {
  const uint64_t base = twl_mix64(seed ^ (stream * 0xD1B54A32D192ED03ULL));
  return twl_mix64(base + (idx + 1) * 0x9E3779B97F4A7C15ULL);
}
static inline double twl_u01(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

typedef struct {
  int cfg;
  int64_t S; /* edited states */
  int U, Q;
  uint64_t N; /* key domain (cfg 1,2,4,5) or distinct-key count D (cfg 3) */
  uint64_t seed;
  /* generated */
  int32_t* parent;  /* S+1, parent[0] = -1 */
  uint32_t* key;    /* S*U raw keys */
  int64_t* delta;   /* S*U */
  uint32_t* qlo;    /* S*Q raw endpoints */
  uint32_t* qhi;
  uint32_t* dkey;   /* cfg3: D sorted distinct keys */
  uint64_t D;
} twl_t;

static double* twl_zipf_table(uint64_t M, double s) {
  double* cum = (double*)malloc(sizeof(double) * M);
  double acc = 0.0;
  for (uint64_t r = 0; r < M; ++r) {
    acc += pow((double)(r + 1), -s);
    cum[r] = acc;
  }
  return cum;
}
static inline uint64_t twl_zipf_pick(const double* cum, uint64_t M, double u) {
  const double t = u * cum[M - 1];
  uint64_t lo = 0, hi = M; /* first index with cum > t */
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (cum[mid] > t) hi = mid; else lo = mid + 1;
  }
  return lo < M ? lo : M - 1;
}
static int twl_cmp_u32(const void* a, const void* b) {
  const uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return (x > y) - (x < y);
}

static void twl_loglen_range(uint64_t dlen, uint64_t dpos, uint64_t n, uint32_t* lo, uint32_t* hi) {
  double l = floor(exp(twl_u01(dlen) * log((double)n)));
  uint64_t len = (l < 1.0) ? 1 : (uint64_t)l;
  if (len > n) len = n;
  const uint64_t a = dpos % (n - len + 1);
  *lo = (uint32_t)a;
  *hi = (uint32_t)(a + len);
}

static int twl_generate(twl_t* w) {
  const int64_t S = w->S;
  const uint64_t seed = w->seed;
  const int U = w->U, Q = w->Q;
  w->parent = (int32_t*)malloc(sizeof(int32_t) * (size_t)(S + 1));
  w->key = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)S * U);
  w->delta = (int64_t*)malloc(sizeof(int64_t) * (size_t)S * U);
  w->qlo = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)S * Q);
  w->qhi = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)S * Q);
  w->dkey = NULL;
  w->D = 0;
  if (!w->parent || !w->key || !w->delta || !w->qlo || !w->qhi) return -1;
  w->parent[0] = -1;
  int64_t run = 1;
  for (int64_t v = 1; v <= S; ++v) {
    int64_t p;
    if (w->cfg == 1) {
      const uint64_t win = (uint64_t)(v < 32 ? v : 32);
      p = v - 1 - (int64_t)(twl_draw(seed, 1, (uint64_t)v) % win);
    } else if (w->cfg == 4) {
      if (run < 4096 && twl_u01(twl_draw(seed, 1, (uint64_t)v)) < 63.0 / 64.0) {
        p = v - 1;
        ++run;
      } else {
        const uint64_t win = (uint64_t)(v < 1024 ? v : 1024);
        p = v - 1 - (int64_t)(twl_draw(seed, 2, (uint64_t)v) % win);
        run = 1;
      }
    } else {
      if (twl_u01(twl_draw(seed, 1, (uint64_t)v)) < 0.9) {
        p = v - 1;
      } else {
        const uint64_t win = (uint64_t)(v < 1024 ? v : 1024);
        p = v - 1 - (int64_t)(twl_draw(seed, 2, (uint64_t)v) % win);
      }
    }
    w->parent[v] = (int32_t)p;
  }
  /* keys */
  const size_t nu = (size_t)S * (size_t)U;
  if (w->cfg == 1) {
    for (size_t i = 0; i < nu; ++i) w->key[i] = (uint32_t)(twl_draw(seed, 3, i) % w->N);
  } else if (w->cfg == 3) {
    const uint64_t D = w->N;
    uint64_t cap = 1;
    while (cap < 2 * D) cap <<= 1;
    uint32_t* tab = (uint32_t*)calloc(cap, sizeof(uint32_t)); /* stores key+1, 0 = empty */
    uint8_t* zero_seen = (uint8_t*)calloc(1, 1);
    w->dkey = (uint32_t*)malloc(sizeof(uint32_t) * D);
    uint64_t got = 0;
    for (uint64_t j = 0; got < D; ++j) {
      const uint32_t c = (uint32_t)(twl_draw(seed, 7, j) >> 32);
      if (c == 0xFFFFFFFFu) { /* key+1 would wrap: tracked separately */
        if (!zero_seen[0]) { zero_seen[0] = 1; w->dkey[got++] = c; }
        continue;
      }
      uint64_t h = (twl_mix64((uint64_t)c) >> 7) & (cap - 1);
      int fresh = 1;
      while (tab[h] != 0) {
        if (tab[h] == c + 1u) { fresh = 0; break; }
        h = (h + 1) & (cap - 1);
      }
      if (fresh) { tab[h] = c + 1u; w->dkey[got++] = c; }
    }
    free(tab);
    free(zero_seen);
    qsort(w->dkey, D, sizeof(uint32_t), twl_cmp_u32);
    w->D = D;
    double* cum = twl_zipf_table(D, 0.9);
    for (size_t i = 0; i < nu; ++i)
      w->key[i] = w->dkey[twl_zipf_pick(cum, D, twl_u01(twl_draw(seed, 3, i)))];
    free(cum);
  } else {
    double* cum = twl_zipf_table(w->N, 1.1);
    for (size_t i = 0; i < nu; ++i)
      w->key[i] = (uint32_t)twl_zipf_pick(cum, w->N, twl_u01(twl_draw(seed, 3, i)));
    free(cum);
  }
  /* deltas */
  const int64_t R = (w->cfg == 1) ? 1000 : (w->cfg == 3 ? 10000 : 1000000);
  for (size_t i = 0; i < nu; ++i)
    w->delta[i] = (int64_t)(twl_draw(seed, 4, i) % (uint64_t)(2 * R + 1)) - R;
  /* queries */
  const size_t nq = (size_t)S * (size_t)Q;
  for (size_t i = 0; i < nq; ++i) {
    if (w->cfg == 1) {
      const uint32_t a = (uint32_t)(twl_draw(seed, 5, i) % w->N);
      const uint32_t b = (uint32_t)(twl_draw(seed, 6, i) % w->N);
      w->qlo[i] = a < b ? a : b;
      w->qhi[i] = (a < b ? b : a) + 1u;
    } else if (w->cfg == 3) {
      twl_loglen_range(twl_draw(seed, 5, i), twl_draw(seed, 6, i), 0xFFFFFFFFull, &w->qlo[i], &w->qhi[i]);
    } else {
      twl_loglen_range(twl_draw(seed, 5, i), twl_draw(seed, 6, i), w->N, &w->qlo[i], &w->qhi[i]);
    }
  }
  return 0;
}

static void twl_free(twl_t* w) {
  free(w->parent); free(w->key); free(w->delta); free(w->qlo); free(w->qhi); free(w->dkey);
  w->parent = NULL; w->key = NULL; w->delta = NULL; w->qlo = NULL; w->qhi = NULL; w->dkey = NULL;
}

/* Standard sizes of the five benchmark configs (S, U, Q, N). */
static void twl_defaults(int cfg, twl_t* w) {
  memset(w, 0, sizeof(*w));
  w->cfg = cfg;
  w->seed = 20261203ULL + (uint64_t)cfg;
  switch (cfg) {
    case 1: w->S = 1000; w->U = 64; w->Q = 16; w->N = 65536; break;
    case 3: w->S = 300000; w->U = 256; w->Q = 32; w->N = 4194304; break;
    default: w->S = 300000; w->U = 128; w->Q = 32; w->N = 1048576; break;
  }
}
#endif
