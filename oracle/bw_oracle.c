/* bw_oracle.c -- CPU oracle for the bwgpu batch path.  TEST INFRASTRUCTURE ONLY.
 *
 * The oracle the GPU path is checked against.  Only tests/, the package's smoke check and
 * bench.py's cpu_baseline leg run it; no product code links or calls it.
 *
 * Modes
 *   sha FILE                    SHA-256 of a file (FIPS 180-4, bw_sha256.h).
 *   ref  --fixture F --topology T --seed S --factor-period P --q Q --states K
 *        [--per-state] [--factor-pct X] --out ANS
 *       Restatement of the reference oracle (src/oracle.cpp:225-451) over the reference's
 *       declared workload (src/common.hpp:279-374): iterative DFS over the version tree,
 *       apply the node's edit on entry / restore it on exit, answer each state's first q ranges
 *       by a direct scan over all base + inserted rows accumulating amount*factor in __int128,
 *       checked into int64 (exit 6 on overflow, src/oracle.cpp:203-211).  K is the reference's
 *       compile-time kStates (src/common.hpp:184-192) given at run time.
 *   spec --fixture BASE --spec SPEC --out ANS
 *       Answers of a BWHHIST1 declared history spec (chain or branching; both header layouts)
 *       over its version-0 base fixture, by the same DFS (the reference's spec oracles,
 *       src/panel_real_history_oracle.cpp, src/panel_real_branch_oracle.cpp).
 *   tree --cfg C [--S n --U n --Q n --N n --seed n] [--algo fenwick|brute] --out ANS
 *       Oracle for the branching multi-update point-update workloads of the benchmark configs
 *       (bw_twl.h).  A point update (key, d) is the reference's OP_INSERT of a fresh row with
 *       t=key, factor 1 (src/oracle.cpp:280-330, 403), so answer[v][k] = sum of d over every
 *       update on the root->v path with lo <= key < hi.  "fenwick" is the rollback-Fenwick design
 *       of the reference's independent rf arm (src/arms_classic.hpp:326-340): DFS apply / query /
 *       undo with a Fenwick tree over keys (ranks for cfg 3).  "brute" walks the ancestor chain
 *       of every state and sums in __int128 (tiny inputs only; it shares no code with "fenwick").
 *       --workload-sha prints the SHA-256 of the generated arrays instead of solving.
 *   stream --form 4|5 [--cfg 3] [--S --U --N --seed (tree)] [--requests R --max-batch M --append-frac f
 *          --append-U n --rate r] [--algo fenwick|brute] [--threads T] [--stream-sha] --out ANS
 *       Answers of the request streams of the batched (4) and served (5) forms (bw_stream.h)
 *       over the tree extended by the appended states, in (request, query) order: the same
 *       rollback-Fenwick DFS with queries grouped by state, or the ancestor walk ("brute").  The
 *       served form runs over a config-2 tree, or a config-3 tree with --cfg 3 (raw 32-bit keys
 *       and endpoints; the Fenwick solvers rank them, the ancestor walk compares them raw).
 *
 * Output: answers as little-endian int64 in the reference's canonical order (state, query)
 * and one JSON line {"n_answers":..,"answers_sha256":..,"wall_s":..}.
 */
#include <errno.h>
#include <inttypes.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "bw_sha256.h"
#include "bw_twl.h"
#include "bw_stream.h"

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void die(int code, const char* msg) {
  fprintf(stderr, "bw_oracle: %s\n", msg);
  exit(code);
}

static int64_t checked_i64(__int128 x) {
  if (x < (__int128)INT64_MIN || x > (__int128)INT64_MAX) die(6, "query answer overflowed int64");
  return (int64_t)x;
}

static double g_solve_s = -1.0;
static void write_answers(const char* path, const int64_t* ans, size_t n, double t0) {
  if (path) {
    FILE* fp = fopen(path, "wb");
    if (!fp) die(4, "cannot write answers");
    if (n && fwrite(ans, sizeof(int64_t), n, fp) != n) die(4, "short write");
    if (fclose(fp) != 0) die(4, "close failed");
  }
  char hex[65];
  bw_sha256_hex(ans, n * sizeof(int64_t), hex);
  printf("{\"n_answers\":%zu,\"answers_sha256\":\"%s\",\"wall_s\":%.6f,\"solve_s\":%.6f}\n", n, hex, now_s() - t0,
         g_solve_s);
}

/* ------------------------------------------------------------------ reference semantics */
/* splitmix64 stream, src/common.hpp:40-54 */
typedef struct { uint64_t s; } ref_rng;
static void rr_init(ref_rng* r, uint64_t seed) { r->s = seed ? seed : 0x9E3779B97F4A7C15ULL; }
static uint64_t rr_next(ref_rng* r) {
  r->s += 0x9E3779B97F4A7C15ULL;
  uint64_t z = r->s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static uint32_t rr_bounded(ref_rng* r, uint32_t n) { return n ? (uint32_t)(rr_next(r) % n) : 0; }
static int64_t rr_range(ref_rng* r, int64_t lo, int64_t hi) {
  if (hi <= lo) return lo;
  return lo + (int64_t)(rr_next(r) % (uint64_t)(hi - lo));
}

typedef struct { int64_t t, amount; int32_t g, pad; } base_rec;
typedef struct {
  int64_t n, tmin, tmax;
  int32_t G;
  base_rec* rows;
  int64_t* factor;
} fixture_t;

static int load_fixture(const char* path, fixture_t* f) {
  FILE* fp = fopen(path, "rb");
  if (!fp) return 0;
  unsigned char h[64];
  if (fread(h, 1, 64, fp) != 64 || memcmp(h, "BWBASE02", 8) != 0) { fclose(fp); return 0; }
  memcpy(&f->n, h + 8, 8);
  memcpy(&f->G, h + 16, 4);
  memcpy(&f->tmin, h + 40, 8);
  memcpy(&f->tmax, h + 48, 8);
  if (f->n <= 0 || f->G <= 0) { fclose(fp); return 0; }
  f->rows = (base_rec*)malloc(sizeof(base_rec) * (size_t)f->n);
  f->factor = (int64_t*)malloc(sizeof(int64_t) * (size_t)f->G);
  if (fread(f->rows, sizeof(base_rec), (size_t)f->n, fp) != (size_t)f->n) { fclose(fp); return 0; }
  if (fread(f->factor, sizeof(int64_t), (size_t)f->G, fp) != (size_t)f->G) { fclose(fp); return 0; }
  fclose(fp);
  for (int32_t g = 0; g < f->G; ++g)
    if (f->factor[g] <= 0) return 0;
  return 1;
}

enum { OP_REPLACE = 0, OP_INSERT = 1, OP_DELETE = 2, OP_MOVE = 3, OP_FACTOR = 4 };
typedef struct { uint8_t op; int32_t g; int64_t id, val, t; } edit_t;

/* nested ranges of one state, src/common.hpp:229-246 */
static void gen_nested(ref_rng* r, int64_t tmin, int64_t tmax, int nq, int64_t* lo_out, int64_t* hi_out) {
  int64_t lo = tmin, hi = tmax + 1;
  const int64_t span0 = hi - lo;
  for (int k = 0; k < nq; ++k) {
    int64_t span = hi - lo;
    if (span <= 4096) {
      lo = tmin + span0 / 2;
      hi = lo + 2048;
      span = hi - lo;
    }
    const int64_t a = span * (int64_t)(3 + rr_bounded(r, 16)) / 100;
    const int64_t b = span * (int64_t)(3 + rr_bounded(r, 16)) / 100;
    lo += a;
    hi -= b;
    if (hi - lo < 1024) hi = lo + 1024;
    lo_out[k] = lo;
    hi_out[k] = hi;
  }
}

static int ref_parent(int topology, int32_t v) {
  if (v <= 0) return -1;
  if (topology == 0) return v - 1;
  if (topology == 1) return (v - 1) / 2;
  return 0;
}

/* The reference oracle's solve (src/oracle.cpp:225-451): iterative DFS over the version tree
 * (children in ascending order), apply a version's edit on entry, answer its q ranges by a direct
 * scan of every base + inserted row accumulating amount*factor in __int128 (checked into int64),
 * restore the edit on exit.  Inserted row k (id N + k) starts absent at (ins_t[k], ins_g[k],
 * amount 0), or at (0, 0, 0) when ins_t is NULL.  Ranges of version v at (v*q + k) when
 * per_state, else shared.  Writes ans[(v-1)*q + k] for v = 1..K-1. */
static void ref_solve(int64_t K, const int32_t* parent, const edit_t* ed, int64_t q, const int64_t* qlo,
                      const int64_t* qhi, int per_state, const fixture_t* fx, int64_t next_ins, const int64_t* ins_t,
                      const int32_t* ins_g, int64_t* ans) {
  const int64_t N = fx->n;
  const int32_t G = fx->G;
  /* mutable row state (src/oracle.cpp:240-278) */
  const int64_t M = N + next_ins;
  int64_t* t = (int64_t*)malloc(sizeof(int64_t) * (size_t)M);
  int64_t* amt = (int64_t*)malloc(sizeof(int64_t) * (size_t)M);
  int32_t* g = (int32_t*)malloc(sizeof(int32_t) * (size_t)M);
  uint8_t* ex = (uint8_t*)malloc((size_t)M);
  for (int64_t i = 0; i < N; ++i) { t[i] = fx->rows[i].t; amt[i] = fx->rows[i].amount; g[i] = fx->rows[i].g; ex[i] = 1; }
  for (int64_t k = 0; k < next_ins; ++k) {
    t[N + k] = ins_t ? ins_t[k] : 0; g[N + k] = ins_g ? ins_g[k] : 0; amt[N + k] = 0; ex[N + k] = 0;
  }
  int64_t* fac = (int64_t*)malloc(sizeof(int64_t) * (size_t)G);
  for (int32_t gg = 0; gg < G; ++gg) fac[gg] = fx->factor[gg];
  /* children in ascending order (CSR) */
  int64_t* koff = (int64_t*)calloc((size_t)K + 1, sizeof(int64_t));
  int32_t* kid = (int32_t*)malloc(sizeof(int32_t) * (size_t)K);
  for (int64_t v = 1; v < K; ++v) koff[parent[v] + 1]++;
  for (int64_t v = 0; v < K; ++v) koff[v + 1] += koff[v];
  {
    int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (size_t)K);
    memcpy(cur, koff, sizeof(int64_t) * (size_t)K);
    for (int64_t v = 1; v < K; ++v) kid[cur[parent[v]]++] = (int32_t)v;
    free(cur);
  }
  /* undo records */
  typedef struct { int64_t amt, t, fac; int32_t g; uint8_t ex; } undo_t;
  undo_t* un = (undo_t*)calloc((size_t)K, sizeof(undo_t));
  int64_t* stk_v = (int64_t*)malloc(sizeof(int64_t) * (size_t)K);
  int64_t* stk_c = (int64_t*)malloc(sizeof(int64_t) * (size_t)K);
  __int128* acc = (__int128*)malloc(sizeof(__int128) * (size_t)q);
  int64_t sp = 0;
  stk_v[0] = 0; stk_c[0] = koff[0]; sp = 1;
  while (sp > 0) {
    const int64_t v = stk_v[sp - 1];
    if (stk_c[sp - 1] < koff[v + 1]) {
      const int32_t c = kid[stk_c[sp - 1]++];
      /* apply (src/oracle.cpp:280-330) */
      const edit_t* e = &ed[c];
      undo_t* u = &un[c];
      switch (e->op) {
        case OP_REPLACE: u->amt = amt[e->id]; u->ex = ex[e->id]; amt[e->id] = e->val; ex[e->id] = 1; break;
        case OP_INSERT:
        case OP_MOVE:
          u->amt = amt[e->id]; u->t = t[e->id]; u->g = g[e->id]; u->ex = ex[e->id];
          t[e->id] = e->t; g[e->id] = e->g; amt[e->id] = e->val; ex[e->id] = 1; break;
        case OP_DELETE: u->ex = ex[e->id]; ex[e->id] = 0; break;
        case OP_FACTOR: u->fac = fac[e->g]; fac[e->g] = e->val; break;
        default: die(7, "unknown op");
      }
      /* answer: direct scan of every row (src/oracle.cpp:359-422) */
      const int64_t* lo = per_state ? &qlo[(size_t)c * q] : qlo;
      const int64_t* hi = per_state ? &qhi[(size_t)c * q] : qhi;
      for (int64_t k = 0; k < q; ++k) acc[k] = 0;
      for (int64_t i = 0; i < M; ++i) {
        if (!ex[i]) continue;
        const int64_t ti = t[i];
        const __int128 w = (__int128)amt[i] * (__int128)fac[g[i]];
        for (int64_t k = 0; k < q; ++k)
          if (ti >= lo[k] && ti < hi[k]) acc[k] += w;
      }
      for (int64_t k = 0; k < q; ++k) ans[(size_t)(c - 1) * q + k] = checked_i64(acc[k]);
      stk_v[sp] = c; stk_c[sp] = koff[c]; ++sp;
    } else {
      if (v > 0) {
        /* revert (src/oracle.cpp:332-355) */
        const edit_t* e = &ed[v];
        const undo_t* u = &un[v];
        switch (e->op) {
          case OP_REPLACE: amt[e->id] = u->amt; ex[e->id] = u->ex; break;
          case OP_INSERT:
          case OP_MOVE: t[e->id] = u->t; g[e->id] = u->g; amt[e->id] = u->amt; ex[e->id] = u->ex; break;
          case OP_DELETE: ex[e->id] = u->ex; break;
          case OP_FACTOR: fac[e->g] = u->fac; break;
          default: break;
        }
      }
      --sp;
    }
  }
  free(t); free(amt); free(g); free(ex); free(fac); free(koff); free(kid); free(un); free(stk_v); free(stk_c);
  free(acc);
}

static int mode_ref(int argc, char** argv) {
  const char* fx = NULL; const char* out = NULL;
  int64_t topo = -1, seed = 0, fp = -1, q = -1, K = 2049, fpct = -1;
  int per_state = 0;
  for (int i = 2; i < argc; ++i) {
    if (!strcmp(argv[i], "--per-state")) { per_state = 1; continue; }
    if (i + 1 >= argc) die(2, "missing value");
    const char* k = argv[i]; const char* v = argv[++i];
    if (!strcmp(k, "--fixture")) fx = v;
    else if (!strcmp(k, "--out")) out = v;
    else if (!strcmp(k, "--topology")) topo = atoll(v);
    else if (!strcmp(k, "--seed")) seed = atoll(v);
    else if (!strcmp(k, "--factor-period")) fp = atoll(v);
    else if (!strcmp(k, "--q")) q = atoll(v);
    else if (!strcmp(k, "--states")) K = atoll(v);
    else if (!strcmp(k, "--factor-pct")) fpct = atoll(v);
    else die(2, "unknown argument");
  }
  if (!fx || topo < 0 || topo > 2 || fp < 0 || q < 1 || K < 2) die(2, "usage");
  const double t0 = now_s();
  fixture_t f;
  if (!load_fixture(fx, &f)) die(1, "cannot load fixture");
  const int64_t N = f.n;
  const int32_t G = f.G;
  const uint64_t useed = (uint64_t)seed, utopo = (uint64_t)topo, ufp = (uint64_t)fp;
  /* declared queries, src/common.hpp:292-303 */
  const size_t nqs = per_state ? (size_t)K : 1;
  int64_t* qlo = (int64_t*)calloc(nqs * (size_t)q, sizeof(int64_t));
  int64_t* qhi = (int64_t*)calloc(nqs * (size_t)q, sizeof(int64_t));
  if (per_state) {
    for (int64_t v = 1; v < K; ++v) {
      ref_rng qr;
      rr_init(&qr, useed * 1000000007ULL + utopo * 65537ULL + ufp * 131ULL + (uint64_t)v * 2654435761ULL + 17ULL);
      gen_nested(&qr, f.tmin, f.tmax, (int)q, &qlo[(size_t)v * q], &qhi[(size_t)v * q]);
    }
  } else {
    ref_rng qr;
    rr_init(&qr, useed * 1000000007ULL + utopo * 65537ULL + ufp * 131ULL + 17ULL);
    gen_nested(&qr, f.tmin, f.tmax, (int)q, qlo, qhi);
  }
  /* edits, src/common.hpp:304-355 */
  edit_t* ed = (edit_t*)calloc((size_t)K, sizeof(edit_t));
  int64_t* gfac = (int64_t*)malloc(sizeof(int64_t) * (size_t)G);
  for (int32_t g = 0; g < G; ++g) gfac[g] = 1000;
  ref_rng rng;
  rr_init(&rng, useed * 1000003ULL + utopo * 7919ULL + ufp * 104729ULL + 20261025ULL);
  int32_t next_group = 0;
  int64_t next_ins = 0;
  int64_t* ins_t = (int64_t*)malloc(sizeof(int64_t) * (size_t)K);
  int32_t* ins_g = (int32_t*)malloc(sizeof(int32_t) * (size_t)K);
  for (int64_t v = 1; v < K; ++v) {
    edit_t e;
    memset(&e, 0, sizeof(e));
    int factor_step;
    if (fpct >= 0) factor_step = rr_bounded(&rng, 100) < (uint32_t)fpct;
    else factor_step = (fp > 0) && (v % fp == 0);
    if (factor_step) {
      const int32_t g = next_group++ % G;
      const int64_t cur = gfac[g];
      const int64_t nf = (cur == 950) ? 1000 : (cur == 1000 ? 1050 : 950);
      gfac[g] = nf;
      e.op = OP_FACTOR; e.g = g; e.val = nf;
    } else {
      const uint32_t roll = rr_bounded(&rng, 100);
      if (roll < 55) {
        e.op = OP_REPLACE;
        e.id = rr_range(&rng, 0, N);
        e.val = f.rows[e.id].amount + rr_range(&rng, -100000, 100000);
      } else if (roll < 75) {
        e.op = OP_INSERT;
        e.id = N + next_ins;
        e.t = rr_range(&rng, f.tmin, f.tmax + 1);
        e.g = (int32_t)rr_bounded(&rng, (uint32_t)G);
        const int64_t base_amt = f.rows[rr_bounded(&rng, (uint32_t)N)].amount;
        e.val = base_amt + rr_range(&rng, -100000, 100000);
        ins_t[next_ins] = e.t;
        ins_g[next_ins] = e.g;
        ++next_ins;
      } else if (roll < 90) {
        e.op = OP_DELETE;
        e.id = rr_range(&rng, 0, N);
      } else {
        e.op = OP_MOVE;
        e.id = rr_range(&rng, 0, N);
        e.t = rr_range(&rng, f.tmin, f.tmax + 1);
        e.g = (int32_t)rr_bounded(&rng, (uint32_t)G);
        e.val = f.rows[e.id].amount + rr_range(&rng, -100000, 100000);
      }
    }
    ed[v] = e;
  }
  int32_t* par = (int32_t*)malloc(sizeof(int32_t) * (size_t)K);
  par[0] = -1;
  for (int64_t v = 1; v < K; ++v) par[v] = ref_parent((int)topo, (int32_t)v);
  const size_t nans = (size_t)(K - 1) * (size_t)q;
  int64_t* ans = (int64_t*)calloc(nans ? nans : 1, sizeof(int64_t));
  ref_solve(K, par, ed, q, qlo, qhi, per_state, &f, next_ins, ins_t, ins_g, ans);
  write_answers(out, ans, nans, t0);
  return 0;
}

/* ------------------------------------------------------------------ BWHHIST1 specs */
/* A declared history spec (chain histories: src/real_history_spec.hpp:41-71, 4-field header;
 * branching histories: src/real_branch_spec.hpp:45-71, 5-field header with a topology flag):
 * "BWHHIST1", i32 kstates, q, fp, G [, topology], 12 i64 (n_base, n_ins, t_min, t_max, n_ep,
 * period_lo, period_hi, bucket, n_inserts, n_deletes, n_credits, n_factor_edits), parent[K] i32,
 * edit[K] (40-byte records: op u8, pad[3], g i32, pad i32, id i64, val i64, t i64), queries[K*q]
 * (lo, hi i64), ins_t[n_ins] i64, ins_g[n_ins] i32, endpoints[n_ep] i64.  Both layouts share the
 * magic: the one whose arrays add up to the file's size is taken. */
typedef struct {
  int32_t K, q, fp, G, topology, header_fields;
  int64_t n_base, n_ins, n_ep;
  int32_t* parent;
  edit_t* ed;
  int64_t *qlo, *qhi;
} spec_t;

static int load_spec(const char* path, spec_t* s) {
  FILE* fp = fopen(path, "rb");
  if (!fp) return 0;
  fseek(fp, 0, SEEK_END);
  const long size = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  unsigned char h[128];
  if (size < 124 || fread(h, 1, 124, fp) != 124 || memcmp(h, "BWHHIST1", 8) != 0) { fclose(fp); return 0; }
  int found = 0;
  for (int nf = 4; nf <= 5 && !found; ++nf) {
    int32_t i4[5] = {0, 0, 0, 0, 0};
    int64_t i8[12];
    memcpy(i4, h + 8, 4 * (size_t)nf);
    memcpy(i8, h + 8 + 4 * nf, 96);
    const int64_t K = i4[0], q = i4[1], nins = i8[1], nep = i8[4];
    if (K < 2 || q < 1 || q > 1024 || nins < 0 || nep < 0 || i4[3] < 1) continue;
    const int64_t want = 8 + 4 * nf + 96 + K * 4 + K * 40 + K * q * 16 + nins * 12 + nep * 8;
    if (want != size) continue;
    found = 1;
    s->K = i4[0]; s->q = i4[1]; s->fp = i4[2]; s->G = i4[3]; s->topology = nf == 5 ? i4[4] : 0;
    s->header_fields = nf;
    s->n_base = i8[0]; s->n_ins = nins; s->n_ep = nep;
  }
  if (!found) { fclose(fp); return 0; }
  fseek(fp, 8 + 4 * s->header_fields + 96, SEEK_SET);
  const size_t K = (size_t)s->K, nq = K * (size_t)s->q;
  s->parent = (int32_t*)malloc(4 * K);
  s->ed = (edit_t*)calloc(K, sizeof(edit_t));
  s->qlo = (int64_t*)malloc(8 * nq);
  s->qhi = (int64_t*)malloc(8 * nq);
  unsigned char* raw = (unsigned char*)malloc(40 * K);
  int64_t* qr = (int64_t*)malloc(16 * nq);
  int ok = fread(s->parent, 4, K, fp) == K && fread(raw, 40, K, fp) == K && fread(qr, 16, nq, fp) == nq;
  fclose(fp);
  for (size_t v = 0; ok && v < K; ++v) {
    const unsigned char* r = raw + 40 * v;
    s->ed[v].op = r[0];
    memcpy(&s->ed[v].g, r + 4, 4);
    memcpy(&s->ed[v].id, r + 16, 8);
    memcpy(&s->ed[v].val, r + 24, 8);
    memcpy(&s->ed[v].t, r + 32, 8);
  }
  for (size_t i = 0; ok && i < nq; ++i) { s->qlo[i] = qr[2 * i]; s->qhi[i] = qr[2 * i + 1]; }
  free(raw);
  free(qr);
  return ok;
}

/* spec --fixture BASE --spec SPEC --out ANS: the answers of every version of a declared history
 * (both layouts), the reference's spec oracles' semantics (src/panel_real_history_oracle.cpp:
 * 63-110 replays the chain; src/panel_real_branch_oracle.cpp:36-83,120-170 the enter/exit DFS):
 * inserted rows start absent at (0, 0, 0); a version's value is the sum over present rows with
 * lo <= t < hi of factor[g] * amount. */
static int mode_spec(int argc, char** argv) {
  const char* fx = NULL; const char* sp = NULL; const char* out = NULL;
  for (int i = 2; i < argc; ++i) {
    if (i + 1 >= argc) die(2, "missing value");
    const char* k = argv[i]; const char* v = argv[++i];
    if (!strcmp(k, "--fixture")) fx = v;
    else if (!strcmp(k, "--spec")) sp = v;
    else if (!strcmp(k, "--out")) out = v;
    else die(2, "unknown argument");
  }
  if (!fx || !sp) die(2, "usage: spec --fixture BASE --spec SPEC --out ANS");
  const double t0 = now_s();
  fixture_t f;
  spec_t s;
  memset(&s, 0, sizeof(s));
  if (!load_fixture(fx, &f)) die(1, "cannot load fixture");
  if (!load_spec(sp, &s)) die(1, "cannot load spec");
  if (s.n_base != f.n || s.G != f.G) die(3, "spec/fixture base mismatch");
  const int64_t M = f.n + s.n_ins;
  for (int64_t v = 1; v < s.K; ++v) {
    if (s.parent[v] < 0 || s.parent[v] >= v) die(3, "spec parent out of order");
    const edit_t* e = &s.ed[v];
    if (e->op > OP_FACTOR) die(3, "spec edit op");
    if (e->op == OP_FACTOR ? (e->g < 0 || e->g >= f.G) : (e->id < 0 || e->id >= M)) die(3, "spec edit target");
    if ((e->op == OP_INSERT || e->op == OP_MOVE) && (e->g < 0 || e->g >= f.G)) die(3, "spec edit group");
  }
  const size_t nans = (size_t)(s.K - 1) * (size_t)s.q;
  int64_t* ans = (int64_t*)calloc(nans ? nans : 1, sizeof(int64_t));
  ref_solve(s.K, s.parent, s.ed, s.q, s.qlo, s.qhi, 1, &f, s.n_ins, NULL, NULL, ans);
  write_answers(out, ans, nans, t0);
  return 0;
}

/* ------------------------------------------------------------------ tree workloads */
static uint64_t lower_bound_u32(const uint32_t* a, uint64_t n, uint64_t x) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if ((uint64_t)a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

static int mode_tree(int argc, char** argv) {
  int cfg = 2;
  for (int i = 2; i + 1 < argc; ++i)
    if (!strcmp(argv[i], "--cfg")) cfg = atoi(argv[i + 1]);
  twl_t w;
  twl_defaults(cfg, &w);
  const char* out = NULL;
  const char* algo = "fenwick";
  const char* dump = NULL;
  int wsha = 0;
  for (int i = 2; i < argc; ++i) {
    if (!strcmp(argv[i], "--workload-sha")) { wsha = 1; continue; }
    if (i + 1 >= argc) die(2, "missing value");
    const char* k = argv[i]; const char* v = argv[++i];
    if (!strcmp(k, "--cfg")) continue;
    else if (!strcmp(k, "--S")) w.S = atoll(v);
    else if (!strcmp(k, "--U")) w.U = atoi(v);
    else if (!strcmp(k, "--Q")) w.Q = atoi(v);
    else if (!strcmp(k, "--N")) w.N = strtoull(v, NULL, 10);
    else if (!strcmp(k, "--seed")) w.seed = strtoull(v, NULL, 10);
    else if (!strcmp(k, "--algo")) algo = v;
    else if (!strcmp(k, "--out")) out = v;
    else if (!strcmp(k, "--dump")) dump = v;
    else die(2, "unknown argument");
  }
  if (w.S < 1 || w.U < 0 || w.Q < 1 || w.N < 1 || w.N > 0xFFFFFFFFull) die(2, "bad sizes");
  if (cfg < 1 || cfg > 5) die(2, "bad cfg");
  const double t0 = now_s();
  if (twl_generate(&w) != 0) die(3, "out of memory");
  const int64_t S = w.S;
  const int U = w.U, Q = w.Q;
  const size_t nu = (size_t)S * U, nq = (size_t)S * Q;
  if (dump) {
    /* plain arrays for oracle/ref_tree_harness.cpp: S,U,Q (i64) then the generated arrays */
    FILE* fp = fopen(dump, "wb");
    if (!fp) die(4, "cannot write dump");
    const int64_t hdr[3] = {S, U, Q};
    int ok = fwrite(hdr, sizeof(hdr), 1, fp) == 1;
    ok &= fwrite(w.parent + 1, sizeof(int32_t), (size_t)S, fp) == (size_t)S;
    ok &= fwrite(w.key, sizeof(uint32_t), nu, fp) == nu;
    ok &= fwrite(w.delta, sizeof(int64_t), nu, fp) == nu;
    ok &= fwrite(w.qlo, sizeof(uint32_t), nq, fp) == nq;
    ok &= fwrite(w.qhi, sizeof(uint32_t), nq, fp) == nq;
    if (fclose(fp) != 0 || !ok) die(4, "short dump write");
  }
  if (wsha) {
    bw_sha256_ctx c;
    bw_sha256_init(&c);
    bw_sha256_update(&c, w.parent + 1, sizeof(int32_t) * (size_t)S);
    bw_sha256_update(&c, w.key, sizeof(uint32_t) * nu);
    bw_sha256_update(&c, w.delta, sizeof(int64_t) * nu);
    bw_sha256_update(&c, w.qlo, sizeof(uint32_t) * nq);
    bw_sha256_update(&c, w.qhi, sizeof(uint32_t) * nq);
    if (w.dkey) bw_sha256_update(&c, w.dkey, sizeof(uint32_t) * w.D);
    char hex[65];
    bw_sha256_final_hex(&c, hex);
    printf("{\"workload_sha256\":\"%s\"}\n", hex);
    twl_free(&w);
    return 0;
  }
  /* key -> Fenwick coordinate: the key itself, or its rank among the distinct keys (cfg 3) */
  const uint64_t NK = (cfg == 3) ? w.D : w.N;
  int64_t* ans = (int64_t*)calloc(nq, sizeof(int64_t));
  const double ts0 = now_s();
  if (!strcmp(algo, "brute")) {
    /* independent: walk the ancestor chain, compare raw keys against raw endpoints */
    __int128* acc = (__int128*)malloc(sizeof(__int128) * (size_t)Q);
    for (int64_t v = 1; v <= S; ++v) {
      for (int k = 0; k < Q; ++k) acc[k] = 0;
      for (int64_t a = v; a > 0; a = w.parent[a]) {
        const size_t base = (size_t)(a - 1) * U;
        for (int i = 0; i < U; ++i) {
          const uint64_t key = w.key[base + i];
          for (int k = 0; k < Q; ++k) {
            const size_t qi = (size_t)(v - 1) * Q + k;
            if (key >= w.qlo[qi] && key < w.qhi[qi]) acc[k] += w.delta[base + i];
          }
        }
      }
      for (int k = 0; k < Q; ++k) ans[(size_t)(v - 1) * Q + k] = checked_i64(acc[k]);
    }
    free(acc);
  } else {
    uint32_t* krank = (uint32_t*)malloc(sizeof(uint32_t) * (nu ? nu : 1));
    uint32_t* lrank = (uint32_t*)malloc(sizeof(uint32_t) * nq);
    uint32_t* hrank = (uint32_t*)malloc(sizeof(uint32_t) * nq);
    for (size_t i = 0; i < nu; ++i)
      krank[i] = (cfg == 3) ? (uint32_t)lower_bound_u32(w.dkey, w.D, w.key[i]) : w.key[i];
    for (size_t i = 0; i < nq; ++i) {
      lrank[i] = (cfg == 3) ? (uint32_t)lower_bound_u32(w.dkey, w.D, w.qlo[i]) : w.qlo[i];
      hrank[i] = (cfg == 3) ? (uint32_t)lower_bound_u32(w.dkey, w.D, w.qhi[i]) : w.qhi[i];
    }
    int64_t* fen = (int64_t*)calloc((size_t)NK + 1, sizeof(int64_t));
    int64_t* koff = (int64_t*)calloc((size_t)S + 2, sizeof(int64_t));
    int32_t* kid = (int32_t*)malloc(sizeof(int32_t) * (size_t)S);
    for (int64_t v = 1; v <= S; ++v) koff[w.parent[v] + 1]++;
    for (int64_t v = 0; v <= S; ++v) koff[v + 1] += koff[v];
    {
      int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (size_t)(S + 1));
      memcpy(cur, koff, sizeof(int64_t) * (size_t)(S + 1));
      for (int64_t v = 1; v <= S; ++v) kid[cur[w.parent[v]]++] = (int32_t)v;
      free(cur);
    }
    int64_t* stk_v = (int64_t*)malloc(sizeof(int64_t) * (size_t)(S + 1));
    int64_t* stk_c = (int64_t*)malloc(sizeof(int64_t) * (size_t)(S + 1));
    int64_t sp = 1;
    stk_v[0] = 0; stk_c[0] = koff[0];
    while (sp > 0) {
      const int64_t v = stk_v[sp - 1];
      if (stk_c[sp - 1] < koff[v + 1]) {
        const int64_t c = kid[stk_c[sp - 1]++];
        const size_t base = (size_t)(c - 1) * U;
        for (int i = 0; i < U; ++i)
          for (uint64_t j = (uint64_t)krank[base + i] + 1; j <= NK; j += j & (~j + 1)) fen[j] += w.delta[base + i];
        for (int k = 0; k < Q; ++k) {
          const size_t qi = (size_t)(c - 1) * Q + k;
          int64_t s = 0;
          for (uint64_t j = hrank[qi]; j > 0; j -= j & (~j + 1)) s += fen[j];
          for (uint64_t j = lrank[qi]; j > 0; j -= j & (~j + 1)) s -= fen[j];
          ans[qi] = s;
        }
        stk_v[sp] = c; stk_c[sp] = koff[c]; ++sp;
      } else {
        if (v > 0) {
          const size_t base = (size_t)(v - 1) * U;
          for (int i = 0; i < U; ++i)
            for (uint64_t j = (uint64_t)krank[base + i] + 1; j <= NK; j += j & (~j + 1)) fen[j] -= w.delta[base + i];
        }
        --sp;
      }
    }
    free(fen); free(koff); free(kid); free(stk_v); free(stk_c); free(krank); free(lrank); free(hrank);
  }
  g_solve_s = now_s() - ts0;
  write_answers(out, ans, nq, t0);
  free(ans);
  twl_free(&w);
  return 0;
}

/* ------------------------------------------------------------------ request streams */
/* Rollback-Fenwick DFS over a tree given by parent[] (states 1..ns-1, parent[v] < v), state v's
 * updates being [upd0[v], upd0[v+1]) of (krank, delta); query q asks state qstate[q] over the
 * coordinate range [lrank[q], hrank[q]).  Same algebra as mode_tree's solver; it serves the
 * streams, whose queries are arbitrary (state, range) pairs and whose trees grow by appends. */
static void solve_states(int64_t ns, const int32_t* parent, const int64_t* upd0, const uint32_t* krank,
                         const int64_t* delta, uint64_t NK, int64_t nq, const int32_t* qstate,
                         const uint32_t* lrank, const uint32_t* hrank, int64_t* ans) {
  int64_t* qoff = (int64_t*)calloc((size_t)ns + 1, sizeof(int64_t));
  int64_t* qid = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nq ? nq : 1));
  for (int64_t q = 0; q < nq; ++q) qoff[qstate[q] + 1]++;
  for (int64_t v = 0; v < ns; ++v) qoff[v + 1] += qoff[v];
  {
    int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (size_t)ns);
    memcpy(cur, qoff, sizeof(int64_t) * (size_t)ns);
    for (int64_t q = 0; q < nq; ++q) qid[cur[qstate[q]]++] = q;
    free(cur);
  }
  int64_t* fen = (int64_t*)calloc((size_t)NK + 1, sizeof(int64_t));
  int64_t* koff = (int64_t*)calloc((size_t)ns + 1, sizeof(int64_t));
  int32_t* kid = (int32_t*)malloc(sizeof(int32_t) * (size_t)ns);
  for (int64_t v = 1; v < ns; ++v) koff[parent[v] + 1]++;
  for (int64_t v = 0; v < ns; ++v) koff[v + 1] += koff[v];
  {
    int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (size_t)ns);
    memcpy(cur, koff, sizeof(int64_t) * (size_t)ns);
    for (int64_t v = 1; v < ns; ++v) kid[cur[parent[v]]++] = (int32_t)v;
    free(cur);
  }
  int64_t* stk_v = (int64_t*)malloc(sizeof(int64_t) * (size_t)ns);
  int64_t* stk_c = (int64_t*)malloc(sizeof(int64_t) * (size_t)ns);
  int64_t sp = 1;
  stk_v[0] = 0;
  stk_c[0] = koff[0];
  while (sp > 0) {
    const int64_t v = stk_v[sp - 1];
    if (stk_c[sp - 1] < koff[v + 1]) {
      const int64_t c = kid[stk_c[sp - 1]++];
      for (int64_t i = upd0[c]; i < upd0[c + 1]; ++i)
        for (uint64_t j = (uint64_t)krank[i] + 1; j <= NK; j += j & (~j + 1)) fen[j] += delta[i];
      for (int64_t k = qoff[c]; k < qoff[c + 1]; ++k) {
        const int64_t q = qid[k];
        int64_t r = 0;
        for (uint64_t j = hrank[q]; j > 0; j -= j & (~j + 1)) r += fen[j];
        for (uint64_t j = lrank[q]; j > 0; j -= j & (~j + 1)) r -= fen[j];
        ans[q] = r;
      }
      stk_v[sp] = c;
      stk_c[sp] = koff[c];
      ++sp;
    } else {
      for (int64_t i = upd0[v]; i < upd0[v + 1]; ++i)
        for (uint64_t j = (uint64_t)krank[i] + 1; j <= NK; j += j & (~j + 1)) fen[j] -= delta[i];
      --sp;
    }
  }
  free(qoff); free(qid); free(fen); free(koff); free(kid); free(stk_v); free(stk_c);
}

/* The same DFS cut into `threads` contiguous pieces of its step sequence, one per OpenMP thread.
 * A first pass walks the tree without arithmetic and records, at the start of every piece, the
 * DFS stack (the root path of the state being entered and the next child of each); a piece's
 * Fenwick tree starts as the prefix sums of its root path's updates (built in O(NK)), then the
 * piece runs its steps as the sequential solver does.  Every query is answered by exactly one
 * piece, with the same arithmetic, so the answers equal solve_states' bit for bit. */
static void solve_states_par(int64_t ns, const int32_t* parent, const int64_t* upd0, const uint32_t* krank,
                             const int64_t* delta, uint64_t NK, int64_t nq, const int32_t* qstate,
                             const uint32_t* lrank, const uint32_t* hrank, int64_t* ans, int threads) {
  int64_t* qoff = (int64_t*)calloc((size_t)ns + 1, sizeof(int64_t));
  int64_t* qid = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nq ? nq : 1));
  for (int64_t q = 0; q < nq; ++q) qoff[qstate[q] + 1]++;
  for (int64_t v = 0; v < ns; ++v) qoff[v + 1] += qoff[v];
  {
    int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (size_t)ns);
    memcpy(cur, qoff, sizeof(int64_t) * (size_t)ns);
    for (int64_t q = 0; q < nq; ++q) qid[cur[qstate[q]]++] = q;
    free(cur);
  }
  int64_t* koff = (int64_t*)calloc((size_t)ns + 1, sizeof(int64_t));
  int32_t* kid = (int32_t*)malloc(sizeof(int32_t) * (size_t)ns);
  for (int64_t v = 1; v < ns; ++v) koff[parent[v] + 1]++;
  for (int64_t v = 0; v < ns; ++v) koff[v + 1] += koff[v];
  {
    int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (size_t)ns);
    memcpy(cur, koff, sizeof(int64_t) * (size_t)ns);
    for (int64_t v = 1; v < ns; ++v) kid[cur[parent[v]]++] = (int32_t)v;
    free(cur);
  }
  /* the DFS loop runs 2 * (ns - 1) + 1 iterations (an entry per child, an exit per state) */
  const int64_t iters = 2 * (ns - 1) + 1;
  if (threads < 1) threads = 1;
  int64_t* cut = (int64_t*)malloc(sizeof(int64_t) * (size_t)(threads + 1));
  for (int t = 0; t <= threads; ++t) cut[t] = iters * t / threads;
  int64_t** snap_v = (int64_t**)calloc((size_t)threads, sizeof(int64_t*));
  int64_t** snap_c = (int64_t**)calloc((size_t)threads, sizeof(int64_t*));
  int64_t* snap_sp = (int64_t*)calloc((size_t)threads, sizeof(int64_t));
  {
    int64_t* stk_v = (int64_t*)malloc(sizeof(int64_t) * (size_t)ns);
    int64_t* stk_c = (int64_t*)malloc(sizeof(int64_t) * (size_t)ns);
    int64_t sp = 1, it = 0;
    int t = 0;
    stk_v[0] = 0;
    stk_c[0] = koff[0];
    while (sp > 0) {
      while (t < threads && cut[t] == it) {
        snap_sp[t] = sp;
        snap_v[t] = (int64_t*)malloc(sizeof(int64_t) * (size_t)sp);
        snap_c[t] = (int64_t*)malloc(sizeof(int64_t) * (size_t)sp);
        memcpy(snap_v[t], stk_v, sizeof(int64_t) * (size_t)sp);
        memcpy(snap_c[t], stk_c, sizeof(int64_t) * (size_t)sp);
        ++t;
      }
      const int64_t v = stk_v[sp - 1];
      if (stk_c[sp - 1] < koff[v + 1]) {
        const int64_t c = kid[stk_c[sp - 1]++];
        stk_v[sp] = c;
        stk_c[sp] = koff[c];
        ++sp;
      } else {
        --sp;
      }
      ++it;
    }
    free(stk_v);
    free(stk_c);
  }
#pragma omp parallel for num_threads(threads) schedule(static, 1)
  for (int t = 0; t < threads; ++t) {
    if (!snap_v[t]) continue; /* an empty piece */
    int64_t* fen = (int64_t*)calloc((size_t)NK + 1, sizeof(int64_t));
    int64_t* stk_v = (int64_t*)malloc(sizeof(int64_t) * (size_t)ns);
    int64_t* stk_c = (int64_t*)malloc(sizeof(int64_t) * (size_t)ns);
    int64_t sp = snap_sp[t];
    memcpy(stk_v, snap_v[t], sizeof(int64_t) * (size_t)sp);
    memcpy(stk_c, snap_c[t], sizeof(int64_t) * (size_t)sp);
    /* the root path's updates as point values, then the Fenwick tree over them in O(NK) */
    for (int64_t d = 1; d < sp; ++d)
      for (int64_t i = upd0[stk_v[d]]; i < upd0[stk_v[d] + 1]; ++i) fen[(uint64_t)krank[i] + 1] += delta[i];
    for (uint64_t j = 1; j <= NK; ++j) {
      const uint64_t up = j + (j & (~j + 1));
      if (up <= NK) fen[up] += fen[j];
    }
    for (int64_t it = cut[t]; it < cut[t + 1] && sp > 0; ++it) {
      const int64_t v = stk_v[sp - 1];
      if (stk_c[sp - 1] < koff[v + 1]) {
        const int64_t c = kid[stk_c[sp - 1]++];
        for (int64_t i = upd0[c]; i < upd0[c + 1]; ++i)
          for (uint64_t j = (uint64_t)krank[i] + 1; j <= NK; j += j & (~j + 1)) fen[j] += delta[i];
        for (int64_t k = qoff[c]; k < qoff[c + 1]; ++k) {
          const int64_t q = qid[k];
          int64_t r = 0;
          for (uint64_t j = hrank[q]; j > 0; j -= j & (~j + 1)) r += fen[j];
          for (uint64_t j = lrank[q]; j > 0; j -= j & (~j + 1)) r -= fen[j];
          ans[q] = r;
        }
        stk_v[sp] = c;
        stk_c[sp] = koff[c];
        ++sp;
      } else {
        for (int64_t i = upd0[v]; i < upd0[v + 1]; ++i)
          for (uint64_t j = (uint64_t)krank[i] + 1; j <= NK; j += j & (~j + 1)) fen[j] -= delta[i];
        --sp;
      }
    }
    free(fen);
    free(stk_v);
    free(stk_c);
  }
  for (int t = 0; t < threads; ++t) {
    free(snap_v[t]);
    free(snap_c[t]);
  }
  free(snap_v); free(snap_c); free(snap_sp); free(cut);
  free(qoff); free(qid); free(koff); free(kid);
}

/* Independent check for small streams: walk each query's ancestor chain over raw keys. */
static void solve_states_brute(const int32_t* parent, const int64_t* upd0, const uint32_t* key, const int64_t* delta,
                               int64_t nq, const int32_t* qstate, const uint32_t* qlo, const uint32_t* qhi,
                               int64_t* ans) {
  for (int64_t q = 0; q < nq; ++q) {
    __int128 acc = 0;
    for (int64_t a = qstate[q]; a > 0; a = parent[a])
      for (int64_t i = upd0[a]; i < upd0[a + 1]; ++i)
        if (key[i] >= qlo[q] && key[i] < qhi[q]) acc += delta[i];
    ans[q] = checked_i64(acc);
  }
}

static int mode_stream(int argc, char** argv) {
  int form = 4;
  for (int i = 2; i + 1 < argc; ++i)
    if (!strcmp(argv[i], "--form")) form = atoi(argv[i + 1]);
  if (form != 4 && form != 5) die(2, "--form must be 4 or 5");
  /* the served form's tree: config 2 (keys are their own ranks) or config 3 (--cfg 3: distinct
   * 32-bit keys, answered through ranks) */
  int tcfg = form == 4 ? 4 : 2;
  for (int i = 2; i + 1 < argc; ++i)
    if (!strcmp(argv[i], "--cfg")) tcfg = atoi(argv[i + 1]);
  if ((form == 4 && tcfg != 4) || (form == 5 && tcfg != 2 && tcfg != 3)) die(2, "--cfg: 4 (batched), 2 or 3 (served)");
  twl_t w;
  twl_defaults(tcfg, &w);
  twl_stream_t st;
  twl_stream_defaults(form, &st);
  const char* out = NULL;
  const char* algo = "fenwick";
  int ssha = 0, threads = 1;
  for (int i = 2; i < argc; ++i) {
    if (!strcmp(argv[i], "--stream-sha")) { ssha = 1; continue; }
    if (i + 1 >= argc) die(2, "missing value");
    const char* k = argv[i]; const char* v = argv[++i];
    if (!strcmp(k, "--form") || !strcmp(k, "--cfg")) continue;
    else if (!strcmp(k, "--S")) w.S = atoll(v);
    else if (!strcmp(k, "--U")) w.U = atoi(v);
    else if (!strcmp(k, "--N")) w.N = strtoull(v, NULL, 10);
    else if (!strcmp(k, "--seed")) w.seed = strtoull(v, NULL, 10);
    else if (!strcmp(k, "--requests")) st.nreq = atoll(v);
    else if (!strcmp(k, "--max-batch")) st.max_batch = atoi(v);
    else if (!strcmp(k, "--append-frac")) st.append_frac = atof(v);
    else if (!strcmp(k, "--append-U")) st.append_U = atoi(v);
    else if (!strcmp(k, "--rate")) st.rate = atof(v);
    else if (!strcmp(k, "--algo")) algo = v;
    else if (!strcmp(k, "--threads")) threads = atoi(v);
    else if (!strcmp(k, "--out")) out = v;
    else die(2, "unknown argument");
  }
  w.Q = 1;
  if (w.S < 1 || w.U < 1 || w.N < 2 || w.N > 0xFFFFFFFFull || st.nreq < 1 || st.max_batch < 1 ||
      (form == 5 && (st.append_U < 1 || st.rate <= 0.0 || st.append_frac < 0.0 || st.append_frac > 1.0)))
    die(2, "bad sizes");
  const double t0 = now_s();
  if (twl_generate(&w) != 0 || twl_stream_generate(&w, &st) != 0) die(3, "out of memory");
  if (ssha) {
    bw_sha256_ctx c;
    bw_sha256_init(&c);
    bw_sha256_update(&c, st.req_q0, sizeof(int64_t) * (size_t)(st.nreq + 1));
    bw_sha256_update(&c, st.req_kind, (size_t)st.nreq);
    bw_sha256_update(&c, st.qstate, sizeof(int32_t) * (size_t)st.nq);
    bw_sha256_update(&c, st.qlo, sizeof(uint32_t) * (size_t)st.nq);
    bw_sha256_update(&c, st.qhi, sizeof(uint32_t) * (size_t)st.nq);
    bw_sha256_update(&c, st.app_parent, sizeof(int32_t) * (size_t)st.napp);
    bw_sha256_update(&c, st.app_key, sizeof(uint32_t) * (size_t)st.napp * (size_t)st.append_U);
    bw_sha256_update(&c, st.app_delta, sizeof(int64_t) * (size_t)st.napp * (size_t)st.append_U);
    char hex[65];
    bw_sha256_final_hex(&c, hex);
    printf("{\"stream_sha256\":\"%s\",\"nq\":%lld,\"napp\":%lld}\n", hex, (long long)st.nq, (long long)st.napp);
    twl_stream_free(&st);
    twl_free(&w);
    return 0;
  }
  /* the tree extended by the appended states: state S+a+1 = append a */
  const int64_t S = w.S, ns = S + st.napp + 1;
  int32_t* parent = (int32_t*)malloc(sizeof(int32_t) * (size_t)ns);
  int64_t* upd0 = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ns + 1));
  const size_t nu = (size_t)S * (size_t)w.U, na = (size_t)st.napp * (size_t)st.append_U;
  uint32_t* key = (uint32_t*)malloc(sizeof(uint32_t) * (nu + na + 1));
  int64_t* delta = (int64_t*)malloc(sizeof(int64_t) * (nu + na + 1));
  memcpy(parent, w.parent, sizeof(int32_t) * (size_t)(S + 1));
  for (int64_t a = 0; a < st.napp; ++a) parent[S + 1 + a] = st.app_parent[a];
  upd0[0] = 0;
  upd0[1] = 0;
  for (int64_t v = 1; v <= S; ++v) upd0[v + 1] = upd0[v] + w.U;
  for (int64_t a = 0; a < st.napp; ++a) upd0[S + 2 + a] = upd0[S + 1 + a] + st.append_U;
  memcpy(key, w.key, sizeof(uint32_t) * nu);
  memcpy(delta, w.delta, sizeof(int64_t) * nu);
  if (na) {
    memcpy(key + nu, st.app_key, sizeof(uint32_t) * na);
    memcpy(delta + nu, st.app_delta, sizeof(int64_t) * na);
  }
  int64_t* ans = (int64_t*)calloc((size_t)(st.nq ? st.nq : 1), sizeof(int64_t));
  if (tcfg == 3 && strcmp(algo, "brute")) {
    /* the Fenwick solvers work on ranks: every key is a distinct key (its index), an endpoint x
     * becomes #keys < x; the brute walk compares raw values and needs neither */
    for (size_t i = 0; i < nu + na; ++i) key[i] = (uint32_t)lower_bound_u32(w.dkey, w.N, key[i]);
    for (int64_t g = 0; g < st.nq; ++g) {
      st.qlo[g] = (uint32_t)lower_bound_u32(w.dkey, w.N, st.qlo[g]);
      st.qhi[g] = (uint32_t)lower_bound_u32(w.dkey, w.N, st.qhi[g]);
    }
  }
  const double ts0 = now_s();
  if (!strcmp(algo, "brute"))
    solve_states_brute(parent, upd0, key, delta, st.nq, st.qstate, st.qlo, st.qhi, ans);
  else if (threads > 1)
    solve_states_par(ns, parent, upd0, key, delta, w.N, st.nq, st.qstate, st.qlo, st.qhi, ans, threads);
  else
    solve_states(ns, parent, upd0, key, delta, w.N, st.nq, st.qstate, st.qlo, st.qhi, ans);
  g_solve_s = now_s() - ts0;
  write_answers(out, ans, (size_t)st.nq, t0);
  free(ans); free(parent); free(upd0); free(key); free(delta);
  twl_stream_free(&st);
  twl_free(&w);
  return 0;
}

static int mode_sha(int argc, char** argv) {
  if (argc < 3) die(2, "usage: sha FILE");
  FILE* fp = fopen(argv[2], "rb");
  if (!fp) die(1, "cannot open");
  bw_sha256_ctx c;
  bw_sha256_init(&c);
  unsigned char buf[1 << 16];
  size_t n;
  while ((n = fread(buf, 1, sizeof(buf), fp)) > 0) bw_sha256_update(&c, buf, n);
  fclose(fp);
  char hex[65];
  bw_sha256_final_hex(&c, hex);
  printf("%s\n", hex);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) die(2, "usage: bw_oracle sha|ref|tree|stream ...");
  if (!strcmp(argv[1], "sha")) return mode_sha(argc, argv);
  if (!strcmp(argv[1], "ref")) return mode_ref(argc, argv);
  if (!strcmp(argv[1], "spec")) return mode_spec(argc, argv);
  if (!strcmp(argv[1], "tree")) return mode_tree(argc, argv);
  if (!strcmp(argv[1], "stream")) return mode_stream(argc, argv);
  die(2, "unknown mode");
  return 2;
}
