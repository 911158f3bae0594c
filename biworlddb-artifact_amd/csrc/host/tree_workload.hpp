// tree_workload.hpp -- branching multi-update point-update workloads of the benchmark configs.
//
// The reference's declared workload has one edit per state on chain/bushy/star trees
// (src/common.hpp:194-374).  These workloads keep its contract (states answer their own range
// queries after their own updates, answers in (state, query) order, src/oracle.cpp:10-18) and
// its counter-based splitmix64 generation style (src/batch_engine_v7c.hpp:240-249), with the
// shapes of the five benchmark configs:
//   cfg 1  N=65,536, 1,000 states, parent uniform over the 32 most recent, 64 updates
//          (uniform keys, deltas in [-1000,1000]), 16 uniform [lo,hi) queries per state
//   cfg 2  N=2^20, 300,000 states, parent = previous w.p. 0.9 else uniform over the prior
//          1,024, 128 Zipf(1.1) updates (deltas in [-10^6,10^6]), 32 log-uniform ranges
//   cfg 3  4,194,304 distinct 32-bit keys, 300,000 states, 256 Zipf(0.9) updates over the
//          distinct set (deltas in [-10^4,10^4]), 32 queries with raw 32-bit endpoints
//   cfg 4  cfg 2 with geometric runs (mean 64, max 4,096)
//   cfg 5  cfg 2 (served)
// The exact draw definitions are restated in the header comment of the test oracle's
// generator; the two implementations are checked byte for byte by the test suite.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace bwg {

struct TreeSpec {
  int cfg = 2;
  int64_t S = 300000;  // edited states (1..S); state 0 is the empty root
  int U = 128, Q = 32;
  uint64_t N = 1u << 20;  // key domain, or distinct-key count for cfg 3
  uint64_t seed = 0;
};

struct TreeWorkload {
  TreeSpec spec;
  std::vector<int32_t> parent;  // S+1, parent[0] = -1
  std::vector<uint32_t> key;    // S*U raw keys
  std::vector<int64_t> delta;   // S*U
  std::vector<uint32_t> qlo, qhi;  // S*Q raw endpoints
  std::vector<uint32_t> dkey;   // cfg 3: sorted distinct keys
  uint64_t coord_n() const { return spec.cfg == 3 ? dkey.size() : spec.N; }
};

// Cumulative Zipf(s) weights of ranks 1..m, summed in rank order: the inverse-CDF table of the
// Zipf draws (rank r is picked for u when it is the first whose weight exceeds u * total).  Built
// once on the host in this order, so every generator -- host or device -- reads the same doubles.
std::vector<double> zipf_cumulative(uint64_t m, double s);

TreeSpec tree_defaults(int cfg);
void tree_generate(const TreeSpec& spec, TreeWorkload& w);
// SHA-256 (hex) over parent[1..S], key, delta, qlo, qhi (and dkey for cfg 3), little-endian.
// Matches the oracle's --workload-sha.
void tree_workload_sha(const TreeWorkload& w, char out_hex[65]);

// Request streams of the batched form (over a cfg-4 tree) and of the served form (over a cfg-2
// tree that grows by appended branch states).  Batched: nreq batches of uniform size in
// [1, max_batch], each query a uniform state and a cfg-2 log-uniform range.  Served: nreq
// requests, each an append (probability append_frac: a new state under a uniform existing state
// with append_U Zipf(1.1) updates) or a query batch of log-uniform size in [1, max_batch] over
// the states existing at its arrival, half point queries [k, k+1) with Zipf(1.1) k, half
// log-uniform ranges; Poisson arrivals at `rate` per second.  Over a cfg-3 tree the appended keys
// and point queries are distinct keys drawn by Zipf(0.9) over the distinct set, and the ranges
// span the 32-bit domain (raw endpoints, resolved by rank).  The exact draws are restated in
// the test oracle's stream generator; the two are checked byte for byte by the test suite.
struct StreamSpec {
  int form = 4;  // 4 = batched, 5 = served
  int64_t nreq = 100000;
  int max_batch = 2048;
  double append_frac = 0.0;
  int append_U = 128;
  double rate = 20000.0;
};
StreamSpec stream_defaults(int form);

struct RequestStream {
  StreamSpec spec;
  std::vector<int64_t> req_q0;    // nreq+1 offsets into the queries
  std::vector<int8_t> req_kind;   // 0 = query batch, 1 = append
  std::vector<int64_t> req_app;   // nreq: appends before request r (an append r creates S+req_app[r]+1)
  std::vector<double> arrival;    // seconds (served form)
  std::vector<int32_t> qstate;
  std::vector<uint32_t> qlo, qhi;
  std::vector<int32_t> app_parent;  // napp
  std::vector<uint32_t> app_key;    // napp * append_U
  std::vector<int64_t> app_delta;
  int64_t nq() const { return int64_t(qstate.size()); }
  int64_t napp() const { return int64_t(app_parent.size()); }
};
// tree: the generated tree the stream runs over (its seed, S and N are used).
void stream_generate(const TreeWorkload& tree, const StreamSpec& spec, RequestStream& rs);
// The sequential part of stream_generate alone: request kinds, sizes (req_q0), appends before
// each request (req_app) and arrival times; qstate / qlo / qhi and the append arrays stay empty.
// Every element of those is a function of its own index (counter-based draws), so they can be
// materialised anywhere afterwards (the device generator does).
void stream_scan(const TreeSpec& tree, const StreamSpec& spec, RequestStream& rs);
// SHA-256 (hex) over req_q0, req_kind, qstate, qlo, qhi, app_parent, app_key, app_delta.
// Matches the oracle's stream --stream-sha.
void stream_sha(const RequestStream& rs, char out_hex[65]);

// cfg 3: CSR rank-LUT over the distinct keys (row = high 16 bits): 65,537 row pointers and the
// low 16 bits of every key, plus host ranks of raw values (#keys < x).
struct RankLut {
  std::vector<uint32_t> row_ptr;
  std::vector<uint16_t> low16;
};
RankLut build_rank_lut(const std::vector<uint32_t>& sorted_keys);
void host_ranks(const RankLut& lut, const uint32_t* in, uint32_t* out, size_t n);

}  // namespace bwg
