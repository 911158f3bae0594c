// ref_plan.hpp -- host side of the reference drop-in path: the BWBASE02 fixture, the declared
// reference-native workload (one edit per state) and the v7 block plan that feeds gpu_v7.h /
// gpu_v7b.h / gpu_session.h.  Restates, in this project's own code, src/common.hpp:57-374 and
// BatchEngineV7::build / build_v7_blocks (src/batch_engine_v7.hpp:576-830,1069-1415,1578-1629;
// compact metadata src/batch_engine_v7b.hpp:773-806) for chain version trees.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace bwg {

struct BaseRow {
  int64_t t, amount;
  int32_t g, pad;
};
static_assert(sizeof(BaseRow) == 24, "BWBASE02 row");

struct RefFixture {
  int64_t n = 0, t_min = 0, t_max = 0;
  int32_t G = 0;
  std::vector<BaseRow> rows;
  std::vector<int64_t> base_factor;
};
// BWBASE02: 64-byte header, n rows, G i64 factors, G i32 labels (src/common.hpp:57-155).
bool load_ref_fixture(const std::string& path, RefFixture& f);

// The reference's edit ops (src/common.hpp:157-175), plus kRestore, which has no reference
// counterpart: it sets a row back to an exact earlier (existence ex, group g, time t, amount val)
// when a depth-first walk of a branching history leaves a version (spec_workload below).
enum RefOp : uint8_t { kReplace = 0, kInsert = 1, kDelete = 2, kMove = 3, kFactor = 4, kRestore = 5 };
struct RefEdit {
  uint8_t op = 0, ex = 0;
  int32_t g = 0;
  int64_t id = 0, val = 0, t = 0;
};

struct RefWorkload {
  int K = 0, topology = 0, seed = 0, factor_period = 0, q = 0;
  bool per_state = false;
  std::vector<int32_t> parent;      // K
  std::vector<RefEdit> edit;        // K (edit[0] unused)
  std::vector<int64_t> qlo, qhi;    // per_state: K*q (state v at v*q+k); shared: q
  std::vector<int64_t> ins_t;       // inserted row k (id = n + k)
  std::vector<int32_t> ins_g;
  int64_t qlo_of(int v, int k) const { return per_state ? qlo[size_t(v) * q + k] : qlo[size_t(k)]; }
  int64_t qhi_of(int v, int k) const { return per_state ? qhi[size_t(v) * q + k] : qhi[size_t(k)]; }
};
// Declared workload with the reference's exact draw sequence (src/common.hpp:279-374).
// factor_pct < 0 keeps the factor-period pattern (BWH_FACTOR_PCT unset).
void make_ref_workload(const RefFixture& f, int K, int topology, int seed, int factor_period, int q,
                       bool per_state, int factor_pct, RefWorkload& w);

struct RefNode {  // resolved edit of one state (NodePlan, src/batch_engine_v7.hpp:544-551)
  uint8_t kind = 0, ex_old = 0, ex_new = 0;
  int32_t g_old = 0, c_old = 0, ti_old = 0, g_new = 0, c_new = 0, ti_new = 0;
  int64_t amt_old = 0, amt_new = 0, fac_val = 0, fac_prev = 0;
};

struct RefPlan {
  int K = 0, q = 0, B = 0, G = 0, ncls = 0, Cpfx = 0, n_active = 0;
  std::vector<int64_t> endpoints;
  std::vector<int32_t> qlb, qhb;      // K*q
  std::vector<RefNode> node;          // K
  std::vector<int32_t> g2row;         // G
  std::vector<int32_t> ms_off;        // n_active+1
  std::vector<int64_t> ms_cls, ms_base;
  std::vector<int32_t> ms_dev;
  std::vector<int64_t> vbase;         // Cpfx
  std::vector<int32_t> bv0, bv1, bidx;
  std::vector<int64_t> fac_start;     // B*G
  std::vector<int64_t> need_off;      // B*G
  std::vector<int32_t> gb_off, gb_block;
  std::vector<int64_t> gb_reg;
  int64_t devbuf_n = 0;
  std::vector<int64_t> snap;          // block-start deviation snapshots (filled when requested)
};

// A BWHHIST1 declared history spec: chain histories (src/real_history_spec.hpp:41-71, 4-field
// header) and branching histories (src/real_branch_spec.hpp:45-71, 5-field header with a
// topology flag).  "BWHHIST1", i32 kstates, q, fp, G [, topology], 12 i64 (n_base, n_ins, t_min,
// t_max, n_ep, period_lo, period_hi, bucket, n_inserts, n_deletes, n_credits, n_factor_edits),
// parent[K] i32, edit[K] (the reference's 40-byte Edit), queries[K*q] (lo, hi i64), ins_t[n_ins]
// i64, ins_g[n_ins] i32, endpoints[n_ep] i64.  Both layouts share the magic; the one whose arrays
// add up to the file's size is taken.  The load checks the tree (parent[v] in [0, v)), the ops
// and every edit's row / group against the base fixture f.
struct HistorySpec {
  int K = 0, q = 0, fp = 0, G = 0, topology = 0, header_fields = 0;
  int64_t n_base = 0, n_ins = 0, t_min = 0, t_max = 0, n_ep = 0, period_lo = 0, period_hi = 0, bucket = 0;
  int64_t n_inserts = 0, n_deletes = 0, n_credits = 0, n_factor_edits = 0;
  std::vector<int32_t> parent;     // K (parent[0] = -1)
  std::vector<RefEdit> edit;       // K (edit[0] unused)
  std::vector<int64_t> qlo, qhi;   // K*q, version v at v*q + k
  std::vector<int64_t> ins_t;      // inserted row k has id n_base + k
  std::vector<int32_t> ins_g;
  std::vector<int64_t> endpoints;
  bool chain() const;
};
bool load_history_spec(const std::string& path, const RefFixture& f, HistorySpec& s, std::string& err);

// The spec as a chain workload for the v7 block path.  The version tree is walked depth first
// (children ascending): entering version v appends a chain state with v's edit and ranges;
// leaving v appends one that undoes the edit (kRestore with the row's exact previous state, or
// kFactor with the previous factor) and repeats v's ranges (its answers are not used).  States
// after the last entry are dropped, so a chain spec maps onto itself.  entry[v] = chain state of
// version v (entry[0] = 0).
void spec_workload(const RefFixture& f, const HistorySpec& s, RefWorkload& w, std::vector<int32_t>& entry);

// Chain trees only (topology 0), like the reference's v7 path (src/batch_engine_v7.hpp:708).
bool build_ref_plan(const RefFixture& f, const RefWorkload& w, int B, bool cpu_snapshots, RefPlan& p,
                    std::string& err);

}  // namespace bwg
