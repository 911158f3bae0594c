// dense_tables.cpp -- see dense_tables.hpp.
#include "dense_tables.hpp"

#include <cstddef>

namespace bwg {

using std::size_t;

bool dense_rank_tables(int64_t S, const std::vector<int32_t>& steps, const std::vector<int32_t>& blk_step0,
                       std::vector<int32_t>& state_erank, std::vector<int32_t>& erank_step,
                       std::vector<int32_t>& blk_e0) {
  state_erank.assign(size_t(S) + 1, -1);
  erank_step.assign(size_t(S), -1);
  std::vector<int32_t> before(steps.size() + 1, 0);  // enter steps before step s
  int64_t cnt = 0;
  for (size_t s = 0; s < steps.size(); ++s) {
    before[s] = int32_t(cnt);
    const int32_t v = steps[s];
    if (v > 0) {
      if (v > S || state_erank[size_t(v)] != -1) return false;
      state_erank[size_t(v)] = int32_t(cnt);
      erank_step[size_t(cnt)] = int32_t(s);
      ++cnt;
    }
  }
  before[steps.size()] = int32_t(cnt);
  if (cnt != S) return false;
  blk_e0.resize(blk_step0.size());
  for (size_t b = 0; b < blk_step0.size(); ++b) {
    const int32_t s = blk_step0[b];
    if (s < 0 || size_t(s) > steps.size()) return false;
    blk_e0[b] = before[size_t(s)];
  }
  state_erank[0] = 0;
  return true;
}

}  // namespace bwg
