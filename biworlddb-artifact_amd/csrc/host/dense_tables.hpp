// dense_tables.hpp -- host tables of the dense-carry form (include/bwgpu_dense.h) that depend only
// on the Euler walk: no HIP here, so test emulations link it too.
#pragma once
#include <cstdint>
#include <vector>

namespace bwg {

// For the walk `steps` (+v enter, -v exit) cut at blk_step0: state_erank[v] = index of v's enter
// step among the enter steps (its preorder rank; entry 0 is 0), erank_step[e] = the step of the
// e-th enter, blk_e0[b] = enter steps before block b.  False if some state is not entered exactly
// once.
bool dense_rank_tables(int64_t S, const std::vector<int32_t>& steps, const std::vector<int32_t>& blk_step0,
                       std::vector<int32_t>& state_erank, std::vector<int32_t>& erank_step,
                       std::vector<int32_t>& blk_e0);

}  // namespace bwg
