/* bwgpu_carry.h -- device carry pass for the v7 / v7b / session drop-in paths.
 *
 * No counterpart in the reference ABI: the reference copies dV to the host, runs the serial
 * carry there and copies carrypref back (src/panel_batch_gpu.cpp:191-204, the CPU loop of
 * src/batch_engine_v7.hpp:936-945).  This entry point computes the same carrypref on the GPU,
 * removing both PCIe crossings:
 *   vstart_0 = vbase;  carrypref[b][0] = 0;  carrypref[b][k] = sum_{c=1..k} vstart_b[c];
 *   vstart_{b+1}[k] = vstart_b[k] + dV[b][k]   (k >= 1; index 0 is never read)
 * Exact int64.  `stream` is an opaque hipStream_t (NULL = default stream).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

typedef struct BwCarryArgs {
  int B, Cpfx;
  const int64_t* vbase; /* Cpfx: V leaf at the root (index c+1 for class c) */
  const int64_t* dV;    /* B*Cpfx */
  int64_t* carrypref;   /* B*Cpfx */
} BwCarryArgs;

#ifdef __cplusplus
extern "C" {
#endif
int bw_carry_launch(const BwCarryArgs* a, void* stream);
#ifdef __cplusplus
}
#endif
