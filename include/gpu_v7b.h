/* gpu_v7b.h -- drop-in C-ABI of the compact-node v7 executor with on-device block-start
 * snapshots.
 *
 * Replaces the reference interface src/gpu_v7b.h:6-52 (entry points :50-52) with byte-identical
 * structs (filled by src/panel_batch_gpu2.cpp:89-101,186-244).  Ownership, stream and error
 * conventions as in gpu_v7.h.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

/* 40 bytes.  kind bits: 1 = row, 2 = factor, 4 = ex_old, 8 = ex_new; a,b = amt_old,amt_new (row)
 * or fac_val,fac_prev (factor) (src/gpu_v7b.h:6-12). */
struct GpuV7bNode {
  uint8_t kind;
  int16_t g_old, g_new;
  int32_t c_old, c_new;
  int32_t ti_old, ti_new;
  int64_t a, b;
};

/* Snapshot scatter (src/gpu_v7b.h:14-22): every row-edit delta is added into the snapshot region
 * of every later block that needs its group.  nstates = kStates. */
struct GpuV7bSnapArgs {
  int nstates;
  const struct GpuV7bNode* nodes;
  const int32_t* bidx;
  const int32_t* gb_off;   /* G+1 */
  const int32_t* gb_block;
  const int64_t* gb_reg;
  int64_t* devbuf;
};

struct GpuV7bArgs {
  int B, G, Cpfx, ncls, q;
  const struct GpuV7bNode* nodes;
  const int32_t *bv0, *bv1;
  const int64_t* fac_start;
  const int32_t* g2row;
  const int32_t* ms_off;
  const int64_t *ms_cls, *ms_base;
  const int32_t* ms_dev;
  int64_t* devbuf;
  const int64_t* need_off;
  const int32_t *qlb, *qhb;
  int64_t *lvals, *dV;
  int threads;
  int shared_bytes;
};

struct GpuV7bCombineArgs {
  int nstates, q, Cpfx;
  const int32_t *qlb, *qhb;
  const int64_t* carrypref;
  const int32_t* bidx;
  const int64_t* lvals;
  int64_t* answers;
};

#ifdef __cplusplus
extern "C" {
#endif
int gpu_v7b_snap_launch(const struct GpuV7bSnapArgs* a);
int gpu_v7b_launch(const struct GpuV7bArgs* a);
int gpu_v7b_combine_launch(const struct GpuV7bCombineArgs* a);
#ifdef __cplusplus
}
#endif
