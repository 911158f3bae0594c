/* gpu_v7.h -- drop-in C-ABI of the v7 block-parallel chain executor.
 *
 * Replaces the reference interface src/gpu_v7.h:7-50 (entry points :49-50).  The structs are
 * laid out byte for byte like the reference's (the drivers src/panel_batch_gpu.cpp:86-110,172-221
 * fill them directly), so a driver compiled against the reference header links against this
 * library unchanged.  No HIP or CUDA types: every pointer is a device pointer owned by the caller;
 * the library allocates nothing.  Launches go to the default stream, like the reference
 * (src/gpu_v7.cu:107,120); the return value is 0 or a hipError_t code.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

/* One edited state (node 0 unused): kind 1 = row edit, 2 = factor edit (src/gpu_v7.h:7-14). */
struct GpuV7Node {
  int32_t kind;
  int32_t g_old, c_old, ti_old;
  int32_t g_new, c_new, ti_new;
  int64_t amt_old, amt_new, fac_val, fac_prev;
  uint8_t ex_old, ex_new;
  uint8_t pad[6];
};

/* Block replay (src/gpu_v7.h:16-35).  `threads` is the workgroup size (rounded up to a multiple
 * of 64 and clamped to [64,1024]); `shared_bytes` is accepted for compatibility -- the kernel
 * sizes its LDS itself (dense leaf + prefix over Cpfx classes + the G factors + a row-edit
 * window) and returns hipErrorInvalidValue when that exceeds the 160 KiB of one CU. */
struct GpuV7Args {
  int B, G, Cpfx, ncls, q;
  const struct GpuV7Node* nodes;
  const int32_t* block_v0;
  const int32_t* block_v1;
  const int64_t* fac_start; /* B*G */
  const int32_t* g2row;     /* G */
  const int32_t* ms_off;    /* n_active+1 */
  const int64_t* ms_cls;
  const int64_t* ms_base;
  const int32_t* ms_dev;
  int64_t* devbuf;          /* block-start deviation snapshots, updated in place */
  const int64_t* need_off;  /* B*G, -1 when the group is not needed in the block */
  const int32_t* qlb;       /* kStates*q */
  const int32_t* qhb;
  int64_t* lvals;           /* (kStates-1)*2q */
  int64_t* dV;              /* B*Cpfx */
  int threads;
  int shared_bytes;
};

/* Combine (src/gpu_v7.h:37-47): answers[(v-1)q+k] = carrypref[b][qhb]-carrypref[b][qlb] +
 * lvals[2((v-1)q+k)+1] - lvals[2((v-1)q+k)], b = bidx[v]. */
struct GpuV7CombineArgs {
  int nstates;
  int q;
  int Cpfx;
  const int32_t* qlb;
  const int32_t* qhb;
  const int64_t* carrypref;
  const int32_t* bidx;
  const int64_t* lvals;
  int64_t* answers;
};

#ifdef __cplusplus
extern "C" {
#endif
int gpu_v7_launch(const struct GpuV7Args* a);
int gpu_combine_launch(const struct GpuV7CombineArgs* a);
#ifdef __cplusplus
}
#endif
