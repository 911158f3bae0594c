/* gpu_session.h -- drop-in C-ABI of the session-resident executor (no class-domain limit).
 *
 * Replaces the reference interface src/gpu_session.h:4-24 (entry points :21-24) with
 * byte-identical structs (filled by src/panel_gpu_session.cpp:80-140).  `fen` and `leaf` are the
 * caller's B*Cpfx global arrays.  Unlike the reference's sess_run_kernel, which never writes dV
 * (src/gpu_session.cu:32-106; its multi-block carry is therefore wrong), gpu_sess_run_launch
 * writes every block's dense dV, so B > 1 is exact.  The run kernel also clears its block's rows
 * of fen/leaf itself; gpu_sess_reset_launch is kept for callers that reset explicitly.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

struct GpuSessNode {
  uint8_t kind;
  int16_t g_old, g_new;
  int32_t c_old, c_new, ti_old, ti_new;
  int64_t a, b;
};
struct GpuSessResetArgs {
  int B, Cpfx;
  int64_t* fen;
  int64_t* leaf;
};
struct GpuSessRunArgs {
  int B, G, Cpfx, ncls, q;
  const struct GpuSessNode* nodes;
  const int32_t *bv0, *bv1;
  const int64_t* fac_start;
  const int32_t* g2row;
  const int32_t* ms_off;
  const int64_t *ms_cls, *ms_base;
  const int32_t* ms_dev;
  int64_t* devbuf;
  const int64_t* need_off;
  const int32_t *qlb, *qhb;
  int64_t* lvals;
  int64_t* dV;
  int64_t* fen;
  int64_t* leaf;
  int threads;
};
struct GpuSessCombineArgs {
  int nstates, q, Cpfx;
  const int32_t *qlb, *qhb;
  const int64_t* carrypref;
  const int32_t* bidx;
  const int64_t* lvals;
  int64_t* answers;
};
struct GpuSessSnapArgs {
  int nstates;
  const struct GpuSessNode* nodes;
  const int32_t* bidx;
  const int32_t* gb_off;
  const int32_t* gb_block;
  const int64_t* gb_reg;
  int64_t* devbuf;
};

#ifdef __cplusplus
extern "C" {
#endif
int gpu_sess_reset_launch(const struct GpuSessResetArgs* a);
int gpu_sess_run_launch(const struct GpuSessRunArgs* a);
int gpu_sess_snap_launch(const struct GpuSessSnapArgs* a);
int gpu_sess_combine_launch(const struct GpuSessCombineArgs* a);
#ifdef __cplusplus
}
#endif
