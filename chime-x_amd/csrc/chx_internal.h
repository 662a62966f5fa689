// chx_internal.h -- shared between the kernel files and the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "chx.h"

struct chx_plan {
  chx_config cfg;
  int32_t nb;        // N / 32: blocks along one side
  int64_t NB;        // upper-triangle block count
  uint32_t *d_maps;  // [2][NB]: id_x_map then id_y_map (main_wrapper.c:528-552)
  uint32_t *d_sums;  // [F][N][2] plan-owned offset sums (methods 0/1)
  int32_t ntiles;    // 128-element tiles per side (int8 matrix-core path)
  uint8_t *d_work;   // int8 planes of one time chunk (int8 matrix-core path)
  size_t work_bytes;
  // Lazily allocated buffers of chx_correlate_host.
  uint8_t *d_in;
  int32_t *d_out;
  size_t in_bytes, out_bytes, sums_bytes;
};

namespace chx {

// Records a HIP error for chx_last_error_detail() and maps it to a status.
int hip_status(hipError_t e);

#define CHX_HIP(expr)                              \
  do {                                             \
    hipError_t e_ = (expr);                        \
    if (e_ != hipSuccess) return chx::hip_status(e_); \
  } while (0)

// Fill n uint32 words with a value (graph-capture safe, unlike a memset node).
int launch_fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t s);

// Packed-integer path (chx_packed.hip).
int launch_offset_accumulate(const chx_plan *p, const uint8_t *in,
                             uint32_t *sums, hipStream_t s);
int launch_preseed(const chx_plan *p, const uint32_t *sums, int32_t *vis,
                   hipStream_t s);
int launch_corr_packed(const chx_plan *p, const uint8_t *in, int32_t *vis,
                       hipStream_t s);
int launch_blocks_to_triangle(int32_t F, int32_t N, const int32_t *blocks,
                              int32_t *tri, hipStream_t s);

// int8 matrix-core path (chx_mfma.hip).  mode: 0 assign, 1 add.
// The plane workspace of one time chunk is at most kWorkspaceBytes.
constexpr int64_t kWorkspaceBytes = 4LL << 30;
size_t mfma_workspace_bytes(const chx_plan *p);
int launch_corr_mfma(const chx_plan *p, const uint8_t *in, int32_t *vis,
                     int mode, hipStream_t s);

}  // namespace chx
