// chx_internal.h -- shared between the kernel files and the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "chx.h"

struct chx_plan {
  chx_config cfg;
  int32_t nb;        // N / 32: blocks along one side
  int64_t NB;        // upper-triangle block count
  uint32_t *d_maps;  // [2][NB]: id_x_map then id_y_map (main_wrapper.c:528-552)
  uint32_t *d_sums;  // [F][N][2] plan-owned offset sums (methods 0/1)
  int32_t ntiles;    // 128-element tiles per side (int8 matrix-core path)
  uint8_t *d_work;   // planes of one time chunk (matrix-core paths; int8: + code sums, corrections)
  size_t work_bytes;
  // Lazily allocated buffers of chx_correlate_host.
  uint8_t *d_in;
  int32_t *d_out;
  size_t in_bytes, out_bytes, sums_bytes;
  // Kernel timing (chx_plan_profile): event triples (before unpack, before
  // GEMM, after GEMM) of each matrix-core chunk launch, not yet read.
  int prof;
  std::vector<hipEvent_t> *prof_ev;
};

namespace chx {

// Records a HIP error for chx_last_error_detail() and maps it to a status.
int hip_status(hipError_t e);
// Sets the text chx_last_error_detail() returns on this thread.
void set_error_detail(const char *msg);
// The shape and time_accum rules of a chx_config (a chx_status).
int validate_config(const chx_config *c);

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
int launch_blocks_to_full(int32_t F, int32_t N, const int32_t *blocks, int32_t *full,
                          hipStream_t s);

// int8 matrix-core path (chx_mfma.hip).  mode: 0 assign, 1 add.
// The plane workspace of one time chunk is at most kWorkspaceBytes.
constexpr int64_t kWorkspaceBytes = 16LL << 30;
constexpr int kMfmaTile = 128;   // elements per output tile side
constexpr int kMfmaStage = 64;   // samples per pipeline stage
constexpr int kMfmaRing = 4;     // stages in flight; stage counts are padded to it
size_t mfma_workspace_bytes(const chx_plan *p);
int launch_corr_mfma(const chx_plan *p, const uint8_t *in, int32_t *vis,
                     int mode, hipStream_t s);
// Offset corrections of offset-binary code planes (both matrix-core plan
// paths): partials [nrun][F][Np] of code sums SR | SI << 16 over a chunk of
// n valid samples -> corr [2][F][Np] int32 pairs, rows (-8 (SR + SI) + 128 n,
// -8 SR - 7 SI) then columns (-8 (SR + SI), 8 (SR - SI)).
int launch_u4_corrections(const uint32_t *partials, int32_t *corr, int F, int Np, int nrun,
                          int32_t n, hipStream_t s);

// FP6 (E2M3) matrix-core path (chx_fp6.hip): same modes, bit-identical
// results.
constexpr int kFp6Ring = 4;  // 64-sample stages in the LDS ring
// Both matrix-core GEMMs use output tiles of 128 rows x 256 columns: row
// tile ti holds ceil((ntiles - ti) / 2) tiles, the columns starting at row
// tile ti.  An int8 work item covers at most kMaxItemStages 64-sample
// stages: its int32 sums of offset-binary code products (at most 450 * 64
// per stage) stay < 2^26.  FP6 items: kMaxItemStagesFp6 (below).
constexpr int kMaxItemStages = 2048;
__host__ __device__ inline int64_t wide_tiles(int64_t ntiles) {
  const int64_t m = ntiles / 2;
  return ntiles % 2 ? (m + 1) * (m + 1) : m * (m + 1);
}
// Work items per channel: the (ntiles + 1) / 2 half tiles (the last tile
// of each row tile ti with ntiles - ti odd, whose right half lies past the
// array) go two to an item.
__host__ __device__ inline int64_t paired_tiles(int64_t ntiles) {
  return wide_tiles(ntiles) - (ntiles + 1) / 2 / 2;
}
// Row tiles per band of the full-tile order.
constexpr int kTileBand = 8;
// Work item tb of one channel (both matrix-core GEMMs, 128 x 256 tiles).
// Full tiles (rows from tile ti, columns from tiles ct and ct + 1, ct = ti,
// ti + 2, ... <= ntiles - 2: row tile ti holds floor((ntiles - ti) / 2) of
// them) come first, band by band: bands of kTileBand row tiles in order; in
// a band, column tile by column tile (ct from the band's first row tile
// on), and in a column its rows in order (the band's rows of ct's parity
// up to ct).  The 32 items one XCD runs together then read the planes of
// about 8 row tiles and 9 column tiles whatever the array size (row by row
// over a channel they read every column).  Then the half tiles, the last
// tile of each row tile ti with ntiles - ti odd (ti = ntiles-1, ntiles-3,
// ...), whose 128 columns are column tile ntiles-1: two of them form one
// item (rows ti_a for waves 0-1, ti_b = ti_a - 2 for waves 2-3); an odd last
// one stays alone (ti_b < 0; its waves 2-3 compute past the array).  Full
// tiles return ti_b = ti_a.
__host__ __device__ inline void tile_coords(int tb, int ntiles, int &ti_a, int &ti_b,
                                            int &ct) {
  const int nfull = (int)wide_tiles(ntiles) - (ntiles + 1) / 2;
  if (tb >= nfull) {
    tb -= nfull;
    ct = ntiles - 1;
    ti_a = ntiles - 1 - 4 * tb;
    ti_b = ti_a - 2;
    return;
  }
  int r0 = 0;
  for (;; r0 += kTileBand) {  // the band holding tb
    const int r1 = r0 + kTileBand < ntiles ? r0 + kTileBand : ntiles;
    int band = 0;
    for (int ti = r0; ti < r1; ti++) band += (ntiles - ti) / 2;
    if (tb < band) break;
    tb -= band;
  }
  const int rl = (r0 + kTileBand < ntiles ? r0 + kTileBand : ntiles) - 1;  // last row
  for (int c = r0;; c++) {  // column tile c holds the band's rows of c's parity up to c
    const int a = r0 + ((c - r0) & 1);  // first such row
    const int b = c < rl ? c : rl;
    const int n = b >= a ? (b - a) / 2 + 1 : 0;
    if (tb < n) {
      ti_a = ti_b = a + 2 * tb;
      ct = c;
      return;
    }
    tb -= n;
  }
}
size_t fp6_workspace_bytes(const chx_plan *p);
int launch_corr_fp6(const chx_plan *p, const uint8_t *in, int32_t *vis, int mode,
                    hipStream_t s);

// Plane formats of the matrix-core GEMMs.  A stage is 64 samples in both;
// stage counts of a job are multiples of the format's ring.
//  int8: planes [2][F][KB][Np][16], KB = 16-sample blocks, 2 bytes per
//        (element, sample).
//  fp6:  P0 [2][F][KB][2 halves][Np][16] then P1 [2][F][KB][2][Np][8], KB =
//        64-sample blocks, 1.5 bytes per (element, sample).
enum PlaneFormat { kPlanesInt8 = 0, kPlanesFp6 = 1 };
inline int format_ring(int fmt) { return fmt == kPlanesFp6 ? kFp6Ring : kMfmaRing; }
// work items of one channel (both formats: tiles of 128 x 256, half tiles
// in pairs)
inline int64_t format_tiles(int, int64_t ntiles) { return paired_tiles(ntiles); }
// work items of one channel on 128 x 128 tiles: the upper triangle of tiles
__host__ __device__ inline int64_t small_tiles(int64_t ntiles) { return ntiles * (ntiles + 1) / 2; }
// The blocks of wave `wave` that no flush stores, bit 2i + j for its row
// group i and column group j (32-element blocks): those below the diagonal
// of a diagonal tile, and every block of a wave whose columns lie past the
// array.  Wave layouts as the GEMM tiles assign them: a 128 x 256 tile
// (rows ti_a, columns from kMfmaTile ct, wave w: all 4 row groups x column
// groups 2w, 2w+1); a pair of half tiles (waves 0-1 rows ti_a, 2-3 rows
// ti_b, columns from kMfmaTile ct + 64 (w%2)); a 128 x 128 tile (wave w:
// row groups 2(w/2) + i, i < 2, column groups 2(w%2) + j).
__host__ __device__ inline int tile_skip_mask(bool pair, bool small, int ti_a, int ti_b, int ct,
                                              int Np, int wave) {
  if (!pair && !small && wave >= 2 && (ct + 1) * kMfmaTile >= Np) return 0xFF;
  const int ti = pair && wave >= 2 ? ti_b : ti_a;
  const int gr0 = small ? 2 * (wave >> 1) : 0;
  const int cw = pair || small ? wave & 1 : wave;
  const int ng = small ? 2 : 4;
  int m = 0;
  for (int i = 0; i < ng; i++)
    for (int j = 0; j < 2; j++)
      if (4 * ct + 2 * cw + j < 4 * ti + gr0 + i) m |= 1 << (2 * i + j);
  return m;
}
// the most stages one work item may cover: FP6 work items hold offset-binary
// code products, at most 450 per sample, from -2^24: 1024 stages keep every
// partial sum within 2^24 (int8: 2048)
constexpr int kMaxItemStagesFp6 = 1024;
inline int format_max_stages(int fmt) {
  return fmt == kPlanesFp6 ? kMaxItemStagesFp6 : kMaxItemStages;
}
// stages per work item of an nst-stage chunk cut at max_st (0 = whole)
inline int format_item_stages(int fmt, int nst, int max_st) {
  int per = (max_st > 0 && max_st < nst) ? max_st : nst;
  return per < format_max_stages(fmt) ? per : format_max_stages(fmt);
}
// plane bytes of one job chunk of `stages` 64-sample stages (both planes)
inline int64_t format_plane_bytes(int fmt, int64_t F, int64_t Np, int64_t stages) {
  return fmt == kPlanesFp6 ? 96 * F * Np * stages : 128 * F * Np * stages;
}
// KB (the layout's block count) of `stages` stages
inline int format_kb(int fmt, int stages) { return fmt == kPlanesFp6 ? stages : 4 * stages; }

// One job as the matrix-core kernels see it (per time chunk).
struct GemmJob {
  const uint8_t *planes;   // int8 planes, or fp6 P0
  const uint8_t *planes1;  // fp6 P1 (unused for int8)
  int32_t *vis;            // block-order output
  int64_t NB;              // blocks per channel
  int F, Np, KB, ntiles, nb, U, mode, stages_per_split;  // mode 0 =, 1 +=, 2 atomic +=
  int small;  // fp6 batch only: 128 x 128 tiles (four waves of 64 x 64) instead of 128 x 256
  // int8 plan path only (offset-binary planes): per-element offset
  // corrections [2][F][Np] of int32 pairs, rows then columns, added by the
  // piece that starts the chunk; nullptr elsewhere.
  const int32_t *corr;
};
// work items of one channel of a GEMM job
inline int64_t job_tiles(int fmt, const GemmJob &g) {
  return g.small ? small_tiles(g.ntiles) : format_tiles(fmt, g.ntiles);
}
// One work item of the persistent batch kernel: tile `tile` (row-major over
// the upper triangle of tiles) of channel f of jobs[job], over stages
// [stage0, stage0 + nstages) of the chunk (nstages a multiple of the ring).
struct GemmItem {
  int job, f, tile, stage0, nstages;
};
// One job's chunk for the batched unpack (offset-binary codes), starting at
// global workgroup unit0 (increasing over jobs): ceil(Np/1024) * runs * F
// workgroups.  The job's units (int8: 1024 elements x one 16-sample block;
// FP6: 1024 elements x one 32-sample half, or for arrays under 1024
// elements several halves of all elements) go round-robin over its runs,
// and run r writes its elements' code sums into row prow0 + r of
// `partials` (rows [F][Np] u32, SR | SI << 16), the partial rows of the
// GEMM job the chunk belongs to.  Its KB
// blocks land at blocks [kb_dst0, kb_dst0 + KB) of planes laid out with
// KB_dst blocks (format layouts above).  t_valid = 0 writes zeros and reads
// nothing.
struct UnpackJob {
  const uint8_t *in;
  uint8_t *planes, *planes1;
  int64_t row_bytes, t0, t_valid, unit0;
  int N, F, Np, KB, KB_dst, kb_dst0;
  uint32_t *partials;
  int prow0, runs;
};
// 32-sample halves per fp6 unpack workgroup: arrays under 1024 elements
// take several, so that all 1024 element slots work
__host__ __device__ inline int fp6_halves_per_unit(int Np) { return Np >= 1024 ? 1 : 1024 / Np; }
// FP6 unpack units of a chunk of KB 64-sample blocks
__host__ __device__ inline int64_t f6_u4_units(int64_t KB, int Np) {
  const int hpu = fp6_halves_per_unit(Np);
  return (2 * KB + hpu - 1) / hpu;
}
// Unpack units of a chunk of KB layout blocks (per channel and element
// group): int8 one per 16-sample block, FP6 f6_u4_units.
__host__ __device__ inline int64_t u4_units(int fmt, int64_t KB, int Np) {
  return fmt == kPlanesFp6 ? f6_u4_units(KB, Np) : KB;
}
// Runs (workgroups per channel and element group) of a batch unpack job: a
// run takes units/8 units, at least 1 and at most 16 (so that its code
// sums, at most 15 * 32 * 8 * 16 per element, fit 16 bits); a small job (a
// streamed frame: 32 FP6 units) still spreads over many workgroups.
__host__ __device__ inline int u4_job_runs(int fmt, int64_t KB, int Np) {
  const int64_t units = u4_units(fmt, KB, Np);
  int64_t per = units / 8;
  per = per < 1 ? 1 : (per > 16 ? 16 : per);
  return (int)((units + per - 1) / per);
}
inline int64_t unpack_units(int, const UnpackJob &u) {
  return (u.Np + 1023) / 1024 * u.runs * u.F;
}
// The offset corrections of one GEMM job (batch and serve): partials
// [nrun][F][Np] of code sums over a chunk of n valid samples -> corr [2][F]
// [Np] int32 pairs (rows, then columns), as launch_u4_corrections; the
// job's workgroups start at blk0 (64 elements each, increasing over jobs).
struct CorrJob {
  const uint32_t *partials;
  int32_t *corr;
  int64_t blk0;
  int F, Np, nrun, n;
};
int launch_u4_corrections_batch(const CorrJob *d_jobs, int njobs, int64_t blocks,
                                hipStream_t s);
int launch_unpack_batch(int fmt, const UnpackJob *d_jobs, int njobs, int64_t units,
                        hipStream_t s);
// Work items of one (job, channel) group; groups are spread over the 8
// per-XCD queues by spread_over_queues (chx_batch.hip).
struct ItemGroup {
  int64_t work;
  std::vector<GemmItem> items;
};
// Appends the groups of GEMM job gj (F channels, ntiles 128-element tiles
// per side, nst stages cut into pieces of format_item_stages(fmt, nst,
// max_st) stages; small: 128 x 128 tiles): one group per channel.  Returns
// the number of pieces per tile.
int add_job_items(int fmt, int gj, int F, int ntiles, int nst, int max_st,
                  std::vector<ItemGroup> &groups, bool small = false);
// Longest-processing-time-first assignment of the groups to 8 queues (long
// items first inside a queue); appends the items to `items` and the 9 queue
// bounds to `qbound`.
void spread_over_queues(std::vector<ItemGroup> &groups, std::vector<GemmItem> &items,
                        std::vector<int> &qbound);
// After spread_over_queues with one round of 8 queue bounds: in each queue
// whose n items are equal and whole, with n % (nwg / 8) = r > 0 and n >
// nwg / 8, the last r items are cut into time pieces (the count that
// minimises the partial round's length, >= 16 stages each) owned by copies
// of their jobs with mode 2 (atomic add), so that the last round keeps every
// workgroup of the queue busy.  Appends the job copies to gjobs; returns how
// many.  The caller's outputs must take atomic adds (zeroed, or
// accumulating).
int cut_partial_rounds(int fmt, std::vector<GemmItem> &items, std::vector<int> &qbound,
                       std::vector<GemmJob> &gjobs, int nwg);
// Invariants the kernels rely on, checked on the host before a launch:
// every plane region inside the workspace [base, base + work_bytes), every
// item inside its job's chunk (whole rings), `rounds` sets of 9 queue bounds
// covering the items in order.  CHX_ERR_HIP if any fails.
int check_launch_tables(int fmt, const std::vector<UnpackJob> &ujobs,
                        const std::vector<GemmJob> &gjobs, const std::vector<GemmItem> &items,
                        const std::vector<int> &qbound, int rounds, const uint8_t *base,
                        size_t work_bytes);
// Persistent batch GEMM over 8 per-XCD item queues (queue q = items
// [qbound[q], qbound[q+1])); the 8 counters must be 0.
int launch_gemm_batch(int fmt, const GemmJob *d_jobs, const GemmItem *d_items,
                      const int *d_qbound, int *d_counters, int nwg, hipStream_t s);
// Per-format kernels behind the two launchers above.
int launch_unpack_batch_int8(const UnpackJob *d_jobs, int njobs, int64_t units, hipStream_t s);
int launch_gemm_batch_int8(const GemmJob *d_jobs, const GemmItem *d_items, const int *d_qbound,
                           int *d_counters, int nwg, hipStream_t s);
int launch_unpack_batch_fp6(const UnpackJob *d_jobs, int njobs, int64_t units, hipStream_t s);
int launch_gemm_batch_fp6(const GemmJob *d_jobs, const GemmItem *d_items, const int *d_qbound,
                          int *d_counters, int nwg, hipStream_t s);

}  // namespace chx
