// chx_mfma.hip -- the correlator on CDNA4 int8 matrix cores.
//
// Each channel's visibilities are a Hermitian rank-T update.  With the
// samples split into a real plane r = R-8 and an imaginary plane i = I-8
// (one int8 per value), a 32x32 reference block (rows y, columns x,
// pairwise_correlator_UT.cl:161-201) is two int8 accumulations over time:
//   RE = Ar.Br^T + Ai.Bi^T      (sum y_r x_r + y_i x_i)
//   IM = Ar.Bi^T - Ai.Br^T      (sum y_r x_i - y_i x_r)
//   im = -IM for upper_triangle_convention 1 (cpu_corr_test.c:189),
//        IM for 0 (cpu_corr_test.c:92).
// There is no time_accum cap.
//
// Two passes per chunk of time:
//  1. unpack (HBM-bound; unpack_u4_kernel for a plan, unpack_u4_batch_kernel
//     for the batch and serve engines): [time][freq][element] bytes ->
//     MFMA-ready planes [plane][freq][16-sample block][element][16 B] of the
//     nibbles' offset-binary codes u in [0, 15] (value u - 8), elements past
//     N and samples past T code 0.  The GEMM then accumulates code products,
//     and -Br becomes the complement 15 - u (one XOR):
//       RE' = sum ur_a ur_b + ui_a ui_b,  IM' = sum ur_a ui_b + ui_a (15 - ur_b),
//     which the per-element code sums SR, SI over the chunk's n valid
//     samples turn into the signed values (padding adds nothing):
//       RE = RE' - 8 (SR_a + SI_a) + 128 n - 8 (SR_b + SI_b)
//       IM = IM' - 8 SR_a - 7 SI_a + 8 (SR_b - SI_b).
//     The unpack writes each workgroup's code sums, a small kernel adds them
//     into one row and one column correction per element
//     (u4_corrections_kernel, u4_corrections_batch_kernel), and the GEMM
//     adds them as it flushes.  Codes switch four bits per operand byte and
//     no sign bits, and every product is >= 0, so the power-limited matrix
//     cores hold a higher clock than with signed operands.
//  2. corr_mfma_kernel / corr_mfma_batch_kernel: a workgroup = 4 wavefronts
//     = a 128-row x 256-column tile of one channel, each wave all 128 rows
//     x 64 columns (4x2 blocks x 2 int32 accumulators: RE' and IM').  The
//     row operands come through a ring of 4 LDS k-steps of 32 samples (8 KB
//     each) filled by buffer->LDS DMA; each wave loads its own column
//     operands straight into VGPRs, three k-steps ahead.  On a tile that
//     straddles the diagonal, the waves whose blocks lie partly below it
//     run a k-loop variant without those blocks' MFMAs.
// A work item covers at most kMaxItemStages 64-sample stages (131072
// samples): code products add at most 450 per sample, < 2^26 per
// accumulator.
#include "chx_internal.h"

namespace chx {
namespace {

typedef int v2i __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;

constexpr int kTile = kMfmaTile;  // 128: tile rows; a tile has 2 * kTile columns
constexpr int kRing = 4;           // LDS slots of one 32-sample k-step
constexpr int kRowGroups = 4;     // 32-element row groups per k-step (in LDS)
// LDS k-step slot: [plane 2][row group 4][1 KB operand].  A 32-row,
// 32-sample operand is 1 KB contiguous: lanes 0-31 read the k-step's even
// 16-sample block, lanes 32-63 the odd one, 16 lanes = 256 contiguous bytes
// (no bank conflicts); one 64-lane buffer->LDS load fills exactly one
// operand.
constexpr int kSlotBytes = 2 * kRowGroups * 1024;  // 8 KB
constexpr int kRingBytes = kRing * kSlotBytes;     // a ring of one row block

// 4x4 byte transpose: r[s] holds byte e of sample s; c[e] gets byte s of
// element e.  v_perm_b32 selects bytes of {s0:s1} (selector 0-3 = s1,
// 4-7 = s0).
__device__ __forceinline__ void transpose4(uint32_t r0, uint32_t r1, uint32_t r2,
                                           uint32_t r3, uint32_t &c0, uint32_t &c1,
                                           uint32_t &c2, uint32_t &c3) {
  const uint32_t t01l = __builtin_amdgcn_perm(r1, r0, 0x05010400u);
  const uint32_t t01h = __builtin_amdgcn_perm(r1, r0, 0x07030602u);
  const uint32_t t23l = __builtin_amdgcn_perm(r3, r2, 0x05010400u);
  const uint32_t t23h = __builtin_amdgcn_perm(r3, r2, 0x07030602u);
  c0 = __builtin_amdgcn_perm(t23l, t01l, 0x05040100u);
  c1 = __builtin_amdgcn_perm(t23l, t01l, 0x07060302u);
  c2 = __builtin_amdgcn_perm(t23h, t01h, 0x05040100u);
  c3 = __builtin_amdgcn_perm(t23h, t01h, 0x07060302u);
}

// Maps the dispatch-order workgroup id so that the workgroups resident on
// one XCD (dispatch is round-robin over the 8 XCDs) take consecutive
// (channel, tile) work items and share their input rows in that XCD's L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t id, int64_t total) {
  const int64_t q = total / 8, r = total % 8;
  const int64_t xcd = id % 8, slot = id / 8;
  return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// Flush one block of code-product accumulators (re = RE', im = IM') into
// int32 visibilities, adding the offset corrections: rc[j] = (row RE, row
// IM) of register j's row, cc = (column RE, column IM) of the lane's column
// (zeros for a piece that does not start the chunk).  Lane l owns column n
// = l%32 and rows m = 8*(j/4) + 4*(l/32) + j%4 of register j
// (v_mfma_i32_32x32x32_i8 D layout).  mode 0 assigns, 1 adds, 2 adds
// atomically.
__device__ __forceinline__ void flush_block(int32_t *blk, const v16i &re, const v16i &im,
                                               int U, int mode, int lane, const v2i (&rc)[16],
                                               const v2i cc) {
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc(blk, (short)0, 2048 * 4, 0x00020000);
  const int voff = ((lane >> 5) * 4 * 32 + (lane & 31)) * 8;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int32_t vr = re[j] + rc[j].x + cc.x;
    const int32_t m = im[j] + rc[j].y + cc.y;
    const int32_t vi = U ? -m : m;
    const int soff = (8 * (j >> 2) + (j & 3)) * 32 * 8;
    if (mode == 0) {
      v2i v = {vr, vi};
      __builtin_amdgcn_raw_buffer_store_b64(v, rsrc, voff, soff, 0);
    } else if (mode == 1) {
      v2i a = __builtin_amdgcn_raw_buffer_load_b64(rsrc, voff, soff, 0);
      a.x += vr;
      a.y += vi;
      __builtin_amdgcn_raw_buffer_store_b64(a, rsrc, voff, soff, 0);
    } else {
      __builtin_amdgcn_raw_ptr_buffer_atomic_add_i32(vr, rsrc, voff, soff, 0);
      __builtin_amdgcn_raw_ptr_buffer_atomic_add_i32(vi, rsrc, voff + 4, soff, 0);
    }
  }
}

// Pass 1: offset-binary code planes.  One workgroup = one run of 256
// element quads (1024 elements) of channel f, element group qb: the
// 16-sample blocks run, run + R, run + 2R, ... of the chunk, so that at any
// time the resident workgroups unpack neighbouring blocks of every channel,
// as a grid of one block per workgroup would (workgroups running together
// read whole input rows and write neighbouring plane rows).  Per block: 16
// word loads per thread (one per sample row; 256 contiguous bytes per wave
// and row; the next block's issued before this one is written), four 4x4
// byte transposes, the codes split into the R and I planes, LDS staging so
// that every global store writes 1 KB of consecutive 16-byte element chunks
// (full cache lines).  Loads and stores go through buffer descriptors,
// branch-free: a block's descriptor covers its valid sample rows only, and
// lanes past N (or past the block's last element) use an offset beyond
// every descriptor, so they read zeros and their stores are dropped;
// streamed once: non-temporal (aux 2).  Source block kb (relative to t0)
// lands at destination block kb_dst0 + kb of a layout with KB_dst blocks.
// Each thread also adds up its four elements' R and I codes over its
// blocks and stores them as one word per element (SR | SI << 16; at most
// 16 * 15 * 256 = 61440 with at most 256 blocks per run) into prow[f]
// [element].  Samples past t_valid and elements past N are code 0.
__device__ __forceinline__ void unpack_u4_run(const uint8_t *__restrict__ in,
                                              uint8_t *__restrict__ out_planes,
                                              uint32_t *__restrict__ prow, int N, int F, int Np,
                                              int64_t row_bytes, int64_t t0, int64_t t_valid,
                                              int KB, int KB_dst, int kb_dst0, int f, int qb,
                                              int run, int R, v4u (*stage)[1024 + 64]) {
  const int tid = threadIdx.x;
  const int q = qb * 256 + tid;  // element quad
  const int rb = (int)row_bytes;  // 16 rows of a block fit in an int
  // lanes that read nothing: an offset past every descriptor (the row
  // offsets added below keep it there, modulo 2^32)
  const uint32_t voff_in = 4 * q < N ? (uint32_t)(f * N + 4 * q) : 0x80000000u;
  auto load = [&](uint32_t (&w)[16], int kb) {
    const int64_t t = 16 * (int64_t)kb;
    const int64_t rows = t_valid - t < 0 ? 0 : (t_valid - t > 16 ? 16 : t_valid - t);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(in + (t0 + t) * row_bytes), (short)0, (int)(rows * row_bytes), 0x00020000);
#pragma unroll
    for (int s = 0; s < 16; s++)
      w[s] = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)(voff_in + (uint32_t)(s * rb)), 0, 2);
  };
  const int64_t plane_stride = (int64_t)F * KB_dst * Np * 16;
  const int e_base = qb * 1024;
  const int n_el = Np - e_base < 1024 ? Np - e_base : 1024;
  uint32_t sr[4] = {0, 0, 0, 0}, si[4] = {0, 0, 0, 0};
  uint32_t w[16];
  load(w, run);
  for (int kb = run; kb < KB; kb += R) {
    uint32_t c[4][4];  // [element][4-sample group]
#pragma unroll
    for (int m = 0; m < 4; m++)
      transpose4(w[4 * m], w[4 * m + 1], w[4 * m + 2], w[4 * m + 3], c[0][m], c[1][m], c[2][m],
                 c[3][m]);
    if (kb + R < KB) load(w, kb + R);
#pragma unroll
    for (int e = 0; e < 4; e++) {
      uint32_t r[4], i[4];
#pragma unroll
      for (int m = 0; m < 4; m++) {
        r[m] = (c[e][m] >> 4) & 0x0F0F0F0Fu;
        i[m] = c[e][m] & 0x0F0F0F0Fu;
      }
      // bytes <= 15: four words add bytewise without carries (<= 60)
      sr[e] = __builtin_amdgcn_sad_u8(r[0] + r[1] + r[2] + r[3], 0u, sr[e]);
      si[e] = __builtin_amdgcn_sad_u8(i[0] + i[1] + i[2] + i[3], 0u, si[e]);
      const int el = 4 * tid + e;
      const int slot = el + (el >> 4);  // one pad chunk per 16: conflict-free b128 writes
      stage[0][slot] = (v4u){r[0], r[1], r[2], r[3]};
      stage[1][slot] = (v4u){i[0], i[1], i[2], i[3]};
    }
    __syncthreads();
    uint8_t *dst = out_planes + (((int64_t)f * KB_dst + kb_dst0 + kb) * Np + e_base) * 16;
    const __amdgpu_buffer_rsrc_t d0 =
        __builtin_amdgcn_make_buffer_rsrc(dst, (short)0, n_el * 16, 0x00020000);
    const __amdgpu_buffer_rsrc_t d1 =
        __builtin_amdgcn_make_buffer_rsrc(dst + plane_stride, (short)0, n_el * 16, 0x00020000);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int el = k * 256 + tid;
      const int slot = el + (el >> 4);
      __builtin_amdgcn_raw_buffer_store_b128(stage[0][slot], d0, el * 16, 0, 2);
      __builtin_amdgcn_raw_buffer_store_b128(stage[1][slot], d1, el * 16, 0, 2);
    }
    __syncthreads();
  }
  if (4 * q < Np) {
    const v4u v = {sr[0] | si[0] << 16, sr[1] | si[1] << 16, sr[2] | si[2] << 16,
                   sr[3] | si[3] << 16};
    *(v4u *)(prow + (int64_t)f * Np + 4 * q) = v;
  }
}

// Plan chunk: grid x = channel, y = element group, z = run (about
// kU4Workgroups workgroups in all); run r writes partials row r.
constexpr int kU4Workgroups = 1024;
constexpr int kU4MaxBlocks = 256;  // blocks per workgroup: the partials fit 16 bits
__host__ __device__ inline int64_t u4_runs(int64_t KB, int64_t F, int64_t Np) {
  const int64_t per = F * ((Np + 1023) / 1024);
  int64_t r = (kU4Workgroups + per - 1) / per;
  const int64_t need = (KB + kU4MaxBlocks - 1) / kU4MaxBlocks;
  if (r < need) r = need;
  return r < KB ? r : KB;
}
__global__ __launch_bounds__(256) void unpack_u4_kernel(
    const uint8_t *__restrict__ in, uint8_t *__restrict__ out_planes,
    uint32_t *__restrict__ partials, int N, int F, int Np, int64_t row_bytes, int64_t t0,
    int64_t t_valid, int KB) {
  __shared__ __attribute__((aligned(16))) v4u stage[2][1024 + 64];  // [plane][element], padded
  unpack_u4_run(in, out_planes, partials + (int64_t)blockIdx.z * F * Np, N, F, Np, row_bytes, t0,
                t_valid, KB, KB, 0, blockIdx.x, blockIdx.y, blockIdx.z, gridDim.z, stage);
}

// Batch: one launch unpacks the chunks of many jobs.  Workgroup g belongs
// to the job j with jobs[j].unit0 <= g < jobs[j+1].unit0 (binary search);
// its index within the job is (run, element group, channel), channel
// fastest so that neighbours read whole input rows; run r writes the job's
// partial row prow0 + r.
__global__ __launch_bounds__(256) void unpack_u4_batch_kernel(const UnpackJob *__restrict__ jobs,
                                                              int njobs) {
  __shared__ __attribute__((aligned(16))) v4u stage[2][1024 + 64];
  const int64_t g = blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].unit0 <= g) lo = mid; else hi = mid - 1;
  }
  const UnpackJob &J = jobs[lo];
  const int64_t u = g - J.unit0;
  const int nq = (J.Np + 1023) / 1024;
  if (u >= (int64_t)nq * J.runs * J.F) return;  // grid padding (uniform per workgroup)
  const int f = (int)(u % J.F);
  const int qb = (int)((u / J.F) % nq);
  const int run = (int)(u / ((int64_t)J.F * nq));
  unpack_u4_run(J.in, J.planes, J.partials + (int64_t)(J.prow0 + run) * J.F * J.Np, J.N, J.F,
                J.Np, J.row_bytes, J.t0, J.t_valid, J.KB, J.KB_dst, J.kb_dst0, f, qb, run, J.runs,
                stage);
}

// The offset corrections of one chunk (n valid samples): per channel f and
// element e, SR and SI = its code sums over the chunk (the nrun partials),
// then corr[0][f][e] = (-8 (SR + SI) + 128 n, -8 SR - 7 SI) for its rows and
// corr[1][f][e] = (-8 (SR + SI), 8 (SR - SI)) for its columns.  Workgroup =
// 64 elements x 4 run quarters (lanes of a wave read 64 consecutive
// partials), added up through LDS.
__device__ __forceinline__ void u4_corrections_block(const uint32_t *__restrict__ partials,
                                                     int32_t *__restrict__ corr, int F, int Np,
                                                     int nrun, int32_t n, int64_t blk) {
  __shared__ uint32_t part[2][4][64];
  const int64_t FN = (int64_t)F * Np;
  const int lane = threadIdx.x & 63, quarter = threadIdx.x >> 6;
  const int64_t i = blk * 64 + lane;  // (f, e) flattened
  uint32_t sr = 0, si = 0;
  if (i < FN) {
    for (int r = quarter; r < nrun; r += 4) {
      const uint32_t v = partials[(int64_t)r * FN + i];
      sr += v & 0xFFFFu;
      si += v >> 16;
    }
  }
  part[0][quarter][lane] = sr;
  part[1][quarter][lane] = si;
  __syncthreads();
  if (quarter == 0 && i < FN) {
    const int32_t SR = (int32_t)(part[0][0][lane] + part[0][1][lane] + part[0][2][lane] +
                                 part[0][3][lane]);
    const int32_t SI = (int32_t)(part[1][0][lane] + part[1][1][lane] + part[1][2][lane] +
                                 part[1][3][lane]);
    ((v2i *)corr)[i] = (v2i){-8 * (SR + SI) + 128 * n, -8 * SR - 7 * SI};
    ((v2i *)corr)[FN + i] = (v2i){-8 * (SR + SI), 8 * (SR - SI)};
  }
}

__global__ __launch_bounds__(256) void u4_corrections_kernel(const uint32_t *__restrict__ partials,
                                                             int32_t *__restrict__ corr, int F,
                                                             int Np, int nrun, int32_t n) {
  u4_corrections_block(partials, corr, F, Np, nrun, n, blockIdx.x);
}

// Many jobs in one launch: workgroup g belongs to the job j with
// jobs[j].blk0 <= g < jobs[j+1].blk0 (binary search).
__global__ __launch_bounds__(256) void u4_corrections_batch_kernel(
    const CorrJob *__restrict__ jobs, int njobs) {
  const int64_t g = blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].blk0 <= g) lo = mid; else hi = mid - 1;
  }
  const CorrJob &J = jobs[lo];
  const int64_t blk = g - J.blk0;
  if (blk * 64 >= (int64_t)J.F * J.Np) return;  // grid padding (uniform per workgroup)
  u4_corrections_block(J.partials, J.corr, J.F, J.Np, J.nrun, J.n, blk);
}

// Pass 2, one output item of channel f over 64-sample stages [st_begin,
// st_begin + nst) (nst a multiple of kMfmaRing), run as 32-sample k-steps.
// Wave w computes 128 rows x 64 columns as 4x2 blocks of two int32
// accumulators (256 AGPRs):
//   !PAIR: rows from kTile ti_a, columns from kTile ct + 64 w (a 128 x 256
//          tile; the four waves share the rows);
//    PAIR: two half tiles of column tile ct: waves 0-1 rows from kTile
//          ti_a, waves 2-3 rows from kTile ti_b, columns kTile ct + 64 (w%2);
// RE' += Ar.Br + Ai.Bi,  IM' += Ar.Bi + Ai.(15 - Br) on offset-binary codes
// (bytes in [0, 15]; 15 - u per byte in registers, one XOR per word), and
// the flush adds the offset corrections J.corr.  The
// row operands, shared by the four waves, come through the LDS ring by
// buffer->LDS DMA; each wave loads its own column operands straight into
// VGPRs, three k-steps ahead, so every operand byte crosses the CU's load
// path once.
// SKIP: bit 2i + j set = the wave's block (row group i, column group j) is
// never stored (below the diagonal of a diagonal tile), so its MFMAs are
// left out, and so are the row operand reads of a row group with both
// blocks left out.  Every column group must keep an MFMA: a column load
// that no MFMA uses would be deleted, and the vmcnt waits below count on
// every wave issuing the same loads per k-step.  Loads, DMAs and barriers
// are then the same in every variant, so the waves of one workgroup may
// run different ones.
template <bool PAIR, int SKIP = 0>
__device__ __forceinline__ void gemm_tile(const GemmJob &J, const int f, const int ti_a,
                                          const int ti_b, const int ct, const int st_begin,
                                          const int nst, uint8_t (*lds)[kSlotBytes]) {
  const int F = J.F, Np = J.Np, KB = J.KB, nb = J.nb, U = J.U, mode = J.mode;
  const int64_t NB = J.NB;
  int32_t *vis = J.vis;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int k0 = 2 * st_begin, nk = 2 * nst;  // k-steps
  auto skipped = [](int i, int j) constexpr { return ((SKIP >> (2 * i + j)) & 1) != 0; };
  static_assert((SKIP & 0x55) != 0x55 && (SKIP & 0xAA) != 0xAA,
                "every column group keeps an MFMA (its loads stay in the counted waits)");

  // An operand (32 elements x 32 samples) is 1 KB: lanes 0-31 the k-step's
  // even 16-sample block, 32-63 the odd one.
  const int64_t blk_bytes = (int64_t)Np * 16;  // one 16-sample block of a channel
  __amdgpu_buffer_rsrc_t r[2];
#pragma unroll
  for (int pl = 0; pl < 2; pl++)
    r[pl] = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(J.planes + ((int64_t)pl * F + f) * KB * blk_bytes), (short)0,
        (int)(KB * blk_bytes), 0x00020000);
  // Row DMAs.  !PAIR: wave w fills plane w/2, row groups 2(w&1), 2(w&1)+1
  // of the one row block.  PAIR: wave w fills plane w&1, all four row
  // groups of row block w/2 (rows ti_a, ti_b; ring w/2 of two).  Row group
  // g holds elements kTile ti + 32 g (always inside the array).
  constexpr int kDma = PAIR ? 4 : 2;  // row DMAs per wave and k-step
  const int rb = PAIR ? wave >> 1 : 0;          // this wave's row block
  const int lp = PAIR ? wave & 1 : wave >> 1;   // plane it fills
  const int g0 = PAIR ? 0 : 2 * (wave & 1);     // first row group it fills
  const int ti = PAIR && rb ? ti_b : ti_a;      // this wave's row tile
  int voff[kDma];
#pragma unroll
  for (int gg = 0; gg < kDma; gg++)
    voff[gg] = (ti * kTile + 32 * (g0 + gg) + (lane & 31)) * 16 + (lane >> 5) * (int)blk_bytes;
  auto issue_rows = [&](int ks, int slot) {
    const int soff = (k0 + ks) * 2 * (int)blk_bytes;
#pragma unroll
    for (int gg = 0; gg < kDma; gg++) {
      lds_void *dst = (lds_void *)&lds[rb * kRing + slot][(lp * kRowGroups + g0 + gg) * 1024];
      __builtin_amdgcn_raw_ptr_buffer_load_lds(r[lp], dst, 16, voff[gg], soff, 0, 0);
    }
  };
  // Column operands of this wave: column groups 2 cw + j of the tile.
  // Columns past the array (waves 2 and 3 of an unpaired half tile) load
  // through an empty descriptor, which fetches nothing and returns zeros;
  // their products are never stored.
  const bool past = !PAIR && wave >= 2 && (ct + 1) * kTile >= Np;
  const int cw = PAIR ? wave & 1 : wave;  // column quarter of the tile
  const __amdgpu_buffer_rsrc_t rz = __builtin_amdgcn_make_buffer_rsrc(
      (void *)J.planes, (short)0, 0, 0x00020000);
  const __amdgpu_buffer_rsrc_t c0 = past ? rz : r[0], c1 = past ? rz : r[1];
  int boff[2];
#pragma unroll
  for (int j = 0; j < 2; j++) {
    int e = ct * kTile + 64 * cw + 32 * j;
    if (e >= Np) e -= kTile;
    boff[j] = (e + (lane & 31)) * 16 + (lane >> 5) * (int)blk_bytes;
  }
  auto load_cols = [&](v4i &R0, v4i &R1, v4i &I0, v4i &I1, int ks) {
    const int soff = (k0 + ks) * 2 * (int)blk_bytes;
    R0 = (v4i)__builtin_amdgcn_raw_buffer_load_b128(c0, boff[0], soff, 0);
    R1 = (v4i)__builtin_amdgcn_raw_buffer_load_b128(c0, boff[1], soff, 0);
    I0 = (v4i)__builtin_amdgcn_raw_buffer_load_b128(c1, boff[0], soff, 0);
    I1 = (v4i)__builtin_amdgcn_raw_buffer_load_b128(c1, boff[1], soff, 0);
  };

  v16i re[4][2], im[4][2];
  const v16i zero = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) re[i][j] = im[i][j] = zero;
  const uint8_t *lds0 = &lds[0][0];
  const int bA = rb * kRingBytes + lane * 16;
  auto load_op = [&](int slot, int plane, int k) -> v4i {
    return *(const v4i *)(lds0 + bA + (slot * kSlotBytes + plane * kRowGroups * 1024 + k * 1024));
  };
  // Row operands of the current k-step (row group 3 double-buffered: its
  // next operands are read before its MFMAs end).
  v4i ar[4], ai[4], ar3n, ai3n;
  // One k-step with column operands (br0, br1, bi0, bi1): for each row
  // group i its 8 MFMAs, after which that group's operands of the next
  // k-step (slot nslot) replace it.
  auto step_ops = [&](const v4i &br0, const v4i &br1, const v4i &bi0, const v4i &bi1,
                      int nslot) {
    const v4i br[2] = {br0, br1}, bi[2] = {bi0, bi1};
    // 15 - Br per byte: IM' += Ar.Bi + Ai.(15 - Br)
    const v4i nbr[2] = {br0 ^ 0x0F0F0F0F, br1 ^ 0x0F0F0F0F};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const v4i xr = i == 3 ? ar3n : ar[i], xi = i == 3 ? ai3n : ai[i];
#pragma unroll
      for (int j = 0; j < 2; j++) {
        if (skipped(i, j)) continue;
        re[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(xr, br[j], re[i][j], 0, 0, 0);
        im[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(xr, bi[j], im[i][j], 0, 0, 0);
      }
#pragma unroll
      for (int j = 0; j < 2; j++) {
        if (skipped(i, j)) continue;
        re[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(xi, bi[j], re[i][j], 0, 0, 0);
        im[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(xi, nbr[j], im[i][j], 0, 0, 0);
      }
      // (the reads of a row group no MFMA uses are dead and left out)
      if (i < 3) {
        ar[i] = load_op(nslot, 0, i);
        ai[i] = load_op(nslot, 1, i);
      }
      if (i == 1) {  // the next k-step's row group 3, into the second buffer
        ar[3] = load_op(nslot, 0, 3);
        ai[3] = load_op(nslot, 1, 3);
      }
    }
    ar3n = ar[3];
    ai3n = ai[3];
  };
  const int by0 = ti * kTile / 32, bx0 = (ct * kTile + 64 * cw) / 32;

  // Issue groups: G(ks), issued during k-step ks, = the row DMAs of k-step
  // ks+4 (kDma) and the column loads of ks+3 (4).  Prologue: G(-4) = DMAs
  // of k-step 0, G(-3..-1) = DMAs of 1..3 and column loads of 0..2 (sched
  // barriers keep the issue order the waits count on).  Wait for this
  // wave's row DMAs of k-step 0 (3 |G| later operations may stay pending),
  // barrier, read them.  Column operands of k-step s live in set s % 4:
  // k-step ks loads ks+3 into the set k-step ks-1 has finished.
  static_assert(kRing == 4, "column loads three k-steps ahead, row DMAs four");
  auto cl = [&](int s) { return s < nk ? s : nk - 1; };
  v4i BR[4][2], BI[4][2];
  issue_rows(0, 0);
  __builtin_amdgcn_sched_barrier(0);
  issue_rows(cl(1), 1);
  load_cols(BR[0][0], BR[0][1], BI[0][0], BI[0][1], 0);
  __builtin_amdgcn_sched_barrier(0);
  issue_rows(cl(2), 2);
  load_cols(BR[1][0], BR[1][1], BI[1][0], BI[1][1], cl(1));
  __builtin_amdgcn_sched_barrier(0);
  issue_rows(cl(3), 3);
  load_cols(BR[2][0], BR[2][1], BI[2][0], BI[2][1], cl(2));
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (PAIR)
    asm volatile("s_waitcnt vmcnt(24)\n\ts_barrier" ::: "memory");
  else
    asm volatile("s_waitcnt vmcnt(18)\n\ts_barrier" ::: "memory");
#pragma unroll
  for (int i = 0; i < 3; i++) {
    ar[i] = load_op(0, 0, i);
    ai[i] = load_op(0, 1, i);
  }
  ar3n = load_op(0, 0, 3);
  ai3n = load_op(0, 1, 3);

  // K-step ks (slot ks % kRing):
  //   wait until G(ks-3) landed (this wave's DMAs of ks+1 and column loads
  //   of ks; G(ks-2), G(ks-1), 2 |G| operations, may stay pending), barrier: every wave's DMAs
  //   of ks+1 landed and every wave's reads of slot ks are done; refill
  //   slot ks with k-step ks+4, load the columns of ks+3; MFMAs of ks,
  //   reading ks+1's row operands as ks's die.
  for (int k00 = 0; k00 < nk; k00 += kRing) {
#pragma unroll
    for (int d = 0; d < kRing; d++) {
      const int ks = k00 + d, n3 = (d + 3) % kRing;
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (PAIR)
        asm volatile("s_waitcnt vmcnt(16)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
      else
        asm volatile("s_waitcnt vmcnt(12)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
      issue_rows(cl(ks + 4), d);
      load_cols(BR[n3][0], BR[n3][1], BI[n3][0], BI[n3][1], cl(ks + 3));
      step_ops(BR[d][0], BR[d][1], BI[d][0], BI[d][1], (d + 1) % kRing);
      // 8 MFMAs per row group i (4 per block not left out); VMEM ops (kDma
      // DMAs, 4 column loads) over groups 0-1, LDS reads of the next row
      // operands from group 1 on (group 1: row group 0's, 2: 1's and 3's,
      // 3: 2's)
      constexpr int m0 = 4 * (!skipped(0, 0) + !skipped(0, 1));
      constexpr int m1 = 4 * (!skipped(1, 0) + !skipped(1, 1));
      constexpr int m2 = 4 * (!skipped(2, 0) + !skipped(2, 1));
      constexpr int m3 = 4 * (!skipped(3, 0) + !skipped(3, 1));
#pragma unroll
      for (int i = 0; i < 8; i++) {
        if (i < m0) __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        if (i < (PAIR ? 4 : 3)) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
      }
#pragma unroll
      for (int i = 0; i < 8; i++) {
        if (i < m1) __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        if (i < (m0 ? 2 : 0)) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        if (i >= 4 && i < (PAIR ? 8 : 7)) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
      }
#pragma unroll
      for (int i = 0; i < 8; i++) {
        if (i < m2) __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        if (i < (m1 ? 2 : 0) + (m3 ? 2 : 0)) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      }
#pragma unroll
      for (int i = 0; i < 8; i++) {
        if (i < m3) __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        if (i < (m2 ? 2 : 0)) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // Code products add at most 450 per sample: < 2^26 per item, no int32
  // overflow.
  // Corrections: the piece that starts the chunk adds them, once.
  const bool first = st_begin == 0;
  const v2i *rowc = (const v2i *)J.corr + (int64_t)f * Np;
  const v2i *colc = (const v2i *)J.corr + ((int64_t)F + f) * Np;
  v2i cc[2];
#pragma unroll
  for (int j = 0; j < 2; j++)
    cc[j] = first ? colc[(bx0 + j) * 32 + (lane & 31)] : (v2i){0, 0};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int by = by0 + i;
    // rows of register j: 32 by + 8 (j/4) + 4 (lane/32) + j%4
    v2i rc[16];
#pragma unroll
    for (int g = 0; g < 4; g++) {
      v4i lo = {0, 0, 0, 0}, hi = {0, 0, 0, 0};
      if (first) {
        const v2i *p = rowc + by * 32 + 8 * g + 4 * (lane >> 5);
        lo = *(const v4i *)p;
        hi = *(const v4i *)(p + 2);
      }
      rc[4 * g] = (v2i){lo.x, lo.y};
      rc[4 * g + 1] = (v2i){lo.z, lo.w};
      rc[4 * g + 2] = (v2i){hi.x, hi.y};
      rc[4 * g + 3] = (v2i){hi.z, hi.w};
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int bx = bx0 + j;
      if (bx >= by && bx < nb) {
        const int64_t b = (int64_t)by * nb - (int64_t)by * (by - 1) / 2 + (bx - by);
        flush_block(vis + ((int64_t)f * NB + b) * 2048, re[i][j], im[i][j], U, mode, lane,
                       rc, cc[j]);
      }
    }
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// A tile with each wave on its own variant: the waves whose blocks include
// some below the diagonal (tile_skip_mask) leave out their MFMAs.  A wave
// past the array (mask 0xFF) runs the full variant: its column loads must
// stay.
template <bool PAIR>
__device__ __forceinline__ void gemm_tile_masked(const GemmJob &J, const int f, const int ti_a,
                                                 const int ti_b, const int ct,
                                                 const int st_begin, const int nst,
                                                 uint8_t (*lds)[kSlotBytes]) {
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int m = tile_skip_mask(PAIR, false, ti_a, ti_b, ct, J.Np, wave);
  if (m == 0xF4) return gemm_tile<PAIR, 0xF4>(J, f, ti_a, ti_b, ct, st_begin, nst, lds);
  if (m == 0x40) return gemm_tile<PAIR, 0x40>(J, f, ti_a, ti_b, ct, st_begin, nst, lds);
  gemm_tile<PAIR, 0>(J, f, ti_a, ti_b, ct, st_begin, nst, lds);
}

// One plan chunk: items = channels x tiles.  Workgroups 0..head-1 take
// items 0..head-1 whole (all stages, mode J.mode); the rest take the tail
// items head.. cut into time pieces of J.stages_per_split stages (the last
// piece the rest), mode tmode.  Each range is XCD-remapped.
__global__ __launch_bounds__(256, 1) void corr_mfma_kernel(GemmJob J, int nst, int head,
                                                           int tail, int tmode) {
  __shared__ __attribute__((aligned(1024))) uint8_t lds[2 * kRing][kSlotBytes];  // two rings
  const int64_t per_f = paired_tiles(J.ntiles);
  int64_t item;
  int s0, ns;
  if ((int)blockIdx.x < head) {
    item = xcd_remap(blockIdx.x, head);
    s0 = 0;
    ns = nst;
  } else {
    const int64_t pieces = (int64_t)gridDim.x - head;
    const int64_t r = xcd_remap((int64_t)blockIdx.x - head, pieces);
    item = head + r % tail;
    s0 = (int)(r / tail) * J.stages_per_split;
    ns = nst - s0 < J.stages_per_split ? nst - s0 : J.stages_per_split;
    J.mode = tmode;
  }
  int ti_a, ti_b, ct;
  tile_coords((int)(item % per_f), J.ntiles, ti_a, ti_b, ct);
  if (ti_b >= 0 && ti_b != ti_a)
    gemm_tile_masked<true>(J, (int)(item / per_f), ti_a, ti_b, ct, s0, ns, lds);
  else
    gemm_tile_masked<false>(J, (int)(item / per_f), ti_a, ti_a, ct, s0, ns, lds);
}

// Batch: persistent workgroups take work items from 8 per-XCD queues: a
// workgroup on XCD x = blockIdx % 8 drains queue x, then moves on to the
// others; queues never refill, so every workgroup ends.
__global__ __launch_bounds__(256, 1) void corr_mfma_batch_kernel(
    const GemmJob *__restrict__ jobs, const GemmItem *__restrict__ items,
    const int *__restrict__ qbound, int *__restrict__ counters) {
  // One LDS object: the ring and, past it, the item handed to the
  // workgroup.  (A second __shared__ variable makes the compiler wait for
  // every buffer->LDS load before each LDS read in the loop.)
  __shared__ __attribute__((aligned(1024))) uint8_t lds[2 * kRing + 1][kSlotBytes];
  int &next = *(int *)&lds[2 * kRing][0];
  int q = blockIdx.x % 8, moves = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      int it = -1;
      while (moves < 8) {
        const int k = atomicAdd(&counters[q], 1) + qbound[q];
        if (k < qbound[q + 1]) {
          it = k;
          break;
        }
        q = (q + 1) % 8;
        moves++;
      }
      next = it;
    }
    __syncthreads();
    const int it = next;
    __syncthreads();
    if (it < 0) break;
    const GemmItem w = items[it];
    const GemmJob J = jobs[w.job];
    int ti_a, ti_b, ct;
    tile_coords(w.tile, J.ntiles, ti_a, ti_b, ct);
    if (ti_b >= 0 && ti_b != ti_a)
      gemm_tile_masked<true>(J, w.f, ti_a, ti_b, ct, w.stage0, w.nstages, lds);
    else
      gemm_tile_masked<false>(J, w.f, ti_a, ti_a, ct, w.stage0, w.nstages, lds);
  }
}

}  // namespace

// Samples per time chunk so that the plane workspace stays within
// kWorkspaceBytes (a multiple of the stage length).
static int64_t chunk_samples(const chx_plan *p) {
  const int64_t Np = (int64_t)p->ntiles * kTile;
  const int64_t per_sample = 2 * (int64_t)p->cfg.num_freq * Np;
  int64_t tc = kWorkspaceBytes / per_sample;
  // one (plane, channel) region must stay addressable by a buffer descriptor
  if (tc > 0x7FFFFFFFLL / Np) tc = 0x7FFFFFFFLL / Np;
  if (tc > 16LL * 65535) tc = 16LL * 65535;  // unpack grid.y = KB <= 65535
  tc = tc / (kMfmaStage * kMfmaRing) * (kMfmaStage * kMfmaRing);
  const int64_t T = (p->cfg.time_steps + kMfmaStage * kMfmaRing - 1) / (kMfmaStage * kMfmaRing) *
                    (kMfmaStage * kMfmaRing);
  if (tc > T) tc = T;
  if (tc < kMfmaStage * kMfmaRing) tc = kMfmaStage * kMfmaRing;
  return tc;
}

// The workspace of one chunk: code planes [2][F][KB][Np][16 B], then the
// unpack's code-sum partials [run][F][Np] u32, then the offset corrections
// [2][F][Np] int32 pairs.
static int64_t u4_plane_bytes(const chx_plan *p) {
  return 2 * (int64_t)p->cfg.num_freq * p->ntiles * kTile * chunk_samples(p);
}
static int64_t u4_partial_bytes(const chx_plan *p) {
  const int64_t Np = (int64_t)p->ntiles * kTile;
  return u4_runs(chunk_samples(p) / 16, p->cfg.num_freq, Np) * p->cfg.num_freq * Np * 4;
}

size_t mfma_workspace_bytes(const chx_plan *p) {
  const int64_t Np = (int64_t)p->ntiles * kTile;
  return (size_t)(u4_plane_bytes(p) + u4_partial_bytes(p) + 2 * (int64_t)p->cfg.num_freq * Np * 8);
}

int launch_corr_mfma(const chx_plan *p, const uint8_t *in, int32_t *vis, int mode,
                     hipStream_t s) {
  const int N = p->cfg.num_elements, F = p->cfg.num_freq, U = p->cfg.upper_triangle_convention;
  const int64_t T = p->cfg.time_steps;
  const int ntiles = p->ntiles, Np = ntiles * kTile;
  const int64_t per_f = paired_tiles(ntiles);
  const int64_t items = (int64_t)F * per_f;
  if (items > 0x7FFFFFFF || !p->d_work) return CHX_ERR_SHAPE;
  int cus = 0;
  CHX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->cfg.device));
  if (cus < 1) cus = 1;
  const int64_t tc = chunk_samples(p);
  const int64_t row_bytes = (int64_t)F * N;
  if (row_bytes > 0x7FFFFFFF / 16) {  // the unpack addresses 16 input rows with 32-bit offsets
    set_error_detail("int8 matrix-core method: num_freq * num_elements must be below 2^27");
    return CHX_ERR_SHAPE;
  }
  const int ring = kMfmaRing;  // stage counts are multiples of it
  for (int64_t t0 = 0; t0 < T; t0 += tc) {
    const int64_t tv = T - t0 < tc ? T - t0 : tc;  // valid samples
    // 64-sample stages, padded to a multiple of the ring (the padding is 0)
    const int nst = (int)((tv + kMfmaStage * ring - 1) / (kMfmaStage * ring)) * ring;
    const int KB = nst * (kMfmaStage / 16);
    // Launch shape (one dispatch).  Whole rounds of one workgroup per CU
    // take whole items (the chunk in one piece, plain stores); the items of
    // a last partial round are cut in time into the fewest pieces that
    // minimise its length and add atomically into zeroed output.  A chunk
    // longer than kMaxItemStages cuts every item into pieces of at most
    // kMaxItemStages stages (atomic).
    int64_t head = 0;
    int split = 1;
    if (nst <= kMaxItemStages) head = items / cus * cus;
    const int64_t tail = items - head;
    if (tail > 0) {
      double best = (double)((tail + cus - 1) / cus);
      for (int sp = 2; sp <= 64 && nst / sp >= 16; sp++) {
        const double r = (double)((tail * sp + cus - 1) / cus) / sp;
        if (r < best - 1e-9) {
          best = r;
          split = sp;
        }
      }
    }
    int sps = ((nst + split - 1) / split + ring - 1) / ring * ring;
    if (sps > kMaxItemStages) sps = kMaxItemStages;
    split = (nst + sps - 1) / sps;
    const int kmode = (t0 == 0) ? mode : 1;
    const int tmode = split > 1 ? 2 : kmode;
    if (tail > 0 && split > 1 && kmode == 0) {
      const int64_t f0 = head / per_f;  // zero the channels the tail touches
      int rc = launch_fill_u32((uint32_t *)vis + f0 * p->NB * 2048, 0u,
                               (int64_t)(F - f0) * p->NB * 2048, s);
      if (rc) return rc;
    }
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    if (p->prof) {
      for (hipEvent_t &e : ev) CHX_HIP(hipEventCreate(&e));
      p->prof_ev->insert(p->prof_ev->end(), ev, ev + 3);
      CHX_HIP(hipEventRecord(ev[0], s));
    }
    uint32_t *partials = (uint32_t *)(p->d_work + u4_plane_bytes(p));
    int32_t *corr = (int32_t *)(p->d_work + u4_plane_bytes(p) + u4_partial_bytes(p));
    const int64_t nrun = u4_runs(KB, F, Np);
    unpack_u4_kernel<<<dim3((unsigned)F, (unsigned)((Np + 1023) / 1024), (unsigned)nrun),
                       dim3(256), 0, s>>>(in, p->d_work, partials, N, F, Np, row_bytes, t0, tv,
                                          KB);
    CHX_HIP(hipGetLastError());
    int rc1 = launch_u4_corrections(partials, corr, F, Np, (int)nrun, (int32_t)tv, s);
    if (rc1) return rc1;
    if (p->prof) CHX_HIP(hipEventRecord(ev[1], s));
    GemmJob J;
    J.planes = p->d_work;
    J.corr = corr;
    J.planes1 = nullptr;
    J.vis = vis;
    J.NB = p->NB;
    J.F = F;
    J.Np = Np;
    J.KB = KB;
    J.ntiles = ntiles;
    J.nb = p->nb;
    J.U = U;
    J.mode = kmode;
    J.stages_per_split = sps;
    J.small = 0;
    const int64_t grid = head + tail * split;
    if (grid > 0x7FFFFFFF) return CHX_ERR_SHAPE;
    corr_mfma_kernel<<<dim3((unsigned)grid), dim3(256), 0, s>>>(J, nst, (int)head, (int)tail,
                                                                  tmode);
    CHX_HIP(hipGetLastError());
    if (p->prof) CHX_HIP(hipEventRecord(ev[2], s));
  }
  return CHX_SUCCESS;
}

int launch_u4_corrections(const uint32_t *partials, int32_t *corr, int F, int Np, int nrun,
                          int32_t n, hipStream_t s) {
  const int64_t FN = (int64_t)F * Np;
  u4_corrections_kernel<<<dim3((unsigned)((FN + 63) / 64)), dim3(256), 0, s>>>(partials, corr, F,
                                                                             Np, nrun, n);
  CHX_HIP(hipGetLastError());
  return CHX_SUCCESS;
}

int launch_u4_corrections_batch(const CorrJob *d_jobs, int njobs, int64_t blocks,
                                hipStream_t s) {
  if (blocks <= 0) return CHX_SUCCESS;
  const int64_t gx = blocks < 65536 ? blocks : 65536;
  const int64_t gy = (blocks + gx - 1) / gx;
  u4_corrections_batch_kernel<<<dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s>>>(d_jobs,
                                                                                       njobs);
  CHX_HIP(hipGetLastError());
  return CHX_SUCCESS;
}

int launch_unpack_batch_int8(const UnpackJob *d_jobs, int njobs, int64_t units, hipStream_t s) {
  if (units <= 0) return CHX_SUCCESS;
  const int64_t gx = units < 65536 ? units : 65536;
  const int64_t gy = (units + gx - 1) / gx;
  unpack_u4_batch_kernel<<<dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s>>>(d_jobs, njobs);
  CHX_HIP(hipGetLastError());
  return CHX_SUCCESS;
}

int launch_gemm_batch_int8(const GemmJob *d_jobs, const GemmItem *d_items, const int *d_qbound,
                           int *d_counters, int nwg, hipStream_t s) {
  corr_mfma_batch_kernel<<<dim3((unsigned)nwg), dim3(256), 0, s>>>(d_jobs, d_items, d_qbound,
                                                                   d_counters);
  CHX_HIP(hipGetLastError());
  return CHX_SUCCESS;
}

}  // namespace chx
