// chx_mfma.hip -- the correlator on CDNA4 int8 matrix cores.
//
// Each channel's visibilities are a Hermitian rank-T update.  With the
// samples split into a real plane r = R-8 and an imaginary plane i = I-8
// (one int8 per value), a 32x32 reference block (rows y, columns x,
// pairwise_correlator_UT.cl:161-201) is four int8 GEMMs over time:
//   RE = Ar.Br^T + Ai.Bi^T      (sum y_r x_r + y_i x_i)
//   IX = Ai.Br^T                (sum y_i x_r)
//   IY = Ar.Bi^T                (sum y_r x_i)
//   im = IX - IY for upper_triangle_convention 1 (cpu_corr_test.c:189),
//        IY - IX for 0 (cpu_corr_test.c:92).
// No offset/preseed correction is needed and there is no time_accum cap.
//
// Two passes per chunk of time:
//  1. unpack_planes_kernel (HBM-bound): [time][freq][element] bytes ->
//     MFMA-ready planes [plane][freq][16-sample block][element][16 B].  A
//     byte with nibbles (R, I) is XORed with 0x88 so each nibble holds its
//     value's 4-bit two's complement; placed in bits 6..3 of an int8 with
//     its sign extended into bit 7 it is 8*value exactly.  Elements past N
//     and samples past T are 0.
//  2. corr_mfma_kernel (matrix-core-bound): a workgroup = 4 wavefronts = a
//     128x128-element tile of one channel (16 reference blocks), each wave a
//     64x64 quarter (2x2 blocks x 3 int32 accumulators), fed by buffer loads
//     that go straight to LDS (no VGPR staging, no VALU) through a ring of
//     4 stages of 64 samples; waves only read operands and issue MFMAs.
// Every product is 64x the true one; each accumulator holds one product
// term over at most kChunk samples (|sum| <= 64*64*kChunk = 2^29) and is
// divided by 64 (exact) when flushed to the int32 output.
#include "chx_internal.h"

namespace chx {
namespace {

typedef int v2i __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;

constexpr int kTile = kMfmaTile;   // elements per tile side
constexpr int kStage = kMfmaStage; // samples per LDS stage (2 MFMA k-steps)
constexpr int kRing = kMfmaRing;    // LDS stages in flight
constexpr int kChunk = 131072;  // samples accumulated before a flush
// LDS stage layout: [plane 2][row group 8][16-sample block 4][32 rows][16 B].
// A 32-row, 32-sample operand is 1 KB contiguous: lanes 0-31 read block
// 2k, lanes 32-63 block 2k+1, 16 lanes = 256 contiguous bytes (no bank
// conflicts); one 64-lane direct load fills exactly one operand.
constexpr int kGroupBytes = 4 * 512;
constexpr int kStageBytes = 2 * 8 * kGroupBytes;  // 32 KB

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

// Real and imaginary planes of four packed samples, each value times 8.
__device__ __forceinline__ void planes(uint32_t w, uint32_t &re, uint32_t &im) {
  const uint32_t t = w ^ 0x88888888u;  // per byte: R two's complement in 7..4, I in 3..0
  re = ((t >> 1) & 0x78787878u) | (t & 0x80808080u);         // 8 R, per byte
  im = ((t << 3) & 0x78787878u) | ((t << 4) & 0x80808080u);  // 8 I, per byte
}

// Maps the dispatch-order workgroup id so that the workgroups resident on
// one XCD (dispatch is round-robin over the 8 XCDs) take consecutive
// (channel, tile) work items and share their input rows in that XCD's L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t id, int64_t total) {
  const int64_t q = total / 8, r = total % 8;
  const int64_t xcd = id % 8, slot = id / 8;
  return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// Flush one block's accumulators.  Lane l owns column n = l%32 and rows
// m = 8*(j/4) + 4*(l/32) + j%4 of register j (v_mfma_i32_32x32x32_i8 D
// layout).  mode 0 assigns, 1 adds, 2 adds atomically.
__device__ __forceinline__ void flush_block(int32_t *blk, const v16i &rea, const v16i &reb,
                                            const v16i &ix, const v16i &iy, int U, int mode,
                                            int lane) {
  // Buffer ops on a descriptor of the block (a wave-uniform base): one VGPR
  // of lane offset, the register's row as a constant scalar offset.
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc(blk, (short)0, 2048 * 4, 0x00020000);
  const int voff = ((lane >> 5) * 4 * 32 + (lane & 31)) * 8;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int32_t vr = (rea[j] >> 6) + (reb[j] >> 6);
    const int32_t vi = U ? (ix[j] >> 6) - (iy[j] >> 6) : (iy[j] >> 6) - (ix[j] >> 6);
    const int soff = (8 * (j >> 2) + (j & 3)) * 32 * 8;  // row m = 8(j/4) + 4(l/32) + j%4
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

// Pass 1.  Workgroup = 256 element quads (1024 elements) x 16 samples of
// one channel.  Thread: 16 word loads (one per sample row; 256 contiguous
// bytes per wave and row), four 4x4 byte transposes, plane split; the
// planes go through LDS so that every global store writes 1 KB of
// consecutive 16-byte element chunks (full cache lines).
// KB is the destination's block count per channel (its layout), kb the
// destination block written (kb - kb_src0 the source's 16-sample block).
__device__ __forceinline__ void unpack_unit(const uint8_t *__restrict__ in,
                                            uint8_t *__restrict__ out_planes, int N, int F,
                                            int Np, int64_t row_bytes, int64_t t0,
                                            int64_t t_valid, int KB, int qb, int kb,
                                            int kb_src0, int f, v4u (*stage)[1024 + 64]) {
  const int tid = threadIdx.x;
  const int q = qb * 256 + tid;  // element quad
  const bool evalid = 4 * q < N;
  uint32_t w[16];
  const int64_t ks = kb - kb_src0;  // source block
  const uint8_t *src = in + (t0 + 16 * ks) * row_bytes + (int64_t)f * N + 4 * q;
#pragma unroll
  for (int s = 0; s < 16; s++) {
    const bool ok = evalid && (16 * ks + s < t_valid);
    w[s] = ok ? *(const uint32_t *)(src + s * row_bytes) : 0x88888888u;
  }
  uint32_t c[4][4];  // [element][4-sample group]
#pragma unroll
  for (int m = 0; m < 4; m++)
    transpose4(w[4 * m], w[4 * m + 1], w[4 * m + 2], w[4 * m + 3], c[0][m], c[1][m], c[2][m],
               c[3][m]);
#pragma unroll
  for (int e = 0; e < 4; e++) {
    uint32_t r[4], i[4];
#pragma unroll
    for (int m = 0; m < 4; m++) planes(c[e][m], r[m], i[m]);
    const int el = 4 * tid + e;
    const int slot = el + (el >> 4);  // one pad chunk per 16: conflict-free b128 writes
    stage[0][slot] = (v4u){r[0], r[1], r[2], r[3]};
    stage[1][slot] = (v4u){i[0], i[1], i[2], i[3]};
  }
  __syncthreads();
  const int64_t plane_stride = (int64_t)F * KB * Np * 16;
  const int e_base = qb * 1024;
  uint8_t *dst = out_planes + (((int64_t)f * KB + kb) * Np + e_base) * 16;
  const int n_el = Np - e_base < 1024 ? Np - e_base : 1024;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int el = k * 256 + tid;
    if (el < n_el) {
      const int slot = el + (el >> 4);
      *(v4u *)(dst + el * 16) = stage[0][slot];
      *(v4u *)(dst + plane_stride + el * 16) = stage[1][slot];
    }
  }
}

// Grid: x = channel, y = element group, z = block, so that workgroups
// running together read whole input rows.
__global__ __launch_bounds__(256) void unpack_planes_kernel(
    const uint8_t *__restrict__ in, uint8_t *__restrict__ out_planes, int N, int F, int Np,
    int64_t row_bytes, int64_t t0, int64_t t_valid, int KB) {
  __shared__ __attribute__((aligned(16))) v4u stage[2][1024 + 64];  // [plane][element], padded
  unpack_unit(in, out_planes, N, F, Np, row_bytes, t0, t_valid, KB, blockIdx.y, blockIdx.z, 0,
              blockIdx.x, stage);
}

// Batch: one launch unpacks the chunks of many jobs.  Workgroup g belongs
// to the job j with jobs[j].unit0 <= g < jobs[j+1].unit0 (binary search);
// its unit index within the job is (block, element group, channel), channel
// fastest so that neighbours read whole input rows.  A job
// writes blocks [kb_dst0, kb_dst0 + KB) of a destination laid out with
// KB_dst blocks per channel (several frames of one stream side by side).
__global__ __launch_bounds__(256) void unpack_batch_kernel(const UnpackJob *__restrict__ jobs,
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
  if (u >= (int64_t)nq * J.KB * J.F) return;  // grid padding (uniform per workgroup)
  const int f = (int)(u % J.F);
  const int qb = (int)((u / J.F) % nq);
  const int kb = (int)(u / ((int64_t)J.F * nq));
  unpack_unit(J.in, J.planes, J.N, J.F, J.Np, J.row_bytes, J.t0, J.t_valid, J.KB_dst, qb,
              J.kb_dst0 + kb, J.kb_dst0, f, stage);
}

// Pass 2, one output tile: the (ti, tj) 128x128-element tile of channel f
// of one job, over stages [st_begin, st_begin + nst) (nst a multiple of
// kRing).
// Shared by the per-plan kernel (one tile per workgroup) and the persistent
// batch kernel (tiles taken from a work list).  J.mode: 0 assign, 1 add
// (first flush; later flushes of the tile add), 2 atomic add.
__device__ __forceinline__ void gemm_tile(const GemmJob &J, const int f, const int ti,
                                          const int tj, const int st_begin, const int nst,
                                          uint8_t (*lds)[kStageBytes]) {
  const int F = J.F, Np = J.Np, KB = J.KB, nb = J.nb, U = J.U, mode = J.mode;
  const int64_t NB = J.NB;
  const uint8_t *planes = J.planes;
  int32_t *vis = J.vis;
  // wave index made provably uniform: descriptors and LDS destinations built
  // from it stay in scalar registers.
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

  // Loads: wave w fills the 8 operands (plane w/2, row groups 4(w&1)..+3,
  // k-steps 0/1) of each stage.  One descriptor per (plane, channel); the
  // lane offset picks the element and the 16-sample block parity, the
  // scalar offset the stage's block.
  const int lp = wave >> 1;
  const int64_t blk_bytes = (int64_t)Np * 16;  // one 16-sample block of a channel
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(planes + ((int64_t)lp * F + f) * KB * blk_bytes), (short)0,
      (int)(KB * blk_bytes), 0x00020000);
  int voff[4];
#pragma unroll
  for (int gg = 0; gg < 4; gg++) {
    const int g = 4 * (wave & 1) + gg;
    const int e0 = (g < 4 ? ti : tj) * kTile + 32 * (g & 3);
    voff[gg] = (e0 + (lane & 31)) * 16 + (lane >> 5) * (int)blk_bytes;
  }
  auto issue_stage = [&](int st, int slot) {
    const int kb0 = 4 * (st_begin + st);
#pragma unroll
    for (int gg = 0; gg < 4; gg++)
#pragma unroll
      for (int k = 0; k < 2; k++) {
        const int g = 4 * (wave & 1) + gg;
        lds_void *dst = (lds_void *)&lds[slot][((lp * 8 + g) * 4 + 2 * k) * 512];
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, dst, 16, voff[gg],
                                                 (kb0 + 2 * k) * (int)blk_bytes, 0, 0);
      }
  };

  // Compute role: wave (wr, wc) = rows 64*wr.., columns 64*wc.. of the tile.
  // 16 accumulators (RE split in two), so each is updated once per 16 MFMAs.
  const int wr = wave >> 1, wc = wave & 1;
  v16i rea[2][2], reb[2][2], acc_ix[2][2], acc_iy[2][2];
  const v16i zero = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) rea[i][j] = reb[i][j] = acc_ix[i][j] = acc_iy[i][j] = zero;
  struct Ops {
    v4i ar[2], ai[2], br[2], bi[2];
  };
  auto read_ops = [&](Ops &o, int slot, int k) {
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const int ga = 2 * wr + i, gb = 4 + 2 * wc + i;
      o.ar[i] = *(const v4i *)&lds[slot][((0 * 8 + ga) * 4 + 2 * k) * 512 + lane * 16];
      o.br[i] = *(const v4i *)&lds[slot][((0 * 8 + gb) * 4 + 2 * k) * 512 + lane * 16];
      o.ai[i] = *(const v4i *)&lds[slot][((1 * 8 + ga) * 4 + 2 * k) * 512 + lane * 16];
      o.bi[i] = *(const v4i *)&lds[slot][((1 * 8 + gb) * 4 + 2 * k) * 512 + lane * 16];
    }
  };
  auto mfma_ops = [&](const Ops &o) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) {
        rea[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(o.ar[i], o.br[j], rea[i][j], 0, 0, 0);
        acc_iy[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(o.ar[i], o.bi[j], acc_iy[i][j], 0, 0, 0);
        acc_ix[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(o.ai[i], o.br[j], acc_ix[i][j], 0, 0, 0);
        reb[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(o.ai[i], o.bi[j], reb[i][j], 0, 0, 0);
      }
  };
  int cur_mode = mode;
  const int by0 = (ti * kTile + 64 * wr) / 32, bx0 = (tj * kTile + 64 * wc) / 32;

  // Prologue: stages 0..kRing-2 in flight; wait for stage 0 (the 8 loads of
  // each later stage may still be pending), read its first operands.
#pragma unroll
  for (int s = 0; s < kRing - 1; s++) issue_stage(s < nst ? s : nst - 1, s);
  asm volatile("s_waitcnt vmcnt(16)\n\ts_barrier" ::: "memory");
  Ops o0, o1;
  read_ops(o0, 0, 0);

  // Per stage st (slot st % kRing; the stage count is a multiple of kRing):
  //   refill the slot of stage st-1 with stage st+kRing-1;
  //   read k-step 1 of st, MFMAs of k-step 0;
  //   wait for this wave's loads of st+1, barrier (all waves' loads landed,
  //   all reads of the refilled slot done);
  //   read k-step 0 of st+1, MFMAs of k-step 1.
  const int stages_per_chunk = kChunk / kStage;
  for (int c0 = 0; c0 < nst; c0 += stages_per_chunk) {
    const int c1 = c0 + stages_per_chunk < nst ? c0 + stages_per_chunk : nst;
    for (int st0 = c0; st0 < c1; st0 += kRing) {
#pragma unroll
      for (int d = 0; d < kRing; d++) {
        const int st = st0 + d;
        const int nx = st + kRing - 1;
        issue_stage(nx < nst ? nx : nst - 1, (d + kRing - 1) % kRing);
        read_ops(o1, d, 1);
        mfma_ops(o0);
        asm volatile("s_waitcnt vmcnt(16)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        read_ops(o0, (d + 1) % kRing, 0);
        mfma_ops(o1);
      }
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) {
        const int by = by0 + i, bx = bx0 + j;
        if (bx >= by && bx < nb) {
          const int64_t b = (int64_t)by * nb - (int64_t)by * (by - 1) / 2 + (bx - by);
          flush_block(vis + ((int64_t)f * NB + b) * 2048, rea[i][j], reb[i][j], acc_ix[i][j],
                      acc_iy[i][j], U, cur_mode, lane);
        }
        rea[i][j] = reb[i][j] = acc_ix[i][j] = acc_iy[i][j] = zero;
      }
    if (cur_mode == 0) cur_mode = 1;
  }
  // Drain the trailing (unused) loads: they must not land in a ring slot
  // after the next tile's first loads.
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// Tile index tb (row-major over the upper triangle of ntiles x ntiles tiles)
// -> (ti, tj), tj >= ti.
__device__ __forceinline__ void tile_coords(int tb, int ntiles, int &ti, int &tj) {
  int r = 0;
  while (tb >= ntiles - r) {
    tb -= ntiles - r;
    r++;
  }
  ti = r;
  tj = r + tb;
}

// One plan: grid.x = channels x upper-triangle tiles (XCD-remapped), grid.y
// = time splits.
// One plan chunk: items = channels x tiles.  Workgroups 0..head-1 take
// items 0..head-1 whole (all stages, mode J.mode); the rest take the tail
// items head.. cut into time pieces of J.stages_per_split stages (the last
// piece the rest), mode tmode.  Each range is XCD-remapped.
__global__ __launch_bounds__(256, 1) void corr_mfma_kernel(GemmJob J, int nst, int head,
                                                           int tail, int tmode) {
  __shared__ __attribute__((aligned(1024))) uint8_t lds[kRing][kStageBytes];
  const int64_t per_f = (int64_t)J.ntiles * (J.ntiles + 1) / 2;
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
  int ti, tj;
  tile_coords((int)(item % per_f), J.ntiles, ti, tj);
  gemm_tile(J, (int)(item / per_f), ti, tj, s0, ns, lds);
}

// Batch: persistent workgroups take work items (job, channel, tile, stage
// range) from 8 queues, one per XCD (dispatch places workgroup i on XCD
// i % 8), so that the tiles of one channel run on one XCD and share its L2.
// A workgroup drains its own queue, then moves on to the next ones; queues
// never refill, so every workgroup ends after at most nitems items and 8
// queue moves.  Queue q holds items [qbound[q], qbound[q+1]).
__global__ __launch_bounds__(256, 1) void corr_mfma_batch_kernel(
    const GemmJob *__restrict__ jobs, const GemmItem *__restrict__ items,
    const int *__restrict__ qbound, int *__restrict__ counters) {
  __shared__ __attribute__((aligned(1024))) uint8_t lds[kRing][kStageBytes];
  __shared__ int next;
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
    int ti, tj;
    tile_coords(w.tile, J.ntiles, ti, tj);
    gemm_tile(J, w.f, ti, tj, w.stage0, w.nstages, lds);
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
  tc = tc / (kStage * kRing) * (kStage * kRing);
  const int64_t T = (p->cfg.time_steps + kStage * kRing - 1) / (kStage * kRing) * (kStage * kRing);
  if (tc > T) tc = T;
  if (tc < kStage * kRing) tc = kStage * kRing;
  return tc;
}

size_t mfma_workspace_bytes(const chx_plan *p) {
  const int64_t Np = (int64_t)p->ntiles * kTile;
  return (size_t)(2 * (int64_t)p->cfg.num_freq * Np * chunk_samples(p));
}

int launch_corr_mfma(const chx_plan *p, const uint8_t *in, int32_t *vis, int mode,
                     hipStream_t s) {
  const int N = p->cfg.num_elements, F = p->cfg.num_freq, U = p->cfg.upper_triangle_convention;
  const int64_t T = p->cfg.time_steps;
  const int ntiles = p->ntiles, Np = ntiles * kTile;
  const int64_t items = (int64_t)F * ntiles * (ntiles + 1) / 2;
  if (items > 0x7FFFFFFF || !p->d_work) return CHX_ERR_SHAPE;
  int cus = 0;
  CHX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->cfg.device));
  if (cus < 1) cus = 1;
  const int64_t tc = chunk_samples(p);
  const int64_t row_bytes = (int64_t)F * N;
  for (int64_t t0 = 0; t0 < T; t0 += tc) {
    const int64_t tv = T - t0 < tc ? T - t0 : tc;          // valid samples
    // stages, padded to a multiple of the ring (the padding samples are 0)
    const int nst_chunk = (int)((tv + kStage * kRing - 1) / (kStage * kRing)) * kRing;
    const int KB = nst_chunk * (kStage / 16);
    // Launch shape (one dispatch): whole rounds of one workgroup per CU take
    // whole items with plain stores; the items of a last partial round are
    // cut in time into the fewest pieces (whole rings, >= 16 stages) that
    // minimise its length, adding atomically into zeroed output.
    const int64_t head = items / cus * cus, tail = items - head;
    int split = 1;
    if (tail > 0) {
      double best = (double)((tail + cus - 1) / cus);
      for (int sp = 2; sp <= 64 && nst_chunk / sp >= 16; sp++) {
        const double r = (double)((tail * sp + cus - 1) / cus) / sp;
        if (r < best - 1e-9) {
          best = r;
          split = sp;
        }
      }
    }
    const int sps = ((nst_chunk + split - 1) / split + kRing - 1) / kRing * kRing;
    split = (nst_chunk + sps - 1) / sps;
    const int kmode = (t0 == 0) ? mode : 1;
    const int tmode = split > 1 ? 2 : kmode;
    if (tail > 0 && split > 1 && kmode == 0) {
      const int64_t f0 = head / ((int64_t)ntiles * (ntiles + 1) / 2);
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
    int rc0 = launch_unpack_planes(in, p->d_work, N, F, Np, row_bytes, t0, tv, KB, s);
    if (rc0) return rc0;
    if (p->prof) CHX_HIP(hipEventRecord(ev[1], s));
    GemmJob J;
    J.planes = p->d_work;
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
    const int64_t grid = head + tail * split;
    if (grid > 0x7FFFFFFF) return CHX_ERR_SHAPE;
    corr_mfma_kernel<<<dim3((unsigned)grid), dim3(256), 0, s>>>(J, nst_chunk, (int)head,
                                                                  (int)tail, tmode);
    CHX_HIP(hipGetLastError());
    if (p->prof) CHX_HIP(hipEventRecord(ev[2], s));
  }
  return CHX_SUCCESS;
}

int launch_unpack_planes(const uint8_t *in, uint8_t *planes, int N, int F, int Np,
                         int64_t row_bytes, int64_t t0, int64_t t_valid, int KB,
                         hipStream_t s) {
  unpack_planes_kernel<<<dim3((unsigned)F, (unsigned)((Np + 1023) / 1024), (unsigned)KB),
                         dim3(256), 0, s>>>(in, planes, N, F, Np, row_bytes, t0, t_valid, KB);
  CHX_HIP(hipGetLastError());
  return CHX_SUCCESS;
}

int launch_unpack_batch_int8(const UnpackJob *d_jobs, int njobs, int64_t units, hipStream_t s) {
  if (units <= 0) return CHX_SUCCESS;
  const int64_t gx = units < 65536 ? units : 65536;
  const int64_t gy = (units + gx - 1) / gx;
  unpack_batch_kernel<<<dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s>>>(d_jobs, njobs);
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
