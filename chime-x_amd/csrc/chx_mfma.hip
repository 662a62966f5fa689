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
// No offset/preseed correction is needed and there is no time_accum cap.
//
// Two passes per chunk of time:
//  1. unpack_planes_kernel (HBM-bound): [time][freq][element] bytes ->
//     MFMA-ready planes [plane][freq][16-sample block][element][16 B].  A
//     byte with nibbles (R, I) is XORed with 0x88 so each nibble holds its
//     value's 4-bit two's complement; placed in bits 6..3 of an int8 with
//     its sign extended into bit 7 it is 8*value exactly.  Elements past N
//     and samples past T are 0.
//  2. corr_mfma_kernel: a workgroup = 4 wavefronts = a 128-row x 256-column
//     tile of one channel, each wave 64 x 128 (2x4 blocks x 2 int32
//     accumulators: RE = Ar.Br + Ai.Bi and IM = Ar.Bi + (-Ai).Br, -Ai
//     negated per byte in registers), fed by buffer loads that go straight
//     to LDS through a ring of 4 k-steps of 32 samples (24 KB each), three
//     in flight.
// Every product is 64x the true one; a work item covers at most
// kMaxItemStages 64-sample stages, so every accumulator stays a multiple of
// 64 of magnitude <= 2^30, divided by 64 (exact) when flushed to the int32
// output.
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
constexpr int kGroups = 12;        // 32-element groups per k-step: 4 of rows, 8 of columns
// LDS k-step slot: [plane 2][group 12][1 KB operand].  A 32-row, 32-sample
// operand is 1 KB contiguous: lanes 0-31 read the k-step's even 16-sample
// block, lanes 32-63 the odd one, 16 lanes = 256 contiguous bytes (no bank
// conflicts); one 64-lane direct load fills exactly one operand.
constexpr int kSlotBytes = 2 * kGroups * 1024;  // 24 KB

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
// layout).  re holds 64 (Ar.Br + Ai.Bi), im 64 (Ar.Bi - Ai.Br).  mode 0
// assigns, 1 adds, 2 adds atomically.
__device__ __forceinline__ void flush_block(int32_t *blk, const v16i &re, const v16i &im, int U,
                                            int mode, int lane) {
  // Buffer ops on a descriptor of the block (a wave-uniform base): one VGPR
  // of lane offset, the register's row as a constant scalar offset.
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc(blk, (short)0, 2048 * 4, 0x00020000);
  const int voff = ((lane >> 5) * 4 * 32 + (lane & 31)) * 8;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int32_t vr = re[j] >> 6;  // exact: every product is a multiple of 64
    const int32_t vi = U ? -(im[j] >> 6) : im[j] >> 6;
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

// Tile tb of one channel: rows [kTile ti, +kTile), columns [kTile ct,
// +2 kTile), row-major over ti, then ct = ti, ti + 2, ...  Row ti holds
// ceil((ntiles - ti) / 2) tiles; the last one's second column half lies past
// the array when ntiles - ti is odd.
__device__ __forceinline__ void tile_coords(int tb, int ntiles, int &ti, int &ct) {
  int r = 0;
  while (tb >= (ntiles - r + 1) / 2) {
    tb -= (ntiles - r + 1) / 2;
    r++;
  }
  ti = r;
  ct = r + 2 * tb;
}

// Pass 2, one output tile (rows from kTile ti, columns from kTile ct) of
// channel f over 64-sample stages [st_begin, st_begin + nst) (nst a multiple
// of kMfmaRing), run as 32-sample k-steps.  Wave w computes rows 64 (w / 2)..
// x columns 128 (w % 2).. as 2x4 blocks of two int32 accumulators (256
// AGPRs):  RE += Ar.Br + Ai.Bi,  IM += Ar.Bi + (-Ai).Br,  where -Ai is
// negated per byte in registers (exact: every byte is 8 x value, in
// [-64, 56]).
__device__ __forceinline__ void gemm_tile(const GemmJob &J, const int f, const int ti,
                                          const int ct, const int st_begin, const int nst,
                                          uint8_t (*lds)[kSlotBytes]) {
  const int F = J.F, Np = J.Np, KB = J.KB, nb = J.nb, U = J.U, mode = J.mode;
  const int64_t NB = J.NB;
  int32_t *vis = J.vis;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int k0 = 2 * st_begin, nk = 2 * nst;  // k-steps

  // Loads: wave w fills plane w/2, groups 6(w&1)..+5 of each k-step (one
  // 1 KB operand each: lanes 0-31 the k-step's even 16-sample block, 32-63
  // the odd one).  Group g < 4 holds rows kTile ti + 32 g, g >= 4 columns
  // kTile ct + 32 (g - 4); groups past the array (the right half of a row's
  // last tile when the tile count is odd) load through an empty descriptor:
  // every offset is out of range, nothing is fetched, their products are
  // never stored.
  const int lp = wave >> 1, gh = wave & 1;
  const int64_t blk_bytes = (int64_t)Np * 16;  // one 16-sample block of a channel
  const __amdgpu_buffer_rsrc_t r0 = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(J.planes + ((int64_t)lp * F + f) * KB * blk_bytes), (short)0,
      (int)(KB * blk_bytes), 0x00020000);
  const __amdgpu_buffer_rsrc_t rz = __builtin_amdgcn_make_buffer_rsrc(
      (void *)J.planes, (short)0, 0, 0x00020000);
  const bool past = (ct + 1) * kTile >= Np;  // groups 8-11 lie past the array
  int voff[6];
#pragma unroll
  for (int gg = 0; gg < 6; gg++) {
    const int g = 6 * gh + gg;
    int e0 = g < 4 ? ti * kTile + 32 * g : ct * kTile + 32 * (g - 4);
    if (e0 >= Np) e0 -= kTile;  // (loaded through the empty descriptor)
    voff[gg] = (e0 + (lane & 31)) * 16 + (lane >> 5) * (int)blk_bytes;
  }
  auto issue = [&](int ks, int slot) {
    const int soff = (k0 + ks) * 2 * (int)blk_bytes;
#pragma unroll
    for (int gg = 0; gg < 6; gg++) {
      const int g = 6 * gh + gg;
      lds_void *dst = (lds_void *)&lds[slot][(lp * kGroups + g) * 1024];
      __builtin_amdgcn_raw_ptr_buffer_load_lds(g >= 8 && past ? rz : r0, dst, 16, voff[gg], soff,
                                               0, 0);
    }
  };

  const int wr = wave >> 1, wc = wave & 1;
  v16i re[2][4], im[2][4];
  const v16i zero = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) re[i][j] = im[i][j] = zero;
  // LDS read addresses: one lane/wave base per operand side; the slot,
  // plane and group enter as immediate offsets.
  const uint8_t *lds0 = &lds[0][0];
  const int bA = lane * 16 + wr * 2048, bB = lane * 16 + 4096 + wc * 4096;
  auto load_op = [&](int base, int slot, int plane, int k) -> v4i {
    return *(const v4i *)(lds0 + base + (slot * kSlotBytes + plane * kGroups * 1024 + k * 1024));
  };
  v4i ar[2], ai[2], br[4], bi[4];  // this k-step's operands
  auto read_a = [&](v4i (&xr)[2], v4i (&xi)[2], int slot) {
#pragma unroll
    for (int i = 0; i < 2; i++) {
      xr[i] = load_op(bA, slot, 0, i);
      xi[i] = load_op(bA, slot, 1, i);
    }
  };
  // One k-step: for each column group j, its 8 MFMAs, after which that
  // group's operands of the next k-step (slot nslot) replace it; the next
  // k-step's row operands are read during group 0.
  auto step_ops = [&](int nslot) {
    v4i nai[2], nar[2], nxi[2];
#pragma unroll
    for (int i = 0; i < 2; i++)  // -x per byte: bytes are multiples of 8
      nai[i] = ((ai[i] ^ (int)0xF8F8F8F8u) + 0x08080808) & (int)0xF8F8F8F8u;
    read_a(nar, nxi, nslot);
#pragma unroll
    for (int j = 0; j < 4; j++) {
#pragma unroll
      for (int i = 0; i < 2; i++) {
        re[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ar[i], br[j], re[i][j], 0, 0, 0);
        im[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ar[i], bi[j], im[i][j], 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 2; i++) {
        re[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ai[i], bi[j], re[i][j], 0, 0, 0);
        im[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(nai[i], br[j], im[i][j], 0, 0, 0);
      }
      br[j] = load_op(bB, nslot, 0, j);
      bi[j] = load_op(bB, nslot, 1, j);
    }
#pragma unroll
    for (int i = 0; i < 2; i++) {
      ar[i] = nar[i];
      ai[i] = nxi[i];
    }
  };
  const int by0 = (ti * kTile + 64 * wr) / 32, bx0 = (ct * kTile + 128 * wc) / 32;

  // Prologue: k-steps 0..kRing-1 in flight (6 loads each); wait for this
  // wave's k-step 0, barrier, read it.
#pragma unroll
  for (int k = 0; k < kRing; k++) issue(k < nk ? k : nk - 1, k);
  asm volatile("s_waitcnt vmcnt(18)\n\ts_barrier" ::: "memory");
  read_a(ar, ai, 0);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    br[j] = load_op(bB, 0, 0, j);
    bi[j] = load_op(bB, 0, 1, j);
  }

  // K-step ks (slot ks % kRing):
  //   wait for this wave's loads of ks+1 (those of ks+2, ks+3 may stay
  //   pending), barrier: every wave's loads of ks+1 landed; every wave's
  //   reads of slot ks are done (they were waited for before the barrier);
  //   refill slot ks with k-step ks+kRing; MFMAs of ks, reading ks+1's
  //   operands as ks's die.
  for (int k00 = 0; k00 < nk; k00 += kRing) {
#pragma unroll
    for (int d = 0; d < kRing; d++) {
      const int ks = k00 + d;
      const int nx = ks + kRing;
      __builtin_amdgcn_sched_barrier(0);
      asm volatile("s_waitcnt vmcnt(12)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
      step_ops((d + 1) % kRing);
      issue(nx < nk ? nx : nk - 1, d);
      // group 0: 8 MFMAs with the 4 reads of the next row operands
#pragma unroll
      for (int i = 0; i < 4; i++) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // LDS read
      }
      __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
      // groups 1-3: 2 reads (the previous group's next operands) and 2 DMAs
      // over each group's 8 MFMAs
#pragma unroll
      for (int g = 0; g < 3; g++) {
#pragma unroll
        for (int i = 0; i < 2; i++) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
#pragma unroll
        for (int i = 0; i < 2; i++) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM (LDS DMA)
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
      }
      __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);  // the last group's next operands
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // Every partial sum is a multiple of 64 of magnitude <= 2 * 64 * 64 *
  // 32 * (2 * kMaxItemStages) = 2^30: no int32 overflow.
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int by = by0 + i, bx = bx0 + j;
      if (bx >= by && bx < nb) {
        const int64_t b = (int64_t)by * nb - (int64_t)by * (by - 1) / 2 + (bx - by);
        flush_block(vis + ((int64_t)f * NB + b) * 2048, re[i][j], im[i][j], U, mode, lane);
      }
    }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// One plan chunk: items = channels x tiles.  Workgroups 0..head-1 take
// items 0..head-1 whole (all stages, mode J.mode); the rest take the tail
// items head.. cut into time pieces of J.stages_per_split stages (the last
// piece the rest), mode tmode.  Each range is XCD-remapped.
__global__ __launch_bounds__(256, 1) void corr_mfma_kernel(GemmJob J, int nst, int head,
                                                           int tail, int tmode) {
  __shared__ __attribute__((aligned(1024))) uint8_t lds[kRing][kSlotBytes];
  const int64_t per_f = wide_tiles(J.ntiles);
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
  int ti, ct;
  tile_coords((int)(item % per_f), J.ntiles, ti, ct);
  gemm_tile(J, (int)(item / per_f), ti, ct, s0, ns, lds);
}

// Batch: persistent workgroups take work items from 8 per-XCD queues: a
// workgroup on XCD x = blockIdx % 8 drains queue x, then moves on to the
// others; queues never refill, so every workgroup ends.
__global__ __launch_bounds__(256, 1) void corr_mfma_batch_kernel(
    const GemmJob *__restrict__ jobs, const GemmItem *__restrict__ items,
    const int *__restrict__ qbound, int *__restrict__ counters) {
  __shared__ __attribute__((aligned(1024))) uint8_t lds[kRing][kSlotBytes];
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
    int ti, ct;
    tile_coords(w.tile, J.ntiles, ti, ct);
    gemm_tile(J, w.f, ti, ct, w.stage0, w.nstages, lds);
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

size_t mfma_workspace_bytes(const chx_plan *p) {
  const int64_t Np = (int64_t)p->ntiles * kTile;
  return (size_t)(2 * (int64_t)p->cfg.num_freq * Np * chunk_samples(p));
}

int launch_corr_mfma(const chx_plan *p, const uint8_t *in, int32_t *vis, int mode,
                     hipStream_t s) {
  const int N = p->cfg.num_elements, F = p->cfg.num_freq, U = p->cfg.upper_triangle_convention;
  const int64_t T = p->cfg.time_steps;
  const int ntiles = p->ntiles, Np = ntiles * kTile;
  const int64_t per_f = wide_tiles(ntiles);
  const int64_t items = (int64_t)F * per_f;
  if (items > 0x7FFFFFFF || !p->d_work) return CHX_ERR_SHAPE;
  int cus = 0;
  CHX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->cfg.device));
  if (cus < 1) cus = 1;
  const int64_t tc = chunk_samples(p);
  const int64_t row_bytes = (int64_t)F * N;
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
    corr_mfma_kernel<<<dim3((unsigned)grid), dim3(256), 0, s>>>(J, nst, (int)head, (int)tail,
                                                                  tmode);
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
