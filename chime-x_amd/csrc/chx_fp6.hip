// chx_fp6.hip -- the correlator on CDNA4 FP6 matrix cores, exactly.
//
// v_mfma_scale_f32_32x32x64_f8f6f4 with both operands E2M3 (1 sign, 2
// exponent bits with bias 1, 3 mantissa bits) and a block scale of 2^3 on
// each issues a 32x32x64 product of small integers in the cycles an int8
// 32x32x32 one takes, and accumulates in FP32.  Per 32x32 block two
// accumulators:
//   RE = Ar.Br^T + Ai.Bi^T,  IM = Ar.Bi^T + Ai.(-Br)^T,  im = U ? -IM : IM.
// The nibble u (value v = u - 8) is carried as u/8, whose E2M3 code is u
// itself (u < 8: the subnormal u/8; u >= 8: exponent 1, mantissa u - 8);
// -Br is the complement 15 - u.  The accumulators hold code products (the
// RE', IM' of the int8 plan path, chx_mfma.hip), at most 450 per sample:
// they start at -2^24 and a work item covers at most kMaxItemStagesFp6 =
// 1024 stages of 64 samples, so every partial sum is an integer of
// magnitude <= 2^24, exact in FP32, and the flush adds 2^24 and the
// per-element offset corrections.  Codes switch four bits and every product
// is >= 0, so the power-limited matrix cores hold a higher clock than with
// signed values (and subnormal E2M3 operands a higher one than normal E3M2
// integers).  The int32 visibilities are bit-identical to the int8 path's.
//
// Two passes per time chunk:
//  1. unpack_f6_u4_kernel / unpack_f6_u4_batch_kernel (HBM-bound):
//     [time][freq][element] bytes -> E2M3
//     codes, bit-packed 32 per 24 bytes (code j at bits 6j..6j+5, little
//     endian), as two arrays per plane: P0 = bytes 0-15 and P1 = bytes 16-23
//     of each element's 32-sample half of a 64-sample block:
//       P0[plane][F][KB][half][Np][16 B],  P1[plane][F][KB][half][Np][8 B].
//     A lane's MFMA operand (row l%32, samples 32(l/32)..+31 of the block)
//     is then one 16-byte and one 8-byte read at lane * 16 / lane * 8 of a
//     32-element group (in LDS or in the planes).
//  2. corr_f6_kernel: a workgroup = 4 waves = a 128-row x 256-column tile
//     of one channel, each wave all 128 rows x 64 columns (4x2 blocks x 2
//     FP32 accumulators = 256 AGPRs).  The row operands, shared by the four
//     waves, come through a ring of 4 LDS stages of one 64-sample k-step
//     (12 KB each) filled by buffer->LDS DMA; each wave loads its own
//     column operands straight into VGPRs, three stages ahead, so every
//     operand byte crosses the CU's load path once.  The power-limited
//     clock and the load path bound the kernel: each load instruction holds
//     up the wave's in-order issue, and a DMA costs about 1.5x a direct load
//     of the same bytes.  The half tiles at the array's right edge go two
//     to a workgroup (one 128-row block per pair of waves, each with its
//     own ring).  On a tile that straddles the diagonal, the waves whose
//     blocks lie partly below it run a k-loop variant without those
//     blocks' MFMAs (less power drawn, so a higher clock for every CU).
//     The batch kernel can also run 128 x 128 tiles (each wave 64 x 64) for
//     small arrays.
#include "chx_internal.h"

namespace chx {
namespace {

typedef int v2i __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef unsigned int v2u __attribute__((ext_vector_type(2)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;

constexpr int kTile = kMfmaTile;     // 128: tile rows; a tile has 2 * kTile columns
constexpr int kK = 64;               // samples per k-step (= per stage)
constexpr int kRing = kFp6Ring;      // LDS stages in the ring
// stages of one work item: code products add at most 450 per sample, and
// 450 * 64 * 1024 = 29491200 < 2^25, so from -2^24 every partial sum stays
// an integer of magnitude <= 2^24
constexpr int kMaxStages = kMaxItemStagesFp6;
constexpr int kRowGroups = 4;        // 32-element row groups per stage (in LDS)
constexpr int kP0Bytes = 2 * kRowGroups * 1024;  // [plane 2][group 4][64 lanes][16 B]
constexpr int kP1Bytes = 2 * kRowGroups * 512;   // [plane 2][group 4][64 lanes][8 B]
constexpr int kSlotBytes = kP0Bytes + kP1Bytes;  // one 128-row block of one stage
constexpr int kRingBytes = kRing * kSlotBytes;   // a ring of one row block
static_assert(kMaxStages % kRing == 0, "work items hold whole rings");

// MFMA operand format and block scale: E2M3 (format 2 of the f8f6f4
// instruction) holding v/8, scaled back by the E8M0 exponent 127 + 3.
constexpr int kFmtE2M3 = 2;
constexpr int kScale2p3 = 127 + 3;

// Four 6-bit codes (one per byte) -> 24 contiguous bits.
__device__ __forceinline__ uint32_t pack24(uint32_t c) {
  return (c & 0x3Fu) | ((c >> 2) & 0xFC0u) | ((c >> 4) & 0x3F000u) | ((c >> 6) & 0xFC0000u);
}

// 32 codes (c[k] = samples 4k..4k+3) -> 24 bytes: P0 = o[0..3], P1 = o[4..5].
__device__ __forceinline__ void pack32(const uint32_t *c, v4u &p0, v2u &p1) {
  uint32_t g[8];
#pragma unroll
  for (int k = 0; k < 8; k++) g[k] = pack24(c[k]);
  p0 = (v4u){g[0] | (g[1] << 24), (g[1] >> 8) | (g[2] << 16), (g[2] >> 16) | (g[3] << 8),
             g[4] | (g[5] << 24)};
  p1 = (v2u){(g[5] >> 8) | (g[6] << 16), (g[6] >> 16) | (g[7] << 8)};
}

// 4x4 byte transpose: r[s] holds byte e of sample s; c[e] gets byte s of
// element e (v_perm_b32 selects bytes of {s0:s1}: 0-3 = s1, 4-7 = s0).
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

// Pass 1: offset-binary codes.  The nibble u (value u - 8) is carried as
// u/8, whose E2M3 code is u itself (u < 8: the subnormal u/8; u >= 8:
// exponent 1, mantissa u - 8), so the unpack only transposes and packs the
// nibbles.  A unit is 1024 elements x one 32-sample half, or for an array
// under 1024 elements fp6_halves_per_unit(Np) consecutive halves of all its
// elements.  One workgroup = one run of channel f and element group qb: the
// units run, run + R, ... of the chunk, so that the resident workgroups
// work on neighbouring units at any time.  Per unit: 32 word loads per
// thread (one per sample row; 256 contiguous bytes per wave and row), 8
// byte transposes, then per plane: codes, bit packing, LDS staging and
// full-run stores (1 KB of P0 or 512 B of P1 per wave instruction; the
// halves of one unit are consecutive destination rows).  Loads and stores
// go through buffer descriptors (branch-free: lanes past N or past the
// valid samples read zeros, stores past the unit are dropped),
// non-temporal; the workgroups resident on a CU overlap each other's loads
// and stores.  Source half hk (relative to t0) lands at destination half 2
// kb_dst0 + hk of a layout with KB_dst blocks.  Besides the planes the run
// writes its elements' code sums over its units, SR | SI << 16 (at most
// 15 * 32 * 136 = 65280 with at most 136 halves per element and run), into
// prow[f][element].  Samples past t_valid and elements past N are code 0.
__device__ __forceinline__ void unpack_f6_u4_run(
    const uint8_t *__restrict__ in, uint8_t *__restrict__ p0, uint8_t *__restrict__ p1,
    uint32_t *__restrict__ prow, int N, int F, int Np, int64_t row_bytes, int64_t t0,
    int64_t t_valid, int KB, int KB_dst, int kb_dst0, int f, int qb, int run, int R,
    v4u *s0, v2u *s1) {
  const int tid = threadIdx.x;
  const bool multi = Np < 1024;
  const int hpu = fp6_halves_per_unit(Np);
  const int epq = multi ? Np / 4 : 256;  // element quads per half
  const int sub = multi ? tid / epq : 0;  // which half of the unit
  const int q = multi ? tid % epq : qb * 256 + tid;  // element quad
  const bool lane_ok = sub < hpu && 4 * q < N;
  const int rb = (int)row_bytes;  // the rows of a unit fit 32-bit offsets
  // lanes that read nothing: an offset past every descriptor (the row
  // offsets added below keep it there, modulo 2^32)
  const uint32_t voff = lane_ok ? (uint32_t)(32 * sub * rb + f * N + 4 * q) : 0x80000000u;
  const int64_t units = f6_u4_units(KB, Np);
  const int e_base = qb * 1024;
  auto load = [&](uint32_t (&w)[32], int64_t u) {
    // the row offsets are formed here, per unit, not kept live across the
    // loop in 32 registers
    uint32_t vo = voff;
    asm volatile("" : "+v"(vo));
    const int64_t hk0 = u * hpu;
    const int64_t nh = 2 * (int64_t)KB - hk0 < hpu ? 2 * (int64_t)KB - hk0 : hpu;
    const int64_t t = 32 * hk0, span = 32 * nh;
    const int64_t rows = t_valid - t < 0 ? 0 : (t_valid - t > span ? span : t_valid - t);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(in + (t0 + t) * row_bytes), (short)0, (int)(rows * row_bytes), 0x00020000);
#pragma unroll
    for (int s = 0; s < 32; s++)
      w[s] = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)(vo + (uint32_t)(s * rb)), 0, 2);
  };
  uint32_t sr[4] = {0, 0, 0, 0}, si[4] = {0, 0, 0, 0};
  for (int64_t u = run; u < units; u += R) {
    const int64_t hk0 = u * hpu;
    const int nh = 2 * (int64_t)KB - hk0 < hpu ? (int)(2 * (int64_t)KB - hk0) : hpu;
    uint32_t c[4][8];  // [element][4-sample group]
    {
      uint32_t w[32];
      load(w, u);
#pragma unroll
      for (int m = 0; m < 8; m++)
        transpose4(w[4 * m], w[4 * m + 1], w[4 * m + 2], w[4 * m + 3], c[0][m], c[1][m],
                   c[2][m], c[3][m]);
    }
    const int n_el = multi ? nh * Np : (Np - e_base < 1024 ? Np - e_base : 1024);
    const int el0 = multi ? sub * Np + 4 * q : 4 * tid;  // this thread's entries
#pragma unroll
    for (int pl = 0; pl < 2; pl++) {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        uint32_t code[8];
#pragma unroll
        for (int m = 0; m < 8; m++)
          code[m] = pl == 0 ? (c[e][m] >> 4) & 0x0F0F0F0Fu : c[e][m] & 0x0F0F0F0Fu;
        // bytes <= 15: eight words add bytewise without carries (<= 120)
        const uint32_t bs = code[0] + code[1] + code[2] + code[3] + code[4] + code[5] +
                            code[6] + code[7];
        if (pl == 0)
          sr[e] = __builtin_amdgcn_sad_u8(bs, 0u, sr[e]);
        else
          si[e] = __builtin_amdgcn_sad_u8(bs, 0u, si[e]);
        const int el = el0 + e;  // = 4 tid + e < 1024
        const int slot = el + (el >> 4);  // one pad entry per 16
        v4u a;
        v2u b;
        pack32(code, a, b);
        s0[slot] = a;
        s1[slot] = b;
      }
      __syncthreads();
      // (plane, channel, destination half) of the unit's first half
      const int64_t row = ((int64_t)pl * F + f) * 2 * KB_dst + 2 * (int64_t)kb_dst0 + hk0;
      const __amdgpu_buffer_rsrc_t d0 = __builtin_amdgcn_make_buffer_rsrc(
          p0 + (row * Np + e_base) * 16, (short)0, n_el * 16, 0x00020000);
      const __amdgpu_buffer_rsrc_t d1 = __builtin_amdgcn_make_buffer_rsrc(
          p1 + (row * Np + e_base) * 8, (short)0, n_el * 8, 0x00020000);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int el = k * 256 + tid;
        const int slot = el + (el >> 4);
        __builtin_amdgcn_raw_buffer_store_b128(s0[slot], d0, el * 16, 0, 2);
        __builtin_amdgcn_raw_buffer_store_b64(s1[slot], d1, el * 8, 0, 2);
      }
      __syncthreads();
    }
  }
  uint32_t *dst = prow + (int64_t)f * Np;
  if (!multi) {
    if (4 * q < Np)
      *(v4u *)(dst + 4 * q) = (v4u){sr[0] | si[0] << 16, sr[1] | si[1] << 16,
                                    sr[2] | si[2] << 16, sr[3] | si[3] << 16};
  } else {
    // the halves of a unit hold the same elements: add them up through LDS
    uint32_t *red = (uint32_t *)&s0[0];  // [half][element]
    if (sub < hpu) {
#pragma unroll
      for (int e = 0; e < 4; e++) red[sub * Np + 4 * q + e] = sr[e] | si[e] << 16;
    }
    __syncthreads();
    for (int e = tid; e < Np; e += 256) {
      uint32_t v = 0;
      for (int h = 0; h < hpu; h++) v += red[h * Np + e];
      dst[e] = v;
    }
  }
}

// Plan chunk: grid x = channel, y = element group, z = run (about
// kF6U4Workgroups workgroups in all); run r writes partials row r.
constexpr int kF6U4Workgroups = 1024;
constexpr int kF6U4MaxHalves = 136;
__host__ __device__ inline int64_t f6_u4_runs(int64_t KB, int64_t F, int Np) {
  const int64_t units = f6_u4_units(KB, Np);
  const int64_t per = F * ((Np + 1023) / 1024);
  int64_t r = (kF6U4Workgroups + per - 1) / per;
  const int64_t upw = kF6U4MaxHalves / fp6_halves_per_unit(Np);  // units per workgroup
  const int64_t need = (units + upw - 1) / upw;
  if (r < need) r = need;
  return r < units ? r : units;
}
__global__ __launch_bounds__(256, 4) void unpack_f6_u4_kernel(
    const uint8_t *__restrict__ in, uint8_t *__restrict__ p0, uint8_t *__restrict__ p1,
    uint32_t *__restrict__ partials, int N, int F, int Np, int64_t row_bytes, int64_t t0,
    int64_t t_valid, int KB) {
  __shared__ __attribute__((aligned(16))) v4u s0[1024 + 64];  // [element] P0, padded
  __shared__ __attribute__((aligned(16))) v2u s1[1024 + 64];  // [element] P1, padded
  unpack_f6_u4_run(in, p0, p1, partials + (int64_t)blockIdx.z * F * Np, N, F, Np, row_bytes,
                   t0, t_valid, KB, KB, 0, blockIdx.x, blockIdx.y, blockIdx.z, gridDim.z, s0,
                   s1);
}

// Batch: one launch unpacks the chunks of many jobs; workgroup g belongs to
// the job j with jobs[j].unit0 <= g < jobs[j+1].unit0 (binary search); its
// index within the job is (run, element group, channel), channel fastest:
// neighbouring workgroups read the same input rows.
__global__ __launch_bounds__(256, 4) void unpack_f6_u4_batch_kernel(
    const UnpackJob *__restrict__ jobs, int njobs) {
  __shared__ __attribute__((aligned(16))) v4u s0[1024 + 64];
  __shared__ __attribute__((aligned(16))) v2u s1[1024 + 64];
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
  unpack_f6_u4_run(J.in, J.planes, J.planes1,
                   J.partials + (int64_t)(J.prow0 + run) * J.F * J.Np, J.N, J.F, J.Np,
                   J.row_bytes, J.t0, J.t_valid, J.KB, J.KB_dst, J.kb_dst0, f, qb, run, J.runs,
                   s0, s1);
}

// Flush one block of code-product accumulators (re = RE' - 2^24, im = IM'
// - 2^24, exact integers) into int32 visibilities, adding back 2^24 and the
// offset corrections: rc[j] = (row RE, row IM) of register j's row, cc =
// (column RE, column IM) of the lane's column (zeros for a piece that does
// not start the chunk).  Lane l owns column n = l%32 and rows m = 8*(j/4)
// + 4*(l/32) + j%4 of register j.  mode 0 assigns, 1 adds, 2 adds
// atomically.
__device__ __forceinline__ void flush_block_f32(int32_t *blk, const v16f &re, const v16f &im,
                                                   int U, int mode, int lane,
                                                   const v2i (&rc)[16], const v2i cc) {
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc(blk, (short)0, 2048 * 4, 0x00020000);
  const int voff = ((lane >> 5) * 4 * 32 + (lane & 31)) * 8;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int32_t vr = (int32_t)re[j] + (1 << 24) + rc[j].x + cc.x;
    const int32_t m = (int32_t)im[j] + (1 << 24) + rc[j].y + cc.y;
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

__device__ __forceinline__ int64_t xcd_remap(int64_t id, int64_t total) {
  const int64_t q = total / 8, r = total % 8;
  const int64_t xcd = id % 8, slot = id / 8;
  return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// Small tile tb of one channel: row-major over the upper triangle of
// 128 x 128 tiles (row tile ti holds ntiles - ti of them).
__device__ __forceinline__ void tile_coords_small(int tb, int ntiles, int &ti, int &ct) {
  int r = 0;
  while (r < ntiles - 1 && tb >= ntiles - r) {
    tb -= ntiles - r;
    r++;
  }
  ti = r;
  ct = r + tb;
}

// Pass 2, one output item of channel f over k-steps [st_begin, st_begin +
// nst) (nst a multiple of kRing).  Wave w computes 128 rows x 64 columns as
// 4x2 blocks of two FP32 accumulators (256 AGPRs):
//   !PAIR: rows from kTile ti_a, columns from kTile ct + 64 w (a 128 x 256
//          tile; the four waves share the rows);
//    PAIR: two half tiles of column tile ct: waves 0-1 rows from kTile
//          ti_a, waves 2-3 rows from kTile ti_b, columns kTile ct + 64 (w%2).
//   SMALL: a 128 x 128 tile (rows kTile ti_a, columns kTile ct): wave w
//          computes rows 64 (w/2) .. + 63 x columns 64 (w%2) .. + 63 as
//          2x2 blocks of two FP32 accumulators (the four waves share the
//          rows' LDS ring).
// Row operands (4 groups x 2 planes per 128-row block) go through the LDS
// ring by buffer->LDS DMA; each wave loads its own column operands (2
// groups x 2 planes) straight into VGPRs, three stages ahead.
// Operands are offset-binary codes: -Br is the complement 15 - u of every
// code, the accumulators start at -2^24 and hold code products (at most 450
// per sample, so an item of at most kMaxStages stages ends below 2^24:
// every partial sum is an integer of magnitude <= 2^24, exact), and the
// flush adds 2^24 and the offset corrections J.corr (the piece that starts
// the chunk adds them, once).
// SKIP: bit 2i + j set = the wave's block (row group i, column group j) is
// never stored (below the diagonal of a diagonal tile), so its MFMAs are
// left out, and a row group with both blocks left out reads no row
// operands.  Every column group must keep an MFMA: a column load that no
// MFMA uses would be deleted, and the vmcnt waits below count on every
// wave issuing the same loads per stage.  Loads, DMAs and barriers are
// then the same in every variant, so the waves of one workgroup may run
// different ones.
template <bool PAIR, bool SMALL = false, int SKIP = 0>
__device__ __forceinline__ void gemm_tile_f6(const GemmJob &J, const int f, const int ti_a,
                                             const int ti_b, const int ct, const int st_begin,
                                             const int nst, uint8_t (*lds)[kSlotBytes]) {
  const int F = J.F, Np = J.Np, KB = J.KB, nb = J.nb, U = J.U, mode = J.mode;
  const int64_t NB = J.NB;
  int32_t *vis = J.vis;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

  const int64_t p0_blk = (int64_t)2 * Np * 16;  // one 64-sample block of one channel
  const int64_t p1_blk = (int64_t)2 * Np * 8;
  __amdgpu_buffer_rsrc_t r0[2], r1[2];
#pragma unroll
  for (int pl = 0; pl < 2; pl++) {
    r0[pl] = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(J.planes + ((int64_t)pl * F + f) * KB * p0_blk), (short)0, (int)(KB * p0_blk),
        0x00020000);
    r1[pl] = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(J.planes1 + ((int64_t)pl * F + f) * KB * p1_blk), (short)0, (int)(KB * p1_blk),
        0x00020000);
  }
  // Row DMAs.  !PAIR: wave w fills plane w/2, row groups 2(w&1), 2(w&1)+1
  // of the one row block: two P0 operands and one P1 group pair per stage.
  // PAIR: wave w fills plane w&1, all four row groups of row block w/2
  // (rows ti_a, ti_b; ring w/2 of two): four P0 operands and two P1 group
  // pairs.  Row group g holds elements kTile ti + 32 g (always inside the
  // array).
  constexpr int kDma = PAIR ? 6 : 3;  // row DMAs per wave and stage
  constexpr int kP0n = PAIR ? 4 : 2, kP1n = PAIR ? 2 : 1;
  static_assert(kDma == kP0n + kP1n, "the waits below count kDma + 8 loads per stage");
  static_assert(!(PAIR && SMALL), "a small tile is one 128-row block");
  constexpr int kG = SMALL ? 2 : 4;                  // row groups this wave computes
  auto skipped = [](int i, int j) constexpr { return ((SKIP >> (2 * i + j)) & 1) != 0; };
  auto row_used = [&](int i) constexpr { return !skipped(i, 0) || !skipped(i, 1); };
  static_assert((SKIP & 0x55 & (SMALL ? 0x05 : 0x55)) != (SMALL ? 0x05 : 0x55) &&
                    (SKIP & 0xAA & (SMALL ? 0x0A : 0xAA)) != (SMALL ? 0x0A : 0xAA),
                "every column group keeps an MFMA (its loads stay in the counted waits)");
  const int gr0 = SMALL ? 2 * (wave >> 1) : 0;       // the first of them
  const int rb = PAIR ? wave >> 1 : 0;               // this wave's row block
  const int lp = PAIR ? wave & 1 : wave >> 1;        // plane it fills
  const int g0 = PAIR ? 0 : 2 * (wave & 1);          // first row group it fills
  const int ti = PAIR && rb ? ti_b : ti_a;           // this wave's row tile
  int voff0[kP0n], voff1[kP1n];
#pragma unroll
  for (int gg = 0; gg < kP0n; gg++)
    voff0[gg] = ((lane >> 5) * Np + ti * kTile + 32 * (g0 + gg) + (lane & 31)) * 16;
#pragma unroll
  for (int pp = 0; pp < kP1n; pp++) {
    const int g = g0 + 2 * pp + (lane >> 5);  // lanes 32-63: the pair's second group
    const int i = lane & 31;
    voff1[pp] = ((i >> 4) * Np + ti * kTile + 32 * g + 2 * (i & 15)) * 8;
  }
  auto issue_rows = [&](int st, int slot) {
    const int kb = st_begin + st;
#pragma unroll
    for (int gg = 0; gg < kP0n; gg++) {
      lds_void *dst =
          (lds_void *)&lds[rb * kRing + slot][(lp * kRowGroups + g0 + gg) * 1024];
      __builtin_amdgcn_raw_ptr_buffer_load_lds(r0[lp], dst, 16, voff0[gg], kb * (int)p0_blk, 0,
                                               0);
    }
#pragma unroll
    for (int pp = 0; pp < kP1n; pp++) {
      lds_void *dst =
          (lds_void *)&lds[rb * kRing + slot][kP0Bytes + (lp * kRowGroups + g0 + 2 * pp) * 512];
      __builtin_amdgcn_raw_ptr_buffer_load_lds(r1[lp], dst, 16, voff1[pp], kb * (int)p1_blk, 0,
                                               0);
    }
  };
  // Column operands of this wave: column groups 2 cw + j of the tile, lane
  // -> (half lane/32, element + lane%32).  Columns past the array (waves 2
  // and 3 of an unpaired half tile) load through an empty descriptor, which
  // fetches nothing and returns zeros; their products are never stored.
  const bool past = !PAIR && !SMALL && wave >= 2 && (ct + 1) * kTile >= Np;
  const int cw = PAIR || SMALL ? wave & 1 : wave;  // column quarter of the tile
  const __amdgpu_buffer_rsrc_t rz = __builtin_amdgcn_make_buffer_rsrc(
      (void *)J.planes, (short)0, 0, 0x00020000);
  __amdgpu_buffer_rsrc_t c0[2], c1[2];
#pragma unroll
  for (int pl = 0; pl < 2; pl++) {
    c0[pl] = past ? rz : r0[pl];
    c1[pl] = past ? rz : r1[pl];
  }
  int boff0[2], boff1[2];
#pragma unroll
  for (int j = 0; j < 2; j++) {
    int e = ct * kTile + 64 * cw + 32 * j;
    if (e >= Np) e -= kTile;
    e += lane & 31;
    boff0[j] = ((lane >> 5) * Np + e) * 16;
    boff1[j] = ((lane >> 5) * Np + e) * 8;
  }
  auto load_op_g = [&](int pl, int j, int s0, int s1) -> v8i {
    const v4u x = __builtin_amdgcn_raw_buffer_load_b128(c0[pl], boff0[j], s0, 0);
    const v2u y = __builtin_amdgcn_raw_buffer_load_b64(c1[pl], boff1[j], s1, 0);
    const v4u y4 = __builtin_shufflevector(y, y, 0, 1, -1, -1);
    return (v8i)__builtin_shufflevector(x, y4, 0, 1, 2, 3, 4, 5, -1, -1);
  };
  auto load_cols = [&](v8i &R0, v8i &R1, v8i &I0, v8i &I1, int st) {
    const int kb = st_begin + st;
    const int s0 = kb * (int)p0_blk, s1 = kb * (int)p1_blk;
    R0 = load_op_g(0, 0, s0, s1);
    R1 = load_op_g(0, 1, s0, s1);
    I0 = load_op_g(1, 0, s0, s1);
    I1 = load_op_g(1, 1, s0, s1);
  };

  v16f re[kG][2], im[kG][2];
  const float z0 = -16777216.0f;  // -2^24
  const v16f zero = {z0, z0, z0, z0, z0, z0, z0, z0, z0, z0, z0, z0, z0, z0, z0, z0};
#pragma unroll
  for (int i = 0; i < kG; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) re[i][j] = im[i][j] = zero;
  const uint8_t *lds0 = &lds[0][0];
  // LDS read addresses.  The 8-byte P1 reads of the eight row operands get
  // one opaque base register each, so that the compiler does not merge two
  // of them into one read2 (whose halves it would then have to move next
  // to their P0 parts).
  const int bA0 = rb * kRingBytes + lane * 16;
  int bA1[2][kRowGroups];
#pragma unroll
  for (int pl = 0; pl < 2; pl++)
#pragma unroll
    for (int k = 0; k < kRowGroups; k++)
      asm volatile("v_mov_b32 %0, %1"
                   : "=v"(bA1[pl][k])
                   : "v"(rb * kRingBytes + kP0Bytes + lane * 8 + pl * kRowGroups * 512 +
                         k * 512));
  auto load_op = [&](int slot, int plane, int k) -> v8i {
    const v4i a = *(const v4i *)(lds0 + bA0 + (slot * kSlotBytes + plane * kRowGroups * 1024 + k * 1024));
    const v2i b = *(const v2i *)(lds0 + bA1[plane][k] + slot * kSlotBytes);
    return (v8i){a.x, a.y, a.z, a.w, b.x, b.y, 0, 0};
  };
  // the four value bits of every 6-bit code: u -> 15 - u
  const v8i sign = {(int)0xCF3CF3CFu, (int)0xF3CF3CF3u, 0x3CF3CF3C, (int)0xCF3CF3CFu,
                    (int)0xF3CF3CF3u, 0x3CF3CF3C, 0, 0};
  // Row operands: two sets, alternating by stage parity: a stage computes
  // with one and reads the next stage's into the other.
  v8i AR[2][kG], AI[2][kG];
  // One stage with column operands (br0, br1, bi0, bi1) and row set p:
  // RE += Ar.Br + Ai.Bi, IM += Ar.Bi + Ai.(-Br), -Br being Br with the
  // sign bit of every code flipped.  The next stage's row operands (slot
  // nslot) are read into set p ^ 1.
  auto stage_ops = [&](const v8i &br0, const v8i &br1, const v8i &bi0, const v8i &bi1,
                       int nslot, int p) {
    const v8i br[2] = {br0, br1}, bi[2] = {bi0, bi1};
    const v8i nbr[2] = {br0 ^ sign, br1 ^ sign};
#pragma unroll
    for (int i = 0; i < kG; i++) {
      if (!row_used(i)) continue;
      const v8i xr = AR[p][i], xi = AI[p][i];
#pragma unroll
      for (int j = 0; j < 2; j++) {
        if (skipped(i, j)) continue;
        re[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
            xr, br[j], re[i][j], kFmtE2M3, kFmtE2M3, 0, kScale2p3, 0, kScale2p3);
        im[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
            xr, bi[j], im[i][j], kFmtE2M3, kFmtE2M3, 0, kScale2p3, 0, kScale2p3);
      }
#pragma unroll
      for (int j = 0; j < 2; j++) {
        if (skipped(i, j)) continue;
        re[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
            xi, bi[j], re[i][j], kFmtE2M3, kFmtE2M3, 0, kScale2p3, 0, kScale2p3);
        im[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
            xi, nbr[j], im[i][j], kFmtE2M3, kFmtE2M3, 0, kScale2p3, 0, kScale2p3);
      }
      AR[p ^ 1][i] = load_op(nslot, 0, gr0 + i);
      AI[p ^ 1][i] = load_op(nslot, 1, gr0 + i);
    }
  };
  const int by0 = ti * kTile / 32 + gr0, bx0 = (ct * kTile + 64 * cw) / 32;

  // Issue groups: G(st), issued during stage st, = the row DMAs of stage
  // st+4 (kDma) and the column loads of stage st+3 (8).  Prologue: G(-4) =
  // DMAs of stage 0, G(-3..-1) = DMAs of stages 1..3 and column loads of
  // 0..2 (sched barriers keep the issue order the waits count on).  Wait
  // for this wave's row DMAs of stage 0 (3 |G| later operations may stay
  // pending), barrier, read them.  Column operands of stage s live in set
  // s % 4: stage st loads st+3 into the set stage st-1 has finished.
  static_assert(kRing == 4, "column loads three stages ahead, row DMAs four");
  auto cl = [&](int s) { return s < nst ? s : nst - 1; };
  v8i BR[4][2], BI[4][2];
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
    asm volatile("s_waitcnt vmcnt(42)\n\ts_barrier" ::: "memory");
  else
    asm volatile("s_waitcnt vmcnt(33)\n\ts_barrier" ::: "memory");
#pragma unroll
  for (int i = 0; i < kG; i++) {
    if (!row_used(i)) continue;
    AR[0][i] = load_op(0, 0, gr0 + i);
    AI[0][i] = load_op(0, 1, gr0 + i);
  }

  // Stage st (slot st % kRing):
  //   wait until G(st-3) landed (this wave's DMAs of st+1 and column loads
  //   of st; G(st-2), G(st-1), 2 |G| operations, may stay pending), barrier: every wave's DMAs
  //   of st+1 landed and every wave's reads of slot st are done; refill
  //   slot st with stage st+4, load the columns of st+3; MFMAs of st,
  //   reading st+1's row operands into the other row set meanwhile.
  for (int st0 = 0; st0 < nst; st0 += kRing) {
#pragma unroll
    for (int d = 0; d < kRing; d++) {
      const int st = st0 + d, n3 = (d + 3) % kRing;
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (PAIR)
        asm volatile("s_waitcnt vmcnt(28)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
      else
        asm volatile("s_waitcnt vmcnt(22)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
      issue_rows(cl(st + 4), d);
      load_cols(BR[n3][0], BR[n3][1], BI[n3][0], BI[n3][1], cl(st + 3));
      stage_ops(BR[d][0], BR[d][1], BI[d][0], BI[d][1], (d + 1) % kRing, d & 1);
      // 8 MFMAs per row group i, each group with 4 LDS reads of the next
      // stage's row operands; VMEM ops (kDma DMAs, 8 column loads) over
      // groups 0-2 (SMALL: over its two groups)
#pragma unroll
      for (int g = 0; g < kG; g++)
#pragma unroll
        for (int i = 0; i < 8; i++) {
          // (a row group's MFMAs: 4 per block not left out)
          if (i < 4 * (!skipped(g, 0) + !skipped(g, 1)))
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
          if (i % 2 == 1 && row_used(g)) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
          if (SMALL) {
            if (i < (g == 0 ? 6 : 5)) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
          } else if (g < 3 && i < (PAIR ? (g == 2 ? 4 : 5) : (g == 2 ? 3 : 4))) {
            __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
          }
        }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // Corrections: the piece that starts the chunk adds them, once.
  const bool first = st_begin == 0;
  const v2i *rowc = (const v2i *)J.corr + (int64_t)f * Np;
  const v2i *colc = (const v2i *)J.corr + ((int64_t)F + f) * Np;
  v2i cc[2];
#pragma unroll
  for (int j = 0; j < 2; j++)
    cc[j] = first ? colc[(bx0 + j) * 32 + (lane & 31)] : (v2i){0, 0};
#pragma unroll
  for (int i = 0; i < kG; i++) {
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
        flush_block_f32(vis + ((int64_t)f * NB + b) * 2048, re[i][j], im[i][j], U, mode, lane,
                        rc, cc[j]);
      }
    }
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// A tile with each wave on its own variant: the waves whose blocks include
// some below the diagonal (tile_skip_mask) leave out their MFMAs.  Masks
// that would leave a column group without MFMAs (a wave past the array, or
// wholly below the diagonal of a 128 x 128 tile) run the full variant.
template <bool PAIR, bool SMALL = false>
__device__ __forceinline__ void gemm_tile_f6_masked(const GemmJob &J, const int f,
                                                    const int ti_a, const int ti_b,
                                                    const int ct, const int st_begin,
                                                    const int nst,
                                                    uint8_t (*lds)[kSlotBytes]) {
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int m = tile_skip_mask(PAIR, SMALL, ti_a, ti_b, ct, J.Np, wave);
  if constexpr (SMALL) {
    if (m == 0x04) return gemm_tile_f6<false, true, 0x04>(J, f, ti_a, ti_b, ct, st_begin, nst, lds);
  } else {
    if (m == 0xF4) return gemm_tile_f6<PAIR, false, 0xF4>(J, f, ti_a, ti_b, ct, st_begin, nst, lds);
    if (m == 0x40) return gemm_tile_f6<PAIR, false, 0x40>(J, f, ti_a, ti_b, ct, st_begin, nst, lds);
  }
  gemm_tile_f6<PAIR, SMALL, 0>(J, f, ti_a, ti_b, ct, st_begin, nst, lds);
}

// One plan chunk: items = channels x tiles.  Workgroups 0..head-1 take
// items 0..head-1 whole (all KB stages, mode J.mode); the rest take the
// tail items head.. cut into time pieces of J.stages_per_split stages (the
// last piece the rest of KB), mode tmode.  Each range is XCD-remapped (with
// a CU count that is a multiple of 8, head is too, and both ranges keep
// blockIdx % 8 = XCD).
__global__ __launch_bounds__(256, 1) void corr_f6_kernel(GemmJob J, int head, int tail,
                                                         int tmode) {
  __shared__ __attribute__((aligned(1024))) uint8_t lds[2 * kRing][kSlotBytes];  // two rings
  const int64_t per_f = paired_tiles(J.ntiles);
  int64_t item;
  int s0, ns;
  if ((int)blockIdx.x < head) {
    item = xcd_remap(blockIdx.x, head);
    s0 = 0;
    ns = J.KB;
  } else {
    const int64_t pieces = (int64_t)gridDim.x - head;
    const int64_t r = xcd_remap((int64_t)blockIdx.x - head, pieces);
    item = head + r % tail;
    s0 = (int)(r / tail) * J.stages_per_split;
    ns = J.KB - s0 < J.stages_per_split ? J.KB - s0 : J.stages_per_split;
    J.mode = tmode;
  }
  int ti_a, ti_b, ct;
  tile_coords((int)(item % per_f), J.ntiles, ti_a, ti_b, ct);
  if (ti_b >= 0 && ti_b != ti_a)
    gemm_tile_f6_masked<true>(J, (int)(item / per_f), ti_a, ti_b, ct, s0, ns, lds);
  else
    gemm_tile_f6_masked<false>(J, (int)(item / per_f), ti_a, ti_a, ct, s0, ns, lds);
}

// Batch: persistent workgroups take work items from 8 per-XCD queues (the
// same scheme as the int8 batch kernel: drain the own queue, then move on;
// queues never refill, so every workgroup ends).
__global__ __launch_bounds__(256, 1) void corr_f6_batch_kernel(
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
    if (J.small) {
      tile_coords_small(w.tile, J.ntiles, ti_a, ct);
      gemm_tile_f6_masked<false, true>(J, w.f, ti_a, ti_a, ct, w.stage0, w.nstages, lds);
      continue;
    }
    tile_coords(w.tile, J.ntiles, ti_a, ti_b, ct);
    if (ti_b >= 0 && ti_b != ti_a)
      gemm_tile_f6_masked<true>(J, w.f, ti_a, ti_b, ct, w.stage0, w.nstages, lds);
    else
      gemm_tile_f6_masked<false>(J, w.f, ti_a, ti_a, ct, w.stage0, w.nstages, lds);
  }
}

}  // namespace

int launch_unpack_batch_fp6(const UnpackJob *d_jobs, int njobs, int64_t units, hipStream_t s) {
  if (units <= 0) return CHX_SUCCESS;
  const int64_t gx = units < 65536 ? units : 65536;
  const int64_t gy = (units + gx - 1) / gx;
  unpack_f6_u4_batch_kernel<<<dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s>>>(d_jobs,
                                                                                     njobs);
  CHX_HIP(hipGetLastError());
  return CHX_SUCCESS;
}

int launch_gemm_batch_fp6(const GemmJob *d_jobs, const GemmItem *d_items, const int *d_qbound,
                          int *d_counters, int nwg, hipStream_t s) {
  corr_f6_batch_kernel<<<dim3((unsigned)nwg), dim3(256), 0, s>>>(d_jobs, d_items, d_qbound,
                                                                 d_counters);
  CHX_HIP(hipGetLastError());
  return CHX_SUCCESS;
}

// Samples per chunk: planes of one chunk within kWorkspaceBytes, a
// multiple of the ring (kRing k-steps), one (array, plane, channel) region
// addressable by a buffer descriptor.
static int64_t f6_chunk_samples(const chx_plan *p) {
  const int64_t Np = (int64_t)p->ntiles * kTile;
  const int64_t unit = (int64_t)kK * kRing;
  const int64_t per_sample = 2 * (int64_t)p->cfg.num_freq * Np * 3 / 4;  // 6 bits per value
  int64_t tc = kWorkspaceBytes / per_sample;
  const int64_t desc = 0x7FFFFFFFLL / (Np / 2);  // KB * 2 * Np * 16 = tc * Np / 2 bytes
  if (tc > desc) tc = desc;
  if (tc > (int64_t)kK * 32767) tc = (int64_t)kK * 32767;  // unpack runs <= 2 KB <= 65535
  // An array with at least a round of items (one per CU) takes chunks of at
  // most kMaxStages stages, so that whole items cover a chunk (plain
  // stores); a longer chunk would cut every item into atomic pieces.
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->cfg.device) !=
          hipSuccess ||
      cus < 1)
    cus = 1;
  if ((int64_t)p->cfg.num_freq * paired_tiles(p->ntiles) >= cus &&
      tc > (int64_t)kK * kMaxStages)
    tc = (int64_t)kK * kMaxStages;
  tc = tc / unit * unit;
  const int64_t T = (p->cfg.time_steps + unit - 1) / unit * unit;
  if (tc > T) tc = T;
  if (tc < unit) tc = unit;
  return tc;
}

// The workspace of one chunk: P0 and P1 code planes, then the unpack's
// code-sum partials [run][F][Np] u32, then the offset corrections
// [2][F][Np] int32 pairs.
static int64_t f6_plane_bytes(const chx_plan *p) {
  return 2 * (int64_t)p->cfg.num_freq * p->ntiles * kTile * f6_chunk_samples(p) * 3 / 4;
}
static int64_t f6_partial_bytes(const chx_plan *p) {
  const int64_t Np = (int64_t)p->ntiles * kTile;
  return f6_u4_runs(f6_chunk_samples(p) / kK, p->cfg.num_freq, (int)Np) * p->cfg.num_freq *
         Np * 4;
}

size_t fp6_workspace_bytes(const chx_plan *p) {
  const int64_t Np = (int64_t)p->ntiles * kTile;
  return (size_t)(f6_plane_bytes(p) + f6_partial_bytes(p) + 2 * (int64_t)p->cfg.num_freq * Np * 8);
}

int launch_corr_fp6(const chx_plan *p, const uint8_t *in, int32_t *vis, int mode,
                    hipStream_t s) {
  const int N = p->cfg.num_elements, F = p->cfg.num_freq, U = p->cfg.upper_triangle_convention;
  const int64_t T = p->cfg.time_steps;
  const int ntiles = p->ntiles, Np = ntiles * kTile;
  const int64_t items = (int64_t)F * paired_tiles(ntiles);
  if (items > 0x7FFFFFFF || !p->d_work) return CHX_ERR_SHAPE;
  int cus = 0;
  CHX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->cfg.device));
  if (cus < 1) cus = 1;
  const int64_t tc = f6_chunk_samples(p);
  const int64_t row_bytes = (int64_t)F * N;
  // the unpack addresses the rows of a unit (at most 32 * 8) with 32-bit offsets
  if (row_bytes > 0x7FFFFFFF / (32 * fp6_halves_per_unit(Np) + 1)) {
    set_error_detail("FP6 matrix-core method: num_freq * num_elements too large for one unpack unit");
    return CHX_ERR_SHAPE;
  }
  uint32_t *partials = (uint32_t *)(p->d_work + f6_plane_bytes(p));
  int32_t *corr = (int32_t *)(p->d_work + f6_plane_bytes(p) + f6_partial_bytes(p));
  for (int64_t t0 = 0; t0 < T; t0 += tc) {
    const int64_t tv = T - t0 < tc ? T - t0 : tc;
    const int nst = (int)((tv + (int64_t)kK * kRing - 1) / ((int64_t)kK * kRing)) * kRing;
    const int KB = nst;  // one 64-sample block per stage
    uint8_t *p0 = p->d_work;
    uint8_t *p1 = p0 + (int64_t)2 * F * KB * 2 * Np * 16;
    // Launch shape (one dispatch).  Whole rounds of one workgroup per CU
    // take whole items (the chunk in one piece, plain stores); the items of
    // a last partial round are cut in time into the fewest pieces that
    // minimise its length and add atomically into zeroed output.  A chunk
    // longer than kMaxStages cuts every item into pieces of at most
    // kMaxStages stages (atomic).
    int64_t head = 0;  // items of the whole rounds
    int split = 1;
    if (nst <= kMaxStages) head = items / cus * cus;
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
    int sps = ((nst + split - 1) / split + kRing - 1) / kRing * kRing;
    if (sps > kMaxStages) sps = kMaxStages;
    split = (nst + sps - 1) / sps;
    const int kmode = (t0 == 0) ? mode : 1;
    const int tmode = split > 1 ? 2 : kmode;  // the tail's pieces add atomically
    if (tail > 0 && split > 1 && kmode == 0) {
      // zero the channels the tail touches (before the head's stores)
      const int64_t f0 = head / paired_tiles(ntiles);
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
    const int64_t nrun = f6_u4_runs(KB, F, Np);
    unpack_f6_u4_kernel<<<dim3((unsigned)F, (unsigned)((Np + 1023) / 1024), (unsigned)nrun),
                          dim3(256), 0, s>>>(in, p0, p1, partials, N, F, Np, row_bytes, t0, tv,
                                             KB);
    CHX_HIP(hipGetLastError());
    int rc1 = launch_u4_corrections(partials, corr, F, Np, (int)nrun, (int32_t)tv, s);
    if (rc1) return rc1;
    if (p->prof) CHX_HIP(hipEventRecord(ev[1], s));
    GemmJob J;
    J.planes = p0;
    J.planes1 = p1;
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
    J.corr = corr;
    const int64_t grid = head + tail * split;
    if (grid > 0x7FFFFFFF) return CHX_ERR_SHAPE;
    corr_f6_kernel<<<dim3((unsigned)grid), dim3(256), 0, s>>>(J, (int)head, (int)tail, tmode);
    CHX_HIP(hipGetLastError());
    if (p->prof) CHX_HIP(hipEventRecord(ev[2], s));
  }
  return CHX_SUCCESS;
}

}  // namespace chx
