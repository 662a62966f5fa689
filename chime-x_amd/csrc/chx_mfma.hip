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
// Operand encoding: a byte with nibbles (R, I) is XORed with 0x88 so each
// nibble holds its value's 4-bit two's complement; masking a nibble into
// the high half of an int8 gives 16*value exactly.  Every product is then
// 256x the true one, accumulated in int32 for at most kChunk samples
// (|RE| <= 256*128*kChunk = 2^30) and divided by 256 exactly when the chunk
// is flushed to the int32 output.
//
// Tiling: workgroup = 4 wavefronts = a 128x128-element output tile of one
// channel (16 reference blocks), wavefront = 64x64 (2x2 blocks, 3 int32
// accumulators of 16 registers each per block), v_mfma_i32_32x32x32_i8 with
// K = 32 samples per stage.  Stages are loaded as [time][element] bytes,
// transposed 4x4 in registers (v_perm_b32), split into planes and written to
// a double-buffered LDS ring in [element][time] rows, so each MFMA operand
// is one 16-byte LDS read per lane.
#include "chx_internal.h"

namespace chx {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kTile = 128;       // elements per tile side
constexpr int kK = 32;           // samples per stage (MFMA K)
constexpr int kChunk = 32768;    // samples accumulated before a flush
constexpr int kRows = 2 * kTile; // LDS rows per plane: row tile, then column tile
// LDS row of element row e: one spare 32-byte row after every 8 keeps both the
// 8-byte transposed writes and the 16-byte operand reads free of bank conflicts.
constexpr int kPlaneBytes = (kRows + kRows / 8) * 32;

__device__ __forceinline__ int lds_row(int e) { return (e + (e >> 3)) * 32; }

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

// Real and imaginary planes of four packed samples, each value times 16.
__device__ __forceinline__ void planes(uint32_t w, uint32_t &re, uint32_t &im) {
  const uint32_t t = w ^ 0x88888888u;
  re = t & 0xF0F0F0F0u;
  im = (t << 4) & 0xF0F0F0F0u;
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
__device__ __forceinline__ void flush_block(int32_t *blk, const v16i &re, const v16i &ix,
                                            const v16i &iy, int U, int mode, int lane) {
  const int n = lane & 31, mhi = (lane >> 5) * 4;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int m = 8 * (j >> 2) + mhi + (j & 3);
    const int32_t vr = re[j] >> 8;
    const int32_t vi = U ? (ix[j] - iy[j]) >> 8 : (iy[j] - ix[j]) >> 8;
    int32_t *o = blk + (m * 32 + n) * 2;
    if (mode == 0) {
      *(int2 *)o = make_int2(vr, vi);
    } else if (mode == 1) {
      int2 a = *(int2 *)o;
      a.x += vr;
      a.y += vi;
      *(int2 *)o = a;
    } else {
      atomicAdd(o, vr);
      atomicAdd(o + 1, vi);
    }
  }
}

// grid.x = channels x upper-triangle tiles (XCD-remapped), grid.y = time
// splits.  mode: 0 assign, 1 add (applies to the first flush; later flushes
// of the same workgroup add); splits > 1 always add atomically.
__global__ __launch_bounds__(256, 1) void corr_mfma_kernel(
    const uint8_t *__restrict__ in, int32_t *__restrict__ vis, int N, int F,
    int64_t row_bytes, int64_t t_split, int ntiles, int64_t NB, int nb, int U,
    int mode) {
  __shared__ __attribute__((aligned(16))) uint8_t lds[2][2][kPlaneBytes];  // [buf][re/im]

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t total = (int64_t)gridDim.x;
  const int64_t item = xcd_remap(blockIdx.x, total);
  const int f = (int)(item / ((int64_t)ntiles * (ntiles + 1) / 2));
  int tb = (int)(item % ((int64_t)ntiles * (ntiles + 1) / 2));
  int ti = 0;
  while (tb >= ntiles - ti) {
    tb -= ntiles - ti;
    ti++;
  }
  const int tj = ti + tb;  // tile column >= tile row
  const int64_t t_begin = (int64_t)blockIdx.y * t_split;

  // Loader role: wave w loads quads 16w..16w+15 (quads 0-31: row tile,
  // 32-63: column tile), lane = (quad, 8-sample group).
  const int lq = 16 * wave + (lane >> 2), lg = lane & 3;
  const int le = (lq & 31) * 4;                       // element within its tile
  const int lelem = (lq < 32 ? ti : tj) * kTile + le; // absolute element
  const bool lvalid = lelem < N;
  const uint8_t *src = in + (t_begin + 8 * lg) * row_bytes + (int64_t)f * N + lelem;
  const int lrow = (lq < 32 ? 0 : kTile) + le;        // first LDS row of the quad

  // Compute role: wave (wr, wc) = rows 64*wr.., columns 64*wc.. of the tile.
  const int wr = wave >> 1, wc = wave & 1;
  const int kb = (lane >> 5) * 16;
  int aoff[2], boff[2];
#pragma unroll
  for (int i = 0; i < 2; i++) {
    aoff[i] = lds_row(64 * wr + 32 * i + (lane & 31)) + kb;
    boff[i] = lds_row(kTile + 64 * wc + 32 * i + (lane & 31)) + kb;
  }

  v16i acc_re[2][2], acc_ix[2][2], acc_iy[2][2];
  const v16i zero = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc_re[i][j] = acc_ix[i][j] = acc_iy[i][j] = zero;

  uint32_t w[8];
  auto load_stage = [&](int64_t ts) {
    const uint8_t *p = src + ts * row_bytes;
#pragma unroll
    for (int s = 0; s < 8; s++)
      w[s] = lvalid ? *(const uint32_t *)(p + s * row_bytes) : 0x88888888u;
  };
  auto store_stage = [&](int buf) {
    uint32_t lo[4], hi[4];
    transpose4(w[0], w[1], w[2], w[3], lo[0], lo[1], lo[2], lo[3]);
    transpose4(w[4], w[5], w[6], w[7], hi[0], hi[1], hi[2], hi[3]);
#pragma unroll
    for (int e = 0; e < 4; e++) {
      uint32_t rl, il, rh, ih;
      planes(lo[e], rl, il);
      planes(hi[e], rh, ih);
      const int off = lds_row(lrow + e) + 8 * lg;
      *(uint2 *)&lds[buf][0][off] = make_uint2(rl, rh);
      *(uint2 *)&lds[buf][1][off] = make_uint2(il, ih);
    }
  };

  const int64_t nst = t_split / kK;
  int cur_mode = mode;

  // Output blocks of this wave (upper triangle, inside N).
  const int by0 = (ti * kTile + 64 * wr) / 32, bx0 = (tj * kTile + 64 * wc) / 32;

  load_stage(0);
  store_stage(0);
  __syncthreads();
  for (int64_t c0 = 0; c0 < nst; c0 += kChunk / kK) {
    const int64_t c1 = c0 + kChunk / kK < nst ? c0 + kChunk / kK : nst;
    for (int64_t st = c0; st < c1; st++) {
      const int buf = (int)(st & 1);
      if (st + 1 < nst) load_stage((st + 1) * kK);
      v4i ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; i++) {
        ar[i] = *(const v4i *)&lds[buf][0][aoff[i]];
        ai[i] = *(const v4i *)&lds[buf][1][aoff[i]];
        br[i] = *(const v4i *)&lds[buf][0][boff[i]];
        bi[i] = *(const v4i *)&lds[buf][1][boff[i]];
      }
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
          acc_re[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ar[i], br[j], acc_re[i][j], 0, 0, 0);
          acc_re[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ai[i], bi[j], acc_re[i][j], 0, 0, 0);
          acc_ix[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ai[i], br[j], acc_ix[i][j], 0, 0, 0);
          acc_iy[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ar[i], bi[j], acc_iy[i][j], 0, 0, 0);
        }
      if (st + 1 < nst) store_stage(buf ^ 1);
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) {
        const int by = by0 + i, bx = bx0 + j;
        if (bx >= by && bx < nb) {
          const int64_t b = (int64_t)by * nb - (int64_t)by * (by - 1) / 2 + (bx - by);
          flush_block(vis + ((int64_t)f * NB + b) * 2048, acc_re[i][j], acc_ix[i][j],
                      acc_iy[i][j], U, cur_mode, lane);
        }
        acc_re[i][j] = acc_ix[i][j] = acc_iy[i][j] = zero;
      }
    if (cur_mode == 0) cur_mode = 1;
  }
}

}  // namespace

int launch_corr_mfma(const chx_plan *p, const uint8_t *in, int32_t *vis, int mode,
                     hipStream_t s) {
  const int N = p->cfg.num_elements, F = p->cfg.num_freq;
  const int64_t T = p->cfg.time_steps;
  const int ntiles = (N + kTile - 1) / kTile;
  const int64_t items = (int64_t)F * ntiles * (ntiles + 1) / 2;
  // Split time only when the tiles alone cannot fill 256 CUs twice over;
  // keep every split a whole number of stages.
  const int64_t stages = T / kK;
  int64_t split = (512 + items - 1) / items;
  if (split > stages / 32) split = stages / 32;  // >= 1024 samples per split
  if (split < 1) split = 1;
  while (stages % split) split--;
  const int64_t t_split = T / split;
  int kmode = mode;
  if (split > 1) {
    if (mode == 0) {
      int rc = launch_fill_u32((uint32_t *)vis, 0u, (int64_t)F * p->NB * 2048, s);
      if (rc) return rc;
    }
    kmode = 2;
  }
  if (items > 0x7FFFFFFF) return CHX_ERR_SHAPE;
  corr_mfma_kernel<<<dim3((unsigned)items, (unsigned)split), dim3(256), 0, s>>>(
      in, vis, N, F, (int64_t)F * N, t_split, ntiles, p->NB, p->nb,
      p->cfg.upper_triangle_convention, kmode);
  CHX_HIP(hipGetLastError());
  return CHX_SUCCESS;
}

}  // namespace chx
