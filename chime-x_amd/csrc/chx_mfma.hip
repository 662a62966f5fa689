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
// double-buffered LDS planes of 16-sample [element] slots (bank-conflict-free
// for both the 8-byte writes and the 16-byte reads), so each MFMA operand is
// one 16-byte LDS read per lane.
#include "chx_internal.h"

namespace chx {
namespace {

typedef int v2i __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kTile = 128;       // elements per tile side
constexpr int kK = 32;           // samples per stage (MFMA K)
constexpr int kChunk = 32768;    // samples accumulated before a flush
#ifndef CHX_MFMA_RING
#define CHX_MFMA_RING 4
#endif
constexpr int kRing = CHX_MFMA_RING;  // stages of raw input in flight per thread
constexpr int kRows = 2 * kTile; // LDS rows per plane: row tile, then column tile
// LDS plane layout (64 banks x 4 B on gfx950): [k-half][row][16 B], so the
// 16 lanes of an operand read cover 256 contiguous bytes; rows are swizzled
// r ^ ((r >> 3) & 3) and the second half is offset by 128 B mod 256, so the
// transposed 8-byte writes of 32 lanes land on 32 distinct bank pairs too.
constexpr int kHalfBytes = kRows * 16 + 128;
constexpr int kPlaneBytes = 2 * kHalfBytes;

__device__ __forceinline__ int lds_slot(int r) { return (r ^ ((r >> 3) & 3)) * 16; }

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
  // Buffer ops on a descriptor of the block (a wave-uniform base): one VGPR
  // of lane offset, the register's row as a constant scalar offset.
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc(blk, (short)0, 2048 * 4, 0x00020000);
  const int voff = ((lane >> 5) * 4 * 32 + (lane & 31)) * 8;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int32_t vr = re[j] >> 8;
    const int32_t vi = U ? (ix[j] - iy[j]) >> 8 : (iy[j] - ix[j]) >> 8;
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

// Scheduling template for one stage: the 8 operand reads first, then each
// MFMA followed by a slice of the next stage's VALU, LDS-write and VMEM work.
__device__ __forceinline__ void interleave_hints() {
  __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);  // DS read
#pragma unroll
  for (int i = 0; i < 8; i++) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
    __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);  // VALU
    __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // DS write
  }
#pragma unroll
  for (int i = 0; i < 8; i++) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
    __builtin_amdgcn_sched_group_barrier(0x002, 3, 0);  // VALU
    __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read
  }
}

// grid.x = channels x upper-triangle tiles (XCD-remapped), grid.y = time
// splits.  mode: 0 assign, 1 add (applies to the first flush; later flushes
// of the same workgroup add); splits > 1 always add atomically.
//
// Pipeline: the raw words of stage s sit in register ring slot s % D, loaded
// D stages before they are transposed into LDS, so each load has D stages of
// matrix work (D * 512 cycles) to arrive; LDS planes are double-buffered, one
// barrier per stage.
template <int D>
__global__ __launch_bounds__(256, 1) void corr_mfma_kernel(
    const uint8_t *__restrict__ in, int32_t *__restrict__ vis, int N, int F,
    int64_t row_bytes, int64_t t_split, int ntiles, int64_t NB, int nb, int U,
    int mode) {
  static_assert((kChunk / kK) % D == 0, "ring depth");
  __shared__ __attribute__((aligned(16))) uint8_t lds[2][2][kPlaneBytes];  // [buf][re/im]

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t total = (int64_t)gridDim.x;
  const int64_t item = xcd_remap(blockIdx.x, total);
  const int64_t per_f = (int64_t)ntiles * (ntiles + 1) / 2;
  const int f = (int)(item / per_f);
  int tb = (int)(item % per_f);
  int ti = 0;
  while (tb >= ntiles - ti) {
    tb -= ntiles - ti;
    ti++;
  }
  const int tj = ti + tb;  // tile column >= tile row
  const int64_t t_begin = (int64_t)blockIdx.y * t_split;

  // Loader role: wave w loads quads 16w..16w+15 (quads 0-31: row tile,
  // 32-63: column tile), lane = (quad, 8-sample group).  Quads beyond N read
  // element 0 of the same row (always in bounds) and are replaced by the
  // byte 0x88, the value (0, 0).
  const int lq = 16 * wave + (lane >> 2), lg = lane & 3;
  const int le = (lq & 31) * 4;                       // element within its tile
  const int lelem = (lq < 32 ? ti : tj) * kTile + le; // absolute element
  const bool lvalid = lelem < N;
  const int lrow = (lq < 32 ? 0 : kTile) + le;        // first LDS row of the quad

  // Compute role: wave (wr, wc) = rows 64*wr.., columns 64*wc.. of the tile.
  const int wr = wave >> 1, wc = wave & 1;
  const int kb = (lane >> 5) * kHalfBytes;
  int aoff[2], boff[2];
#pragma unroll
  for (int i = 0; i < 2; i++) {
    aoff[i] = lds_slot(64 * wr + 32 * i + (lane & 31)) + kb;
    boff[i] = lds_slot(kTile + 64 * wc + 32 * i + (lane & 31)) + kb;
  }

  v16i acc_re[2][2], acc_ix[2][2], acc_iy[2][2];
  const v16i zero = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc_re[i][j] = acc_ix[i][j] = acc_iy[i][j] = zero;

  uint32_t ring[D][8];
  // Buffer loads: the descriptor is rebased on the stage's first row (a
  // scalar), the lane offset is one VGPR and the sample row is the scalar
  // offset, so the eight loads of a stage need no 64-bit address math.
  const uint8_t *chan = in + t_begin * row_bytes + (int64_t)f * N;
  const int voff = 8 * lg * (int)row_bytes + (lvalid ? lelem : 0);
  const int rows_bytes = (int)(kK * row_bytes);
  auto load_stage = [&](uint32_t (&w)[8], int64_t st) {
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(chan + st * kK * row_bytes), (short)0, rows_bytes, 0x00020000);
#pragma unroll
    for (int s = 0; s < 8; s++)
      w[s] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, voff, s * (int)row_bytes, 0);
  };
  auto store_stage = [&](uint32_t (&w)[8], int buf) {
    uint32_t lo[4], hi[4];
#pragma unroll
    for (int s = 0; s < 8; s++) w[s] = lvalid ? w[s] : 0x88888888u;
    transpose4(w[0], w[1], w[2], w[3], lo[0], lo[1], lo[2], lo[3]);
    transpose4(w[4], w[5], w[6], w[7], hi[0], hi[1], hi[2], hi[3]);
#pragma unroll
    for (int e = 0; e < 4; e++) {
      uint32_t rl, il, rh, ih;
      planes(lo[e], rl, il);
      planes(hi[e], rh, ih);
      const int off = (lg >> 1) * kHalfBytes + lds_slot(lrow + e) + 8 * (lg & 1);
      *(uint2 *)&lds[buf][0][off] = make_uint2(rl, rh);
      *(uint2 *)&lds[buf][1][off] = make_uint2(il, ih);
    }
  };
  auto compute_stage = [&](int buf) {
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
  };

  const int64_t nst = t_split / kK;
  int cur_mode = mode;
  // Output blocks of this wave (upper triangle, inside N).
  const int by0 = (ti * kTile + 64 * wr) / 32, bx0 = (tj * kTile + 64 * wc) / 32;

  // Every load is issued unconditionally (stage index clamped into range)
  // so the wait counts stay exact: stage st+1's words are the oldest D*8
  // loads in flight when they are consumed.
  const int64_t last = nst - 1;
#pragma unroll
  for (int d = 0; d < D; d++) load_stage(ring[d], d < last ? d : last);
  store_stage(ring[0], 0);
  load_stage(ring[0], D < last ? D : last);
  __syncthreads();

  for (int64_t c0 = 0; c0 < nst; c0 += kChunk / kK) {
    const int64_t c1 = c0 + kChunk / kK < nst ? c0 + kChunk / kK : nst;
    for (int64_t st0 = c0; st0 < c1; st0 += D) {
#pragma unroll
      for (int d = 0; d < D; d++) {
        const int64_t st = st0 + d;
        const int buf = (int)(st & 1);
        // One basic block per stage: the operand reads, the 16 MFMAs, the
        // transposition of stage st+1 into the other LDS buffer (harmless
        // after the last stage) and the loads of stage st+1+D, interleaved so
        // the VALU/LDS/VMEM work issues while the matrix core is busy.
        compute_stage(buf);
        store_stage(ring[(d + 1) % D], buf ^ 1);
        const int64_t nxt = st + 1 + D;
        load_stage(ring[(d + 1) % D], nxt < last ? nxt : last);
        interleave_hints();
        __syncthreads();
      }
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
  const dim3 grid((unsigned)items, (unsigned)split), blk(256);
  const int64_t nst = t_split / kK;
  const int U = p->cfg.upper_triangle_convention;
  const int64_t rb = (int64_t)F * N;
  // The stage loop runs in groups of the ring depth: pick the deepest ring
  // that divides the per-split stage count.
  if (nst % kRing == 0)
    corr_mfma_kernel<kRing><<<grid, blk, 0, s>>>(in, vis, N, F, rb, t_split, ntiles, p->NB, p->nb, U, kmode);
  else if (nst % 2 == 0)
    corr_mfma_kernel<2><<<grid, blk, 0, s>>>(in, vis, N, F, rb, t_split, ntiles, p->NB, p->nb, U, kmode);
  else
    corr_mfma_kernel<1><<<grid, blk, 0, s>>>(in, vis, N, F, rb, t_split, ntiles, p->NB, p->nb, U, kmode);
  CHX_HIP(hipGetLastError());
  return CHX_SUCCESS;
}

}  // namespace chx
