// chx_packed.hip -- the packed-integer correlator on CDNA4 VALU.
//
// Re-creates the reference's three-pass method exactly (same integer
// results, same lane algebra) with a CDNA4 launch design:
//   offset accumulate  offset_accumulator.cl:7-49
//   preseed            preseed_multifreq{,_UT}.cl,
//                      preseed_multifreq_highly_packed_correlator_method{,_UT}.cl
//   corr (conference)  pairwise_correlator{,_UT}.cl:22-207
//   corr (highly packed) packed_correlator_overflow_protected_to_2272_iter{,_UT}.cl
// What changes: one wavefront owns a 32x32 block for a whole time slice and
// widens its 16-bit lanes into int32 registers every time_accum samples, so
// there is no global spin lock and no per-chunk flush; slices that share a
// block add with int32 atomics (integer addition is order-free, the result
// is exact).  The highly-packed kernel indexes frequency, which the
// reference's does not (it is valid only for F = 1 there).
#include "chx_internal.h"

namespace chx {
namespace {

constexpr int kStage = 16;  // samples staged in LDS per step
// LDS row lengths of the staged samples.  A 128-byte row padded by 16 bytes
// starts 4 banks further on than the row before it, so the 16-byte stores
// of the 8 lanes that write 2 consecutive rows at once (4 lanes per row, 32
// bytes apart) hit 8 different groups of 4 banks; the same holds for the
// 256-byte column rows padded by 16 bytes.  The column-pair rows keep the pairs (4tx, 4tx+1)
// of the 8 column lanes in their first 128 bytes and (4tx+2, 4tx+3) in the
// next 128, so one read of 8 lanes covers 128 consecutive bytes.
constexpr int kRowU32 = 32 + 4;
constexpr int kColPairs = 32 + 2;
constexpr int kRxU32 = 16 + 4;

typedef unsigned int v4u __attribute__((ext_vector_type(4)));

inline int64_t imin(int64_t a, int64_t b) { return a < b ? a : b; }
inline int64_t imax(int64_t a, int64_t b) { return a > b ? a : b; }

__device__ __forceinline__ uint32_t mad24(uint32_t a, uint32_t b, uint32_t c) {
  return __umul24(a, b) + c;  // v_mad_u32_u24: 24-bit operands, low 32 bits
}

__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------------------
// offsetAccumulateElements.  Workgroup = 64 columns of 16 consecutive bytes
// of a time row (elements of one or more channels; the row is F*N bytes)
// over `tc` samples; its 4 waves take every 4th row (wave w: rows w, w+4,
// ...), so together they read 4 whole 1 KB row pieces at a time.  Nibbles
// are summed in 16-bit lanes (even bytes / odd bytes of each word in
// separate registers); the waves' lanes are added through LDS, 15 * tc <=
// 65535 in all for tc <= 4096, and wave 0 adds the totals into sums with
// atomics, as the reference does (a quarter of the atomics of one thread
// per column and slice).
__global__ __launch_bounds__(256) void offset_accumulate_kernel(
    const v4u *__restrict__ in, uint32_t *__restrict__ sums, int64_t row_u4,
    int64_t T, int tc) {
  __shared__ uint32_t part[3][16][64];  // waves 1-3: [wave - 1][accumulator][lane]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t col = (int64_t)blockIdx.x * 64 + lane;
  const bool valid = col < row_u4;
  const int64_t t_begin = (int64_t)blockIdx.y * tc;
  const int64_t t_end = min(T, t_begin + tc);
  uint32_t acc[4][4] = {};  // [word][R even, R odd, I even, I odd]
  if (valid) {
    const v4u *p = in + (t_begin + wave) * row_u4 + col;
#pragma unroll 4
    for (int64_t t = t_begin + wave; t < t_end; t += 4, p += 4 * row_u4) {
      const v4u v = __builtin_nontemporal_load(p);
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int q = 0; q < 4; q++) {
        acc[q][0] += (w[q] >> 4) & 0x000F000Fu;
        acc[q][1] += (w[q] >> 12) & 0x000F000Fu;
        acc[q][2] += w[q] & 0x000F000Fu;
        acc[q][3] += (w[q] >> 8) & 0x000F000Fu;
      }
    }
  }
  if (wave > 0) {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int k = 0; k < 4; k++) part[wave - 1][4 * q + k][lane] = acc[q][k];
  }
  __syncthreads();
  if (wave != 0 || !valid) return;
#pragma unroll
  for (int w = 0; w < 3; w++)
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int k = 0; k < 4; k++) acc[q][k] += part[w][4 * q + k][lane];
  uint32_t *o = sums + col * 32;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    // bytes 4q, 4q+1, 4q+2, 4q+3
    atomicAdd(o + (4 * q + 0) * 2 + 0, acc[q][0] & 0xFFFFu);
    atomicAdd(o + (4 * q + 0) * 2 + 1, acc[q][2] & 0xFFFFu);
    atomicAdd(o + (4 * q + 1) * 2 + 0, acc[q][1] & 0xFFFFu);
    atomicAdd(o + (4 * q + 1) * 2 + 1, acc[q][3] & 0xFFFFu);
    atomicAdd(o + (4 * q + 2) * 2 + 0, acc[q][0] >> 16);
    atomicAdd(o + (4 * q + 2) * 2 + 1, acc[q][2] >> 16);
    atomicAdd(o + (4 * q + 3) * 2 + 0, acc[q][1] >> 16);
    atomicAdd(o + (4 * q + 3) * 2 + 1, acc[q][3] >> 16);
  }
}

// ---------------------------------------------------------------------------
// preseed.  Workgroup = one (block, channel); thread = 4 consecutive
// entries of one block row.  Assigns the correction of the method/convention
// (uint32 arithmetic, as the reference's kernels do).
template <int METHOD, int U>
__global__ __launch_bounds__(256) void preseed_kernel(
    const uint32_t *__restrict__ sums, int32_t *__restrict__ vis,
    const uint32_t *__restrict__ maps, int64_t NB, int N, uint32_t base) {
  const int64_t b = blockIdx.x;
  const int f = blockIdx.y;
  const int bx = maps[b], by = maps[NB + b];
  const int ly = threadIdx.x >> 3, lx0 = (threadIdx.x & 7) * 4;
  const uint2 sy = *(const uint2 *)(sums + 2 * ((int64_t)f * N + by * 32 + ly));
  const uint4 sx01 = *(const uint4 *)(sums + 2 * ((int64_t)f * N + bx * 32 + lx0));
  const uint4 sx23 = *(const uint4 *)(sums + 2 * ((int64_t)f * N + bx * 32 + lx0 + 2));
  const uint32_t xR[4] = {sx01.x, sx01.z, sx23.x, sx23.z};
  const uint32_t xI[4] = {sx01.y, sx01.w, sx23.y, sx23.w};
  const uint32_t yR = sy.x, yI = sy.y;
  uint32_t out[8];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint32_t re, im;
    if (METHOD == 0) {
      re = base - 8u * ((xR[k] + xI[k]) + (yR + yI));
      im = U ? 8u * ((xI[k] - xR[k]) + (yR - yI)) : 8u * ((xR[k] - xI[k]) + (yI - yR));
    } else {
      re = base + (7u * xI[k] - 8u * xR[k]) - 8u * (yR + yI);
      im = U ? (8u * xI[k] + 7u * xR[k]) + 8u * (yR - yI)
             : (0u - 8u * xI[k] - 7u * xR[k]) - 8u * (yR - yI);
    }
    out[2 * k] = re;
    out[2 * k + 1] = im;
  }
  int32_t *o = vis + (((int64_t)f * NB + b) * 1024 + ly * 32 + lx0) * 2;
  *(uint4 *)o = make_uint4(out[0], out[1], out[2], out[3]);
  *(uint4 *)(o + 4) = make_uint4(out[4], out[5], out[6], out[7]);
}

// Adds four complex values at o[0..7], plainly or with atomics.
__device__ __forceinline__ void add_out4(int32_t *o, const int32_t *v, int atomic) {
  if (atomic) {
#pragma unroll
    for (int k = 0; k < 8; k++) atomicAdd(o + k, v[k]);
  } else {
    int4 a = *(int4 *)o, c = *(int4 *)(o + 4);
    a.x += v[0]; a.y += v[1]; a.z += v[2]; a.w += v[3];
    c.x += v[4]; c.y += v[5]; c.z += v[6]; c.w += v[7];
    *(int4 *)o = a;
    *(int4 *)(o + 4) = c;
  }
}

// ---------------------------------------------------------------------------
// corr, conference packing.  Wavefront = one (block, channel, time slice).
// Lane: rows y = 4*(l>>3) + 0..3, columns x = 4*(l&7) + 0..3.  Per sample
// and pair: a0 += R_x * ((R_y<<16)|I_y), a1 += I_x * ((R_y<<16)|I_y)
// (pairwise_correlator_UT.cl:101-157); every ta samples the lanes are widened
// into int32 (re += hi(a0) + lo(a1); im += lo(a0) - hi(a1) for U=1,
// pairwise_correlator_UT.cl:167-168; im += hi(a1) - lo(a0) for U=0,
// pairwise_correlator.cl:167-168).
template <int U>
__global__ __launch_bounds__(64) void corr_conference_kernel(
    const uint8_t *__restrict__ in, int32_t *__restrict__ vis,
    const uint32_t *__restrict__ maps, int64_t NB, int N, int64_t row_bytes,
    int64_t t_slice, int ta, int atomic) {
  // (R<<16)|I of the row elements; (R, I) of the column elements
  __shared__ __attribute__((aligned(16))) uint32_t py[kStage][kRowU32];
  __shared__ __attribute__((aligned(16))) uint2 px[kStage][kColPairs];
  const int lane = threadIdx.x;
  const int64_t b = blockIdx.x;
  const int f = blockIdx.y;
  const int bx = maps[b], by = maps[NB + b];
  const int ty = lane >> 3, tx = lane & 7;
  const int ls = lane >> 2, lq = lane & 3;  // loader: sample, 8-byte quarter
  const int64_t t0 = (int64_t)blockIdx.z * t_slice;
  const uint8_t *ysrc = in + (t0 + ls) * row_bytes + (int64_t)f * N + by * 32 + 8 * lq;
  const uint8_t *xsrc = in + (t0 + ls) * row_bytes + (int64_t)f * N + bx * 32 + 8 * lq;

  uint32_t a0[16], a1[16];
  int32_t re[16], im[16];
#pragma unroll
  for (int p = 0; p < 16; p++) a0[p] = a1[p] = re[p] = im[p] = 0;
  int cnt = 0;

  for (int64_t ts = 0; ts < t_slice; ts += kStage) {
    const uint2 vy = *(const uint2 *)(ysrc + ts * row_bytes);
    const uint2 vx = *(const uint2 *)(xsrc + ts * row_bytes);
    __syncthreads();
    {
      const uint32_t wy[2] = {vy.x, vy.y};
      const uint32_t wx[2] = {vx.x, vx.y};
#pragma unroll
      for (int q = 0; q < 2; q++) {
        uint32_t pk[4];
        uint2 rx[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const uint32_t by8 = (wy[q] >> (8 * k)) & 0xFFu;
          const uint32_t bx8 = (wx[q] >> (8 * k)) & 0xFFu;
          pk[k] = ((by8 >> 4) << 16) | (by8 & 15u);
          rx[k] = make_uint2(bx8 >> 4, bx8 & 15u);
        }
        *(uint4 *)&py[ls][8 * lq + 4 * q] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
        *(uint4 *)&px[ls][4 * lq + 2 * q] = make_uint4(rx[0].x, rx[0].y, rx[1].x, rx[1].y);
        *(uint4 *)&px[ls][16 + 4 * lq + 2 * q] = make_uint4(rx[2].x, rx[2].y, rx[3].x, rx[3].y);
      }
    }
    __syncthreads();
    for (int s8 = 0; s8 < kStage; s8 += 8) {
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const uint4 yv = *(const uint4 *)&py[s8 + j][4 * ty];
        const uint4 x01 = *(const uint4 *)&px[s8 + j][2 * tx];
        const uint4 x23 = *(const uint4 *)&px[s8 + j][16 + 2 * tx];
        const uint32_t Py[4] = {yv.x, yv.y, yv.z, yv.w};
        const uint32_t Rx[4] = {x01.x, x01.z, x23.x, x23.z};
        const uint32_t Ix[4] = {x01.y, x01.w, x23.y, x23.w};
#pragma unroll
        for (int yy = 0; yy < 4; yy++)
#pragma unroll
          for (int xx = 0; xx < 4; xx++) {
            a0[yy * 4 + xx] = mad24(Rx[xx], Py[yy], a0[yy * 4 + xx]);
            a1[yy * 4 + xx] = mad24(Ix[xx], Py[yy], a1[yy * 4 + xx]);
          }
      }
      cnt += 8;
      if (cnt == ta) {
#pragma unroll
        for (int p = 0; p < 16; p++) {
          re[p] += (int32_t)((a0[p] >> 16) + (a1[p] & 0xFFFFu));
          im[p] += U ? (int32_t)((a0[p] & 0xFFFFu) - (a1[p] >> 16))
                     : (int32_t)((a1[p] >> 16) - (a0[p] & 0xFFFFu));
          a0[p] = a1[p] = 0;
        }
        cnt = 0;
      }
    }
  }
  int32_t *o = vis + (((int64_t)f * NB + b) * 1024 + (4 * ty) * 32 + 4 * tx) * 2;
#pragma unroll
  for (int yy = 0; yy < 4; yy++) {
    const int32_t v[8] = {re[yy * 4 + 0], im[yy * 4 + 0], re[yy * 4 + 1], im[yy * 4 + 1],
                          re[yy * 4 + 2], im[yy * 4 + 2], re[yy * 4 + 3], im[yy * 4 + 3]};
    add_out4(o + yy * 64, v, atomic);
  }
}

// ---------------------------------------------------------------------------
// corr, highly-packed method.  Same ownership as above.  For the lane's
// column pairs (x0, x1) = (4tx+2m, 4tx+2m+1), m = 0, 1, and rows y:
//   P'(y) = (R_y<<16) | (15 - I_y)
//   a0 += P(x0) * P'(y), a1 += P(x1) * P'(y)   (mod 2^32: hi lane
//       R_x*I'_y + I_x*R_y, lo lane I_x*I'_y)
//   a2 += ((R_x0<<16) | R_x1) * R_y
// (packed_correlator_overflow_protected_to_2272_iter_UT.cl:114-159).  After
// every 8-sample step, once 120 samples have gone in, the lanes' top bits are
// moved into a packed overflow word (:166-232); every ta samples the lanes
// and overflow counts are widened into int32 (:250-285).
template <int U>
__global__ __launch_bounds__(64) void corr_highly_packed_kernel(
    const uint8_t *__restrict__ in, int32_t *__restrict__ vis,
    const uint32_t *__restrict__ maps, int64_t NB, int N, int64_t row_bytes,
    int64_t t_slice, int ta, int atomic) {
  // P'(y); R_y; P(x) = (R_x<<16) | I_x; (R_x(2m)<<16) | R_x(2m+1)
  __shared__ __attribute__((aligned(16))) uint32_t hy[kStage][kRowU32];
  __shared__ __attribute__((aligned(16))) uint32_t hry[kStage][kRowU32];
  __shared__ __attribute__((aligned(16))) uint32_t hx[kStage][kRowU32];
  __shared__ __attribute__((aligned(16))) uint32_t hrx[kStage][kRxU32];
  const int lane = threadIdx.x;
  const int64_t b = blockIdx.x;
  const int f = blockIdx.y;
  const int bx = maps[b], by = maps[NB + b];
  const int ty = lane >> 3, tx = lane & 7;
  const int ls = lane >> 2, lq = lane & 3;
  const int64_t t0 = (int64_t)blockIdx.z * t_slice;
  const uint8_t *ysrc = in + (t0 + ls) * row_bytes + (int64_t)f * N + by * 32 + 8 * lq;
  const uint8_t *xsrc = in + (t0 + ls) * row_bytes + (int64_t)f * N + bx * 32 + 8 * lq;

  uint32_t a0[8], a1[8], a2[8], ov[8];  // index yy*2 + m
  int32_t re[16], im[16];               // index yy*4 + xx
#pragma unroll
  for (int p = 0; p < 8; p++) a0[p] = a1[p] = a2[p] = ov[p] = 0;
#pragma unroll
  for (int p = 0; p < 16; p++) re[p] = im[p] = 0;
  int cnt = 0, since_harvest = 0;

  for (int64_t ts = 0; ts < t_slice; ts += kStage) {
    const uint2 vy = *(const uint2 *)(ysrc + ts * row_bytes);
    const uint2 vx = *(const uint2 *)(xsrc + ts * row_bytes);
    __syncthreads();
    {
      const uint32_t wy[2] = {vy.x ^ 0x0F0F0F0Fu, vy.y ^ 0x0F0F0F0Fu};
      const uint32_t wx[2] = {vx.x, vx.y};
#pragma unroll
      for (int q = 0; q < 2; q++) {
        // nibbles one per byte; v_perm_b32 then builds each lane value
        // (selector 0-3: bytes of the low nibbles, 4-7: the high, 12: zero)
        const uint32_t ylo = wy[q] & 0x0F0F0F0Fu, yhi = (wy[q] >> 4) & 0x0F0F0F0Fu;
        const uint32_t xlo = wx[q] & 0x0F0F0F0Fu, xhi = (wx[q] >> 4) & 0x0F0F0F0Fu;
        uint32_t pyv[4], ryv[4], pxv[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const uint32_t pk_sel = 0x0C000C00u | ((4u + k) << 16) | k;  // (hi<<16) | lo
          pyv[k] = __builtin_amdgcn_perm(yhi, ylo, pk_sel);
          ryv[k] = __builtin_amdgcn_perm(yhi, ylo, 0x0C0C0C00u | (4u + k));
          pxv[k] = __builtin_amdgcn_perm(xhi, xlo, pk_sel);
        }
        *(uint4 *)&hy[ls][8 * lq + 4 * q] = make_uint4(pyv[0], pyv[1], pyv[2], pyv[3]);
        *(uint4 *)&hry[ls][8 * lq + 4 * q] = make_uint4(ryv[0], ryv[1], ryv[2], ryv[3]);
        *(uint4 *)&hx[ls][8 * lq + 4 * q] = make_uint4(pxv[0], pxv[1], pxv[2], pxv[3]);
        *(uint2 *)&hrx[ls][4 * lq + 2 * q] =  // (R_x0<<16) | R_x1, (R_x2<<16) | R_x3
            make_uint2(__builtin_amdgcn_perm(xhi, xlo, 0x0C040C05u),
                       __builtin_amdgcn_perm(xhi, xlo, 0x0C060C07u));
      }
    }
    __syncthreads();
    for (int s8 = 0; s8 < kStage; s8 += 8) {
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const uint4 yv = *(const uint4 *)&hy[s8 + j][4 * ty];
        const uint4 ry = *(const uint4 *)&hry[s8 + j][4 * ty];
        const uint4 xv = *(const uint4 *)&hx[s8 + j][4 * tx];
        const uint2 rxx = *(const uint2 *)&hrx[s8 + j][2 * tx];
        const uint32_t Py[4] = {yv.x, yv.y, yv.z, yv.w};
        const uint32_t Ry[4] = {ry.x, ry.y, ry.z, ry.w};
        const uint32_t Px[4] = {xv.x, xv.y, xv.z, xv.w};
        const uint32_t Rxx[2] = {rxx.x, rxx.y};
#pragma unroll
        for (int yy = 0; yy < 4; yy++)
#pragma unroll
          for (int m = 0; m < 2; m++) {
            const int p = yy * 2 + m;
            a0[p] = mad24(Px[2 * m], Py[yy], a0[p]);
            a1[p] = mad24(Px[2 * m + 1], Py[yy], a1[p]);
            a2[p] = mad24(Rxx[m], Ry[yy], a2[p]);
          }
      }
      cnt += 8;
      since_harvest += 8;
      if (since_harvest >= 120) {
#pragma unroll
        for (int p = 0; p < 8; p++) {
          ov[p] += ((a0[p] & 0xE0000000u) >> 5) | ((a0[p] & 0x00008000u) << 5) |
                   ((a1[p] & 0xE0000000u) >> 21) | ((a1[p] & 0x00008000u) >> 11) |
                   ((a2[p] & 0x80008000u) >> 15);
          a0[p] &= 0x1FFF7FFFu;
          a1[p] &= 0x1FFF7FFFu;
          a2[p] &= 0x7FFF7FFFu;
        }
        since_harvest = 0;
      }
      if (cnt == ta) {
#pragma unroll
        for (int yy = 0; yy < 4; yy++)
#pragma unroll
          for (int m = 0; m < 2; m++) {
            const int p = yy * 2 + m, q0 = yy * 4 + 2 * m, q1 = q0 + 1;
            re[q0] += (int32_t)((((a2[p] >> 16) & 0xFFFFu) + ((ov[p] & 0x000F0000u) >> 1)) -
                                ((a0[p] & 0xFFFFu) + ((ov[p] & 0x00F00000u) >> 5)));
            re[q1] += (int32_t)(((a2[p] & 0xFFFFu) + ((ov[p] & 0x0000000Fu) << 15)) -
                                ((a1[p] & 0xFFFFu) + ((ov[p] & 0x000000F0u) << 11)));
            const uint32_t h0 = ((a0[p] >> 16) & 0xFFFFu) + ((ov[p] & 0xFF000000u) >> 11);
            const uint32_t h1 = ((a1[p] >> 16) & 0xFFFFu) + ((ov[p] & 0x0000FF00u) << 5);
            im[q0] += U ? -(int32_t)h0 : (int32_t)h0;
            im[q1] += U ? -(int32_t)h1 : (int32_t)h1;
            a0[p] = a1[p] = a2[p] = ov[p] = 0;
          }
        cnt = 0;
        since_harvest = 0;
      }
    }
  }
  int32_t *o = vis + (((int64_t)f * NB + b) * 1024 + (4 * ty) * 32 + 4 * tx) * 2;
#pragma unroll
  for (int yy = 0; yy < 4; yy++) {
    const int32_t v[8] = {re[yy * 4 + 0], im[yy * 4 + 0], re[yy * 4 + 1], im[yy * 4 + 1],
                          re[yy * 4 + 2], im[yy * 4 + 2], re[yy * 4 + 3], im[yy * 4 + 3]};
    add_out4(o + yy * 64, v, atomic);
  }
}

// gpu_data_reorg.c:76-124 on the device: thread = 4 entries of a block row.
__global__ __launch_bounds__(256) void blocks_to_triangle_kernel(
    const int32_t *__restrict__ blocks, int32_t *__restrict__ tri, int64_t NB, int N,
    int nb) {
  const int64_t b = blockIdx.x;
  const int f = blockIdx.y;
  // invert the row-major upper-triangle enumeration
  int by = 0;
  int64_t start = 0;
  while (start + (nb - by) <= b) {
    start += nb - by;
    by++;
  }
  const int bx = by + (int)(b - start);
  const int ly = threadIdx.x >> 3, lx0 = (threadIdx.x & 7) * 4;
  const int64_t ntri = (int64_t)N * (N + 1) / 2;
  const int32_t *src = blocks + (((int64_t)f * NB + b) * 1024 + ly * 32 + lx0) * 2;
  const int64_t y = by * 32 + ly;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int lx = lx0 + k;
    if (bx == by && lx < ly) continue;
    const int64_t x = bx * 32 + lx;
    const int64_t a = f * ntri + y * N - ((y - 1) * y) / 2 + (x - y);
    *(int2 *)(tri + 2 * a) = *(const int2 *)(src + 2 * k);
  }
}

}  // namespace

int launch_fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return CHX_SUCCESS;
  const int64_t blocks = imin((n + 255) / 256, 4096);
  fill_u32_kernel<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(p, v, n);
  CHX_HIP(hipGetLastError());
  return CHX_SUCCESS;
}

int launch_offset_accumulate(const chx_plan *p, const uint8_t *in, uint32_t *sums,
                             hipStream_t s) {
  const int64_t T = p->cfg.time_steps;
  const int64_t row_u4 = (int64_t)p->cfg.num_freq * p->cfg.num_elements / 16;
  const int64_t gx = (row_u4 + 63) / 64;
  // Enough slices to fill the chip with ~4096 wavefronts, at most 4096
  // samples per slice (16-bit lane bound), at least 32.
  int64_t tc = (T * gx * 4) / 4096;
  tc = imax(32, imin(4096, (tc + 31) / 32 * 32));
  const int64_t gy = (T + tc - 1) / tc;
  offset_accumulate_kernel<<<dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s>>>(
      (const v4u *)in, sums, row_u4, T, (int)tc);
  CHX_HIP(hipGetLastError());
  return CHX_SUCCESS;
}

int launch_preseed(const chx_plan *p, const uint32_t *sums, int32_t *vis, hipStream_t s) {
  const int N = p->cfg.num_elements, F = p->cfg.num_freq;
  const uint32_t base = (uint32_t)((uint64_t)p->cfg.time_steps * 128u);
  const dim3 grid((unsigned)p->NB, (unsigned)F), blk(256);
  const int key = p->cfg.kernel_batch * 2 + p->cfg.upper_triangle_convention;
  switch (key) {
    case 0: preseed_kernel<0, 0><<<grid, blk, 0, s>>>(sums, vis, p->d_maps, p->NB, N, base); break;
    case 1: preseed_kernel<0, 1><<<grid, blk, 0, s>>>(sums, vis, p->d_maps, p->NB, N, base); break;
    case 2: preseed_kernel<1, 0><<<grid, blk, 0, s>>>(sums, vis, p->d_maps, p->NB, N, base); break;
    case 3: preseed_kernel<1, 1><<<grid, blk, 0, s>>>(sums, vis, p->d_maps, p->NB, N, base); break;
    default: return CHX_ERR_INVALID_ARGUMENT;
  }
  CHX_HIP(hipGetLastError());
  return CHX_SUCCESS;
}

int launch_corr_packed(const chx_plan *p, const uint8_t *in, int32_t *vis, hipStream_t s) {
  const int N = p->cfg.num_elements, F = p->cfg.num_freq, ta = p->cfg.time_accum;
  const int64_t T = p->cfg.time_steps;
  const int64_t row_bytes = (int64_t)F * N;
  // Slice length: a multiple of both the stage (16) and ta, so lane windows
  // start where the reference's chunks start.  Split T only as far as needed
  // for ~2048 wavefronts.
  int64_t unit = ta;
  while (unit % kStage) unit += ta;
  const int64_t units = T / unit;
  const int64_t blocks = p->NB * F;
  int64_t split = (2048 + blocks - 1) / blocks;
  if (split > units) split = units;
  while (units % split) split--;
  const int64_t t_slice = T / split;
  const int atomic = split > 1;
  const dim3 grid((unsigned)p->NB, (unsigned)F, (unsigned)split), blk(64);
  const int key = p->cfg.kernel_batch * 2 + p->cfg.upper_triangle_convention;
  switch (key) {
    case 0: corr_conference_kernel<0><<<grid, blk, 0, s>>>(in, vis, p->d_maps, p->NB, N, row_bytes, t_slice, ta, atomic); break;
    case 1: corr_conference_kernel<1><<<grid, blk, 0, s>>>(in, vis, p->d_maps, p->NB, N, row_bytes, t_slice, ta, atomic); break;
    case 2: corr_highly_packed_kernel<0><<<grid, blk, 0, s>>>(in, vis, p->d_maps, p->NB, N, row_bytes, t_slice, ta, atomic); break;
    case 3: corr_highly_packed_kernel<1><<<grid, blk, 0, s>>>(in, vis, p->d_maps, p->NB, N, row_bytes, t_slice, ta, atomic); break;
    default: return CHX_ERR_INVALID_ARGUMENT;
  }
  CHX_HIP(hipGetLastError());
  return CHX_SUCCESS;
}

// gpu_data_reorg.c:36-74 on the device: workgroup = one 32x32 tile (ty, tx)
// of the full matrix of one channel.  The source block -- (ty, tx) on or
// above the diagonal, (tx, ty) below it -- is staged through LDS so reads
// and writes are both whole rows; below the diagonal (x < y) an entry is
// the conjugate of its transpose (imaginary part negated mod 2^32).
__global__ __launch_bounds__(256) void blocks_to_full_kernel(const int32_t *__restrict__ blocks,
                                                             int32_t *__restrict__ full,
                                                             int64_t NB, int N, int nb) {
  __shared__ int2 tile[32][33];
  const int tx = blockIdx.x % nb, ty = blockIdx.x / nb, f = blockIdx.y;
  const int by = ty < tx ? ty : tx, bx = ty < tx ? tx : ty;  // source block, bx >= by
  const int64_t b = (int64_t)by * nb - (int64_t)by * (by - 1) / 2 + (bx - by);
  const int2 *src = (const int2 *)(blocks + ((int64_t)f * NB + b) * 2048);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int e = k * 256 + threadIdx.x;
    tile[e >> 5][e & 31] = src[e];
  }
  __syncthreads();
  int2 *dst = (int2 *)(full + (int64_t)f * N * N * 2);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int e = k * 256 + threadIdx.x;
    const int ly = e >> 5, lx = e & 31;
    const int y = ty * 32 + ly, x = tx * 32 + lx;
    int2 v;
    if (x >= y) {
      v = tile[ly][lx];  // (ty, tx) is the source block itself
    } else {
      v = tile[lx][ly];  // transpose of the source block
      v.y = (int)(0u - (uint32_t)v.y);
    }
    dst[(int64_t)y * N + x] = v;
  }
}

int launch_blocks_to_full(int32_t F, int32_t N, const int32_t *blocks, int32_t *full,
                          hipStream_t s) {
  const int nb = N / 32;
  const int64_t NB = (int64_t)nb * (nb + 1) / 2;
  blocks_to_full_kernel<<<dim3((unsigned)(nb * nb), (unsigned)F), dim3(256), 0, s>>>(
      blocks, full, NB, N, nb);
  CHX_HIP(hipGetLastError());
  return CHX_SUCCESS;
}

int launch_blocks_to_triangle(int32_t F, int32_t N, const int32_t *blocks, int32_t *tri,
                              hipStream_t s) {
  const int nb = N / 32;
  const int64_t NB = (int64_t)nb * (nb + 1) / 2;
  blocks_to_triangle_kernel<<<dim3((unsigned)NB, (unsigned)F), dim3(256), 0, s>>>(
      blocks, tri, NB, N, nb);
  CHX_HIP(hipGetLastError());
  return CHX_SUCCESS;
}

}  // namespace chx
