/*
 * chx_oracle.c -- CPU oracle for the CHIME X-engine correlator.
 *
 * TEST INFRASTRUCTURE ONLY (see chx_oracle.h).  Plain C, no SIMD intrinsics.
 * All accumulation is done in uint32_t and converted to int32_t at the end,
 * which is the two's-complement wrap the reference's `int +=` gets on every
 * target it ran on, without signed-overflow undefined behaviour.
 */
#include "chx_oracle.h"

#include <stdlib.h>
#include <string.h>

#define HI_NIB(b) (((b) >> 4) & 0x0F)
#define LO_NIB(b) ((b)&0x0F)

/* input_generator.c:8-15 */
int chxo_offset_and_clip(int value, int offset, int lo, int hi) {
  int v = value + offset;
  if (v > hi) return hi;
  if (v < lo) return lo;
  return v;
}

/* input_generator.c:17-103.  The quirks kept on purpose:
 *  - single_freq == F is accepted and then matches no channel (:30-31, :88),
 *  - the seed is re-applied at every time step unless no_repeat (:46-54),
 *  - real is drawn before imaginary (:75-76),
 *  - unknown generator types fall back to the default value (:78-81). */
void chxo_generate(int gen_type, int seed, int default_real, int default_imag,
                   int initial_real, int initial_imag, int single_freq,
                   int64_t T, int F, int N, int no_repeat, uint8_t *out) {
  if (single_freq > F || single_freq < 0) single_freq = CHXO_ALL_FREQUENCIES;
  const unsigned dr = (unsigned)chxo_offset_and_clip(default_real, 8, 0, 15);
  const unsigned di = (unsigned)chxo_offset_and_clip(default_imag, 8, 0, 15);
  const unsigned ir = (unsigned)chxo_offset_and_clip(initial_real, 8, 0, 15);
  const unsigned ii = (unsigned)chxo_offset_and_clip(initial_imag, 8, 0, 15);
  const uint8_t default_byte = (uint8_t)(((dr << 4) & 0xF0) + (di & 0x0F));
  const int random = (gen_type == CHXO_GEN_RANDOM_SEEDED);

  if (random) srand((unsigned)seed);
  for (int64_t t = 0; t < T; t++) {
    if (random && !no_repeat) srand((unsigned)seed);
    for (int f = 0; f < F; f++) {
      uint8_t *row = out + (t * F + f) * (int64_t)N;
      for (int e = 0; e < N; e++) {
        unsigned re, im;
        switch (gen_type) {
          case CHXO_GEN_CONSTANT:
            re = ir;
            im = ii;
            break;
          case CHXO_GEN_RAMP_UP:
            re = (unsigned)(f + (int)ir + e) % 16u;
            im = (unsigned)(f + (int)ii + e) % 16u;
            break;
          case CHXO_GEN_RAMP_DOWN:
            re = 15u - (unsigned)(f + (int)ir + e) % 16u;
            im = 15u - (unsigned)(f + (int)ii + e) % 16u;
            break;
          case CHXO_GEN_RANDOM_SEEDED:
            re = (unsigned)(rand() % 16);
            im = (unsigned)(rand() % 16);
            break;
          default:
            re = dr;
            im = di;
            break;
        }
        uint8_t byte = (uint8_t)((((re & 0xFF) << 4) & 0xF0) + (im & 0x0F));
        if (single_freq != CHXO_ALL_FREQUENCIES && f != single_freq)
          byte = default_byte;
        row[e] = byte;
      }
    }
  }
}

/* Decode one channel's time slice [t0, t0+tc) into element-major signed
 * arrays r[e*tc + k], i[e*tc + k] (value = nibble - 8, cpu_corr_test.c:181-186). */
static void decode_slice(const uint8_t *in, int F, int N, int f, int64_t t0,
                         int tc, int16_t *r, int16_t *im) {
  for (int k = 0; k < tc; k++) {
    const uint8_t *row = in + ((t0 + k) * F + f) * (int64_t)N;
    for (int e = 0; e < N; e++) {
      r[(int64_t)e * tc + k] = (int16_t)((int)HI_NIB(row[e]) - 8);
      im[(int64_t)e * tc + k] = (int16_t)((int)LO_NIB(row[e]) - 8);
    }
  }
}

/* One visibility (row element y, column element x) over a decoded slice:
 * re = sum x_r*y_r + x_i*y_i; im(U=1) = sum x_r*y_i - x_i*y_r
 * (cpu_corr_test.c:188-189), im(U=0) = sum x_i*y_r - x_r*y_i (:91-92). */
static inline void vis_slice(const int16_t *xr, const int16_t *xi,
                             const int16_t *yr, const int16_t *yi, int tc,
                             int U, uint32_t *re, uint32_t *im) {
  int32_t sre = 0, sim = 0;
  for (int k = 0; k < tc; k++) {
    sre += xr[k] * yr[k] + xi[k] * yi[k];
    sim += xr[k] * yi[k] - xi[k] * yr[k];
  }
  *re += (uint32_t)sre;
  *im += U ? (uint32_t)sim : (uint32_t)(-(int64_t)sim);
}

/* Slices are at most this many samples: |partial| <= 128 * 4096 < 2^31. */
#define SLICE 4096

void chxo_correlate_triangle(int64_t T, int F, int N, const uint8_t *in, int U,
                             int32_t *out) {
  const int64_t ntri = (int64_t)N * (N + 1) / 2;
  uint32_t *acc = (uint32_t *)out;
  memset(out, 0, sizeof(int32_t) * 2 * ntri * F);
  int16_t *r = (int16_t *)malloc(sizeof(int16_t) * (size_t)N * SLICE);
  int16_t *im = (int16_t *)malloc(sizeof(int16_t) * (size_t)N * SLICE);
  if (!r || !im) abort();
  for (int f = 0; f < F; f++) {
    for (int64_t t0 = 0; t0 < T; t0 += SLICE) {
      int tc = (int)((T - t0) < SLICE ? (T - t0) : SLICE);
      decode_slice(in, F, N, f, t0, tc, r, im);
      int64_t o = (int64_t)f * ntri;
      for (int y = 0; y < N; y++) {
        for (int x = y; x < N; x++, o++) {
          vis_slice(r + (int64_t)x * tc, im + (int64_t)x * tc,
                    r + (int64_t)y * tc, im + (int64_t)y * tc, tc, U,
                    &acc[2 * o], &acc[2 * o + 1]);
        }
      }
    }
  }
  free(r);
  free(im);
}

int64_t chxo_num_blocks(int N) {
  int64_t nb = N / 32;
  return nb * (nb + 1) / 2;
}

/* main_wrapper.c:528-540: for block row y, block columns x = y .. nb-1. */
void chxo_block_maps(int N, uint32_t *x_map, uint32_t *y_map) {
  const int nb = N / 32;
  int64_t b = 0;
  for (int by = 0; by < nb; by++)
    for (int bx = by; bx < nb; bx++, b++) {
      x_map[b] = (uint32_t)bx;
      y_map[b] = (uint32_t)by;
    }
}

void chxo_correlate_blocks(int64_t T, int F, int N, const uint8_t *in, int U,
                           int32_t *out) {
  const int nb = N / 32;
  const int64_t NB = chxo_num_blocks(N);
  uint32_t *acc = (uint32_t *)out;
  memset(out, 0, sizeof(int32_t) * 2048 * NB * F);
  int16_t *r = (int16_t *)malloc(sizeof(int16_t) * (size_t)N * SLICE);
  int16_t *im = (int16_t *)malloc(sizeof(int16_t) * (size_t)N * SLICE);
  if (!r || !im) abort();
  for (int f = 0; f < F; f++) {
    for (int64_t t0 = 0; t0 < T; t0 += SLICE) {
      int tc = (int)((T - t0) < SLICE ? (T - t0) : SLICE);
      decode_slice(in, F, N, f, t0, tc, r, im);
      int64_t b = 0;
      for (int by = 0; by < nb; by++)
        for (int bx = by; bx < nb; bx++, b++) {
          uint32_t *blk = acc + ((int64_t)f * NB + b) * 2048;
          for (int ly = 0; ly < 32; ly++) {
            const int y = by * 32 + ly;
            for (int lx = 0; lx < 32; lx++) {
              const int x = bx * 32 + lx;
              vis_slice(r + (int64_t)x * tc, im + (int64_t)x * tc,
                        r + (int64_t)y * tc, im + (int64_t)y * tc, tc, U,
                        &blk[(ly * 32 + lx) * 2], &blk[(ly * 32 + lx) * 2 + 1]);
            }
          }
        }
    }
  }
  free(r);
  free(im);
}

/* cpu_corr_test.c:156-203 (U=1) / :59-106 (U=0), in the reference's loop
 * order: t, f, y, x >= y, with the first time step assigning. */
int chxo_generate_and_correlate_triangle(int T, int F, int N, int32_t *out,
                                         int gen_type, int seed,
                                         int default_real, int default_imag,
                                         int initial_real, int initial_imag,
                                         int single_freq, int no_repeat, int U) {
  uint8_t *data = (uint8_t *)malloc((size_t)T * F * N);
  if (!data) return -1;
  chxo_generate(gen_type, seed, default_real, default_imag, initial_real,
                initial_imag, single_freq, T, F, N, no_repeat, data);
  uint32_t *acc = (uint32_t *)out;
  for (int64_t t = 0; t < T; t++) {
    int64_t o = 0;
    for (int f = 0; f < F; f++) {
      const uint8_t *row = data + (t * F + f) * (int64_t)N;
      for (int y = 0; y < N; y++) {
        const int yr = (int)HI_NIB(row[y]) - 8, yi = (int)LO_NIB(row[y]) - 8;
        for (int x = y; x < N; x++, o++) {
          const int xr = (int)HI_NIB(row[x]) - 8, xi = (int)LO_NIB(row[x]) - 8;
          const uint32_t re = (uint32_t)(xr * yr + xi * yi);
          const uint32_t im =
              U ? (uint32_t)(xr * yi - xi * yr) : (uint32_t)(xi * yr - xr * yi);
          if (t != 0) {
            acc[2 * o] += re;
            acc[2 * o + 1] += im;
          } else {
            acc[2 * o] = re;
            acc[2 * o + 1] = im;
          }
        }
      }
    }
  }
  free(data);
  return 0;
}

/* gpu_data_reorg.c:76-124 */
void chxo_blocks_to_triangle(int F, int N, const int32_t *blocks, int32_t *tri) {
  const int nb = N / 32;
  const int64_t NB = chxo_num_blocks(N);
  const int64_t ntri = (int64_t)N * (N + 1) / 2;
  for (int f = 0; f < F; f++) {
    int64_t b = 0;
    for (int by = 0; by < nb; by++)
      for (int bx = by; bx < nb; bx++, b++) {
        const int32_t *blk = blocks + ((int64_t)f * NB + b) * 2048;
        for (int ly = 0; ly < 32; ly++)
          for (int lx = 0; lx < 32; lx++) {
            if (bx == by && lx < ly) continue; /* :103-115 */
            const int64_t y = by * 32 + ly, x = bx * 32 + lx;
            const int64_t a = f * ntri + y * N - ((y - 1) * y) / 2 + (x - y); /* :101 */
            tri[2 * a] = blk[(ly * 32 + lx) * 2];
            tri[2 * a + 1] = blk[(ly * 32 + lx) * 2 + 1];
          }
      }
  }
}

/* reorganize_GPU_to_full_Matrix_for_comparison (gpu_data_reorg.c:36-74):
 * every block entry on or above the diagonal (x >= y) goes to full[y][x];
 * strictly above it (x > y) also to full[x][y] with the imaginary part
 * negated (:59-62).  Entries below the diagonal of a diagonal block are not
 * read. */
void chxo_blocks_to_full(int F, int N, const int32_t *blocks, int32_t *full) {
  const int nb = N / 32;
  const int64_t NB = chxo_num_blocks(N);
  for (int f = 0; f < F; f++) {
    int64_t b = 0;
    for (int by = 0; by < nb; by++)
      for (int bx = by; bx < nb; bx++, b++) {
        const int32_t *blk = blocks + ((int64_t)f * NB + b) * 2048;
        for (int ly = 0; ly < 32; ly++)
          for (int lx = 0; lx < 32; lx++) {
            const int64_t y = by * 32 + ly, x = bx * 32 + lx;
            if (x < y) continue;
            const int32_t re = blk[(ly * 32 + lx) * 2], im = blk[(ly * 32 + lx) * 2 + 1];
            int32_t *o = full + ((int64_t)f * N * N + y * N + x) * 2;
            o[0] = re;
            o[1] = im;
            if (x > y) {
              int32_t *t = full + ((int64_t)f * N * N + x * N + y) * 2;
              t[0] = re;
              t[1] = (int32_t)(0u - (uint32_t)im);
            }
          }
      }
  }
}

/* reorganize_32_to_16_feed_GPU_Correlated_Data (gpu_data_reorg.c:6-34): in
 * place over full matrices [F/2][2E][2E][2]; walks every input entry in
 * order (:17-19) and keeps it when both indices lie in the same half
 * (:20-27), so frequency 2g gets the (<E, <E) quadrant and 2g+1 the
 * (>=E, >=E) one.  F/2 rounds down (:12). */
void chxo_split_paired_feeds(int F, int E, int32_t *data) {
  const int fin = F / 2, ein = 2 * E;
  int64_t a = 0, o = 0;
  for (int g = 0; g < fin; g++)
    for (int y = 0; y < ein; y++)
      for (int x = 0; x < ein; x++, a += 2) {
        if ((x < E) != (y < E)) continue; /* :29-30 */
        data[o++] = data[a];
        data[o++] = data[a + 1];
      }
}

/* reorganize_data_16_element_with_triangle_conversion (gpu_data_reorg.c:
 * 126-149): [F_in][16][16][2] full matrices -> [F_out][136][2] upper
 * triangles (row y, x >= y), zero for frequencies F_in .. F_out-1 (:141-145). */
void chxo_feed16_to_triangle(int F_out, int F_in, const int32_t *in, int32_t *out) {
  int64_t o = 0;
  for (int f = 0; f < F_out; f++)
    for (int y = 0; y < 16; y++)
      for (int x = y; x < 16; x++) {
        const int64_t i = ((int64_t)f * 256 + y * 16 + x) * 2; /* :134 */
        out[o++] = f < F_in ? in[i] : 0;
        out[o++] = f < F_in ? in[i + 1] : 0;
      }
}

/* offset_accumulator.cl:7-49 */
void chxo_offset_sums(int64_t T, int F, int N, const uint8_t *in, uint32_t *sums) {
  const int64_t row = (int64_t)F * N;
  for (int64_t t = 0; t < T; t++) {
    const uint8_t *p = in + t * row;
    for (int64_t j = 0; j < row; j++) {
      sums[2 * j] += HI_NIB(p[j]);
      sums[2 * j + 1] += LO_NIB(p[j]);
    }
  }
}

/* preseed kernels.  x is the column element, y the row element of the block
 * entry; S?R/S?I their raw-nibble sums over all T samples. */
void chxo_preseed(int method, int U, int64_t T, int F, int N,
                  const uint32_t *sums, int32_t *blocks) {
  const int nb = N / 32;
  const int64_t NB = chxo_num_blocks(N);
  const uint32_t base = (uint32_t)T * 128u; /* NUM_TIMESAMPLES_x_128 */
  for (int f = 0; f < F; f++) {
    int64_t b = 0;
    for (int by = 0; by < nb; by++)
      for (int bx = by; bx < nb; bx++, b++) {
        uint32_t *blk = (uint32_t *)blocks + ((int64_t)f * NB + b) * 2048;
        for (int ly = 0; ly < 32; ly++)
          for (int lx = 0; lx < 32; lx++) {
            const uint32_t *sx = sums + 2 * ((int64_t)f * N + bx * 32 + lx);
            const uint32_t *sy = sums + 2 * ((int64_t)f * N + by * 32 + ly);
            const uint32_t xR = sx[0], xI = sx[1], yR = sy[0], yI = sy[1];
            uint32_t re, im;
            if (method == 0) {
              /* preseed_multifreq_UT.cl:54-115, preseed_multifreq.cl:54-70 */
              re = base - 8u * ((xR + xI) + (yR + yI));
              im = U ? 8u * ((xI - xR) + (yR - yI)) : 8u * ((xR - xI) + (yI - yR));
            } else {
              /* ..._highly_packed_correlator_method_UT.cl:54-117, ..._method.cl:55-70 */
              re = base + (7u * xI - 8u * xR) - 8u * (yR + yI);
              im = U ? (8u * xI + 7u * xR) + 8u * (yR - yI)
                     : (0u - 8u * xI - 7u * xR) - 8u * (yR - yI);
            }
            blk[(ly * 32 + lx) * 2] = re;
            blk[(ly * 32 + lx) * 2 + 1] = im;
          }
      }
  }
}

/* mad24: 24-bit unsigned operands, low 32 bits of the product, plus c. */
static inline uint32_t mad24(uint32_t a, uint32_t b, uint32_t c) {
  return (uint32_t)((uint64_t)(a & 0xFFFFFFu) * (uint64_t)(b & 0xFFFFFFu)) + c;
}

/* pairwise_correlator{,_UT}.cl: lanes per (x, y) pair over one chunk:
 * a0 = sum mad24(R_x, (R_y<<16)|I_y) -> hi R_x*R_y, lo R_x*I_y (:107-111)
 * a1 = sum mad24(I_x, (R_y<<16)|I_y) -> hi I_x*R_y, lo I_x*I_y (:113-117)
 * flush (UT :167-168, conjugate pairwise_correlator.cl:167-168). */
static void conference_chunk(const uint8_t *in, int F, int N, int f, int64_t t0,
                             int ta, int x, int y, int U, uint32_t *re,
                             uint32_t *im) {
  uint32_t a0 = 0, a1 = 0;
  for (int k = 0; k < ta; k++) {
    const uint8_t *row = in + ((t0 + k) * F + f) * (int64_t)N;
    const uint32_t bx = row[x], by = row[y];
    const uint32_t py = ((uint32_t)HI_NIB(by) << 16) | LO_NIB(by);
    a0 = mad24(HI_NIB(bx), py, a0);
    a1 = mad24(LO_NIB(bx), py, a1);
  }
  *re += (a0 >> 16) + (a1 & 0xFFFFu);
  if (U)
    *im += (a0 & 0xFFFFu) - (a1 >> 16);
  else
    *im += (a1 >> 16) - (a0 & 0xFFFFu);
}

/* packed_correlator_overflow_protected_to_2272_iter_UT.cl:84-285 for one pair
 * of adjacent columns x0 = 2m, x1 = 2m+1 and one row y over one chunk:
 *   P'(y) = (R_y<<16) | (15-I_y)                          (:91-95)
 *   a0 = sum mad24(P(x0), P'(y)), a1 = sum mad24(P(x1), P'(y))  (:121-131)
 *   a2 = sum mad24((R_x0<<16)|R_x1, R_y)                  (:134-137)
 * with the overflow harvest after every 8-sample step once >= 120 samples
 * have accumulated (:166-232) and the flush of :250-253. */
static void hp_chunk(const uint8_t *in, int F, int N, int f, int64_t t0, int ta,
                     int x0, int y, int U, uint32_t *re0, uint32_t *im0,
                     uint32_t *re1, uint32_t *im1) {
  uint32_t a0 = 0, a1 = 0, a2 = 0, ov = 0;
  int counter = 0;
  for (int k = 0; k < ta; k++) {
    const uint8_t *row = in + ((t0 + k) * F + f) * (int64_t)N;
    const uint32_t b0 = row[x0], b1 = row[x0 + 1], by = row[y] ^ 0x0Fu;
    const uint32_t py = ((uint32_t)HI_NIB(by) << 16) | LO_NIB(by);
    const uint32_t ry = HI_NIB(by);
    const uint32_t p0 = ((uint32_t)HI_NIB(b0) << 16) | LO_NIB(b0);
    const uint32_t p1 = ((uint32_t)HI_NIB(b1) << 16) | LO_NIB(b1);
    const uint32_t reals = (p0 & 0x000F0000u) | ((p1 >> 16) & 0xFu);
    a0 = mad24(p0, py, a0);
    a1 = mad24(p1, py, a1);
    a2 = mad24(reals, ry, a2);
    counter++;
    if ((k & 7) == 7 && counter >= 120) {
      ov += ((a0 & 0xE0000000u) >> 5) | ((a0 & 0x00008000u) << 5) |
            ((a1 & 0xE0000000u) >> 21) | ((a1 & 0x00008000u) >> 11) |
            ((a2 & 0x80008000u) >> 15);
      a0 &= 0x1FFF7FFFu;
      a1 &= 0x1FFF7FFFu;
      a2 &= 0x7FFF7FFFu;
      counter = 0;
    }
  }
  *re0 += (((a2 >> 16) & 0xFFFFu) + ((ov & 0x000F0000u) >> 1)) -
          ((a0 & 0xFFFFu) + ((ov & 0x00F00000u) >> 5));
  *re1 += ((a2 & 0xFFFFu) + ((ov & 0x0000000Fu) << 15)) -
          ((a1 & 0xFFFFu) + ((ov & 0x000000F0u) << 11));
  const uint32_t h0 = ((a0 >> 16) & 0xFFFFu) + ((ov & 0xFF000000u) >> 11);
  const uint32_t h1 = ((a1 >> 16) & 0xFFFFu) + ((ov & 0x0000FF00u) << 5);
  if (U) {
    *im0 -= h0;
    *im1 -= h1;
  } else {
    *im0 += h0;
    *im1 += h1;
  }
}

void chxo_packed_raw(int method, int U, int64_t T, int F, int N, int time_accum,
                     const uint8_t *in, int32_t *blocks) {
  const int nb = N / 32;
  const int64_t NB = chxo_num_blocks(N);
  const int ta = time_accum;
  if (ta <= 0 || ta % 8 != 0 || T % ta != 0) return;
  for (int f = 0; f < F; f++) {
    int64_t b = 0;
    for (int by = 0; by < nb; by++)
      for (int bx = by; bx < nb; bx++, b++) {
        uint32_t *blk = (uint32_t *)blocks + ((int64_t)f * NB + b) * 2048;
        for (int64_t t0 = 0; t0 < T; t0 += ta)
          for (int ly = 0; ly < 32; ly++) {
            const int y = by * 32 + ly;
            if (method == 0) {
              for (int lx = 0; lx < 32; lx++) {
                uint32_t *o = blk + (ly * 32 + lx) * 2;
                conference_chunk(in, F, N, f, t0, ta, bx * 32 + lx, y, U, &o[0],
                                 &o[1]);
              }
            } else {
              for (int lx = 0; lx < 32; lx += 2) {
                uint32_t *o = blk + (ly * 32 + lx) * 2;
                hp_chunk(in, F, N, f, t0, ta, bx * 32 + lx, y, U, &o[0], &o[1],
                         &o[2], &o[3]);
              }
            }
          }
      }
  }
}

/* cpu_corr_test.c:269-330 (error count, sum of squared differences, max
 * squared amplitude error). */
int chxo_compare_triangle(int F, int N, const int32_t *gpu, const int32_t *cpu,
                          int64_t *err2, int64_t *max_err) {
  const int64_t n = (int64_t)F * N * (N + 1) / 2;
  int errors = 0;
  int64_t e2 = 0, mx = 0;
  for (int64_t k = 0; k < n; k++) {
    const int64_t dr = (int64_t)gpu[2 * k] - cpu[2 * k];
    const int64_t di = (int64_t)gpu[2 * k + 1] - cpu[2 * k + 1];
    if (dr != 0 || di != 0) {
      errors++;
      const int64_t a = dr * dr + di * di;
      e2 += a;
      if (a > mx) mx = a;
    }
  }
  if (err2) *err2 = e2;
  if (max_err) *max_err = mx;
  return errors;
}
