/*
 * chx_oracle.h -- CPU oracle for the CHIME X-engine correlator.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it, and only as the checker.
 *
 * Every function restates one piece of the reference's behaviour and names
 * the reference file:line it follows.  Sizes are 64-bit where the reference
 * uses `int` byte counts (main_wrapper.c:403, cpu_corr_test.c:161 overflow
 * at 2^31 bytes); the arithmetic on the values is the reference's.
 *
 * Data format (four_bit_macros.h:8-9, input_generator.c:84-96):
 *   one byte per complex sample, real in the high nibble, imaginary in the
 *   low nibble, offset binary: value = nibble - 8, range [-8, 7].
 *   Input layout [time][freq][element].
 * Output block layout (pairwise_correlator_UT.cl:161-201):
 *   [freq][block][y_local 32][x_local 32][re, im] int32, blocks enumerated
 *   row-major over the upper triangle of 32x32 element blocks
 *   (main_wrapper.c:528-540).
 */
#ifndef CHX_ORACLE_H
#define CHX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHXO_GEN_CONSTANT 1
#define CHXO_GEN_RAMP_UP 2
#define CHXO_GEN_RAMP_DOWN 3
#define CHXO_GEN_RANDOM_SEEDED 4
#define CHXO_ALL_FREQUENCIES (-1)

/* input_generator.c:8-15 */
int chxo_offset_and_clip(int value, int offset, int lo, int hi);

/* input_generator.c:17-103: fills out[T][F][N]. Calls libc srand()/rand()
 * (glibc TYPE_3 additive-feedback generator) exactly where the reference
 * does: srand(seed) once, and again at every time step unless no_repeat. */
void chxo_generate(int gen_type, int seed, int default_real, int default_imag,
                   int initial_real, int initial_imag, int single_freq,
                   int64_t T, int F, int N, int no_repeat, uint8_t *out);

/* cpu_corr_test.c:156-203 (U=1) and :59-106 (U=0), without the generator:
 * out[F][N(N+1)/2][2] int32, row y, column x >= y, assigned (not added). */
void chxo_correlate_triangle(int64_t T, int F, int N, const uint8_t *in, int U,
                             int32_t *out);

/* Same arithmetic evaluated at every (y, x) of every upper-triangle 32x32
 * block, including the lower halves of diagonal blocks: the values the
 * reference GPU buffer holds (pairwise_correlator_UT.cl:161-201). */
void chxo_correlate_blocks(int64_t T, int F, int N, const uint8_t *in, int U,
                           int32_t *out);

/* cpu_corr_test.c:156-203 / 59-106 with the reference's signature: generate
 * then correlate the upper triangle.  Returns 0, or -1 on allocation failure. */
int chxo_generate_and_correlate_triangle(int T, int F, int N, int32_t *out,
                                         int gen_type, int seed,
                                         int default_real, int default_imag,
                                         int initial_real, int initial_imag,
                                         int single_freq, int no_repeat, int U);

/* main_wrapper.c:277, 528-540 */
int64_t chxo_num_blocks(int N);
void chxo_block_maps(int N, uint32_t *x_map, uint32_t *y_map);

/* gpu_data_reorg.c:76-124: block order -> packed upper triangle. */
void chxo_blocks_to_triangle(int F, int N, const int32_t *blocks, int32_t *tri);
void chxo_blocks_to_full(int F, int N, const int32_t *blocks, int32_t *full);

/* gpu_data_reorg.c:6-34: two 16-feed frequencies packed as one 2E-element
 * frequency -> [F][E][E][2], in place (F/2 input frequencies). */
void chxo_split_paired_feeds(int F, int E, int32_t *data);
/* gpu_data_reorg.c:126-149: 16x16 full matrices -> 136-entry triangles,
 * zero-padded to F_out frequencies. */
void chxo_feed16_to_triangle(int F_out, int F_in, const int32_t *in, int32_t *out);

/* offset_accumulator.cl:7-49: sums[F][N][2] = (sum_t R, sum_t I) of the raw
 * nibbles, added into sums (the host zeroes it, main_wrapper.c:671-679). */
void chxo_offset_sums(int64_t T, int F, int N, const uint8_t *in, uint32_t *sums);

/* preseed_multifreq{,_UT}.cl:19-116 (method 0) and
 * preseed_multifreq_highly_packed_correlator_method{,_UT}.cl:19-155
 * (method 1): assigns the offset-binary correction to every block entry. */
void chxo_preseed(int method, int U, int64_t T, int F, int N,
                  const uint32_t *sums, int32_t *blocks);

/* Emulation of the packed-lane MAC kernels, one time_accum chunk at a time:
 * method 0 = pairwise_correlator{,_UT}.cl:22-207 (two 16-bit lanes per
 * mad24), method 1 = packed_correlator_overflow_protected_to_2272_iter
 * {,_UT}.cl (inverted Y imaginary, overflow harvest every 120 samples).
 * Adds the raw nibble-space products into blocks.  Lane overflow beyond the
 * method's time_accum cap is reproduced, not prevented. */
void chxo_packed_raw(int method, int U, int64_t T, int F, int N, int time_accum,
                     const uint8_t *in, int32_t *blocks);

/* cpu_corr_test.c:269-330 without the prints: exact compare of two upper
 * triangles; returns the error count. */
int chxo_compare_triangle(int F, int N, const int32_t *gpu, const int32_t *cpu,
                          int64_t *err2, int64_t *max_err);

#ifdef __cplusplus
}
#endif

#endif
