/*
 * correlator_test.c -- command-line driver for the MI355X correlator.
 *
 * Plain C on top of the C ABI (include/chx.h); accepts the reference
 * driver's flags with the same defaults (main_wrapper.c:51-228):
 *   -d device, -i iterations, -f num_freq, -e num_elem, -t time_accum,
 *   -T time_steps, -w timer_without_copies, -U upper_triangle_convention,
 *   -c check_results, -v verbose, -g gen_type, -r random_seed,
 *   -p no_repeat_random, -q generate_frequency, -x/-y default_real/imag,
 *   -X/-Y initial_real/imag, -k kernel_batch (0 conference, 1 highly
 *   packed, 2 int8 matrix cores, 3 FP6 matrix cores), -h help; and one of
 *   its own, -o file: write the visibilities as the packed upper triangle
 *   ([freq][N(N+1)/2][re, im] int32, reorganize_GPU_to_upper_triangle's
 *   order, gpu_data_reorg.c:76-124) to a file.
 * Deliberate differences: -Y sets initial_imaginary (the reference writes
 * initial_real, main_wrapper.c:203-205); -k also accepts 2 and 3; byte counts are
 * 64-bit (main_wrapper.c:403); the data generator and the --check_results
 * CPU correlator below are this program's own restatements of
 * input_generator.c:17-103 and cpu_corr_test.c:59-106/156-203.
 */
#include <getopt.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "chx.h"

#define INT8_PEAK_OPS 5.0e15 /* MI355X dense int8 matrix peak */

static double wall_s(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return (double)tv.tv_sec + tv.tv_usec * 1e-6;
}

static void print_help(void) {
  printf(
      "Usage: ./correlator_test [opts]\n\n"
      "  --help (-h)\n"
      "  --device (-d) [n]                 Default: 0.\n"
      "  --iterations (-i) [n]             Default: 100.\n"
      "  --time_accum (-t) [n]             Default: 256. [8,288] step 8 for -k 0, [8,2272] for -k 1, any multiple of 8 for -k 2 and 3.\n"
      "  --time_steps (-T) [n]             Default: 128*time_accum. Multiple of 32 and of time_accum.\n"
      "  --num_freq (-f) [n]               Default: 1.\n"
      "  --num_elem (-e) [n]               Default: 2048. Multiple of 32.\n"
      "  --timer_without_copies (-w)       Copy the input once, outside the timed loop.\n"
      "  --upper_triangle_convention (-U) [0|1]  Default: 1 (standard); 0 = complex conjugate.\n"
      "  --check_results (-c)              Check against a CPU correlation.\n"
      "  --verbose (-v)                    Print every mismatching product (implies -c).\n"
      "  --gen_type (-g) [n]               Default: 4. 1 constant, 2 ramp up, 3 ramp down, 4 random.\n"
      "  --random_seed (-r) [n]            Default: 42.\n"
      "  --no_repeat_random (-p)           Continuous random stream instead of one repeated sample.\n"
      "  --generate_frequency (-q) [n]     Default: -1 (all). Other channels get the default value.\n"
      "  --default_real (-x) / --default_imaginary (-y) [n]  Default: 0, range [-8,7].\n"
      "  --initial_real (-X) / --initial_imaginary (-Y) [n]  Default: 0, range [-8,7].\n"
      "  --kernel_batch (-k) [n]           Default: 0. 0 conference, 1 highly packed, 2 int8 matrix cores, 3 FP6 matrix cores.\n"
      "  --output_triangle (-o) [file]     Write the upper-triangle visibilities ([freq][N(N+1)/2][re,im] int32).\n");
}

static int clip(int v, int off, int lo, int hi) {
  v += off;
  return v > hi ? hi : (v < lo ? lo : v);
}

/* Data generator: input_generator.c:17-103 semantics. */
static void generate(int gen, int seed, int dre, int dim, int ire, int iim, int single,
                     int64_t T, int F, int N, int no_repeat, uint8_t *out) {
  if (single > F || single < 0) single = -1;
  const unsigned cdr = (unsigned)clip(dre, 8, 0, 15), cdi = (unsigned)clip(dim, 8, 0, 15);
  const unsigned cir = (unsigned)clip(ire, 8, 0, 15), cii = (unsigned)clip(iim, 8, 0, 15);
  const uint8_t dbyte = (uint8_t)(((cdr << 4) & 0xF0) + (cdi & 0x0F));
  if (gen == 4) srand((unsigned)seed);
  for (int64_t t = 0; t < T; t++) {
    if (gen == 4 && !no_repeat) srand((unsigned)seed);
    for (int f = 0; f < F; f++)
      for (int e = 0; e < N; e++) {
        unsigned re, im;
        if (gen == 1) {
          re = cir;
          im = cii;
        } else if (gen == 2) {
          re = (unsigned)(f + (int)cir + e) % 16u;
          im = (unsigned)(f + (int)cii + e) % 16u;
        } else if (gen == 3) {
          re = 15u - (unsigned)(f + (int)cir + e) % 16u;
          im = 15u - (unsigned)(f + (int)cii + e) % 16u;
        } else if (gen == 4) {
          re = (unsigned)(rand() % 16);
          im = (unsigned)(rand() % 16);
        } else {
          re = cdr;
          im = cdi;
        }
        uint8_t b = (uint8_t)(((re << 4) & 0xF0) + (im & 0x0F));
        out[(t * F + f) * (int64_t)N + e] = (single < 0 || f == single) ? b : dbyte;
      }
  }
}

/* Upper-triangle CPU correlation, cpu_corr_test.c:156-203 (U=1) / :59-106. */
static void cpu_correlate(const uint8_t *in, int64_t T, int F, int N, int U, int32_t *tri) {
  const int64_t ntri = (int64_t)N * (N + 1) / 2;
  uint32_t *acc = (uint32_t *)tri;
  memset(tri, 0, sizeof(int32_t) * 2 * ntri * F);
  for (int64_t t = 0; t < T; t++)
    for (int f = 0; f < F; f++) {
      const uint8_t *row = in + (t * F + f) * (int64_t)N;
      int64_t o = f * ntri;
      for (int y = 0; y < N; y++) {
        const int yr = (row[y] >> 4) - 8, yi = (row[y] & 15) - 8;
        for (int x = y; x < N; x++, o++) {
          const int xr = (row[x] >> 4) - 8, xi = (row[x] & 15) - 8;
          acc[2 * o] += (uint32_t)(xr * yr + xi * yi);
          acc[2 * o + 1] += U ? (uint32_t)(xr * yi - xi * yr) : (uint32_t)(xi * yr - xr * yi);
        }
      }
    }
}

#define CHECK(call)                                                                   \
  do {                                                                                \
    int rc_ = (call);                                                                 \
    if (rc_ != CHX_SUCCESS) {                                                         \
      fprintf(stderr, "%s failed: %s %s\n", #call, chx_strerror(rc_),                 \
              chx_last_error_detail());                                               \
      return 1;                                                                       \
    }                                                                                 \
  } while (0)

int main(int argc, char **argv) {
  int device = 0, iterations = 100, num_freq = 1, num_elem = 2048, time_accum = 256;
  int64_t time_steps = 256 * 128;
  int T_changed = 0, without_copies = 0, check = 0, verbose = 0, gen_type = 4;
  int seed = 42, no_repeat = 0, gen_freq = -1, dre = 0, dim = 0, ire = 0, iim = 0;
  int kernel_batch = 0, U = 1;
  const char *out_path = NULL;
  static struct option opts[] = {
      {"device", required_argument, 0, 'd'},
      {"iterations", required_argument, 0, 'i'},
      {"num_freq", required_argument, 0, 'f'},
      {"num_elem", required_argument, 0, 'e'},
      {"time_accum", required_argument, 0, 't'},
      {"time_steps", required_argument, 0, 'T'},
      {"timer_without_copies", no_argument, 0, 'w'},
      {"upper_triangle_convention", required_argument, 0, 'U'},
      {"check_results", no_argument, 0, 'c'},
      {"verbose", no_argument, 0, 'v'},
      {"gen_type", required_argument, 0, 'g'},
      {"random_seed", required_argument, 0, 'r'},
      {"no_repeat_random", no_argument, 0, 'p'},
      {"generate_frequency", required_argument, 0, 'q'},
      {"default_real", required_argument, 0, 'x'},
      {"default_imaginary", required_argument, 0, 'y'},
      {"initial_real", required_argument, 0, 'X'},
      {"initial_imaginary", required_argument, 0, 'Y'},
      {"kernel_batch", required_argument, 0, 'k'},
      {"output_triangle", required_argument, 0, 'o'},
      {"help", no_argument, 0, 'h'},
      {0, 0, 0, 0}};
  for (;;) {
    int idx = 0;
    int c = getopt_long(argc, argv, "d:i:f:e:t:T:wcvg:r:pq:x:y:X:Y:hk:U:o:", opts, &idx);
    if (c == -1) break;
    switch (c) {
      case 'h': print_help(); return 0;
      case 'd': device = atoi(optarg); break;
      case 'i': iterations = atoi(optarg); break;
      case 'f': num_freq = atoi(optarg); break;
      case 'e': num_elem = atoi(optarg); break;
      case 't':
        time_accum = atoi(optarg);
        if (!T_changed) time_steps = 128LL * time_accum;
        break;
      case 'T': time_steps = atoll(optarg); T_changed = 1; break;
      case 'w': without_copies = 1; break;
      case 'c': check = 1; break;
      case 'v': verbose = 1; check = 1; break;
      case 'g':
        gen_type = atoi(optarg);
        if (gen_type < 1 || gen_type > 15) {  /* main_wrapper.c:177-182 accepts 1..15 */
          printf("Invalid parameter for gen_type.  See help for options\n");
          print_help();
          return -1;
        }
        break;
      case 'r': seed = atoi(optarg); break;
      case 'p': no_repeat = 1; break;
      case 'q': gen_freq = atoi(optarg); break;
      case 'x': dre = atoi(optarg); break;
      case 'y': dim = atoi(optarg); break;
      case 'X': ire = atoi(optarg); break;
      case 'Y': iim = atoi(optarg); break;
      case 'k':
        kernel_batch = atoi(optarg);
        if (kernel_batch < 0 || kernel_batch > 3) {
          printf("Invalid parameter for kernel_batch.  See help for options\n");
          print_help();
          return -1;
        }
        break;
      case 'o': out_path = optarg; break;
      case 'U':
        U = atoi(optarg);
        if (U < 0 || U > 1) {
          printf("Invalid parameter for upper_triangle_convention.  See help for options\n");
          print_help();
          return -1;
        }
        break;
      default: print_help(); return -1;
    }
  }

  chx_config cfg = {num_elem, num_freq, time_steps, time_accum, U, kernel_batch, device};
  chx_plan *plan = NULL;
  CHECK(chx_plan_create(&cfg, &plan));
  size_t in_bytes = 0, out_bytes = 0;
  CHECK(chx_plan_sizes(plan, &in_bytes, &out_bytes, NULL));
  const int64_t NB = chx_num_blocks(num_elem);
  printf("%s: method %d, N=%d, F=%d, T=%lld, time_accum=%d, U=%d, %lld blocks/channel\n",
         chx_version(), kernel_batch, num_elem, num_freq, (long long)time_steps, time_accum,
         U, (long long)NB);

  uint8_t *h_in = NULL;
  int32_t *h_out = NULL;
  CHECK(chx_host_alloc((void **)&h_in, in_bytes));
  CHECK(chx_host_alloc((void **)&h_out, out_bytes));
  generate(gen_type, seed, dre, dim, ire, iim, gen_freq, time_steps, num_freq, num_elem,
           no_repeat, h_in);
  /* Two stages of device buffers, a copy queue and a compute queue ordered
   * by events (main_wrapper.c:264-274, 613-785): pass i copies the input of
   * stage i % 2 while stage (i + 1) % 2, copied in pass i - 1, is
   * correlated; a stage's copy waits for its previous correlation.  One
   * plan serves both stages: its work is queued on the compute queue only. */
  uint8_t *d_in[2] = {NULL, NULL};
  int32_t *d_out[2] = {NULL, NULL};
  CHECK(chx_plan_buffers(plan, &d_in[0], &d_out[0]));
  CHECK(chx_device_alloc(device, in_bytes, (void **)&d_in[1]));
  CHECK(chx_device_alloc(device, out_bytes, (void **)&d_out[1]));
  void *qcopy = NULL, *qcomp = NULL, *written[2] = {NULL, NULL}, *computed[2] = {NULL, NULL};
  CHECK(chx_stream_create(device, &qcopy));
  CHECK(chx_stream_create(device, &qcomp));
  for (int k = 0; k < 2; k++) {
    CHECK(chx_event_create(&written[k]));
    CHECK(chx_event_create(&computed[k]));
  }
  if (without_copies) {  /* both stages loaded once, outside the timed loop */
    CHECK(chx_copy_to_device(d_in[0], h_in, in_bytes, qcopy));
    CHECK(chx_copy_to_device(d_in[1], h_in, in_bytes, qcopy));
    CHECK(chx_stream_synchronize(qcopy));
  }
  /* one untimed correlation (warms the device) */
  CHECK(chx_copy_to_device(d_in[0], h_in, in_bytes, qcomp));
  CHECK(chx_correlate(plan, d_in[0], d_out[0], qcomp));
  CHECK(chx_stream_synchronize(qcomp));

  int computed_once[2] = {0, 0}, last = 0;
  double t0 = wall_s();
  for (int i = 0; i <= iterations; i++) {
    const int W = i % 2, K = (i + 1) % 2;
    if (i < iterations && !without_copies) {
      if (computed_once[W]) CHECK(chx_stream_wait_event(qcopy, computed[W]));
      CHECK(chx_copy_to_device(d_in[W], h_in, in_bytes, qcopy));
      CHECK(chx_event_record(written[W], qcopy));
    }
    if (i >= 1) {
      if (!without_copies) CHECK(chx_stream_wait_event(qcomp, written[K]));
      CHECK(chx_correlate(plan, d_in[K], d_out[K], qcomp));
      CHECK(chx_event_record(computed[K], qcomp));
      computed_once[K] = 1;
      last = K;
    }
  }
  CHECK(chx_stream_synchronize(qcopy));
  CHECK(chx_stream_synchronize(qcomp));
  const double dt = wall_s() - t0;
  CHECK(chx_copy_to_host(h_out, d_out[last], out_bytes, qcomp));
  CHECK(chx_stream_synchronize(qcomp));
  int stages_equal = 1;
  if (computed_once[0] && computed_once[1]) {
    /* both stages hold the same input: their visibilities must be equal */
    int32_t *other = (int32_t *)malloc(out_bytes);
    if (!other) {
      printf("failed to allocate memory\n");
      return 1;
    }
    CHECK(chx_copy_to_host(other, d_out[1 - last], out_bytes, qcomp));
    CHECK(chx_stream_synchronize(qcomp));
    stages_equal = memcmp(other, h_out, out_bytes) == 0;
    free(other);
    printf("Stages 0 and 1 %s.\n", stages_equal ? "agree" : "DIFFER");
  }

  const double vis = (double)num_elem * (num_elem + 1) / 2.0;
  const double samples = (double)time_steps * num_freq * iterations;
  const double cmac_s = samples * vis / dt;
  printf("Correlation matrices computation time: %6.4fs on GPU (%.1f kHz of 400 MHz band, "
         "or %.1fx10^3 correlation matrices/s)%s\n",
         dt, samples / dt / 1e3, samples / dt / 1e3,
         without_copies ? "" : " [H2D included, overlapped with computation]");
  printf("    [%.3f T complex-MAC/s; %.1f%% of the int8 matrix peak (%.1f POP/s)]\n",
         cmac_s / 1e12, 100.0 * 8.0 * cmac_s / INT8_PEAK_OPS, INT8_PEAK_OPS / 1e15);

  int rc = stages_equal ? 0 : 1;
  if (out_path) {
    const int64_t ntri = (int64_t)num_elem * (num_elem + 1) / 2;
    int32_t *tri = (int32_t *)malloc(sizeof(int32_t) * 2 * ntri * num_freq);
    if (!tri) {
      printf("failed to allocate memory\n");
      return 1;
    }
    CHECK(chx_blocks_to_triangle(num_freq, num_elem, h_out, tri));
    FILE *fo = fopen(out_path, "wb");
    const size_t n = (size_t)(2 * ntri * num_freq);
    if (!fo || fwrite(tri, sizeof(int32_t), n, fo) != n) {
      printf("failed to write %s\n", out_path);
      rc = 1;
    }
    if (fo) fclose(fo);
    free(tri);
  }
  if (check) {
    printf("Checking results. Please wait...\n");
    double c0 = wall_s();
    const int64_t ntri = (int64_t)num_elem * (num_elem + 1) / 2;
    int32_t *cpu = (int32_t *)malloc(sizeof(int32_t) * 2 * ntri * num_freq);
    int32_t *gpu = (int32_t *)malloc(sizeof(int32_t) * 2 * ntri * num_freq);
    if (!cpu || !gpu) {
      printf("failed to allocate memory\n");
      return 1;
    }
    cpu_correlate(h_in, time_steps, num_freq, num_elem, U, cpu);
    CHECK(chx_blocks_to_triangle(num_freq, num_elem, h_out, gpu));
    int64_t errors = 0, err2 = 0, maxe = 0;
    for (int f = 0; f < num_freq; f++) {
      int64_t k = f * ntri;
      for (int y = 0; y < num_elem; y++)
        for (int x = y; x < num_elem; x++, k++) {
          const int64_t dr = (int64_t)gpu[2 * k] - cpu[2 * k];
          const int64_t di = (int64_t)gpu[2 * k + 1] - cpu[2 * k + 1];
          if (dr || di) {
            errors++;
            err2 += dr * dr + di * di;
            if (dr * dr + di * di > maxe) maxe = dr * dr + di * di;
            if (verbose)
              printf("freq: %6d element_x: %6d element_y: %6d Real CPU/GPU %8d %8d "
                     "Imaginary CPU/GPU %8d %8d ERR: %7lld\n",
                     f, x, y, cpu[2 * k], gpu[2 * k], cpu[2 * k + 1], gpu[2 * k + 1],
                     (long long)errors);
          }
        }
    }
    printf("\nTotal number of errors: %lld, Sum of Squared Differences: %lld \n",
           (long long)errors, (long long)err2);
    printf("Maximum amplitude squared error: %lld\n", (long long)maxe);
    if (errors > 0 || !stages_equal) {
      printf("Error with correlation/accumulation! Num Err: %lld\n", (long long)errors);
      rc = 1;
    } else {
      printf("Correlation/accumulation successful! CPU matches GPU.\n");
    }
    const double cdt = wall_s() - c0;
    printf("Full Corr: %4.2fs on CPU (%.2f kHz)\n", cdt, time_steps / cdt / 1e3);
    free(cpu);
    free(gpu);
  } else {
    printf("\nGPU calculations have not been verified (use -c).\n\n");
  }
  for (int k = 0; k < 2; k++) {
    chx_event_destroy(written[k]);
    chx_event_destroy(computed[k]);
  }
  chx_stream_destroy(qcopy);
  chx_stream_destroy(qcomp);
  chx_device_free(d_in[1]);
  chx_device_free(d_out[1]);
  chx_host_free(h_in);
  chx_host_free(h_out);
  chx_plan_destroy(plan);
  return rc;
}
