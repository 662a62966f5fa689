// chx_gen.hip -- the reference's test-data generator (input_generator.c:17-103)
// on the device, byte-identical to it.
//
// Random mode uses libc rand() after srand(seed) (input_generator.c:46-54,
// 75-76).  glibc's rand() is random_r() with the default TYPE_3 state: an
// additive lagged-Fibonacci generator (published algorithm of glibc
// stdlib/random_r.c):
//   r_0 = seed (1 if seed is 0);
//   r_i = 16807 r_{i-1} mod (2^31 - 1)   for i = 1..30 (Schrage's method on
//         signed 32-bit values, as srandom_r computes it);
//   r_i = r_{i-31}                        for i = 31..33;
//   r_i = r_{i-31} + r_{i-3} mod 2^32     for i >= 34;
//   the k-th rand() after srand returns r_{k+344} >> 1.
// The 31-word window (r_{i-31} .. r_{i-1}) advances by a linear map W over
// Z/2^32, so the window at output k is W^k applied to the window at output
// 0.  The host computes W^L, W^2L, W^4L, ... (L outputs per thread segment);
// a first kernel jumps every thread to its segment start with those
// matrices, a second runs the recurrence in registers and writes the bytes
// (byte b = (o_2b % 16) << 4 | o_2b+1 % 16, in the generator's [t][f][e]
// order).  The repeat mode (srand before every time sample) generates one
// time row and copies it to every row.
#include <string.h>

#include <vector>

#include "chx_internal.h"

namespace chx {
namespace {

constexpr int kW = 31;               // window length
constexpr int kSegOut = 248 * 64;    // outputs per thread segment (31 * 8 * 64)
constexpr int kSegBytes = kSegOut / 2;

struct Mat {
  uint32_t a[kW][kW];
};

void mat_mul(const Mat &x, const Mat &y, Mat &z) {
  Mat t;
  for (int i = 0; i < kW; i++)
    for (int j = 0; j < kW; j++) {
      uint32_t s = 0;
      for (int k = 0; k < kW; k++) s += x.a[i][k] * y.a[k][j];
      t.a[i][j] = s;
    }
  z = t;
}

// One step: new window = (w1, ..., w30, w0 + w28).
void step_matrix(Mat &m) {
  memset(&m, 0, sizeof(m));
  for (int j = 0; j < kW - 1; j++) m.a[j][j + 1] = 1;
  m.a[kW - 1][0] = 1;
  m.a[kW - 1][28] = 1;
}

void mat_pow(const Mat &m, uint64_t e, Mat &out) {
  Mat r, b = m;
  memset(&r, 0, sizeof(r));
  for (int i = 0; i < kW; i++) r.a[i][i] = 1;
  while (e) {
    if (e & 1) mat_mul(r, b, r);
    mat_mul(b, b, b);
    e >>= 1;
  }
  out = r;
}

void mat_vec(const Mat &m, const uint32_t *v, uint32_t *out) {
  uint32_t t[kW];
  for (int i = 0; i < kW; i++) {
    uint32_t s = 0;
    for (int k = 0; k < kW; k++) s += m.a[i][k] * v[k];
    t[i] = s;
  }
  memcpy(out, t, sizeof(t));
}

// Window (r_313 .. r_343) before the first rand() output after srand(seed).
void initial_window(int32_t seed, uint32_t *w) {
  std::vector<uint32_t> r(344);
  int32_t word = seed == 0 ? 1 : seed;
  r[0] = (uint32_t)word;
  for (int i = 1; i < 31; i++) {
    const int32_t hi = word / 127773, lo = word % 127773;
    word = 16807 * lo - 2836 * hi;
    if (word < 0) word += 2147483647;
    r[i] = (uint32_t)word;
  }
  for (int i = 31; i < 34; i++) r[i] = r[i - 31];
  for (int i = 34; i < 344; i++) r[i] = r[i - 31] + r[i - 3];
  for (int j = 0; j < kW; j++) w[j] = r[313 + j];
}

// Thread s: window at output s * kSegOut = prod_b P_b^(bit b of s) w0.
__global__ __launch_bounds__(256) void jump_kernel(const uint32_t *__restrict__ P, int nbits,
                                                   const uint32_t *__restrict__ w0,
                                                   uint32_t *__restrict__ states, int64_t nseg) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  uint32_t v[kW];
#pragma unroll
  for (int j = 0; j < kW; j++) v[j] = w0[j];
  for (int b = 0; b < nbits; b++) {
    if (!((s >> b) & 1)) continue;
    const uint32_t *m = P + (size_t)b * kW * kW;
    uint32_t t[kW];
#pragma unroll
    for (int i = 0; i < kW; i++) {
      uint32_t acc = 0;
#pragma unroll
      for (int k = 0; k < kW; k++) acc += m[i * kW + k] * v[k];
      t[i] = acc;
    }
#pragma unroll
    for (int j = 0; j < kW; j++) v[j] = t[j];
  }
#pragma unroll
  for (int j = 0; j < kW; j++) states[s * kW + j] = v[j];
}

// Thread s: outputs [s*kSegOut, (s+1)*kSegOut) -> bytes [s*kSegBytes, ...),
// stopping at nbytes.  248 outputs (31 words of 4 bytes) per unrolled
// block, so window slots and word positions are compile-time constants.
__global__ __launch_bounds__(256) void rand_bytes_kernel(const uint32_t *__restrict__ states,
                                                         int64_t nseg, uint8_t *__restrict__ out,
                                                         int64_t nbytes) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  uint32_t q[kW];
#pragma unroll
  for (int j = 0; j < kW; j++) q[j] = states[s * kW + j];
  const int64_t base = s * (int64_t)kSegBytes;
  for (int blk = 0; blk < kSegOut / 248; blk++) {
    const int64_t b0 = base + (int64_t)blk * 124;  // first byte of the block
    if (b0 >= nbytes) break;
    uint32_t word = 0;
#pragma unroll
    for (int i = 0; i < 248; i++) {
      const int h = i % kW;  // slot of r_{k-31}
      const uint32_t r = q[h] + q[(h + 28) % kW];
      q[h] = r;
      const uint32_t nib = (r >> 1) & 15u;
      // output i: byte i/2 of the block, high nibble if i is even
      const int byte_in_word = (i / 2) % 4;
      word |= nib << (8 * byte_in_word + ((i & 1) ? 0 : 4));
      if (i % 8 == 7) {
        const int64_t at = b0 + (i / 8) * 4;
        if (at + 4 <= nbytes) {
          *(uint32_t *)(out + at) = word;
        } else {
          for (int k = 0; k < 4 && at + k < nbytes; k++) out[at + k] = (uint8_t)(word >> (8 * k));
        }
        word = 0;
      }
    }
  }
}

// Non-random types (and the default for unknown ones) at (t, f, e).
__global__ void pattern_kernel(uint8_t *__restrict__ out, int64_t T, int F, int N, int type,
                               unsigned ir, unsigned ii, unsigned dr, unsigned di) {
  const int64_t n = T * F * (int64_t)N;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int e = (int)(i % N);
    const int f = (int)((i / N) % F);
    unsigned re, im;
    if (type == 1) {
      re = ir;
      im = ii;
    } else if (type == 2) {
      re = (unsigned)(f + (int)ir + e) % 16u;
      im = (unsigned)(f + (int)ii + e) % 16u;
    } else if (type == 3) {
      re = 15u - (unsigned)(f + (int)ir + e) % 16u;
      im = 15u - (unsigned)(f + (int)ii + e) % 16u;
    } else {
      re = dr;
      im = di;
    }
    out[i] = (uint8_t)(((re << 4) & 0xF0u) + (im & 0x0Fu));
  }
}

// Rows 1..T-1 = row 0 (the repeat mode's srand before every sample).
__global__ void repeat_rows_kernel(uint8_t *__restrict__ out, int64_t T, int64_t row) {
  const int64_t n = (T - 1) * row;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[row + i] = out[i % row];
}

// Channels other than sf get the default byte.
__global__ void single_freq_kernel(uint8_t *__restrict__ out, int64_t T, int F, int N, int sf,
                                   uint8_t dflt) {
  const int64_t n = T * F * (int64_t)N;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if ((int)((i / N) % F) != sf) out[i] = dflt;
}

int clip15(int v) {  // offset_and_clip_value(v, 8, 0, 15), input_generator.c:8-15
  v += 8;
  return v > 15 ? 15 : (v < 0 ? 0 : v);
}

// Jump matrices P_b = W^(kSegOut 2^b) for b < nbits (host).
void jump_matrices(int nbits, std::vector<uint32_t> &P) {
  Mat w, ml;
  step_matrix(w);
  mat_pow(w, kSegOut, ml);
  P.resize((size_t)nbits * kW * kW);
  for (int b = 0; b < nbits; b++) {
    memcpy(&P[(size_t)b * kW * kW], &ml.a[0][0], sizeof(uint32_t) * kW * kW);
    mat_mul(ml, ml, ml);
  }
}

// The first nbytes bytes of the random stream after srand(seed) into out.
int launch_rand_bytes(int32_t seed, int64_t nbytes, uint8_t *out, hipStream_t s) {
  const int64_t nseg = (nbytes + kSegBytes - 1) / kSegBytes;
  int nbits = 1;
  while ((1LL << nbits) < nseg) nbits++;
  std::vector<uint32_t> P;
  jump_matrices(nbits, P);
  uint32_t w0[kW];
  initial_window(seed, w0);
  uint32_t *d_P = nullptr, *d_w0 = nullptr, *d_states = nullptr;
  hipError_t e = hipMalloc((void **)&d_P, P.size() * 4);
  if (e == hipSuccess) e = hipMalloc((void **)&d_w0, sizeof(w0));
  if (e == hipSuccess) e = hipMalloc((void **)&d_states, (size_t)nseg * kW * 4);
  if (e == hipSuccess) e = hipMemcpy(d_P, P.data(), P.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_w0, w0, sizeof(w0), hipMemcpyHostToDevice);
  const unsigned grid = (unsigned)((nseg + 255) / 256);
  if (e == hipSuccess) {
    jump_kernel<<<grid, 256, 0, s>>>(d_P, nbits, d_w0, d_states, nseg);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    rand_bytes_kernel<<<grid, 256, 0, s>>>(d_states, nseg, out, nbytes);
    e = hipGetLastError();
  }
  // the temporaries are freed once the kernels that read them are done
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (d_P) (void)hipFree(d_P);
  if (d_w0) (void)hipFree(d_w0);
  if (d_states) (void)hipFree(d_states);
  return e == hipSuccess ? CHX_SUCCESS : hip_status(e);
}

}  // namespace
}  // namespace chx

using namespace chx;

extern "C" {

int chx_rand_host(int32_t seed, int64_t k0, int64_t n, int32_t *out) {
  if (!out || k0 < 0 || n < 0) return CHX_ERR_INVALID_ARGUMENT;
  uint32_t w[kW];
  initial_window(seed, w);
  Mat m, mk;
  step_matrix(m);
  mat_pow(m, (uint64_t)k0, mk);
  mat_vec(mk, w, w);
  for (int64_t k = 0; k < n; k++) {
    const uint32_t r = w[0] + w[28];
    memmove(w, w + 1, sizeof(uint32_t) * (kW - 1));
    w[kW - 1] = r;
    out[k] = (int32_t)(r >> 1);
  }
  return CHX_SUCCESS;
}

int chx_generate_device(int32_t gen_type, int32_t seed, int32_t default_real,
                        int32_t default_imag, int32_t initial_real, int32_t initial_imag,
                        int32_t single_freq, int64_t T, int32_t F, int32_t N,
                        int32_t no_repeat, uint8_t *d_data, void *stream) {
  if (!d_data || T < 1 || F < 1 || N < 1) return CHX_ERR_INVALID_ARGUMENT;
  if (T > (INT64_MAX / 4) / ((int64_t)F * N)) return CHX_ERR_SHAPE;
  hipStream_t s = (hipStream_t)stream;
  if (single_freq > F || single_freq < 0) single_freq = -1;  // input_generator.c:31-32
  const unsigned dr = (unsigned)clip15(default_real), di = (unsigned)clip15(default_imag);
  const unsigned ir = (unsigned)clip15(initial_real), ii = (unsigned)clip15(initial_imag);
  const int64_t row = (int64_t)F * N, total = T * row;
  const unsigned grid = (unsigned)std::min<int64_t>((total + 255) / 256, 16384);
  if (gen_type == 4) {  // GENERATE_DATASET_RANDOM_SEEDED
    const int64_t nbytes = no_repeat ? total : row;
    int rc = launch_rand_bytes(seed, nbytes, d_data, s);
    if (rc) return rc;
    if (!no_repeat && T > 1) {
      repeat_rows_kernel<<<grid, 256, 0, s>>>(d_data, T, row);
      CHX_HIP(hipGetLastError());
    }
  } else {
    pattern_kernel<<<grid, 256, 0, s>>>(d_data, T, F, N, gen_type, ir, ii, dr, di);
    CHX_HIP(hipGetLastError());
  }
  if (single_freq >= 0) {
    single_freq_kernel<<<grid, 256, 0, s>>>(d_data, T, F, N, single_freq,
                                            (uint8_t)(((dr << 4) & 0xF0u) + (di & 0x0Fu)));
    CHX_HIP(hipGetLastError());
  }
  return CHX_SUCCESS;
}

}  // extern "C"
