// chx_api.hip -- the C ABI (include/chx.h): plans, validation, dispatch and
// the host-side layout helpers.  Replaces the reference host runtime
// (main_wrapper.c:238-572 set-up, :691-778 per-iteration kernel order).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <new>

#include "chx_internal.h"

namespace chx {

static thread_local char g_last_error[256];

int hip_status(hipError_t e) {
  snprintf(g_last_error, sizeof(g_last_error), "%s: %s", hipGetErrorName(e),
           hipGetErrorString(e));
  if (e == hipErrorOutOfMemory) return CHX_ERR_OUT_OF_MEMORY;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return CHX_ERR_NO_DEVICE;
  return CHX_ERR_HIP;
}

void set_error_detail(const char *msg) {
  snprintf(g_last_error, sizeof(g_last_error), "%s", msg);
}

// Shape rules (main_wrapper.c:277-278, 564-572 and the kernels' loop
// steps), enforced here because the reference does not check them.
int validate_config(const chx_config *c) {
  if (!c) return CHX_ERR_INVALID_ARGUMENT;
  if (c->kernel_batch < CHX_METHOD_CONFERENCE || c->kernel_batch > CHX_METHOD_FP6_MFMA)
    return CHX_ERR_INVALID_ARGUMENT;
  if (c->upper_triangle_convention != 0 && c->upper_triangle_convention != 1)
    return CHX_ERR_INVALID_ARGUMENT;
  if (c->num_elements < 32 || c->num_elements % 32 != 0) return CHX_ERR_SHAPE;
  if (c->num_freq < 1) return CHX_ERR_SHAPE;
  if (c->time_steps < 32 || c->time_steps % 32 != 0) return CHX_ERR_SHAPE;
  // Byte count of the input must fit a signed 64-bit size, and the block
  // buffer a 32-bit grid.
  const int64_t nb = c->num_elements / 32;
  if ((int64_t)c->num_freq * nb * (nb + 1) / 2 > 0x7FFFFFFFLL) return CHX_ERR_SHAPE;
  if (c->time_steps > (INT64_MAX / 2) / ((int64_t)c->num_freq * c->num_elements))
    return CHX_ERR_SHAPE;
  const int ta = c->time_accum;
  if (ta < 8 || ta % 8 != 0 || c->time_steps % ta != 0) return CHX_ERR_TIME_ACCUM;
  const int32_t cap = chx_max_time_accum(c->kernel_batch);
  if (cap > 0 && ta > cap) return CHX_ERR_TIME_ACCUM;
  return CHX_SUCCESS;
}

}  // namespace chx

using namespace chx;

extern "C" {

const char *chx_version(void) { return "chx 0.1.0 (gfx950)"; }

const char *chx_strerror(int status) {
  switch (status) {
    case CHX_SUCCESS: return "success";
    case CHX_ERR_INVALID_ARGUMENT: return "invalid argument";
    case CHX_ERR_SHAPE: return "invalid shape (N % 32, N < 32, F < 1, T % 32 or T < 32)";
    case CHX_ERR_TIME_ACCUM: return "invalid time_accum (% 8, T % time_accum, or over the method cap)";
    case CHX_ERR_NO_DEVICE: return "no usable GPU device";
    case CHX_ERR_OUT_OF_MEMORY: return "out of memory";
    case CHX_ERR_HIP: return "HIP runtime error";
    case CHX_ERR_BUDGET: return "request exceeds the HBM budget";
    case CHX_ERR_BUSY: return "no free slot";
    case CHX_ERR_NOT_FOUND: return "unknown handle";
    default: return "unknown status";
  }
}

const char *chx_last_error_detail(void) { return g_last_error; }

int32_t chx_max_time_accum(int32_t kernel_batch) {
  switch (kernel_batch) {
    case CHX_METHOD_CONFERENCE: return CHX_CONFERENCE_MAX_TIME_ACCUM;
    case CHX_METHOD_HIGHLY_PACKED: return CHX_HIGHLY_PACKED_MAX_TIME_ACCUM;
    case CHX_METHOD_INT8_MFMA: return 0;
    case CHX_METHOD_FP6_MFMA: return 0;
    default: return -1;
  }
}

int chx_device_arch(int32_t device, char *name, size_t len) {
  if (!name || len == 0) return CHX_ERR_INVALID_ARGUMENT;
  name[0] = 0;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CHX_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return CHX_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  CHX_HIP(hipGetDeviceProperties(&prop, device));
  size_t k = 0;  // the target name, without its feature suffixes (":sramecc+:xnack-")
  while (k + 1 < len && prop.gcnArchName[k] && prop.gcnArchName[k] != ':') {
    name[k] = prop.gcnArchName[k];
    k++;
  }
  name[k] = 0;
  return CHX_SUCCESS;
}

int chx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int64_t chx_num_blocks(int32_t num_elements) {
  if (num_elements < 32 || num_elements % 32) return -1;
  const int64_t nb = num_elements / 32;
  return nb * (nb + 1) / 2;
}

int chx_block_maps(int32_t num_elements, uint32_t *x_map, uint32_t *y_map) {
  if (!x_map || !y_map || chx_num_blocks(num_elements) < 0) return CHX_ERR_INVALID_ARGUMENT;
  const int nb = num_elements / 32;
  int64_t b = 0;
  for (int by = 0; by < nb; by++)
    for (int bx = by; bx < nb; bx++, b++) {
      x_map[b] = (uint32_t)bx;
      y_map[b] = (uint32_t)by;
    }
  return CHX_SUCCESS;
}

int chx_blocks_to_triangle(int32_t num_freq, int32_t num_elements, const int32_t *blocks,
                           int32_t *triangle) {
  if (!blocks || !triangle || num_freq < 1 || chx_num_blocks(num_elements) < 0)
    return CHX_ERR_INVALID_ARGUMENT;
  const int N = num_elements, nb = N / 32;
  const int64_t NB = chx_num_blocks(N), ntri = (int64_t)N * (N + 1) / 2;
  for (int f = 0; f < num_freq; f++) {
    int64_t b = 0;
    for (int by = 0; by < nb; by++)
      for (int bx = by; bx < nb; bx++, b++) {
        const int32_t *blk = blocks + ((int64_t)f * NB + b) * 2048;
        for (int ly = 0; ly < 32; ly++) {
          const int64_t y = by * 32 + ly;
          for (int lx = (bx == by ? ly : 0); lx < 32; lx++) {
            const int64_t x = bx * 32 + lx;
            const int64_t a = f * ntri + y * N - ((y - 1) * y) / 2 + (x - y);
            triangle[2 * a] = blk[(ly * 32 + lx) * 2];
            triangle[2 * a + 1] = blk[(ly * 32 + lx) * 2 + 1];
          }
        }
      }
  }
  return CHX_SUCCESS;
}

int chx_blocks_to_full(int32_t num_freq, int32_t num_elements, const int32_t *blocks,
                       int32_t *full) {
  if (!blocks || !full || num_freq < 1 || chx_num_blocks(num_elements) < 0)
    return CHX_ERR_INVALID_ARGUMENT;
  const int N = num_elements, nb = N / 32;
  const int64_t NB = chx_num_blocks(N);
  for (int f = 0; f < num_freq; f++) {
    int64_t b = 0;
    int32_t *o = full + (int64_t)f * N * N * 2;
    for (int by = 0; by < nb; by++)
      for (int bx = by; bx < nb; bx++, b++) {
        const int32_t *blk = blocks + ((int64_t)f * NB + b) * 2048;
        for (int ly = 0; ly < 32; ly++) {
          const int64_t y = by * 32 + ly;
          for (int lx = (bx == by ? ly : 0); lx < 32; lx++) {
            const int64_t x = bx * 32 + lx;
            const int32_t re = blk[(ly * 32 + lx) * 2], im = blk[(ly * 32 + lx) * 2 + 1];
            o[(y * N + x) * 2] = re;
            o[(y * N + x) * 2 + 1] = im;
            o[(x * N + y) * 2] = re;  // x == y: the same entry again
            o[(x * N + y) * 2 + 1] = x == y ? im : (int32_t)(0u - (uint32_t)im);
          }
        }
      }
  }
  return CHX_SUCCESS;
}

int chx_split_paired_feeds(int32_t num_freq, int32_t num_elements, int32_t *data) {
  if (!data || num_freq < 0 || num_elements < 0) return CHX_ERR_INVALID_ARGUMENT;
  // Per packed frequency g and row y of its 2E x 2E matrix, the E entries of
  // the row's own half move down to the next free output row. The write
  // cursor never passes the read cursor, so memmove keeps it in place.
  const int64_t E = num_elements, row = 2 * E;  // complex entries per input row
  int64_t w = 0;
  for (int64_t g = 0; g < num_freq / 2; g++)
    for (int64_t y = 0; y < row; y++) {
      const int64_t r = ((g * row + y) * row + (y < E ? 0 : E)) * 2;
      memmove(data + w, data + r, (size_t)E * 2 * sizeof(int32_t));
      w += E * 2;
    }
  return CHX_SUCCESS;
}

int chx_feed16_to_triangle(int32_t num_freq_out, int32_t num_freq, const int32_t *input,
                           int32_t *output) {
  if (!input || !output || num_freq_out < 0 || num_freq < 0) return CHX_ERR_INVALID_ARGUMENT;
  int32_t *o = output;
  for (int64_t f = 0; f < num_freq_out; f++)
    for (int y = 0; y < 16; y++) {
      const size_t n = (size_t)(16 - y) * 2;
      if (f < num_freq)
        memcpy(o, input + (f * 256 + y * 17) * 2, n * sizeof(int32_t));
      else
        memset(o, 0, n * sizeof(int32_t));
      o += n;
    }
  return CHX_SUCCESS;
}

int chx_blocks_to_full_device(int32_t num_freq, int32_t num_elements, const int32_t *d_blocks,
                              int32_t *d_full, void *stream) {
  if (!d_blocks || !d_full || num_freq < 1 || chx_num_blocks(num_elements) < 0)
    return CHX_ERR_INVALID_ARGUMENT;
  if ((int64_t)(num_elements / 32) * (num_elements / 32) > 0x7FFFFFFFLL || num_freq > 65535)
    return CHX_ERR_SHAPE;
  return launch_blocks_to_full(num_freq, num_elements, d_blocks, d_full, (hipStream_t)stream);
}

int chx_blocks_to_triangle_device(int32_t num_freq, int32_t num_elements,
                                  const int32_t *d_blocks, int32_t *d_triangle,
                                  void *stream) {
  if (!d_blocks || !d_triangle || num_freq < 1 || chx_num_blocks(num_elements) < 0)
    return CHX_ERR_INVALID_ARGUMENT;
  return launch_blocks_to_triangle(num_freq, num_elements, d_blocks, d_triangle,
                                   (hipStream_t)stream);
}

int chx_plan_create(const chx_config *cfg, chx_plan **out) {
  if (!out) return CHX_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int rc = validate_config(cfg);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CHX_ERR_NO_DEVICE;
  if (cfg->device < 0 || cfg->device >= ndev) return CHX_ERR_NO_DEVICE;
  CHX_HIP(hipSetDevice(cfg->device));

  chx_plan *p = (chx_plan *)calloc(1, sizeof(chx_plan));
  if (!p) return CHX_ERR_OUT_OF_MEMORY;
  p->cfg = *cfg;
  p->nb = cfg->num_elements / 32;
  p->NB = chx_num_blocks(cfg->num_elements);
  p->in_bytes = (size_t)cfg->time_steps * cfg->num_freq * cfg->num_elements;
  p->out_bytes = (size_t)cfg->num_freq * p->NB * 2048 * sizeof(int32_t);
  p->sums_bytes = (size_t)cfg->num_freq * cfg->num_elements * 2 * sizeof(uint32_t);

  uint32_t *maps = (uint32_t *)malloc(sizeof(uint32_t) * 2 * p->NB);
  if (!maps) {
    free(p);
    return CHX_ERR_OUT_OF_MEMORY;
  }
  chx_block_maps(cfg->num_elements, maps, maps + p->NB);
  hipError_t e = hipMalloc((void **)&p->d_maps, sizeof(uint32_t) * 2 * p->NB);
  if (e == hipSuccess)
    e = hipMemcpy(p->d_maps, maps, sizeof(uint32_t) * 2 * p->NB, hipMemcpyHostToDevice);
  free(maps);
  p->ntiles = (cfg->num_elements + 127) / 128;
  const bool matrix = cfg->kernel_batch == CHX_METHOD_INT8_MFMA ||
                      cfg->kernel_batch == CHX_METHOD_FP6_MFMA;
  if (e == hipSuccess && !matrix) e = hipMalloc((void **)&p->d_sums, p->sums_bytes);
  if (e == hipSuccess && matrix) {
    p->work_bytes = cfg->kernel_batch == CHX_METHOD_FP6_MFMA ? fp6_workspace_bytes(p)
                                                             : mfma_workspace_bytes(p);
    e = hipMalloc((void **)&p->d_work, p->work_bytes);
  }
  if (e != hipSuccess) {
    rc = hip_status(e);
    chx_plan_destroy(p);
    return rc;
  }
  *out = p;
  return CHX_SUCCESS;
}

int chx_plan_profile(chx_plan *p, int32_t enable) {
  if (!p) return CHX_ERR_INVALID_ARGUMENT;
  if (enable && !p->prof_ev) {
    p->prof_ev = new (std::nothrow) std::vector<hipEvent_t>();
    if (!p->prof_ev) return CHX_ERR_OUT_OF_MEMORY;
  }
  p->prof = enable ? 1 : 0;
  return CHX_SUCCESS;
}

int chx_plan_kernel_times(chx_plan *p, double *unpack_ms, double *gemm_ms, int64_t *launches) {
  if (!p) return CHX_ERR_INVALID_ARGUMENT;
  double u = 0, g = 0;
  int64_t n = 0;
  int rc = CHX_SUCCESS;
  if (p->prof_ev) {
    std::vector<hipEvent_t> &ev = *p->prof_ev;
    for (size_t k = 0; k + 2 < ev.size(); k += 3) {
      float a = 0, b = 0;
      hipError_t e = hipEventSynchronize(ev[k + 2]);
      if (e == hipSuccess) e = hipEventElapsedTime(&a, ev[k], ev[k + 1]);
      if (e == hipSuccess) e = hipEventElapsedTime(&b, ev[k + 1], ev[k + 2]);
      if (e != hipSuccess && !rc) rc = hip_status(e);
      u += a;
      g += b;
      n++;
    }
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    ev.clear();
  }
  if (unpack_ms) *unpack_ms = u;
  if (gemm_ms) *gemm_ms = g;
  if (launches) *launches = n;
  return rc;
}

int chx_plan_destroy(chx_plan *p) {
  if (!p) return CHX_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(p->cfg.device);
  if (p->prof_ev) {
    for (hipEvent_t e : *p->prof_ev)
      if (e) (void)hipEventDestroy(e);
    delete p->prof_ev;
  }
  if (p->d_maps) (void)hipFree(p->d_maps);
  if (p->d_sums) (void)hipFree(p->d_sums);
  if (p->d_work) (void)hipFree(p->d_work);
  if (p->d_in) (void)hipFree(p->d_in);
  if (p->d_out) (void)hipFree(p->d_out);
  free(p);
  return CHX_SUCCESS;
}

int chx_plan_get_config(const chx_plan *p, chx_config *cfg) {
  if (!p || !cfg) return CHX_ERR_INVALID_ARGUMENT;
  *cfg = p->cfg;
  return CHX_SUCCESS;
}

int chx_plan_sizes(const chx_plan *p, size_t *in_b, size_t *out_b, size_t *sums_b) {
  if (!p) return CHX_ERR_INVALID_ARGUMENT;
  if (in_b) *in_b = p->in_bytes;
  if (out_b) *out_b = p->out_bytes;
  if (sums_b) *sums_b = p->sums_bytes;
  return CHX_SUCCESS;
}

int chx_offset_accumulate(const chx_plan *p, const uint8_t *d_input, uint32_t *d_sums,
                          void *stream) {
  if (!p || !d_input || !d_sums) return CHX_ERR_INVALID_ARGUMENT;
  return launch_offset_accumulate(p, d_input, d_sums, (hipStream_t)stream);
}

int chx_preseed(const chx_plan *p, const uint32_t *d_sums, int32_t *d_vis, void *stream) {
  if (!p || !d_sums || !d_vis) return CHX_ERR_INVALID_ARGUMENT;
  if (p->cfg.kernel_batch >= CHX_METHOD_INT8_MFMA) return CHX_ERR_INVALID_ARGUMENT;
  return launch_preseed(p, d_sums, d_vis, (hipStream_t)stream);
}

int chx_corr(const chx_plan *p, const uint8_t *d_input, int32_t *d_vis, void *stream) {
  if (!p || !d_input || !d_vis) return CHX_ERR_INVALID_ARGUMENT;
  if (p->cfg.kernel_batch == CHX_METHOD_INT8_MFMA)
    return launch_corr_mfma(p, d_input, d_vis, 1, (hipStream_t)stream);
  if (p->cfg.kernel_batch == CHX_METHOD_FP6_MFMA)
    return launch_corr_fp6(p, d_input, d_vis, 1, (hipStream_t)stream);
  return launch_corr_packed(p, d_input, d_vis, (hipStream_t)stream);
}

int chx_correlate(chx_plan *p, const uint8_t *d_input, int32_t *d_vis, void *stream) {
  if (!p || !d_input || !d_vis) return CHX_ERR_INVALID_ARGUMENT;
  hipStream_t s = (hipStream_t)stream;
  if (p->cfg.kernel_batch == CHX_METHOD_INT8_MFMA)
    return launch_corr_mfma(p, d_input, d_vis, 0, s);
  if (p->cfg.kernel_batch == CHX_METHOD_FP6_MFMA) return launch_corr_fp6(p, d_input, d_vis, 0, s);
  int rc = launch_fill_u32(p->d_sums, 0u, (int64_t)p->sums_bytes / 4, s);
  if (!rc) rc = launch_offset_accumulate(p, d_input, p->d_sums, s);
  if (!rc) rc = launch_preseed(p, p->d_sums, d_vis, s);
  if (!rc) rc = launch_corr_packed(p, d_input, d_vis, s);
  return rc;
}

int chx_plan_buffers(chx_plan *p, uint8_t **d_input, int32_t **d_vis) {
  if (!p) return CHX_ERR_INVALID_ARGUMENT;
  CHX_HIP(hipSetDevice(p->cfg.device));
  if (!p->d_in) CHX_HIP(hipMalloc((void **)&p->d_in, p->in_bytes));
  if (!p->d_out) CHX_HIP(hipMalloc((void **)&p->d_out, p->out_bytes));
  if (d_input) *d_input = p->d_in;
  if (d_vis) *d_vis = p->d_out;
  return CHX_SUCCESS;
}

int chx_device_alloc(int32_t device, size_t bytes, void **ptr) {
  if (!ptr) return CHX_ERR_INVALID_ARGUMENT;
  *ptr = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CHX_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return CHX_ERR_NO_DEVICE;
  CHX_HIP(hipSetDevice(device));
  CHX_HIP(hipMalloc(ptr, bytes));
  return CHX_SUCCESS;
}

int chx_device_free(void *ptr) {
  if (!ptr) return CHX_ERR_INVALID_ARGUMENT;
  CHX_HIP(hipFree(ptr));
  return CHX_SUCCESS;
}

int chx_host_alloc(void **ptr, size_t bytes) {
  if (!ptr) return CHX_ERR_INVALID_ARGUMENT;
  CHX_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
  return CHX_SUCCESS;
}

int chx_host_free(void *ptr) {
  if (!ptr) return CHX_ERR_INVALID_ARGUMENT;
  CHX_HIP(hipHostFree(ptr));
  return CHX_SUCCESS;
}

int chx_copy_to_device(void *d_dst, const void *h_src, size_t bytes, void *stream) {
  if (!d_dst || !h_src) return CHX_ERR_INVALID_ARGUMENT;
  CHX_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  return CHX_SUCCESS;
}

int chx_copy_to_host(void *h_dst, const void *d_src, size_t bytes, void *stream) {
  if (!h_dst || !d_src) return CHX_ERR_INVALID_ARGUMENT;
  CHX_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  return CHX_SUCCESS;
}

int chx_stream_synchronize(void *stream) {
  CHX_HIP(hipStreamSynchronize((hipStream_t)stream));
  return CHX_SUCCESS;
}

int chx_stream_create(int32_t device, void **stream) {
  if (!stream) return CHX_ERR_INVALID_ARGUMENT;
  *stream = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CHX_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return CHX_ERR_NO_DEVICE;
  CHX_HIP(hipSetDevice(device));
  hipStream_t s = nullptr;
  CHX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void *)s;
  return CHX_SUCCESS;
}

int chx_stream_destroy(void *stream) {
  if (!stream) return CHX_ERR_INVALID_ARGUMENT;
  CHX_HIP(hipStreamDestroy((hipStream_t)stream));
  return CHX_SUCCESS;
}

int chx_event_create(void **event) {
  if (!event) return CHX_ERR_INVALID_ARGUMENT;
  *event = nullptr;
  hipEvent_t e = nullptr;
  CHX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  *event = (void *)e;
  return CHX_SUCCESS;
}

int chx_event_destroy(void *event) {
  if (!event) return CHX_ERR_INVALID_ARGUMENT;
  CHX_HIP(hipEventDestroy((hipEvent_t)event));
  return CHX_SUCCESS;
}

int chx_event_record(void *event, void *stream) {
  if (!event) return CHX_ERR_INVALID_ARGUMENT;
  CHX_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
  return CHX_SUCCESS;
}

int chx_stream_wait_event(void *stream, void *event) {
  if (!event) return CHX_ERR_INVALID_ARGUMENT;
  CHX_HIP(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0));
  return CHX_SUCCESS;
}

int chx_correlate_host(chx_plan *p, const uint8_t *h_input, int32_t *h_vis) {
  if (!p || !h_input || !h_vis) return CHX_ERR_INVALID_ARGUMENT;
  int rc = chx_plan_buffers(p, nullptr, nullptr);
  if (rc) return rc;
  CHX_HIP(hipMemcpy(p->d_in, h_input, p->in_bytes, hipMemcpyHostToDevice));
  rc = chx_correlate(p, p->d_in, p->d_out, nullptr);
  if (rc) return rc;
  CHX_HIP(hipMemcpy(h_vis, p->d_out, p->out_bytes, hipMemcpyDeviceToHost));
  return CHX_SUCCESS;
}

}  // extern "C"
