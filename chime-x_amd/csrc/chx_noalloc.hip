// chx_noalloc.hip -- chx_correlate without a plan and without allocating.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc): the
// packed-integer methods' scratch -- the block id maps and the offset sums
// -- lives in a file-scope device array of fixed size instead of plan-owned
// buffers.  The caller passes the input and the output; nothing is
// allocated or freed.
//
// Every call on one GPU uses that GPU's one array, whose contents are those
// of the last call: each call rewrites the maps and zeroes the sums (with
// kernels, not memset nodes) before using them.  Calls that overlap on one
// GPU (on different streams, or from different threads) race on the array.
#include <mutex>

#include "chx_internal.h"

namespace chx {
namespace {

// The scratch of one call: block maps [2][NB] u32, then (256-byte aligned)
// offset sums [F][N][2] u32.  2 MiB holds config 5 (N = 4096, F = 32:
// 1 MiB of sums, 66 KB of maps).
static __device__ __align__(256) uint32_t scratch[CHX_NOALLOC_SCRATCH_BYTES / 4];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint32_t *g_scratch[kMaxDevices];  // device address of `scratch`, per device

// id_x_map / id_y_map (main_wrapper.c:528-540): thread = one row by of
// blocks, bx = by .. nb-1.
__global__ __launch_bounds__(256) void block_maps_kernel(uint32_t *__restrict__ maps, int nb,
                                                         int64_t NB) {
  const int by = blockIdx.x * blockDim.x + threadIdx.x;
  if (by >= nb) return;
  const int64_t b0 = (int64_t)by * nb - (int64_t)by * (by - 1) / 2;
  for (int bx = by; bx < nb; bx++) {
    maps[b0 + bx - by] = (uint32_t)bx;       // x map
    maps[NB + b0 + bx - by] = (uint32_t)by;  // y map
  }
}

// The device address of this device's `scratch` (resolved once per device,
// with the device current).
int scratch_ptr(int device, uint32_t **out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[device]) {
    void *ptr = nullptr;
    CHX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return CHX_ERR_HIP;
    g_scratch[device] = (uint32_t *)ptr;
  }
  *out = g_scratch[device];
  return CHX_SUCCESS;
}

}  // namespace
}  // namespace chx

using namespace chx;

extern "C" int chx_correlate_noalloc(const chx_config *cfg, const uint8_t *d_input,
                                     int32_t *d_vis, void *stream) {
  set_error_detail("");
  if (!d_input || !d_vis) return CHX_ERR_INVALID_ARGUMENT;
  int rc = validate_config(cfg);
  if (rc) return rc;
  if (cfg->kernel_batch != CHX_METHOD_CONFERENCE &&
      cfg->kernel_batch != CHX_METHOD_HIGHLY_PACKED) {
    set_error_detail("chx_correlate_noalloc: the matrix-core methods need a plane workspace; "
                     "use a plan");
    return CHX_ERR_INVALID_ARGUMENT;
  }
  const int64_t NB = chx_num_blocks(cfg->num_elements);
  const int64_t maps_words = (2 * NB + 63) / 64 * 64;  // sums start 256-byte aligned
  const int64_t sums_words = 2 * (int64_t)cfg->num_freq * cfg->num_elements;
  if ((maps_words + sums_words) * 4 > (int64_t)CHX_NOALLOC_SCRATCH_BYTES) {
    char msg[160];
    snprintf(msg, sizeof(msg),
             "chx_correlate_noalloc: needs %lld bytes of scratch, the limit is %lld",
             (long long)((maps_words + sums_words) * 4), (long long)CHX_NOALLOC_SCRATCH_BYTES);
    set_error_detail(msg);
    return CHX_ERR_BUDGET;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CHX_ERR_NO_DEVICE;
  if (cfg->device < 0 || cfg->device >= ndev || cfg->device >= kMaxDevices)
    return CHX_ERR_NO_DEVICE;
  int cur = -1;
  CHX_HIP(hipGetDevice(&cur));
  if (cur != cfg->device) CHX_HIP(hipSetDevice(cfg->device));
  // the caller's current device is restored on every return below
  struct Restore {
    int dev;
    bool on;
    ~Restore() {
      if (on) (void)hipSetDevice(dev);
    }
  } restore{cur, cur != cfg->device};
  uint32_t *base = nullptr;
  rc = scratch_ptr(cfg->device, &base);
  if (rc) return rc;

  // A plan on the stack: the packed launchers read its shape and its two
  // scratch pointers only.
  chx_plan p = {};
  p.cfg = *cfg;
  p.nb = cfg->num_elements / 32;
  p.NB = NB;
  p.d_maps = base;
  p.d_sums = base + maps_words;
  p.sums_bytes = (size_t)sums_words * 4;
  hipStream_t s = (hipStream_t)stream;
  block_maps_kernel<<<dim3((unsigned)((p.nb + 255) / 256)), dim3(256), 0, s>>>(p.d_maps, p.nb,
                                                                                NB);
  CHX_HIP(hipGetLastError());
  rc = launch_fill_u32(p.d_sums, 0u, sums_words, s);
  if (!rc) rc = launch_offset_accumulate(&p, d_input, p.d_sums, s);
  if (!rc) rc = launch_preseed(&p, p.d_sums, d_vis, s);
  if (!rc) rc = launch_corr_packed(&p, d_input, d_vis, s);
  return rc;
}
