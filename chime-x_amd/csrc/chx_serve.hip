// chx_serve.hip -- streaming correlation with continuous batching.
//
// Threads: callers open streams and submit frames (any thread); one worker
// thread per engine forms and runs launches.  A frame takes a slot (at most
// frame_slots frames in flight); a host frame is copied into its slot's
// device buffer on the copy stream, a device frame is read in place.  When
// the GPU is free, the worker takes every queued frame whose copy has
// landed (in submit order, up to max_batch_frames and the workspace),
// groups them by stream and launches, on the compute stream:
//   one batched unpack of all frames (a stream's frames side by side in one
//   plane range), then one persistent matrix-core pass with one GEMM job
//   per stream, which adds the group's products into the stream's
//   device-resident accumulator (the last partial round of each XCD queue
//   cut into time pieces that add atomically, so no workgroup idles).
// Frames that arrive meanwhile are copied while the launch runs and form
// the next launch.  Latency of a frame = submit -> completion of the launch
// that integrated it, measured on the host clock; the engine keeps the
// records of the most recent kLatencyRecords frames.
//
// After a launch fails (a sticky error), the worker still retires every
// queued frame without launching, so that reads and closes, which wait for
// a stream's frames, return (with the error) and never free an accumulator
// a queued frame still names.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "chx_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

// Latency records kept (most recent frames; 16 MB at most).
constexpr size_t kLatencyRecords = 1u << 20;

struct Stream {
  int N, F, T, U;
  int ntiles, nb, KBf;    // layout blocks of one frame: int8 16-sample blocks (frame padded
                          // to whole rings), fp6 64-sample blocks (group padded instead)
  int64_t Np, NB;
  int64_t integ_frames;   // 0 = no cap
  int64_t submitted = 0, done = 0;
  int32_t *acc = nullptr;
  size_t acc_bytes = 0;
  size_t frame_bytes = 0;
  size_t plane_bytes = 0;  // planes of one frame alone (a group may pad to whole rings)
  double cmacs_per_frame = 0;
};

struct Frame {
  int64_t id, stream;
  int slot;
  bool copied;            // true: from host, through the slot buffer
  const uint8_t *src;     // device bytes [T][F][N]
  Clock::time_point t_submit;
};

int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// Layout blocks of k frames of a stream side by side: fp6 pads the group to
// whole rings (zero blocks at its end); int8 frames are ring-aligned already.
int64_t group_blocks(int fmt, const Stream &s, int64_t k) {
  return fmt == chx::kPlanesFp6 ? round_up(k * s.KBf, chx::kFp6Ring) : k * s.KBf;
}
int64_t group_stages(int fmt, const Stream &s, int64_t k) {
  const int64_t kb = group_blocks(fmt, s, k);
  return fmt == chx::kPlanesFp6 ? kb : kb / 4;
}
int64_t group_plane_bytes(int fmt, const Stream &s, int64_t k) {
  return chx::format_plane_bytes(fmt, s.F, s.Np, group_stages(fmt, s, k));
}
// Code-sum rows of k frames side by side: each frame's unpack runs, then
// the ring padding's (FP6).
int64_t group_rows(int fmt, const Stream &s, int64_t k) {
  const int64_t pad = group_blocks(fmt, s, k) - k * s.KBf;
  return k * chx::u4_job_runs(fmt, s.KBf, (int)s.Np) +
         (pad > 0 ? chx::u4_job_runs(fmt, pad, (int)s.Np) : 0);
}
int64_t group_partial_bytes(int fmt, const Stream &s, int64_t k) {
  return round_up(group_rows(fmt, s, k) * s.F * s.Np * 4, 256);
}
// Workspace of k frames of a stream: planes, code-sum rows and offset
// corrections.
int64_t group_bytes(int fmt, const Stream &s, int64_t k) {
  return group_plane_bytes(fmt, s, k) + group_partial_bytes(fmt, s, k) +
         round_up(2 * s.F * s.Np * 8, 256);
}
// One (array, plane, channel) region of a group must stay addressable by a
// 32-bit buffer descriptor.
bool group_fits_descriptor(int fmt, const Stream &s, int64_t k) {
  const int64_t kb = group_blocks(fmt, s, k);
  return (fmt == chx::kPlanesFp6 ? 32 : 16) * kb * s.Np <= 0x7FFFFFFFLL;
}

}  // namespace

struct chx_serve {
  chx_serve_config cfg{};
  int fmt = 1;  // chx::PlaneFormat of the method
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::thread worker;
  bool stop = false;
  int error = CHX_SUCCESS;  // sticky worker error
  bool fail_next_launch = false;  // host-forced failure (chx_serve_debug_fail_next_launch)

  hipStream_t copy_s = nullptr, compute_s = nullptr, io_s = nullptr;
  hipEvent_t done_ev = nullptr;
  std::vector<uint8_t *> slot_buf;
  std::vector<hipEvent_t> slot_ev;
  std::vector<int> free_slots;
  uint8_t *d_work = nullptr;
  // device tables of one launch (grow-only) and their pinned staging
  uint8_t *d_tab = nullptr, *h_tab = nullptr;
  size_t tab_cap = 0;
  int *d_counters = nullptr;
  int workgroups = 0;

  std::unordered_map<int64_t, Stream> streams;
  int64_t next_stream = 1, next_frame = 1;
  std::deque<Frame> pending;                 // submitted, not yet launched
  std::unordered_map<int64_t, int64_t> open_frames;  // frame id -> stream (not done)
  size_t bytes_in_use = 0;

  // statistics
  std::deque<std::pair<int64_t, double>> lat;  // (frame id, ms), completion order, bounded
  int64_t frames_submitted = 0, frames_done = 0, batches = 0, batch_frames = 0;
  double cmacs = 0;
  bool have_first = false;
  Clock::time_point t_first, t_last;
};

namespace chx {
namespace {

// Builds and runs one launch over `batch`; returns a status.
int run_launch(chx_serve *S, const std::vector<Frame> &batch,
               const std::vector<const Stream *> &st) {
  // Group frames by stream, in order of first appearance.
  std::vector<int64_t> order;
  std::unordered_map<int64_t, std::vector<int>> members;
  for (int i = 0; i < (int)batch.size(); i++) {
    auto &m = members[batch[i].stream];
    if (m.empty()) order.push_back(batch[i].stream);
    m.push_back(i);
  }
  std::vector<UnpackJob> ujobs;
  std::vector<CorrJob> cjobs;
  std::vector<GemmJob> gjobs;
  std::vector<ItemGroup> groups;
  int64_t units = 0, off = 0, cblocks = 0;
  for (int64_t sid : order) {
    const std::vector<int> &m = members[sid];
    const Stream &s = *st[m[0]];
    const int fmt = S->fmt;
    const int KBt = (int)group_blocks(fmt, s, (int64_t)m.size());
    uint8_t *planes = S->d_work + off;
    uint8_t *planes1 = fmt == kPlanesFp6 ? planes + 64LL * s.F * KBt * s.Np : nullptr;
    const int64_t pbytes = group_plane_bytes(fmt, s, (int64_t)m.size());
    uint32_t *partials = (uint32_t *)(planes + pbytes);
    int32_t *corr =
        (int32_t *)(planes + pbytes + group_partial_bytes(fmt, s, (int64_t)m.size()));
    off += group_bytes(fmt, s, (int64_t)m.size());
    int prow = 0;
    for (int k = 0; k < (int)m.size(); k++) {
      UnpackJob u;
      u.in = batch[m[k]].src;
      u.planes = planes;
      u.planes1 = planes1;
      u.row_bytes = (int64_t)s.F * s.N;
      u.t0 = 0;
      u.t_valid = s.T;
      u.unit0 = units;
      u.N = s.N;
      u.F = s.F;
      u.Np = (int)s.Np;
      u.KB = s.KBf;
      u.KB_dst = KBt;
      u.kb_dst0 = k * s.KBf;
      u.partials = partials;
      u.prow0 = prow;
      u.runs = u4_job_runs(fmt, u.KB, (int)s.Np);
      prow += u.runs;
      ujobs.push_back(u);
      units += unpack_units(fmt, u);
    }
    const int used = s.KBf * (int)m.size();
    if (KBt > used) {  // ring padding: zero blocks (reads nothing)
      UnpackJob z = ujobs.back();
      z.t_valid = 0;
      z.unit0 = units;
      z.KB = KBt - used;
      z.kb_dst0 = used;
      z.prow0 = prow;
      z.runs = u4_job_runs(fmt, z.KB, (int)s.Np);
      prow += z.runs;
      ujobs.push_back(z);
      units += unpack_units(fmt, z);
    }
    {
      if (prow != group_rows(fmt, s, (int64_t)m.size())) return CHX_ERR_HIP;
      CorrJob c;
      c.partials = partials;
      c.corr = corr;
      c.blk0 = cblocks;
      c.F = s.F;
      c.Np = (int)s.Np;
      c.nrun = prow;
      c.n = (int)((int64_t)m.size() * s.T);  // valid samples (padding adds nothing)
      cjobs.push_back(c);
      cblocks += ((int64_t)s.F * s.Np + 63) / 64;
    }
    const int nst = (int)group_stages(fmt, s, (int64_t)m.size());
    const int pieces = add_job_items(fmt, (int)gjobs.size(), s.F, s.ntiles, nst,
                                     S->cfg.max_item_stages, groups);
    GemmJob g;
    g.planes = planes;
    g.planes1 = planes1;
    g.vis = s.acc;
    g.NB = s.NB;
    g.F = s.F;
    g.Np = (int)s.Np;
    g.KB = KBt;
    g.ntiles = s.ntiles;
    g.nb = s.nb;
    g.U = s.U;
    g.mode = pieces > 1 ? 2 : 1;  // accumulate (atomically when a tile is cut in time)
    g.stages_per_split = nst;
    g.small = 0;
    g.corr = corr;
    gjobs.push_back(g);
  }
  std::vector<GemmItem> items;
  std::vector<int> qbound;
  spread_over_queues(groups, items, qbound);
  // accumulators take atomic adds: a queue's last partial round may be cut
  // in time
  if (S->cfg.max_item_stages == 0) cut_partial_rounds(S->fmt, items, qbound, gjobs, S->workgroups);
  int rc = check_launch_tables(S->fmt, ujobs, gjobs, items, qbound, 1, S->d_work,
                               (size_t)S->cfg.workspace_bytes);
  if (rc) return rc;

  // One table upload: [ujobs][cjobs][gjobs][items][qbound], 256-byte aligned
  // parts.
  const size_t a0 = 0, ac = round_up(a0 + ujobs.size() * sizeof(UnpackJob), 256),
               a1 = round_up(ac + cjobs.size() * sizeof(CorrJob), 256),
               a2 = round_up(a1 + gjobs.size() * sizeof(GemmJob), 256),
               a3 = round_up(a2 + items.size() * sizeof(GemmItem), 256),
               end = a3 + qbound.size() * sizeof(int);
  if (end > S->tab_cap) {
    if (S->d_tab) (void)hipFree(S->d_tab);
    if (S->h_tab) (void)hipHostFree(S->h_tab);
    S->d_tab = nullptr;
    S->h_tab = nullptr;
    S->tab_cap = 0;
    const size_t cap = round_up((int64_t)end * 2, 1 << 16);
    CHX_HIP(hipMalloc((void **)&S->d_tab, cap));
    CHX_HIP(hipHostMalloc((void **)&S->h_tab, cap, hipHostMallocDefault));
    S->tab_cap = cap;
  }
  memcpy(S->h_tab + a0, ujobs.data(), ujobs.size() * sizeof(UnpackJob));
  if (!cjobs.empty()) memcpy(S->h_tab + ac, cjobs.data(), cjobs.size() * sizeof(CorrJob));
  memcpy(S->h_tab + a1, gjobs.data(), gjobs.size() * sizeof(GemmJob));
  memcpy(S->h_tab + a2, items.data(), items.size() * sizeof(GemmItem));
  memcpy(S->h_tab + a3, qbound.data(), qbound.size() * sizeof(int));
  hipStream_t cs = S->compute_s;
  CHX_HIP(hipMemcpyAsync(S->d_tab, S->h_tab, end, hipMemcpyHostToDevice, cs));
  rc = launch_fill_u32((uint32_t *)S->d_counters, 0u, 8, cs);
  if (!rc)
    rc = launch_unpack_batch(S->fmt, (const UnpackJob *)(S->d_tab + a0), (int)ujobs.size(),
                             units, cs);
  if (!rc)
    rc = launch_u4_corrections_batch((const CorrJob *)(S->d_tab + ac), (int)cjobs.size(), cblocks,
                                     cs);
  if (!rc)
    rc = launch_gemm_batch(S->fmt, (const GemmJob *)(S->d_tab + a1),
                           (const GemmItem *)(S->d_tab + a2),
                           (const int *)(S->d_tab + a3), S->d_counters, S->workgroups, cs);
  if (rc) return rc;
  CHX_HIP(hipEventRecord(S->done_ev, cs));
  CHX_HIP(hipEventSynchronize(S->done_ev));  // the staging buffer is reused next launch
  return CHX_SUCCESS;
}

// Worker: forms launches from the queue until stopped and drained.
void worker_main(chx_serve *S) {
  (void)hipSetDevice(S->cfg.device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    S->cv_work.wait(lk, [&] { return S->stop || !S->pending.empty(); });
    if (S->pending.empty()) break;  // stop requested and drained
    // Wait (unlocked) for the oldest frame's copy; copies land in order.
    const Frame &f0 = S->pending.front();
    if (f0.copied) {
      hipEvent_t e = S->slot_ev[f0.slot];
      lk.unlock();
      hipError_t err = hipEventSynchronize(e);
      lk.lock();
      if (err != hipSuccess && S->error == CHX_SUCCESS) S->error = hip_status(err);
    }
    // Take the frames whose data is on the device, in submit order.
    std::vector<Frame> batch;
    std::vector<const Stream *> st;
    std::unordered_map<int64_t, int64_t> per_stream;
    int64_t ws = 0;
    const int maxf = S->cfg.max_batch_frames;
    while (!S->pending.empty() && (int)batch.size() < maxf) {
      const Frame &f = S->pending.front();
      if (f.copied && !batch.empty() && hipEventQuery(S->slot_ev[f.slot]) != hipSuccess) break;
      auto sit = S->streams.find(f.stream);
      if (sit == S->streams.end()) {  // cannot happen (closes drain first): retire it
        S->open_frames.erase(f.id);
        S->free_slots.push_back(f.slot);
        S->pending.pop_front();
        S->cv_done.notify_all();
        continue;
      }
      const Stream &s = sit->second;
      const int64_t k = per_stream[f.stream];  // frames of this stream taken so far
      const int64_t grow = group_bytes(S->fmt, s, k + 1) - (k ? group_bytes(S->fmt, s, k) : 0);
      if (ws + grow > S->cfg.workspace_bytes) break;
      if (!group_fits_descriptor(S->fmt, s, k + 1)) break;
      per_stream[f.stream] = k + 1;
      ws += grow;
      batch.push_back(f);
      st.push_back(&s);
      S->pending.pop_front();
    }
    // Streams cannot close while they have frames in flight, so the
    // Stream pointers stay valid while unlocked.
    int rc = S->error;
    if (rc == CHX_SUCCESS && S->fail_next_launch) {
      S->fail_next_launch = false;
      rc = CHX_ERR_HIP;
    }
    if (rc == CHX_SUCCESS && !batch.empty()) {
      lk.unlock();
      rc = run_launch(S, batch, st);
      lk.lock();
    }
    const Clock::time_point now = Clock::now();
    if (rc != CHX_SUCCESS && S->error == CHX_SUCCESS) S->error = rc;
    for (const Frame &f : batch) {
      auto sit = S->streams.find(f.stream);
      if (sit != S->streams.end()) sit->second.done++;
      S->open_frames.erase(f.id);
      S->free_slots.push_back(f.slot);
      if (rc == CHX_SUCCESS && sit != S->streams.end()) {
        const double ms = std::chrono::duration<double, std::milli>(now - f.t_submit).count();
        S->lat.emplace_back(f.id, ms);
        if (S->lat.size() > kLatencyRecords) S->lat.pop_front();
        S->frames_done++;
        S->cmacs += sit->second.cmacs_per_frame;
      }
    }
    if (!batch.empty()) {
      S->batches++;
      S->batch_frames += (int64_t)batch.size();
      S->t_last = now;
    }
    S->cv_done.notify_all();
  }
}

void free_serve(chx_serve *S) {
  for (auto &kv : S->streams)
    if (kv.second.acc) (void)hipFree(kv.second.acc);
  for (uint8_t *p : S->slot_buf)
    if (p) (void)hipFree(p);
  for (hipEvent_t e : S->slot_ev)
    if (e) (void)hipEventDestroy(e);
  if (S->done_ev) (void)hipEventDestroy(S->done_ev);
  if (S->d_work) (void)hipFree(S->d_work);
  if (S->d_tab) (void)hipFree(S->d_tab);
  if (S->h_tab) (void)hipHostFree(S->h_tab);
  if (S->d_counters) (void)hipFree(S->d_counters);
  if (S->copy_s) (void)hipStreamDestroy(S->copy_s);
  if (S->compute_s) (void)hipStreamDestroy(S->compute_s);
  if (S->io_s) (void)hipStreamDestroy(S->io_s);
  delete S;
}

// Waits (lock held) until every submitted frame of stream sid is done --
// also after a sticky error, since the worker then retires queued frames
// without launching them.
int drain_stream(chx_serve *S, std::unique_lock<std::mutex> &lk, int64_t sid) {
  if (!S->streams.count(sid)) return CHX_ERR_NOT_FOUND;
  S->cv_done.wait(lk, [&] {
    auto it = S->streams.find(sid);
    return it == S->streams.end() || it->second.done == it->second.submitted;
  });
  if (!S->streams.count(sid)) return CHX_ERR_NOT_FOUND;  // closed meanwhile
  return S->error;
}

int submit(chx_serve *S, int64_t sid, const uint8_t *ptr, bool host, int64_t *frame_id) {
  if (!S || !ptr) return CHX_ERR_INVALID_ARGUMENT;
  const Clock::time_point t = Clock::now();
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->error) return S->error;
  auto it = S->streams.find(sid);
  if (it == S->streams.end()) return CHX_ERR_NOT_FOUND;
  Stream &s = it->second;
  if (s.integ_frames > 0 && s.submitted >= s.integ_frames) return CHX_ERR_BUDGET;
  if (host && (int64_t)s.frame_bytes > S->cfg.slot_bytes) return CHX_ERR_BUDGET;
  if (S->free_slots.empty()) return CHX_ERR_BUSY;
  // Reserve the slot and count the frame (a close or read now waits for it).
  const int slot = S->free_slots.back();
  S->free_slots.pop_back();
  Frame f;
  f.id = S->next_frame++;
  f.stream = sid;
  f.slot = slot;
  f.copied = host;
  f.src = host ? S->slot_buf[slot] : ptr;
  f.t_submit = t;
  s.submitted++;
  S->frames_submitted++;
  S->open_frames[f.id] = sid;
  if (!S->have_first) {
    S->have_first = true;
    S->t_first = t;
  }
  const size_t bytes = s.frame_bytes;
  if (host) {
    // Queue the copy without the lock (a pageable source copies synchronously).
    lk.unlock();
    hipError_t e = hipSetDevice(S->cfg.device);
    if (e == hipSuccess)
      e = hipMemcpyAsync(S->slot_buf[slot], ptr, bytes, hipMemcpyHostToDevice, S->copy_s);
    if (e == hipSuccess) e = hipEventRecord(S->slot_ev[slot], S->copy_s);
    lk.lock();
    if (e != hipSuccess) {  // undo the reservation
      const int rc = hip_status(e);
      auto sit = S->streams.find(sid);
      if (sit != S->streams.end()) sit->second.submitted--;
      S->frames_submitted--;
      S->open_frames.erase(f.id);
      S->free_slots.push_back(slot);
      S->cv_done.notify_all();
      return rc;
    }
  }
  S->pending.push_back(f);
  if (frame_id) *frame_id = f.id;
  S->cv_work.notify_one();
  return CHX_SUCCESS;
}

}  // namespace
}  // namespace chx

using namespace chx;

extern "C" {

int chx_serve_create(const chx_serve_config *cfg, chx_serve **out) {
  if (!cfg || !out) return CHX_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  chx_serve_config c = *cfg;
  if (c.frame_slots < 1 || c.slot_bytes < 0 || c.workspace_bytes < 0 ||
      c.hbm_budget_bytes < 0 || c.max_batch_frames < 0 || c.max_item_stages < 0)
    return CHX_ERR_INVALID_ARGUMENT;
  if (c.max_batch_frames == 0 || c.max_batch_frames > c.frame_slots)
    c.max_batch_frames = c.frame_slots;
  if (c.workspace_bytes == 0) c.workspace_bytes = 2 * (int64_t)c.max_batch_frames * c.slot_bytes;
  if (c.method == 0) c.method = CHX_METHOD_FP6_MFMA;
  if (c.method != CHX_METHOD_INT8_MFMA && c.method != CHX_METHOD_FP6_MFMA)
    return CHX_ERR_INVALID_ARGUMENT;
  const int fmt = c.method == CHX_METHOD_FP6_MFMA ? kPlanesFp6 : kPlanesInt8;
  c.max_item_stages = (int32_t)round_up(c.max_item_stages, format_ring(fmt));
  const int64_t fixed = (int64_t)c.frame_slots * c.slot_bytes + c.workspace_bytes;
  if (c.hbm_budget_bytes > 0 && fixed > c.hbm_budget_bytes) return CHX_ERR_BUDGET;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CHX_ERR_NO_DEVICE;
  if (c.device < 0 || c.device >= ndev) return CHX_ERR_NO_DEVICE;
  CHX_HIP(hipSetDevice(c.device));
  int cus = 0;
  CHX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));

  chx_serve *S = new (std::nothrow) chx_serve;
  if (!S) return CHX_ERR_OUT_OF_MEMORY;
  S->cfg = c;
  S->fmt = fmt;
  S->workgroups = std::max(8, cus / 8 * 8);
  hipError_t e = hipStreamCreateWithFlags(&S->copy_s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->compute_s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->io_s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&S->done_ev, hipEventDisableTiming);
  if (e == hipSuccess && c.workspace_bytes > 0) e = hipMalloc((void **)&S->d_work, c.workspace_bytes);
  if (e == hipSuccess) e = hipMalloc((void **)&S->d_counters, 8 * sizeof(int));
  S->slot_buf.assign(c.frame_slots, nullptr);
  S->slot_ev.assign(c.frame_slots, nullptr);
  for (int k = 0; k < c.frame_slots && e == hipSuccess; k++) {
    if (c.slot_bytes > 0) e = hipMalloc((void **)&S->slot_buf[k], c.slot_bytes);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&S->slot_ev[k], hipEventDisableTiming);
    S->free_slots.push_back(c.frame_slots - 1 - k);
  }
  if (e != hipSuccess) {
    const int rc = hip_status(e);
    free_serve(S);
    return rc;
  }
  S->bytes_in_use = (size_t)fixed;
  S->worker = std::thread(worker_main, S);
  *out = S;
  return CHX_SUCCESS;
}

int chx_serve_destroy(chx_serve *S) {
  if (!S) return CHX_ERR_INVALID_ARGUMENT;
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  (void)hipSetDevice(S->cfg.device);
  (void)hipStreamSynchronize(S->copy_s);
  free_serve(S);
  return CHX_SUCCESS;
}

int chx_serve_open(chx_serve *S, int32_t N, int32_t F, int32_t T, int32_t U,
                   int64_t integration_frames, int64_t *stream_id) {
  if (!S || !stream_id) return CHX_ERR_INVALID_ARGUMENT;
  if ((U != 0 && U != 1) || integration_frames < 0) return CHX_ERR_INVALID_ARGUMENT;
  if (N < 32 || N % 32 || F < 1 || T < 1) return CHX_ERR_SHAPE;
  const int64_t nb = N / 32, NB = nb * (nb + 1) / 2, nt = (N + kMfmaTile - 1) / kMfmaTile;
  if ((int64_t)F * NB > 0x7FFFFFFFLL || (int64_t)F * nt * (nt + 1) / 2 > 0x7FFFFFFFLL)
    return CHX_ERR_SHAPE;
  Stream s;
  s.N = N;
  s.F = F;
  s.T = T;
  s.U = U;
  s.ntiles = (int)nt;
  s.nb = (int)nb;
  s.Np = nt * kMfmaTile;
  s.NB = NB;
  const int fmt = S->fmt;
  s.KBf = fmt == kPlanesFp6 ? (int)((T + kMfmaStage - 1) / kMfmaStage)
                            : (int)(round_up(T, (int64_t)kMfmaStage * kMfmaRing) / 16);
  s.integ_frames = integration_frames;
  s.acc_bytes = (size_t)F * NB * 2048 * sizeof(int32_t);
  s.frame_bytes = (size_t)T * F * N;
  s.plane_bytes = (size_t)group_bytes(fmt, s, 1);
  s.cmacs_per_frame = (double)F * T * ((double)N * (N + 1) / 2);
  std::lock_guard<std::mutex> g(S->mu);
  if (S->error) return S->error;
  if ((int64_t)s.plane_bytes > S->cfg.workspace_bytes) return CHX_ERR_BUDGET;
  if (!group_fits_descriptor(fmt, s, 1)) return CHX_ERR_SHAPE;
  if (S->cfg.hbm_budget_bytes > 0 &&
      (int64_t)(S->bytes_in_use + s.acc_bytes) > S->cfg.hbm_budget_bytes)
    return CHX_ERR_BUDGET;
  CHX_HIP(hipSetDevice(S->cfg.device));
  CHX_HIP(hipMalloc((void **)&s.acc, s.acc_bytes));
  int rc = launch_fill_u32((uint32_t *)s.acc, 0u, (int64_t)s.acc_bytes / 4, S->io_s);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(S->io_s);
    if (e != hipSuccess) rc = hip_status(e);
  }
  if (rc) {
    (void)hipFree(s.acc);
    return rc;
  }
  S->bytes_in_use += s.acc_bytes;
  *stream_id = S->next_stream++;
  S->streams.emplace(*stream_id, s);
  return CHX_SUCCESS;
}

int chx_serve_submit(chx_serve *S, int64_t sid, const uint8_t *h_frame, int64_t *frame_id) {
  return submit(S, sid, h_frame, true, frame_id);
}

int chx_serve_submit_device(chx_serve *S, int64_t sid, const uint8_t *d_frame,
                            int64_t *frame_id) {
  return submit(S, sid, d_frame, false, frame_id);
}

int chx_serve_wait_frame(chx_serve *S, int64_t frame_id) {
  if (!S) return CHX_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(S->mu);
  if (frame_id < 1 || frame_id >= S->next_frame) return CHX_ERR_NOT_FOUND;
  S->cv_done.wait(lk, [&] { return !S->open_frames.count(frame_id) || S->error; });
  return S->error;
}

static int read_acc(chx_serve *S, int64_t sid, int32_t *dst, hipMemcpyKind kind) {
  if (!S || !dst) return CHX_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(S->mu);
  int rc = drain_stream(S, lk, sid);
  if (rc) return rc;
  auto sit = S->streams.find(sid);
  if (sit == S->streams.end()) return CHX_ERR_NOT_FOUND;
  const Stream &s = sit->second;
  // The lock is held: no frame of this stream can be submitted meanwhile.
  CHX_HIP(hipSetDevice(S->cfg.device));
  CHX_HIP(hipMemcpyAsync(dst, s.acc, s.acc_bytes, kind, S->io_s));
  CHX_HIP(hipStreamSynchronize(S->io_s));
  return CHX_SUCCESS;
}

int chx_serve_read(chx_serve *S, int64_t sid, int32_t *h_vis) {
  return read_acc(S, sid, h_vis, hipMemcpyDeviceToHost);
}

int chx_serve_read_device(chx_serve *S, int64_t sid, int32_t *d_vis) {
  return read_acc(S, sid, d_vis, hipMemcpyDeviceToDevice);
}

int chx_serve_close(chx_serve *S, int64_t sid) {
  if (!S) return CHX_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(S->mu);
  int rc = drain_stream(S, lk, sid);
  if (rc == CHX_ERR_NOT_FOUND) return rc;
  auto sit = S->streams.find(sid);
  if (sit == S->streams.end()) return CHX_ERR_NOT_FOUND;
  Stream &s = sit->second;
  (void)hipSetDevice(S->cfg.device);
  if (s.acc) (void)hipFree(s.acc);
  S->bytes_in_use -= s.acc_bytes;
  S->streams.erase(sid);
  return rc;
}

int chx_serve_get_stats(chx_serve *S, chx_serve_stats *o) {
  if (!S || !o) return CHX_ERR_INVALID_ARGUMENT;
  std::vector<double> v;
  int err;
  memset(o, 0, sizeof(*o));
  {
    // copy under the lock; sort outside it, so the worker and submitters
    // are not held up
    std::lock_guard<std::mutex> g(S->mu);
    o->frames_submitted = S->frames_submitted;
    o->frames_done = S->frames_done;
    o->batches = S->batches;
    o->streams_open = (int64_t)S->streams.size();
    o->bytes_in_use = (int64_t)S->bytes_in_use;
    o->cmacs = S->cmacs;
    o->mean_batch_frames = S->batches ? (double)S->batch_frames / S->batches : 0.0;
    if (S->have_first && S->frames_done)
      o->wall_seconds = std::chrono::duration<double>(S->t_last - S->t_first).count();
    v.reserve(S->lat.size());
    for (auto &r : S->lat) v.push_back(r.second);
    err = S->error;
  }
  if (!v.empty()) {
    double sum = 0;
    for (double x : v) sum += x;
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) {  // nearest rank
      size_t k = (size_t)(p * (double)v.size() + 0.999999);
      if (k < 1) k = 1;
      if (k > v.size()) k = v.size();
      return v[k - 1];
    };
    o->latency_p50_ms = pct(0.50);
    o->latency_p99_ms = pct(0.99);
    o->latency_max_ms = v.back();
    o->latency_mean_ms = sum / (double)v.size();
  }
  return err;
}

int chx_serve_reset_stats(chx_serve *S) {
  if (!S) return CHX_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> g(S->mu);
  S->lat.clear();
  S->frames_submitted = S->frames_done = S->batches = S->batch_frames = 0;
  S->cmacs = 0;
  S->have_first = false;
  return CHX_SUCCESS;
}

int chx_serve_latencies(chx_serve *S, int64_t *ids, double *ms, int64_t max, int64_t *count) {
  if (!S || !count || max < 0) return CHX_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> g(S->mu);
  const int64_t n = (int64_t)S->lat.size();
  *count = n;
  for (int64_t k = 0; k < n && k < max; k++) {
    if (ids) ids[k] = S->lat[k].first;
    if (ms) ms[k] = S->lat[k].second;
  }
  return CHX_SUCCESS;
}

// Test hook, not in the public header: the next launch the worker forms
// fails (sticky CHX_ERR_HIP) before anything is launched.
int chx_serve_debug_fail_next_launch(chx_serve *S) {
  if (!S) return CHX_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> g(S->mu);
  S->fail_next_launch = true;
  return CHX_SUCCESS;
}

}  // extern "C"
