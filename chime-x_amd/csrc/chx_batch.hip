// chx_batch.hip -- many matrix-core correlations (int8 or FP6 operands) of
// uneven shape (elements, channels, integration length) in one captured
// hipGraph.
//
// Plan (host, at creation):
//  * time is cut into rounds of chunk_samples samples, the same for every
//    job, sized so that one round's planes of all jobs still active fit the
//    workspace budget; a job takes part in the rounds its length reaches;
//  * per round: one batched unpack launch over every active job's chunk, then
//    one persistent GEMM launch over work items (job, channel, 128x256 tile
//    -- or, FP6, 128x128 for arrays up to small_tile_elements -- and stage
//    range); a job's tiles are cut in time when max_item_stages asks
//    or the chunk exceeds the item cap (format_max_stages), in which case
//    their partial sums add atomically; FP6 rounds also launch the jobs'
//    offset corrections (code sums from the unpack) between the two;
//  * items are grouped per (job, channel) and the groups spread over 8
//    per-XCD queues by longest-processing-time first, so a channel's tiles
//    share one XCD's L2; inside a queue, long items run first.
// Run: counters and atomically-filled outputs are cleared with fill kernels
// (no memset nodes), then the rounds' launches, all on one stream.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "chx_internal.h"

struct chx_batch {
  int device = 0;
  int fmt = 1;  // chx::PlaneFormat
  std::vector<chx_job_desc> jobs;
  chx_batch_tuning tune{};
  int64_t chunk = 0;
  int rounds = 0;
  int64_t nitems = 0, cmacs = 0;
  struct Round {
    int ujob0 = 0, nujobs = 0;
    int64_t units = 0;
    int cjob0 = 0, ncjobs = 0;  // FP6: offset-correction jobs
    int64_t cblocks = 0;
  };
  std::vector<Round> round;
  struct Fill {
    int32_t *p;
    int64_t n;
  };
  std::vector<Fill> zero;  // outputs cleared before the first round
  uint8_t *d_work = nullptr;
  size_t work_bytes = 0;
  chx::GemmJob *d_gjobs = nullptr;
  chx::GemmItem *d_items = nullptr;
  chx::UnpackJob *d_ujobs = nullptr;
  chx::CorrJob *d_cjobs = nullptr;
  int *d_qbound = nullptr;    // [rounds][9]
  int *d_counters = nullptr;  // [rounds][8]
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace chx {

int add_job_items(int fmt, int gj, int F, int ntiles, int nst, int max_st,
                  std::vector<ItemGroup> &groups, bool small) {
  const int per = format_item_stages(fmt, nst, max_st);
  const int pieces = (nst + per - 1) / per;
  const int tiles = (int)(small ? small_tiles(ntiles) : format_tiles(fmt, ntiles));
  for (int f = 0; f < F; f++) {
    ItemGroup G;
    G.work = 0;
    G.items.reserve((size_t)tiles * pieces);
    for (int t = 0; t < tiles; t++)
      for (int k = 0; k < pieces; k++) {
        const int s0 = k * per, ns = per < nst - s0 ? per : nst - s0;
        G.items.push_back({gj, f, t, s0, ns});
        // + a flush/prologue allowance; a small tile does half the MFMAs
        G.work += (small ? ns / 2 : ns) + 8;
      }
    groups.push_back(std::move(G));
  }
  return pieces;
}

int cut_partial_rounds(int fmt, std::vector<GemmItem> &items, std::vector<int> &qbound,
                       std::vector<GemmJob> &gjobs, int nwg) {
  if (qbound.size() != 9 || nwg < 8) return 0;
  const int wq = nwg / 8;  // workgroups that start on each queue
  const int ring = format_ring(fmt);
  std::vector<int> tail_job(gjobs.size(), -1);
  std::vector<GemmItem> out;
  std::vector<int> qb2(1, 0);
  int added = 0;
  for (int q = 0; q < 8; q++) {
    const int b0 = qbound[q], b1 = qbound[q + 1];
    const int n = b1 - b0, r = wq > 0 ? n % wq : 0;
    bool equal = n > 0;
    for (int k = b0; k < b1 && equal; k++)
      equal = items[k].nstages == items[b0].nstages && gjobs[items[k].job].mode != 2;
    int best_p = 1;
    if (equal && n > wq && r > 0) {
      const int len = items[b0].nstages;
      double best = 1.0;  // the partial round's length in whole items
      for (int p = 2; p <= 8 && len / p >= 16; p++) {
        const double rounds = (double)((r * p + wq - 1) / wq) / p;
        if (rounds < best - 1e-9) {
          best = rounds;
          best_p = p;
        }
      }
    }
    if (best_p == 1) {
      out.insert(out.end(), items.begin() + b0, items.begin() + b1);
    } else {
      // the queue's last r items become pieces, owned by copies of their
      // jobs that add atomically, after the whole items
      out.insert(out.end(), items.begin() + b0, items.begin() + (b1 - r));
      const int len = items[b0].nstages;
      const int per = ((len + best_p - 1) / best_p + ring - 1) / ring * ring;
      for (int k = b1 - r; k < b1; k++) {
        const GemmItem &w = items[k];
        if (tail_job[w.job] < 0) {
          GemmJob j = gjobs[w.job];
          j.mode = 2;
          tail_job[w.job] = (int)gjobs.size();
          gjobs.push_back(j);
          added++;
        }
        for (int s0 = 0; s0 < w.nstages; s0 += per) {
          const int ns = per < w.nstages - s0 ? per : w.nstages - s0;
          out.push_back({tail_job[w.job], w.f, w.tile, w.stage0 + s0, ns});
        }
      }
    }
    qb2.push_back((int)out.size());
  }
  items.swap(out);
  qbound.swap(qb2);
  return added;
}

void spread_over_queues(std::vector<ItemGroup> &groups, std::vector<GemmItem> &items,
                        std::vector<int> &qbound) {
  // Longest-processing-time first: heaviest group to the lightest queue.
  std::stable_sort(groups.begin(), groups.end(),
                   [](const ItemGroup &a, const ItemGroup &c) { return a.work > c.work; });
  std::vector<std::vector<GemmItem>> q(8);
  int64_t load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (ItemGroup &G : groups) {
    const int k = (int)(std::min_element(load, load + 8) - load);
    load[k] += G.work;
    q[k].insert(q[k].end(), G.items.begin(), G.items.end());
  }
  for (int k = 0; k < 8; k++) {
    std::stable_sort(q[k].begin(), q[k].end(),
                     [](const GemmItem &a, const GemmItem &c) { return a.nstages > c.nstages; });
    qbound.push_back((int)items.size());
    items.insert(items.end(), q[k].begin(), q[k].end());
  }
  qbound.push_back((int)items.size());
}

int check_launch_tables(int fmt, const std::vector<UnpackJob> &ujobs,
                        const std::vector<GemmJob> &gjobs, const std::vector<GemmItem> &items,
                        const std::vector<int> &qbound, int rounds, const uint8_t *base,
                        size_t work_bytes) {
  const uint8_t *end = base + work_bytes;
  const int ring = format_ring(fmt);
  const int spb = fmt == kPlanesFp6 ? 64 : 16;  // samples per layout block
  for (const UnpackJob &u : ujobs) {
    const int64_t stages = fmt == kPlanesFp6 ? u.KB_dst : u.KB_dst / 4;
    const int64_t bytes = format_plane_bytes(fmt, u.F, u.Np, stages);
    if (!u.in || u.planes < base || u.planes + bytes > end || u.t_valid < 0 ||
        u.t_valid > (int64_t)spb * u.KB || u.Np % kMfmaTile || u.N > u.Np || u.kb_dst0 < 0 ||
        u.kb_dst0 + u.KB > u.KB_dst)
      return CHX_ERR_HIP;
    if (fmt == kPlanesFp6 && u.planes1 != u.planes + 64LL * u.F * u.KB_dst * u.Np)
      return CHX_ERR_HIP;
    // code sums: runs of at most 16 units (16-bit sums), inside their
    // region; the sample rows of one unit fit 32-bit offsets
    const int64_t unit_rows = fmt == kPlanesFp6 ? 32LL * fp6_halves_per_unit(u.Np) : 16;
    if (!u.partials || u.prow0 < 0 || u.runs < 1 ||
        (int64_t)u.runs * 16 < u4_units(fmt, u.KB, u.Np) ||
        u.row_bytes * (unit_rows + 1) > 0x7FFFFFFFLL)
      return CHX_ERR_HIP;
    const int64_t rows = (int64_t)(u.prow0 + u.runs) * u.F * u.Np * 4;
    if ((const uint8_t *)u.partials < base || (const uint8_t *)u.partials + rows > end)
      return CHX_ERR_HIP;
  }
  for (const GemmJob &g : gjobs) {
    const int stages = fmt == kPlanesFp6 ? g.KB : g.KB / 4;
    const int64_t bytes = format_plane_bytes(fmt, g.F, g.Np, stages);
    const int64_t region = fmt == kPlanesFp6 ? 32LL * g.KB * g.Np : 16LL * g.KB * g.Np;
    if (!g.vis || g.planes < base || g.planes + bytes > end || stages % ring ||
        format_kb(fmt, stages) != g.KB || g.ntiles * kMfmaTile != g.Np || region > 0x7FFFFFFFLL)
      return CHX_ERR_HIP;
    if (fmt == kPlanesFp6 && g.planes1 != g.planes + 64LL * g.F * g.KB * g.Np) return CHX_ERR_HIP;
    if (!g.corr || (const uint8_t *)g.corr < base ||
        (const uint8_t *)g.corr + 16LL * g.F * g.Np > end)
      return CHX_ERR_HIP;
    if (g.small && fmt != kPlanesFp6) return CHX_ERR_HIP;
  }
  for (const GemmItem &w : items) {
    if (w.job < 0 || w.job >= (int)gjobs.size()) return CHX_ERR_HIP;
    const GemmJob &g = gjobs[w.job];
    const int stages = fmt == kPlanesFp6 ? g.KB : g.KB / 4;
    if (w.f < 0 || w.f >= g.F || w.tile < 0 || w.tile >= job_tiles(fmt, g) ||
        w.stage0 < 0 || w.nstages < ring || w.nstages % ring || w.stage0 + w.nstages > stages ||
        w.nstages > format_max_stages(fmt))
      return CHX_ERR_HIP;
  }
  if ((int)qbound.size() != 9 * rounds || qbound.empty() ||
      qbound.back() != (int)items.size() || qbound[0] != 0)
    return CHX_ERR_HIP;
  for (size_t k = 1; k < qbound.size(); k++)
    if (qbound[k] < qbound[k - 1] || (k % 9 == 0 && qbound[k] != qbound[k - 1]))
      return CHX_ERR_HIP;
  return CHX_SUCCESS;
}

int launch_unpack_batch(int fmt, const UnpackJob *d_jobs, int njobs, int64_t units,
                        hipStream_t s) {
  return fmt == kPlanesFp6 ? launch_unpack_batch_fp6(d_jobs, njobs, units, s)
                           : launch_unpack_batch_int8(d_jobs, njobs, units, s);
}

int launch_gemm_batch(int fmt, const GemmJob *d_jobs, const GemmItem *d_items,
                      const int *d_qbound, int *d_counters, int nwg, hipStream_t s) {
  return fmt == kPlanesFp6
             ? launch_gemm_batch_fp6(d_jobs, d_items, d_qbound, d_counters, nwg, s)
             : launch_gemm_batch_int8(d_jobs, d_items, d_qbound, d_counters, nwg, s);
}

namespace {

constexpr int64_t kDefaultBudget = 16LL << 30;

int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

int validate_job(const chx_job_desc &j) {
  if (!j.d_input || !j.d_vis || j.reserved != 0) return CHX_ERR_INVALID_ARGUMENT;
  if (j.upper_triangle_convention != 0 && j.upper_triangle_convention != 1)
    return CHX_ERR_INVALID_ARGUMENT;
  if (j.num_elements < 32 || j.num_elements % 32 != 0) return CHX_ERR_SHAPE;
  if (j.num_freq < 1 || j.time_steps < 1) return CHX_ERR_SHAPE;
  const int64_t nb = j.num_elements / 32, nt = (j.num_elements + kMfmaTile - 1) / kMfmaTile;
  if ((int64_t)j.num_freq * nb * (nb + 1) / 2 > 0x7FFFFFFFLL) return CHX_ERR_SHAPE;
  if ((int64_t)j.num_freq * nt * (nt + 1) / 2 > 0x7FFFFFFFLL) return CHX_ERR_SHAPE;
  if (j.time_steps > (INT64_MAX / 2) / ((int64_t)j.num_freq * j.num_elements))
    return CHX_ERR_SHAPE;
  return CHX_SUCCESS;
}

void free_batch(chx_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  if (b->d_work) (void)hipFree(b->d_work);
  if (b->d_gjobs) (void)hipFree(b->d_gjobs);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_ujobs) (void)hipFree(b->d_ujobs);
  if (b->d_cjobs) (void)hipFree(b->d_cjobs);
  if (b->d_qbound) (void)hipFree(b->d_qbound);
  if (b->d_counters) (void)hipFree(b->d_counters);
  delete b;
}

template <class T>
hipError_t upload(T **dst, const std::vector<T> &v) {
  if (v.empty()) return hipSuccess;
  hipError_t e = hipMalloc((void **)dst, sizeof(T) * v.size());
  if (e == hipSuccess) e = hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
  return e;
}

// Host tables of one batch (uploaded once at creation).
struct Tables {
  std::vector<UnpackJob> ujobs;
  std::vector<CorrJob> cjobs;
  std::vector<GemmJob> gjobs;
  std::vector<GemmItem> items;
  std::vector<int> qbound;
};

// Workspace offsets of one job: its planes, and (FP6) its code-sum partial
// rows and offset corrections.
struct JobOffsets {
  int64_t planes = 0, partials = 0, corr = 0;
};

// Chunk length, round count, per-job workspace offsets and total size.
int plan_sizes(chx_batch *b, std::vector<JobOffsets> &off) {
  const int nj = (int)b->jobs.size();
  std::vector<int64_t> Np(nj);
  off.assign(nj, JobOffsets{});
  b->cmacs = 0;
  const int64_t unit = (int64_t)kMfmaStage * format_ring(b->fmt);  // samples per ring
  int64_t per_unit = 0, maxT = 0;  // plane bytes of one ring of all jobs
  for (int j = 0; j < nj; j++) {
    const chx_job_desc &d = b->jobs[j];
    Np[j] = round_up(d.num_elements, kMfmaTile);
    per_unit += format_plane_bytes(b->fmt, d.num_freq, Np[j], format_ring(b->fmt));
    maxT = std::max(maxT, d.time_steps);
    const int64_t N = d.num_elements;
    b->cmacs += (int64_t)d.num_freq * d.time_steps * (N * (N + 1) / 2);
  }
  // Chunk: as long as the budget allows (whole rings), no longer than the
  // longest job, and short enough that one (plane, channel) region stays
  // addressable by a 32-bit buffer descriptor (int8: 2 bytes per element and
  // sample, i.e. chunk * Np; fp6: chunk * Np / 2 per array).
  int64_t chunk = b->tune.workspace_bytes / per_unit * unit;
  if (chunk < unit) return CHX_ERR_BUDGET;
  chunk = std::min(chunk, round_up(maxT, unit));
  for (int j = 0; j < nj; j++) {
    const int64_t lim = b->fmt == kPlanesFp6 ? 0x7FFFFFFFLL / (Np[j] / 2) : 0x7FFFFFFFLL / Np[j];
    chunk = std::min<int64_t>(chunk, lim / unit * unit);
  }
  // Layout of a chunk: every job's planes, then its code-sum partial rows
  // (the most any of its rounds writes) and offset corrections.  The chunk
  // shrinks until the whole layout fits the budget.
  auto layout = [&](int64_t c) {
    int64_t ws = 0;
    for (int j = 0; j < nj; j++) {
      off[j].planes = ws;
      const int64_t stages = std::min(c, round_up(b->jobs[j].time_steps, unit)) / kMfmaStage;
      ws += format_plane_bytes(b->fmt, b->jobs[j].num_freq, Np[j], stages);
    }
    for (int j = 0; j < nj; j++) {
      const int64_t F = b->jobs[j].num_freq, T = b->jobs[j].time_steps;
      int64_t rows = 0;
      for (int64_t t0 = 0; t0 < T; t0 += c) {
        const int64_t stages = round_up(std::min(c, T - t0), unit) / kMfmaStage;
        rows = std::max<int64_t>(rows,
                                 u4_job_runs(b->fmt, format_kb(b->fmt, (int)stages), (int)Np[j]));
      }
      off[j].partials = ws;
      ws += round_up(rows * F * Np[j] * 4, 256);
      off[j].corr = ws;
      ws += round_up(2 * F * Np[j] * 8, 256);
    }
    return ws;
  };
  const int64_t budget = b->tune.workspace_bytes;
  int64_t ws = layout(chunk);
  while (ws > budget && chunk > unit) {
    int64_t c = (int64_t)((double)chunk * budget / ws) / unit * unit;
    if (c >= chunk) c = chunk - unit;
    chunk = std::max(c, unit);
    ws = layout(chunk);
  }
  if (ws > budget) return CHX_ERR_BUDGET;
  b->chunk = chunk;
  b->rounds = (int)((maxT + chunk - 1) / chunk);
  b->work_bytes = (size_t)ws;
  return CHX_SUCCESS;
}

// Every round's unpack jobs, GEMM jobs, items and queues, with plane
// pointers into the workspace at `base`.
int plan_tables(chx_batch *b, uint8_t *base, const std::vector<JobOffsets> &off, Tables &t) {
  const int nj = (int)b->jobs.size();
  const int64_t chunk = b->chunk;
  std::vector<UnpackJob> &ujobs = t.ujobs;
  std::vector<CorrJob> &cjobs = t.cjobs;
  std::vector<GemmJob> &gjobs = t.gjobs;
  std::vector<GemmItem> &items = t.items;
  std::vector<int> &qbound = t.qbound;
  b->round.clear();
  b->zero.clear();
  const int max_st = b->tune.max_item_stages;  // multiple of the ring, 0 = none
  const int64_t unit = (int64_t)kMfmaStage * format_ring(b->fmt);
  for (int r = 0; r < b->rounds; r++) {
    chx_batch::Round R;
    R.ujob0 = (int)ujobs.size();
    R.cjob0 = (int)cjobs.size();
    std::vector<ItemGroup> groups;
    for (int j = 0; j < nj; j++) {
      const chx_job_desc &d = b->jobs[j];
      const int64_t t0 = (int64_t)r * chunk;
      if (t0 >= d.time_steps) continue;
      const int64_t tv = std::min(chunk, d.time_steps - t0);
      const int nst = (int)(round_up(tv, unit) / kMfmaStage);
      const int KB = format_kb(b->fmt, nst);
      const int F = d.num_freq, N = d.num_elements;
      const int nt = (N + kMfmaTile - 1) / kMfmaTile;
      const int64_t Npj = round_up(N, kMfmaTile);
      uint8_t *planes = base + off[j].planes;
      uint32_t *partials = (uint32_t *)(base + off[j].partials);
      int32_t *corr = (int32_t *)(base + off[j].corr);
      uint8_t *planes1 = b->fmt == kPlanesFp6 ? planes + 64LL * F * KB * Npj : nullptr;

      UnpackJob u;
      u.in = d.d_input;
      u.planes = planes;
      u.planes1 = planes1;
      u.row_bytes = (int64_t)F * N;
      u.t0 = t0;
      u.t_valid = tv;
      u.unit0 = R.units;
      u.N = N;
      u.F = F;
      u.Np = (int)Npj;
      u.KB = KB;
      u.KB_dst = KB;
      u.kb_dst0 = 0;
      u.partials = partials;
      u.prow0 = 0;
      u.runs = u4_job_runs(b->fmt, KB, (int)Npj);
      ujobs.push_back(u);
      R.units += unpack_units(b->fmt, u);
      {
        CorrJob c;
        c.partials = partials;
        c.corr = corr;
        c.blk0 = R.cblocks;
        c.F = F;
        c.Np = (int)Npj;
        c.nrun = u.runs;
        c.n = (int)tv;
        cjobs.push_back(c);
        R.cblocks += ((int64_t)F * Npj + 63) / 64;
      }

      // Items of this job: whole chunk per tile, or max_st-stage pieces;
      // 128 x 128 tiles for arrays of at most small_tile_elements (FP6).
      const bool small = b->fmt == kPlanesFp6 && N <= b->tune.small_tile_elements;
      const int per = format_item_stages(b->fmt, nst, max_st);
      const int pieces =
          add_job_items(b->fmt, (int)gjobs.size(), F, nt, nst, max_st, groups, small);
      GemmJob g;
      g.planes = planes;
      g.planes1 = planes1;
      g.vis = d.d_vis;
      g.NB = chx_num_blocks(N);
      g.F = F;
      g.Np = (int)Npj;
      g.KB = KB;
      g.ntiles = nt;
      g.nb = N / 32;
      g.U = d.upper_triangle_convention;
      g.mode = pieces > 1 ? 2 : (r == 0 ? 0 : 1);
      g.stages_per_split = per;
      g.small = small ? 1 : 0;
      g.corr = corr;
      if (pieces > 1 && r == 0) b->zero.push_back({d.d_vis, (int64_t)F * g.NB * 2048});
      gjobs.push_back(g);
    }
    R.nujobs = (int)ujobs.size() - R.ujob0;
    R.ncjobs = (int)cjobs.size() - R.cjob0;
    spread_over_queues(groups, items, qbound);
    if (items.size() > 0x7FFFFFFFu) return CHX_ERR_SHAPE;
    b->round.push_back(R);
  }
  b->nitems = (int64_t)items.size();
  return CHX_SUCCESS;
}

int check_tables(const chx_batch *b, const uint8_t *base, const Tables &t) {
  return check_launch_tables(b->fmt, t.ujobs, t.gjobs, t.items, t.qbound, b->rounds, base,
                             b->work_bytes);
}

// Allocates the workspace, builds and checks the tables, uploads them.
int plan_batch(chx_batch *b) {
  std::vector<JobOffsets> off;
  int rc = plan_sizes(b, off);
  if (rc) return rc;
  CHX_HIP(hipMalloc((void **)&b->d_work, b->work_bytes));
  Tables t;
  rc = plan_tables(b, b->d_work, off, t);
  if (!rc) rc = check_tables(b, b->d_work, t);
  if (rc) return rc;
  hipError_t e = upload(&b->d_ujobs, t.ujobs);
  if (e == hipSuccess) e = upload(&b->d_cjobs, t.cjobs);
  if (e == hipSuccess) e = upload(&b->d_gjobs, t.gjobs);
  if (e == hipSuccess) e = upload(&b->d_items, t.items);
  if (e == hipSuccess) e = upload(&b->d_qbound, t.qbound);
  if (e == hipSuccess) e = hipMalloc((void **)&b->d_counters, sizeof(int) * 8 * b->rounds);
  if (e != hipSuccess) return hip_status(e);
  return CHX_SUCCESS;
}

// The run's launches, in order, on stream s.
int enqueue(const chx_batch *b, hipStream_t s) {
  int rc = launch_fill_u32((uint32_t *)b->d_counters, 0u, 8LL * b->rounds, s);
  for (const chx_batch::Fill &z : b->zero)
    if (!rc) rc = launch_fill_u32((uint32_t *)z.p, 0u, z.n, s);
  for (int r = 0; r < b->rounds && !rc; r++) {
    const chx_batch::Round &R = b->round[r];
    rc = launch_unpack_batch(b->fmt, b->d_ujobs + R.ujob0, R.nujobs, R.units, s);
    if (!rc)
      rc = launch_u4_corrections_batch(b->d_cjobs + R.cjob0, R.ncjobs, R.cblocks, s);
    if (!rc)
      rc = launch_gemm_batch(b->fmt, b->d_gjobs, b->d_items, b->d_qbound + 9 * r,
                             b->d_counters + 8 * r, b->tune.workgroups, s);
  }
  return rc;
}

// Validation shared by chx_batch_create and chx_batch_plan (host only).
int prepare(const chx_job_desc *jobs, int32_t num_jobs, const chx_batch_tuning *tuning,
            chx_batch_tuning &t, int &fmt) {
  if (!jobs || num_jobs < 1) return CHX_ERR_INVALID_ARGUMENT;
  for (int j = 0; j < num_jobs; j++) {
    int rc = validate_job(jobs[j]);
    if (rc) return rc;
  }
  t = tuning ? *tuning : chx_batch_tuning{};
  if (t.max_item_stages < 0 || t.workgroups < 0 || t.workgroups % 8 != 0 ||
      t.workspace_bytes < 0 || t.small_tile_elements < 0)
    return CHX_ERR_INVALID_ARGUMENT;
  if (t.method == 0) t.method = CHX_METHOD_FP6_MFMA;
  if (t.method != CHX_METHOD_INT8_MFMA && t.method != CHX_METHOD_FP6_MFMA)
    return CHX_ERR_INVALID_ARGUMENT;
  fmt = t.method == CHX_METHOD_FP6_MFMA ? kPlanesFp6 : kPlanesInt8;
  if (fmt != kPlanesFp6) t.small_tile_elements = 0;  // 128 x 128 tiles: FP6 only
  if (t.workspace_bytes == 0) t.workspace_bytes = kDefaultBudget;
  t.max_item_stages = (int32_t)round_up(t.max_item_stages, format_ring(fmt));
  // The smallest round (one ring of every job) must fit the budget.
  int64_t per_unit = 0;
  for (int j = 0; j < num_jobs; j++)
    per_unit += format_plane_bytes(fmt, jobs[j].num_freq,
                                   round_up(jobs[j].num_elements, kMfmaTile), format_ring(fmt));
  if (per_unit > t.workspace_bytes) return CHX_ERR_BUDGET;
  return CHX_SUCCESS;
}

}  // namespace
}  // namespace chx

using namespace chx;

extern "C" {

int chx_batch_create(int32_t device, const chx_job_desc *jobs, int32_t num_jobs,
                     const chx_batch_tuning *tuning, chx_batch **out) {
  if (!out) return CHX_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  chx_batch_tuning t;
  int fmt = kPlanesFp6;
  int rc = prepare(jobs, num_jobs, tuning, t, fmt);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CHX_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return CHX_ERR_NO_DEVICE;
  CHX_HIP(hipSetDevice(device));
  if (t.workgroups == 0) {
    int cus = 0;
    CHX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    t.workgroups = std::max(8, cus / 8 * 8);
  }
  chx_batch *b = new (std::nothrow) chx_batch;
  if (!b) return CHX_ERR_OUT_OF_MEMORY;
  b->device = device;
  b->fmt = fmt;
  b->jobs.assign(jobs, jobs + num_jobs);
  b->tune = t;
  rc = plan_batch(b);
  if (rc) {
    free_batch(b);
    return rc;
  }
  *out = b;
  return CHX_SUCCESS;
}

int chx_batch_plan(const chx_job_desc *jobs, int32_t num_jobs, const chx_batch_tuning *tuning,
                   chx_batch_info *info) {
  if (!info) return CHX_ERR_INVALID_ARGUMENT;
  chx_batch_tuning t;
  int fmt = kPlanesFp6;
  int rc = prepare(jobs, num_jobs, tuning, t, fmt);
  if (rc) return rc;
  chx_batch b;
  b.fmt = fmt;
  b.jobs.assign(jobs, jobs + num_jobs);
  b.tune = t;
  std::vector<JobOffsets> off;
  rc = plan_sizes(&b, off);
  if (rc) return rc;
  // Plane pointers relative to a stand-in base (never dereferenced).
  uint8_t *base = reinterpret_cast<uint8_t *>((uintptr_t)1 << 44);
  Tables tab;
  rc = plan_tables(&b, base, off, tab);
  if (!rc) rc = check_tables(&b, base, tab);
  if (rc) return rc;
  return chx_batch_get_info(&b, info);
}

int chx_batch_run(chx_batch *b, int32_t flags, void *stream) {
  if (!b || (flags & ~CHX_BATCH_EAGER)) return CHX_ERR_INVALID_ARGUMENT;
  CHX_HIP(hipSetDevice(b->device));
  hipStream_t s = (hipStream_t)stream;
  if (flags & CHX_BATCH_EAGER) return enqueue(b, s);
  if (!b->exec) {
    // Captured on the batch's own stream (a legacy default stream cannot
    // be captured), replayed on the caller's.
    if (!b->cap) CHX_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    CHX_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    int rc = enqueue(b, b->cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? rc : hip_status(e);
    }
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(g);
      b->exec = nullptr;
      return hip_status(e);
    }
    b->graph = g;
  }
  CHX_HIP(hipGraphLaunch(b->exec, s));
  return CHX_SUCCESS;
}

int chx_batch_get_info(const chx_batch *b, chx_batch_info *info) {
  if (!b || !info) return CHX_ERR_INVALID_ARGUMENT;
  memset(info, 0, sizeof(*info));
  info->num_jobs = (int32_t)b->jobs.size();
  info->rounds = b->rounds;
  info->chunk_samples = b->chunk;
  info->items = b->nitems;
  info->workgroups = b->tune.workgroups;
  info->max_item_stages = b->tune.max_item_stages;
  info->method = b->tune.method;
  for (const chx_job_desc &j : b->jobs)
    if (b->fmt == kPlanesFp6 && j.num_elements <= b->tune.small_tile_elements)
      info->small_tile_jobs++;
  info->workspace_bytes = b->work_bytes;
  info->cmacs = b->cmacs;
  return CHX_SUCCESS;
}

int chx_batch_destroy(chx_batch *b) {
  if (!b) return CHX_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(b->device);
  free_batch(b);
  return CHX_SUCCESS;
}

}  // extern "C"
