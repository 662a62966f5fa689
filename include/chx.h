/*
 * chx.h -- C ABI of the MI355X (gfx950) CHIME X-engine correlator.
 *
 * Plain C: fixed-width integers, plain pointers and sizes, `void *` for a
 * hipStream_t (NULL = the device's default stream).  Every entry point
 * returns an int status (CHX_SUCCESS = 0, like SDK_SUCCESS in the
 * reference's generator_defs.h:10) and never exits the process.
 *
 * Data format (four_bit_macros.h:8-9, input_generator.c:84-96):
 *   input  [time][freq][element] bytes, real in the high nibble, imaginary in
 *          the low nibble, offset binary (value = nibble - 8, in [-8, 7]);
 *   output [freq][block][y 32][x 32][re, im] int32 ("block order"), blocks
 *          enumerated row-major over the upper triangle of 32x32 element
 *          blocks (main_wrapper.c:528-540), diagonal blocks complete
 *          (their lower halves hold the (y > x) products).
 * Visibility at row y, column x (cpu_corr_test.c:156-203, :59-106):
 *   re = sum_t x_r y_r + x_i y_i
 *   im = sum_t x_r y_i - x_i y_r   (upper_triangle_convention = 1)
 *   im = sum_t x_i y_r - x_r y_i   (upper_triangle_convention = 0)
 * Sums wrap modulo 2^32 exactly as the reference's int32 accumulation.
 */
#ifndef CHX_H
#define CHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* kernel_batch values.  0 and 1 are the reference's -k values
 * (main_wrapper.c:281-305); 2 and 3 have no reference counterpart: the
 * matrix-core paths, which need no offset/preseed pass and have no
 * time_accum cap.  2 computes on int8 operands with int32 accumulation; 3 on
 * FP6 E2M3 operands (every 4-bit value v is exact as v/8, scaled back by a
 * block scale of 2^3) with FP32 accumulation kept below 2^24, so its results
 * are bit-identical to 2's. */
#define CHX_METHOD_CONFERENCE 0
#define CHX_METHOD_HIGHLY_PACKED 1
#define CHX_METHOD_INT8_MFMA 2
#define CHX_METHOD_FP6_MFMA 3

/* time_accum caps of the packed methods: 16-bit lanes gain <= 225 per sample
 * (README.md:14; 288 is the largest multiple of 8 not above 291), and the
 * 4-bit overflow counters of the highly-packed method
 * (packed_correlator_overflow_protected_to_2272_iter_UT.cl:171). */
#define CHX_CONFERENCE_MAX_TIME_ACCUM 288
#define CHX_HIGHLY_PACKED_MAX_TIME_ACCUM 2272

enum chx_status {
  CHX_SUCCESS = 0,
  CHX_ERR_INVALID_ARGUMENT = 1, /* NULL pointer, bad enum value */
  CHX_ERR_SHAPE = 2,            /* N % 32, N < 32, F < 1, T % 32, T < 32 */
  CHX_ERR_TIME_ACCUM = 3,       /* ta % 8, T % ta, ta over the method cap */
  CHX_ERR_NO_DEVICE = 4,        /* no GPU, or device index out of range */
  CHX_ERR_OUT_OF_MEMORY = 5,    /* device or pinned allocation failed */
  CHX_ERR_HIP = 6,              /* any other HIP runtime error */
  CHX_ERR_BUDGET = 7,           /* request exceeds a configured HBM budget */
  CHX_ERR_BUSY = 8,             /* no free slot (non-blocking admission) */
  CHX_ERR_NOT_FOUND = 9         /* unknown job/stream handle */
};

/* Replaces the -D shape macros of the runtime-compiled OpenCL program
 * (main_wrapper.c:309-312) and the CLI values that pick it
 * (main_wrapper.c:84-103, 281-305). */
typedef struct chx_config {
  int32_t num_elements;              /* N: NUM_ELEMENTS */
  int32_t num_freq;                  /* F: NUM_FREQUENCIES */
  int64_t time_steps;                /* T: NUM_TIMESAMPLES (64-bit here) */
  int32_t time_accum;                /* NUM_TIME_ACCUM */
  int32_t upper_triangle_convention; /* -U: 1 standard, 0 conjugate */
  int32_t kernel_batch;              /* -k: CHX_METHOD_* */
  int32_t device;                    /* -d */
} chx_config;

typedef struct chx_plan chx_plan;

/* Validates cfg and builds a plan: replaces context/queue/program/kernel
 * creation and the block id maps (main_wrapper.c:238-372, 528-552).  On a
 * validation error nothing touches the GPU. */
int chx_plan_create(const chx_config *cfg, chx_plan **plan);
int chx_plan_destroy(chx_plan *plan);
int chx_plan_get_config(const chx_plan *plan, chx_config *cfg);

/* Bytes of the input ([T][F][N] u8), of the block-order output
 * (F * NB * 2048 int32), and of the offset sums (F * N * 2 uint32). */
int chx_plan_sizes(const chx_plan *plan, size_t *input_bytes,
                   size_t *output_bytes, size_t *sums_bytes);

/* ---- Stage operations on device pointers (one per reference kernel). ---- */

/* offsetAccumulateElements (offset_accumulator.cl:7-49): adds
 * (sum_t R, sum_t I) of the raw nibbles into d_sums[F][N][2]; the caller
 * zeroes d_sums first, as main_wrapper.c:671-679 does. */
int chx_offset_accumulate(const chx_plan *plan, const uint8_t *d_input,
                          uint32_t *d_sums, void *stream);

/* preseed (preseed_multifreq{,_UT}.cl:19-116 and
 * preseed_multifreq_highly_packed_correlator_method{,_UT}.cl): assigns the
 * offset-binary correction of the plan's method and convention to every
 * entry of d_vis.  Methods 0 and 1 only. */
int chx_preseed(const chx_plan *plan, const uint32_t *d_sums, int32_t *d_vis,
                void *stream);

/* corr (pairwise_correlator{,_UT}.cl:22-207,
 * packed_correlator_overflow_protected_to_2272_iter{,_UT}.cl): adds the
 * method's products into d_vis.  Methods 0/1 add the raw nibble-space
 * products (d_vis must hold the preseed); methods 2 and 3 add the finished
 * signed visibilities.  No lock buffer: blocks are owned, split time ranges add
 * with atomics. */
int chx_corr(const chx_plan *plan, const uint8_t *d_input, int32_t *d_vis,
             void *stream);

/* The whole per-iteration pipeline (main_wrapper.c:691-778): assigns the
 * final visibilities of d_input to d_vis.  Methods 0/1: zero sums ->
 * offset accumulate -> preseed -> corr, with plan-owned sums; methods 2 and
 * 3: one matrix-core pass per plane chunk. */
int chx_correlate(chx_plan *plan, const uint8_t *d_input, int32_t *d_vis,
                  void *stream);

/* Host-pointer convenience: copies h_input in, correlates, copies the block
 * buffer out; synchronous.  Device buffers are owned by the plan. */
int chx_correlate_host(chx_plan *plan, const uint8_t *h_input, int32_t *h_vis);

/* chx_correlate without a plan and without any allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Methods 0 and 1 only (the matrix-core methods need a plane
 * workspace: CHX_ERR_INVALID_ARGUMENT).  Their scratch -- the block id maps
 * and the offset sums, (2 NB rounded up to 64 + 2 F N) uint32 -- lives in a
 * static device array of CHX_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; a larger request returns CHX_ERR_BUDGET (chx_last_error_detail
 * says how much it needs).  Each call rewrites the scratch before use, but
 * all calls on one GPU share it: calls that overlap on one GPU (other
 * streams, other threads) race.  Assigns d_vis like chx_correlate.  Works
 * on cfg->device, whichever device is current, and leaves the calling
 * thread's current device as it found it. */
#define CHX_NOALLOC_SCRATCH_BYTES (2u << 20)
int chx_correlate_noalloc(const chx_config *cfg, const uint8_t *d_input, int32_t *d_vis,
                          void *stream);

/* Kernel timing (no reference counterpart; for roofline measurements):
 * while enabled, every matrix-core chunk (int8 or FP6) the plan launches records HIP
 * events on its stream before the unpack kernel, between it and the GEMM
 * kernel, and after the GEMM kernel.  chx_plan_kernel_times waits for the
 * recorded launches, returns the summed unpack and GEMM durations and the
 * number of chunk launches, and clears the records. */
int chx_plan_profile(chx_plan *plan, int32_t enable);
int chx_plan_kernel_times(chx_plan *plan, double *unpack_ms, double *gemm_ms,
                          int64_t *launches);

/* ---- Plan-owned device buffers and copies (for C hosts that do not link
 * HIP themselves; replace clCreateBuffer / clEnqueueWrite/ReadBuffer,
 * main_wrapper.c:379-501, 655-667, 798-799). ---- */

/* Allocates (once) and returns the plan's device input and output buffers. */
int chx_plan_buffers(chx_plan *plan, uint8_t **d_input, int32_t **d_vis);
/* Device memory on a device (clCreateBuffer, main_wrapper.c:455-501), e.g.
 * a second stage of input and output buffers. */
int chx_device_alloc(int32_t device, size_t bytes, void **ptr);
int chx_device_free(void *ptr);
/* Pinned host memory (hipHostMalloc), for fast H2D of large inputs. */
int chx_host_alloc(void **ptr, size_t bytes);
int chx_host_free(void *ptr);
/* Asynchronous copies on stream (NULL = default); dst/src by direction. */
int chx_copy_to_device(void *d_dst, const void *h_src, size_t bytes, void *stream);
int chx_copy_to_host(void *h_dst, const void *d_src, size_t bytes, void *stream);
/* Waits for all work queued on stream (NULL = default stream). */
int chx_stream_synchronize(void *stream);
/* Queues and events for hosts that overlap copies with computation: replace
 * the two out-of-order command queues and the event chaining of the
 * reference loop (main_wrapper.c:264-274, 614-785).  A stream is a
 * non-blocking hipStream_t on device; an event orders work across streams
 * (chx_stream_wait_event: later work on stream waits for the work recorded
 * before the event). */
int chx_stream_create(int32_t device, void **stream);
int chx_stream_destroy(void *stream);
int chx_event_create(void **event);
int chx_event_destroy(void *event);
int chx_event_record(void *event, void *stream);
int chx_stream_wait_event(void *stream, void *event);

/* ---- Batch: many correlations of uneven shape in one graph launch. ----
 * No reference counterpart: the reference builds one program per shape and
 * runs one shape per process (main_wrapper.c:309-312).  Every job uses one
 * matrix-core method (FP6 unless the tuning says int8), and a job's
 * visibilities are bit-identical to what chx_correlate assigns for the same
 * job on its own.  Jobs are fixed
 * at creation (device pointers included); each run re-reads the inputs and
 * assigns the outputs.  A batch is not thread-safe; runs of one batch on
 * different streams must not overlap (they share the workspace). */
typedef struct chx_job_desc {
  int32_t num_elements;              /* N: multiple of 32, >= 32 */
  int32_t num_freq;                  /* F >= 1 */
  int64_t time_steps;                /* T >= 1 (no time_accum: matrix cores) */
  int32_t upper_triangle_convention; /* 1 standard, 0 conjugate */
  int32_t reserved;                  /* must be 0 */
  const uint8_t *d_input;            /* [T][F][N] bytes on the batch's device */
  int32_t *d_vis;                    /* [F][NB][32][32][2] int32 (assigned) */
} chx_job_desc;

/* Launch shape of a batch; zeros pick the defaults.  Picked per job list
 * by the autotuner (chime-x_amd/autotune.py). */
typedef struct chx_batch_tuning {
  int32_t max_item_stages;  /* time extent of one work item, in 64-sample
                               stages (rounded up to a multiple of the
                               method's ring, 4 stages); 0 = a whole chunk
                               per item (at most 2048 stages; longer
                               chunks are cut) */
  int32_t workgroups;       /* persistent workgroups (multiple of 8);
                               0 = one per compute unit */
  int64_t workspace_bytes;  /* HBM budget of one round's planes;
                               0 = 16 GiB */
  int32_t method;           /* CHX_METHOD_INT8_MFMA or CHX_METHOD_FP6_MFMA;
                               0 = FP6 */
  int32_t small_tile_elements; /* FP6 only: jobs of at most this many
                               elements use output tiles of 128 x 128 (four
                               waves of 64 x 64) instead of 128 x 256;
                               0 = none (the int8 method ignores it) */
} chx_batch_tuning;

typedef struct chx_batch_info {
  int32_t num_jobs;
  int32_t rounds;          /* time chunks: unpack + matrix-core pass each */
  int64_t chunk_samples;   /* samples of every job per round */
  int64_t items;           /* work items over all rounds */
  int32_t workgroups;
  int32_t max_item_stages;
  size_t workspace_bytes;  /* allocated planes workspace */
  int64_t cmacs;           /* sum over jobs of F * T * N(N+1)/2 */
  int32_t method;          /* matrix-core method used */
  int32_t small_tile_jobs; /* jobs on 128 x 128 tiles */
} chx_batch_info;

typedef struct chx_batch chx_batch;

/* Validates every job (CHX_ERR_SHAPE / CHX_ERR_INVALID_ARGUMENT, nothing
 * touches the GPU), plans the rounds and the work items, allocates the
 * workspace (CHX_ERR_BUDGET when one 256-sample round of all jobs exceeds
 * workspace_bytes). */
int chx_batch_create(int32_t device, const chx_job_desc *jobs, int32_t num_jobs,
                     const chx_batch_tuning *tuning, chx_batch **batch);

/* The same planning on the host only (nothing touches a GPU): validates
 * like chx_batch_create and reports the rounds, work items and workspace a
 * batch of these jobs and tuning would use (workgroups 0 = one per compute
 * unit of the device). */
int chx_batch_plan(const chx_job_desc *jobs, int32_t num_jobs,
                   const chx_batch_tuning *tuning, chx_batch_info *info);

/* flags: 0 = replay the batch's hipGraph (captured on first use) on
 * stream; CHX_BATCH_EAGER = launch the same kernels directly. */
#define CHX_BATCH_EAGER 1
int chx_batch_run(chx_batch *batch, int32_t flags, void *stream);
int chx_batch_get_info(const chx_batch *batch, chx_batch_info *info);
int chx_batch_destroy(chx_batch *batch);

/* ---- Streaming serve: continuous batching of arriving frames. ----
 * Replaces the reference's fixed pipeline loop (main_wrapper.c:613-795,
 * one shape, H2D -> zero -> offset -> preseed -> corr per iteration) with
 * an engine that keeps accumulators device-resident per stream (one
 * integration of one array), admits streams by HBM budget and frames by
 * free frame slots, and integrates every frame that is on the device when
 * the GPU comes free in one launch (frames of many streams together; a
 * stream's frames as one longer chunk).  Frames copy in on their own stream
 * while earlier frames compute.  Results are bit-identical to correlating
 * each stream's frames alone (matrix-core methods, 2^32-wrapping int32
 * sums).  All entry points are thread-safe. */
typedef struct chx_serve_config {
  int32_t device;
  int32_t frame_slots;      /* frames in flight at most (copying, queued or
                               computing); a submit needs a free slot */
  int64_t slot_bytes;       /* device buffer per slot: the largest frame */
  int64_t workspace_bytes;  /* planes of one launch (int8: 2 bytes, FP6: 1.5
                               bytes per input byte when N % 128 == 0 and
                               frames are whole rings); 0 = 2 *
                               max_batch_frames * slot_bytes */
  int64_t hbm_budget_bytes; /* slots + workspace + accumulators; 0 = no cap */
  int32_t max_batch_frames; /* frames per launch at most; 0 = frame_slots */
  int32_t max_item_stages;  /* as in chx_batch_tuning; 0 = whole items, and
                               when a queue's items are equal, its last
                               partial round cut in time */
  int32_t method;           /* CHX_METHOD_INT8_MFMA or CHX_METHOD_FP6_MFMA;
                               0 = FP6 */
  int32_t reserved;         /* 0 */
} chx_serve_config;

typedef struct chx_serve_stats {
  int64_t frames_submitted;
  int64_t frames_done;
  int64_t batches;           /* launches */
  int64_t streams_open;
  int64_t bytes_in_use;      /* slots + workspace + open accumulators */
  double wall_seconds;       /* first submit -> last completion */
  double cmacs;              /* algorithmic complex MACs of the frames done */
  double mean_batch_frames;
  double latency_p50_ms;     /* frame-to-visibility: submit -> its products */
  double latency_p99_ms;     /*   are in the stream's accumulator */
  double latency_max_ms;
  double latency_mean_ms;
} chx_serve_stats;

typedef struct chx_serve chx_serve;

int chx_serve_create(const chx_serve_config *cfg, chx_serve **serve);
/* Finishes every submitted frame, then frees everything. */
int chx_serve_destroy(chx_serve *serve);

/* Admits a stream: frames of frame_samples x num_freq x num_elements bytes,
 * integrated into an accumulator that starts at zero.  CHX_ERR_BUDGET when
 * its accumulator does not fit the budget now (retry after a close), or a
 * frame does not fit a slot or the workspace.  integration_frames > 0 caps
 * the frames the stream accepts (0 = no cap). */
int chx_serve_open(chx_serve *serve, int32_t num_elements, int32_t num_freq,
                   int32_t frame_samples, int32_t upper_triangle_convention,
                   int64_t integration_frames, int64_t *stream_id);
/* Next frame of a stream from host memory: queued for an asynchronous copy
 * into a free slot.  h_frame must stay unchanged until the frame is done
 * (chx_serve_wait_frame); pinned memory (chx_host_alloc) lets the copy
 * overlap computation.  CHX_ERR_BUSY when no slot is free (retry),
 * CHX_ERR_BUDGET when the stream's integration_frames are all submitted. */
int chx_serve_submit(chx_serve *serve, int64_t stream_id, const uint8_t *h_frame,
                     int64_t *frame_id);
/* The same for a frame already in device memory: read in place (no copy).
 * Its data must be complete when submitted (the work writing it finished)
 * and stay valid until the frame is done.  Takes a slot as well. */
int chx_serve_submit_device(chx_serve *serve, int64_t stream_id,
                            const uint8_t *d_frame, int64_t *frame_id);
/* Blocks until the frame's products are in its stream's accumulator. */
int chx_serve_wait_frame(chx_serve *serve, int64_t frame_id);
/* Blocks until every frame submitted to the stream is done, then copies its
 * accumulator ([F][NB][32][32][2] int32, block order) to host or device.
 * After a failed launch (a sticky error) the engine retires queued frames
 * without integrating them; read and close still wait for them, then return
 * the error (close also frees the stream). */
int chx_serve_read(chx_serve *serve, int64_t stream_id, int32_t *h_vis);
int chx_serve_read_device(chx_serve *serve, int64_t stream_id, int32_t *d_vis);
/* Waits for the stream's frames (also after an error), releases its
 * accumulator. */
int chx_serve_close(chx_serve *serve, int64_t stream_id);
int chx_serve_get_stats(chx_serve *serve, chx_serve_stats *stats);
/* Clears the latency records and counters (not the streams). */
int chx_serve_reset_stats(chx_serve *serve);
/* Latency records of the most recent 2^20 completed frames (older ones are
 * dropped; the percentiles of chx_serve_get_stats cover the same records),
 * oldest first: frame ids and submit-to-done milliseconds.  Writes at most
 * max records (either array may be NULL) and the number of records held
 * into *count. */
int chx_serve_latencies(chx_serve *serve, int64_t *frame_ids, double *ms, int64_t max,
                        int64_t *count);

/* ---- Test data on the device. ---- */

/* generate_char_data_set (input_generator.c:17-103) on the GPU: fills
 * d_data ([T][F][N] bytes) byte-identically to the reference generator,
 * random mode included (libc rand() of glibc after srand(seed), re then im
 * per element, srand again before every time sample unless no_repeat).
 * gen_type: 1 constant, 2 ramp up, 3 ramp down, 4 random, other: default
 * values; single_freq -1 = all channels.  Returns when the data is written
 * (allocates temporaries). */
int chx_generate_device(int32_t gen_type, int32_t seed, int32_t default_real,
                        int32_t default_imag, int32_t initial_real,
                        int32_t initial_imag, int32_t single_freq, int64_t T,
                        int32_t F, int32_t N, int32_t no_repeat, uint8_t *d_data,
                        void *stream);

/* The rand() values k0 .. k0+n-1 after srand(seed), as glibc computes them
 * (host only; jumps ahead without generating the first k0). */
int chx_rand_host(int32_t seed, int64_t k0, int64_t n, int32_t *out);

/* ---- Layout helpers (host only, no GPU needed). ---- */

/* NB = (N/32)(N/32+1)/2 (main_wrapper.c:277). Returns -1 on a bad N. */
int64_t chx_num_blocks(int32_t num_elements);

/* id_x_map / id_y_map (main_wrapper.c:528-540). */
int chx_block_maps(int32_t num_elements, uint32_t *x_map, uint32_t *y_map);

/* reorganize_GPU_to_upper_triangle (gpu_data_reorg.c:76-124): block order
 * -> [F][N(N+1)/2][2] packed upper triangle, row-major. */
int chx_blocks_to_triangle(int32_t num_freq, int32_t num_elements,
                           const int32_t *blocks, int32_t *triangle);

/* reorganize_GPU_to_full_Matrix_for_comparison (gpu_data_reorg.c:36-74):
 * block order -> the full Hermitian matrix [F][N][N][2], row y, column x;
 * below the diagonal an entry is the conjugate of its transpose (the
 * imaginary part negated, wrapping mod 2^32). */
int chx_blocks_to_full(int32_t num_freq, int32_t num_elements,
                       const int32_t *blocks, int32_t *full);
/* The same on the device (the reference's version is host-only and unused
 * because TRIANGLE is 1, main_wrapper.c:49, 847-855). */
int chx_blocks_to_full_device(int32_t num_freq, int32_t num_elements,
                              const int32_t *d_blocks, int32_t *d_full, void *stream);

/* reorganize_32_to_16_feed_GPU_Correlated_Data (gpu_data_reorg.c:6-34), the
 * 16-feed (pathfinder) unpacking. Bytes [T][F][E] with E = 16 read as
 * [T][F/2][2E] put frequencies 2g and 2g+1 side by side in one 2E-element
 * frequency; after a run at num_elements 2E, F/2 frequencies, and
 * chx_blocks_to_full, this keeps the two diagonal quadrants of each 2E x 2E
 * matrix: in place, [F/2][2E][2E][2] -> [2(F/2)][E][E][2] (frequency 2g = the
 * rows and columns < E). F/2 rounds down; entries past the output keep their
 * input values. Any E >= 0 (the reference allows any). */
int chx_split_paired_feeds(int32_t num_freq, int32_t num_elements, int32_t *data);

/* reorganize_data_16_element_with_triangle_conversion (gpu_data_reorg.c:
 * 126-149): [num_freq][16][16][2] full matrices -> [num_freq_out][136][2]
 * upper triangles (row y, column x >= y, row-major); frequencies num_freq ..
 * num_freq_out-1 are zero. input and output must not overlap. */
int chx_feed16_to_triangle(int32_t num_freq_out, int32_t num_freq, const int32_t *input,
                           int32_t *output);

/* Device version of the same reorg (no reference counterpart: it keeps the
 * visibilities on the GPU for consumers that want the triangle). */
int chx_blocks_to_triangle_device(int32_t num_freq, int32_t num_elements,
                                  const int32_t *d_blocks, int32_t *d_triangle,
                                  void *stream);

/* Replaces oclGetOpenCLErrorCodeStr
 * (amd_firepro_error_code_list_for_opencl.c:4-132). */
const char *chx_strerror(int status);
/* Text of the last HIP error seen by this thread, or "" if none. */
const char *chx_last_error_detail(void);

/* Largest time_accum a method accepts (0 = no cap), -1 on a bad method. */
int32_t chx_max_time_accum(int32_t kernel_batch);

/* The GPU's target name (e.g. "gfx950") into name[len]; keys the
 * autotuner's cache (no reference counterpart). */
int chx_device_arch(int32_t device, char *name, size_t len);

/* Number of visible GPUs (0 on a machine without one). */
int chx_device_count(void);

/* Library version string. */
const char *chx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CHX_H */
