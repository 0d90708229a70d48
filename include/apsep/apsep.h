/*
 * apsep.h -- C ABI of the MI355X (gfx950) APSEP / ANSV-left engine.
 *
 * Problem: for every d_in[i] (int32), d_out[i] = the largest j < i with
 * d_in[j] < d_in[i] (strict, incl = 0) or d_in[j] <= d_in[i] (incl = 1),
 * else -1.  Output is int32.
 *
 * Every entry point is extern "C" with plain pointers and sizes: no C++ or
 * torch types cross this boundary.  Device pointers come from hipMalloc (or
 * any allocator giving 16-byte aligned device memory); streams are
 * hipStream_t passed as void* (NULL = the default stream).
 *
 * Errors are returned, never raised: the library does not print or exit.
 * The reference exits through gpuAssert (src/apsep.cuh:39-47); the C++ shims
 * in apsep.hpp restore that behaviour for drop-in callers.
 *
 * Threading: one apsep_scratch per concurrent caller; a scratch is bound to
 * the device that was current when it was created.
 */
#ifndef APSEP_APSEP_H
#define APSEP_APSEP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum apsep_status {
    APSEP_OK = 0,
    APSEP_ERR_INVALID_ARG = 1,  /* null pointer, bad enum, negative size   */
    APSEP_ERR_ALIGN = 2,        /* d_in / d_out not 16-byte aligned         */
    APSEP_ERR_CAPACITY = 3,     /* n larger than the scratch was sized for  */
    APSEP_ERR_ALLOC = 4,        /* device allocation failed                 */
    APSEP_ERR_LAUNCH = 5,       /* kernel launch / runtime call failed      */
    APSEP_ERR_DEVICE = 6,       /* no usable HIP device, or wrong device    */
    APSEP_ERR_TOO_LARGE = 7,    /* n > INT32_MAX (int32 indices)            */
    APSEP_ERR_UNSUPPORTED = 8,  /* e.g. a stream in an invalidated capture  */
    APSEP_ERR_BUDGET = 9        /* serving: job exceeds the memory budget   */
} apsep_status;

/* Solver variants.  WSTL = multi-launch pipeline (tile pass, 64-ary tree
 * build, fix-up pass); SPT = one persistent launch with grid barriers
 * among the workgroups that run together; LOOKBACK = one launch with
 * decoupled look-back;
 * AUTO = the fastest measured for the size. */
typedef enum apsep_algo {
    APSEP_ALGO_WSTL = 0,
    APSEP_ALGO_SPT = 1,
    APSEP_ALGO_LOOKBACK = 2,
    APSEP_ALGO_AUTO = 3
} apsep_algo;

typedef struct apsep_scratch apsep_scratch;

/* Human-readable text for a status code (static storage). */
const char* apsep_status_string(apsep_status s);

/* Library version as "major.minor.patch". */
const char* apsep_version(void);

/* What lies behind the calling thread's last failing status: the HIP error
 * (name, text, library function and line) or the device-side error; "none"
 * when nothing has failed.  No counterpart in the reference, whose gpuAssert
 * prints the CUDA error text and exits (src/apsep.cuh:39-47); the apsep.hpp
 * shims print it the same way.  Static per-thread storage, overwritten by
 * the next failure. */
const char* apsep_last_error(void);

/* Replaces allocWSTLScratch (src/apsep.cuh:2520-2537) and allocSPTScratch
 * (src/apsep.cuh:3013-3040): device scratch sized for up to n_cap elements,
 * bound to the current device.  n_cap may be 0 (then only n = 0 runs). */
apsep_status apsep_scratch_create(apsep_algo algo, int64_t n_cap, int incl,
                                  apsep_scratch** out);

/* Replaces freeWSTLScratch (src/apsep.cuh:2539-2546) and freeSPTScratch
 * (src/apsep.cuh:3042-3050).  NULL is accepted. */
apsep_status apsep_scratch_destroy(apsep_scratch* s);

/* Bytes of device memory a scratch for n_cap elements holds. */
apsep_status apsep_scratch_bytes(apsep_algo algo, int64_t n_cap, int incl,
                                 int64_t* bytes);

/* Replaces runWSTL (src/apsep.cuh:2548-2596) and runSPT
 * (src/apsep.cuh:3052-3067): asynchronous solve on `stream`.
 * n <= 0 is a no-op returning APSEP_OK (as the reference returns early);
 * n > capacity -> APSEP_ERR_CAPACITY.  d_in and d_out: n int32 each,
 * 16-byte aligned, not overlapping.  d_out needs no initialisation.
 * May be recorded into a graph (stream capture) and replayed: a captured
 * run takes a fixed epoch and empties its level-1 keys with a kernel, so
 * captured and eager runs share a scratch in any order; a stream in an
 * invalidated capture state gives APSEP_ERR_UNSUPPORTED. */
apsep_status apsep_run(apsep_scratch* s, const int32_t* d_in, int32_t* d_out,
                       int64_t n, void* stream);

/* Replaces launchWSTL (src/apsep.cuh:2598-2605) and launchSPT
 * (src/apsep.cuh:3069-3076): allocate scratch, run on the default stream,
 * synchronise, free. */
apsep_status apsep_launch(apsep_algo algo, const int32_t* d_in, int32_t* d_out,
                          int64_t n, int incl);

/* Same as apsep_run with the predicate chosen per call (incl = 0: strict
 * "<", 1: "<="), whatever the scratch was created with.  Replaces the INCL
 * template argument of runSPT (src/apsep.cuh:3051-3052), which the
 * reference also picks per call on an INCL-agnostic scratch. */
apsep_status apsep_run_mode(apsep_scratch* s, const int32_t* d_in,
                            int32_t* d_out, int64_t n, int incl, void* stream);

/* Solve without allocating: no counterpart in the reference.  For callers
 * that cannot call hipMalloc (code under graph capture, tight serving loops,
 * harnesses that forbid allocation).  The same answers as apsep_run_mode
 * (multi-launch solver, predicate per call), with the scratch in a fixed
 * per-GPU array of the library (about 8.2 MB, held by every GPU the library
 * runs on until the process ends) instead of an apsep_scratch.  Asynchronous
 * on `stream`; n <= 0 is a no-op; n > APSEP_NOALLOC_MAX_N ->
 * APSEP_ERR_CAPACITY.  Calls on one GPU share the array: calls that overlap
 * on the device (on different streams, or from different threads) race and
 * give unspecified answers.  It keeps no state between calls, so it may be
 * recorded into a graph and replayed. */
#define APSEP_NOALLOC_MAX_N 67108864
apsep_status apsep_run_noalloc(const int32_t* d_in, int32_t* d_out, int64_t n, int incl,
                               void* stream);

/* Shape of a scratch: the counterpart of the fields of WSTLScratch /
 * SPTScratch (src/apsep.cuh:2509-2518, 3001-3011). */
typedef struct apsep_scratch_info {
    int64_t n_cap;             /* capacity in elements                      */
    int64_t bytes;             /* device bytes held                         */
    int32_t algo;              /* resolved apsep_algo (never AUTO)          */
    int32_t incl;              /* default predicate                         */
    int32_t device;            /* HIP device the scratch lives on           */
    int32_t tile_elems;        /* elements per tile (workgroup)             */
    int32_t tiles;             /* tiles at capacity (num_blocks)            */
    int32_t levels;            /* levels of the 64-ary tile-min hierarchy   */
    int32_t persistent_blocks; /* SPT persistent grid (num_phys), else 0    */
    int32_t reserved;
} apsep_scratch_info;

apsep_status apsep_scratch_query(const apsep_scratch* s, apsep_scratch_info* info);

/* The SPT persistent grid of later runs on this scratch: blocks workgroups
 * (0 restores the default, occupancy x CUs, persistent_blocks above).  A
 * caller sharing the GPU with other work may ask for fewer; more than fit
 * at once is allowed (the workgroups that start late leave).  Every tile is
 * solved whatever the grid.  APSEP_ERR_UNSUPPORTED for a scratch of another
 * algorithm, APSEP_ERR_INVALID_ARG for blocks < 0.  No counterpart in the
 * reference, whose grid is fixed at num_phys (src/apsep.cuh:3024-3032). */
apsep_status apsep_scratch_set_persistent_blocks(apsep_scratch* s, int32_t blocks);

/* SPT launch mode of later eager runs on this scratch: enable = 1 launches
 * the persistent kernel with hipLaunchCooperativeKernel, as the reference's
 * runSPT does with cudaLaunchCooperativeKernel (src/apsep.cuh:3063-3066),
 * on at most the co-resident grid; 0 (the default) launches it normally,
 * the workgroups finding at run time which of them run together (faster on
 * MI355X, and capturable).  Answers do not depend on it.  Captured runs
 * always launch normally.  A cooperative launch the runtime refuses runs
 * nothing; the run then launches normally and the refusal is counted
 * (apsep_scratch_cooperative_refusals).  APSEP_ERR_UNSUPPORTED for a
 * scratch of another algorithm. */
apsep_status apsep_scratch_set_cooperative(apsep_scratch* s, int32_t enable);
apsep_status apsep_scratch_cooperative_refusals(const apsep_scratch* s, int64_t* refused);

/* Blocking check of the scratch's device-side error word, then clears it.
 * *device_error != 0 means a run on this scratch since the last check was
 * cut short on the device (1: a persistent launch timed out at its grid
 * barrier instead of hanging; 2: a look-back wait gave up); its output is
 * invalid.  Later runs are not affected by an earlier run's error.
 * No counterpart in the reference, which has no device-side error path. */
apsep_status apsep_scratch_check(apsep_scratch* s, int32_t* device_error);

/* Per-kernel device time.  No counterpart in the reference (its benchmarks
 * time whole calls, src/main.cu:125-142); for roofline reports, which need
 * the duration of each kernel of a solve.  Launches each kernel of a solve
 * of d_in into d_out `reps` times back to back, in solve order (every
 * launch of a kernel rewrites the same values, so d_out holds the answer
 * afterwards), times each batch with HIP events on `stream` and blocks
 * until done.  Sums are returned: divide by `runs` (= reps).  For SPT the
 * single persistent kernel counts as the tile pass. */
typedef struct apsep_kernel_times {
    int64_t runs;              /* launches per kernel timed                 */
    double tile_ms;            /* tile pass (SPT: the whole kernel)         */
    double levels_ms;          /* level build (0 when not launched)         */
    double fix_ms;             /* fix-up pass                               */
} apsep_kernel_times;

apsep_status apsep_time_kernels(apsep_scratch* s, const int32_t* d_in, int32_t* d_out,
                                int64_t n, int incl, int reps, void* stream,
                                apsep_kernel_times* out);

/* apsep_time_kernels over `count` (1..8) arrays of n keys in turn: launch r
 * of each kernel solves array r % count with its own scratch (same
 * algorithm, same device), input and output, so consecutive launches stream
 * different data through the caches, as consecutive solves of different
 * arrays do.  Each d_out[k] holds its answer afterwards.  No counterpart in
 * the reference. */
apsep_status apsep_time_kernels_rotating(apsep_scratch* const* s, const int32_t* const* d_in,
                                         int32_t* const* d_out, int32_t count, int64_t n,
                                         int incl, int reps, void* stream,
                                         apsep_kernel_times* out);

/* ---- Segmented (CSR) batches ---------------------------------------------
 * No counterpart in the reference, whose callers solve many arrays with one
 * launchWSTL / launchSPT each (src/apsep.cuh:2598-2605, 3069-3076), paying
 * an allocation, a launch sequence and a device synchronisation per array.
 * A plan holds the device scratch for batches of up to total_cap elements
 * in up to nseg_cap segments, with the predicate fixed at creation. */
typedef struct apsep_batch_plan apsep_batch_plan;

apsep_status apsep_batch_plan_create(int64_t total_cap, int32_t nseg_cap, int incl,
                                     apsep_batch_plan** out);
apsep_status apsep_batch_plan_destroy(apsep_batch_plan* p);

/* Segment k covers d_in[d_offsets[k] .. d_offsets[k+1]) (nseg + 1 int64
 * device offsets, non-decreasing, d_offsets[0] = 0, d_offsets[nseg] =
 * total; total is passed so that no device read is needed to size the
 * launch).  d_out[i] = index of the nearest qualifying predecessor of i
 * inside i's segment, relative to the segment start, or -1.  Async on
 * `stream`.  Malformed offsets give unspecified answers but no access
 * outside d_in, d_out and d_offsets[0 .. nseg]. */
apsep_status apsep_run_segmented(apsep_batch_plan* p, const int32_t* d_in,
                                 const int64_t* d_offsets, int32_t* d_out, int32_t nseg,
                                 int64_t total, void* stream);

/* With graphs on and a non-null stream, the first apsep_run_segmented
 * records its launch sequence into a hipGraph and later calls with the same
 * pointers, nseg, total and stream replay it with one hipGraphLaunch (the
 * offsets and keys may change between calls; other arguments re-record).
 * The sequence holds no memset node. */
apsep_status apsep_batch_plan_use_graph(apsep_batch_plan* p, int enable);

/* Tile shape of a plan's batches: 128, 256 (the default) or 512 lanes of 16
 * keys per workgroup.  Answers do not depend on it.  No counterpart in the
 * reference (its kernels take BLOCK_SIZE / IPT as template parameters,
 * src/apsep.cuh:2548, 3052); set_tile_lanes drops a recorded graph, and it
 * and apsep_batch_plan_use_graph switch the tuned table (below) off.  Until
 * the caller sets a shape, apsep_run_segmented picks it per batch by size
 * (128 lanes up to 2^24 keys, 256 beyond: the faster shapes measured on
 * MI355X); apsep_run_jobs always runs the current shape, whose tile size
 * the caller's job table is built for. */
apsep_status apsep_batch_plan_set_tile_lanes(apsep_batch_plan* p, int32_t lanes);
apsep_status apsep_batch_plan_tile_lanes(const apsep_batch_plan* p, int32_t* lanes);

/* One persistent launch per batch (enable = 1) instead of the kernel
 * sequence (0, the default): co-resident workgroups map the segments, solve
 * the tiles, build the hierarchy and fix up the open elements, with grid
 * barriers between the steps (a captured run is preceded by a
 * one-workgroup kernel that clears its barrier, an eager run needs none;
 * with graphs on, the pair is recorded and replayed).  Answers do not depend on it.  No counterpart in
 * the reference, whose persistent kernel (runSPT, src/apsep.cuh:3052-3067)
 * solves one array.  It drops a recorded graph and switches the tuned table
 * off, like apsep_batch_plan_use_graph.  apsep_run_jobs ignores it (job
 * batches always run the kernel sequence).  Workgroups that cannot start
 * while the others run (a GPU shared with other work) leave at once and the
 * rest share their work; a barrier that still waits ~1 s raises the device
 * error that apsep_batch_plan_check reports, and the output is then
 * invalid. */
apsep_status apsep_batch_plan_set_persistent(apsep_batch_plan* p, int enable);

/* Device-side error of the plan's persistent launches (as
 * apsep_scratch_check: 0 = none; blocking; the word is cleared once read). */
apsep_status apsep_batch_plan_check(apsep_batch_plan* p, int32_t* device_error);

/* Autotune a plan for a batch: for each tile shape and each launch sequence
 * (kernel sequence, one persistent launch), times `reps` direct-launch runs
 * and `reps` graph replays of this batch on `stream` (non-null); keeps the
 * fastest shape, sequence and mode and reports that shape and sequence's two
 * mean times (blocking; d_out holds the answer afterwards).  The choice is
 * also recorded for the batch's size class (c with 2^(c-1) < total <= 2^c):
 * from then on apsep_run_segmented runs every batch of a tuned class with
 * that class's shape, sequence and mode. */
apsep_status apsep_batch_plan_tune(apsep_batch_plan* p, const int32_t* d_in,
                                   const int64_t* d_offsets, int32_t* d_out, int32_t nseg,
                                   int64_t total, void* stream, int32_t reps,
                                   float* eager_ms, float* graph_ms);

/* The tuned entry of size class `size_class_log2` (0..31): tile lanes (0
 * when the class has not been tuned), graph mode and its mean time in ms. */
apsep_status apsep_batch_plan_tuned(const apsep_batch_plan* p, int32_t size_class_log2,
                                    int32_t* lanes, int32_t* graph, float* ms);
/* The tuned launch sequence of a size class: 1 = one persistent launch. */
apsep_status apsep_batch_plan_tuned_persistent(const apsep_batch_plan* p,
                                               int32_t size_class_log2, int32_t* persistent);

/* In-place job batches: many independent arrays solved by one launch
 * sequence, each read where it lies and answered into a buffer of the
 * caller's (no concatenation, no copy).  No counterpart in the reference
 * (its callers run one launchWSTL / launchSPT per array,
 * src/apsep.cuh:2598-2605, 3069-3076).  d_jobs is a device array of
 * `count` apsep_job; job k covers the plan's tiles
 * first_tile .. first_tile + ceil(n / tile_elems) - 1, where tile_elems
 * comes from apsep_batch_plan_tile_elems for the plan's current tile shape,
 * and `tiles` is the total (first_tile of job k = the sum of the earlier
 * jobs' tile counts).  d_answers[i] = index of the nearest qualifying
 * predecessor of key i inside the job, or -1.  Async on `stream`; with
 * graphs on (apsep_batch_plan_use_graph) the sequence is recorded once and
 * replayed for the same d_jobs, count, tiles and stream (the table's
 * contents may change).  count <= the plan's nseg_cap and tiles <=
 * apsep_batch_plan_job_tiles, else APSEP_ERR_CAPACITY.  A tile that no job
 * covers (tiles above the jobs' total) is solved as empty; keys and answers
 * must not overlap.  A table whose first_tile values are not those prefix
 * sums gives unspecified answers and may address memory outside the jobs'
 * buffers. */
typedef struct apsep_job {
    const int32_t* d_keys;    /* n keys, 16-byte aligned                     */
    int32_t* d_answers;       /* n answers, 16-byte aligned                  */
    int64_t n;
    int64_t first_tile;
} apsep_job;

apsep_status apsep_batch_plan_tile_elems(const apsep_batch_plan* p, int32_t* elems);
apsep_status apsep_batch_plan_job_tiles(const apsep_batch_plan* p, int64_t* tiles);
apsep_status apsep_run_jobs(apsep_batch_plan* p, const apsep_job* d_jobs, int32_t count,
                            int64_t tiles, void* stream);

/* ---- Serving: HBM pool, budget admission, continuous batching ----------
 * No counterpart in the reference (its launch* allocate and synchronise per
 * call, src/apsep.cuh:2598-2605, 3069-3076); for callers that solve a
 * stream of independent arrays.  A server owns a pool of pool_bytes of
 * device memory for the answers, two streams and a batch plan per stream;
 * one thread drives it.  Keys are read in place from the caller's device
 * memory (never copied into the pool).
 *   submit   admits a job of n keys if its answer slot fits in the free
 *            part of the pool (APSEP_ERR_BUDGET otherwise: nothing is
 *            queued); the device keys (16-byte aligned, APSEP_ERR_ALIGN
 *            otherwise) must be complete when submit is called and stay
 *            valid until the job is done
 *   poll     non-blocking: retires finished batches, then packs queued jobs
 *            (up to max_batch keys / max_batch_jobs jobs per batch, two
 *            batches queued per stream) into in-place job batches
 *            (apsep_run_jobs), alternating streams so that consecutive
 *            batches overlap on the device
 *   job_info state (0 queued, 1 running, 2 done), submit/done times in
 *            microseconds of a monotonic clock, and the device answers
 *            (valid until release)
 *   release  frees a done job's slot; drain blocks until every job is done */
typedef struct apsep_server apsep_server;

typedef struct apsep_job_info {
    int32_t state;
    int32_t reserved;
    int64_t n;
    double submit_us;
    double done_us;
    const int32_t* d_answers;
} apsep_job_info;

apsep_status apsep_server_create(int64_t pool_bytes, int64_t max_batch,
                                 int32_t max_batch_jobs, int incl, apsep_server** out);
apsep_status apsep_server_destroy(apsep_server* s);
apsep_status apsep_server_submit(apsep_server* s, const int32_t* d_keys, int64_t n,
                                 int64_t* job_id);
apsep_status apsep_server_poll(apsep_server* s, int32_t* batches_launched);
apsep_status apsep_server_job_info(apsep_server* s, int64_t job_id, apsep_job_info* info);
apsep_status apsep_server_release(apsep_server* s, int64_t job_id);
apsep_status apsep_server_drain(apsep_server* s);

/* Bulk forms for high job rates: submit a list (the accepted prefix gets
 * consecutive ids from *first_id; admission stops at the first job that
 * does not fit), the id below which every job is done, submit/done times
 * of a range of ids, and release of every job below an id (all done). */
apsep_status apsep_server_submit_many(apsep_server* s, const int32_t* const* d_keys,
                                      const int64_t* n, int32_t count, int64_t* first_id,
                                      int32_t* accepted);
apsep_status apsep_server_done_upto(apsep_server* s, int64_t* done_upto);
apsep_status apsep_server_job_times(apsep_server* s, int64_t first_id, int32_t count,
                                    double* submit_us, double* done_us);
apsep_status apsep_server_release_upto(apsep_server* s, int64_t id);

#ifdef __cplusplus
}
#endif

#endif /* APSEP_APSEP_H */
