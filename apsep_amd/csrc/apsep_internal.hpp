// apsep_internal.hpp -- host-side state shared by the solver translation
// units (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/apsep/apsep.h"
#include "apsep_device.hpp"

struct apsep_scratch {
    apsep_algo algo;      // resolved (never AUTO)
    int incl;
    int64_t n_cap;
    int device;
    int num_cus;
    int nt_cap;
    int tile;             // elements per tile the arrays are sized for
    void* base;
    int64_t bytes;
    int* mn0;                   // tile mins
    unsigned long long* key[5]; // epoch-keyed levels 1..
    unsigned epoch;             // next run's epoch
    int* cmin;
    int* qcount;
    int2* qlist;
    int* scount;
    int2* spine;
    int spt_grid;        // SPT persistent grid size (occupancy x CUs)
    int spt_blocks;      // SPT grid the caller asked for (0: spt_grid)
    int spt_cooperative; // SPT eager runs launched with hipLaunchCooperativeKernel
    int64_t spt_coop_refused;  // cooperative launches the runtime refused (run normally instead)
    int* spt_link;       // SPT: per tile, the next tile of a workgroup's overflow list
    apsep::GridBarrier* bar;   // BAR_SETS barriers (persistent launches)
    int bar_parity;            // the set the next eager persistent launch uses
    // look-back runs: control words and the published (tagged) words
    apsep::LbCtl* lb_ctl;
    unsigned long long* lb_M;
    unsigned* lb_GC;           // LB_GC_SETS sets of group counters, lb_gc_stride apart
    int64_t lb_gc_stride;
    int lb_ngc_cap;            // group counters of a run at capacity (every run's fit in it)
    int64_t lb_m_bytes, lb_cm_bytes, lb_g_bytes[apsep::LB_LEVELS];  // published arrays
    unsigned lb_tag;           // the last eager run's tag (LB_TAG_CAPTURED_MIN and up: captured runs)
    int lb_parity;             // the counter set the next eager run uses
    unsigned long long* lb_CM;
    unsigned long long* lb_G[apsep::LB_LEVELS];
};

// A segmented batch: its own scratch (sized for the concatenation) plus the
// tile -> segment map, and the captured launch sequence when graphs are on.
struct apsep_batch_plan {
    apsep_scratch* s;
    int64_t total_cap;
    int32_t nseg_cap;
    int* tmap;             // [tiles at capacity]
    int2* tbound;          // [tiles at capacity]: that segment's start and end
    apsep::Page* pages;    // [tiles at capacity]: in-place job batches
    int tile_lanes;        // tile shape the batches run with: 128, 256 or 512 lanes
    int use_graph;
    int persistent;        // segmented batches as one persistent launch
    // autotuned settings per size class (class c: batches of 2^(c-1) + 1 ..
    // 2^c keys), applied to apsep_run_segmented while `tuned_auto` is set
    int tuned_auto;
    int lanes_explicit;    // the caller chose the tile shape (no size rule)
    int cls_lanes[32];     // 0: class not tuned
    int cls_graph[32];
    int cls_persist[32];
    float cls_ms[32];
    hipGraph_t graph;
    hipGraphExec_t exec;
    // arguments the captured graph was recorded with
    const void* g_in;
    const void* g_offs;
    void* g_out;
    int32_t g_nseg;
    int64_t g_total;
    void* g_stream;
    int g_jobs;            // the graph holds a job batch (apsep_run_jobs)
};

namespace apsep {

// Tile lanes for a batch of `total` keys when neither the caller nor the
// tuner chose (measured on MI355X: 128-lane tiles are fastest up to about
// 2^24 keys, where a batch does not fill the GPU with 256-lane tiles; 256
// beyond).
inline int default_batch_lanes(int64_t total) { return total <= (1ll << 24) ? 128 : 256; }

// The HIP error behind the calling thread's last failing status: recorded
// where the library clears HIP's sticky error (apsep_last_error reports it).
void record_hip_error(hipError_t e, const char* func, int line);
// Reads and clears HIP's last error; records it when it is one.
inline hipError_t take_hip_error(const char* func, int line) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) record_hip_error(e, func, line);
    return e;
}
#define APSEP_CLEAR_HIP_ERROR() ((void)::apsep::take_hip_error(__func__, __LINE__))
// true when the launches just issued failed (the error is recorded)
#define APSEP_LAUNCH_FAILED() (::apsep::take_hip_error(__func__, __LINE__) != hipSuccess)


int level_sizes(int nt, int* sizes);
int64_t single_scratch_bytes(int64_t n_cap, apsep_algo algo, int tile);
apsep_status run_lookback(apsep_scratch* s, const int32_t* in, int32_t* out, int n, int incl,
                          hipStream_t st);
apsep_status single_scratch_init(apsep_scratch* s);
apsep_status single_run(apsep_scratch* s, const int32_t* in, int32_t* out, int n,
                        int incl, hipStream_t st);
apsep_status single_error(apsep_scratch* s, int* code);
// apsep_run_noalloc's solve: scratch in a per-device file-scope array
apsep_status noalloc_run(const int32_t* in, int32_t* out, int n, int incl, hipStream_t st);
// The segmented solver, built for tiles of 128, 256 and 512 lanes
// (apsep_segmented_128.hip, _256.hip, _512.hip).  seg_tile_elems is that
// build's tile in elements.
namespace seg128 {
apsep_status seg_run(apsep_batch_plan* p, const int32_t* in, const int64_t* offs,
                     int32_t* out, int nseg, int n, hipStream_t st);
apsep_status jobs_run(apsep_batch_plan* p, const void* jobs, int njobs, int tiles,
                      hipStream_t st);
int seg_tile_elems();
}  // namespace seg128
namespace seg256 {
apsep_status seg_run(apsep_batch_plan* p, const int32_t* in, const int64_t* offs,
                     int32_t* out, int nseg, int n, hipStream_t st);
apsep_status jobs_run(apsep_batch_plan* p, const void* jobs, int njobs, int tiles,
                      hipStream_t st);
int seg_tile_elems();
}  // namespace seg256
namespace seg512 {
apsep_status seg_run(apsep_batch_plan* p, const int32_t* in, const int64_t* offs,
                     int32_t* out, int nseg, int n, hipStream_t st);
apsep_status jobs_run(apsep_batch_plan* p, const void* jobs, int njobs, int tiles,
                      hipStream_t st);
int seg_tile_elems();
}  // namespace seg512

// `reps` back-to-back launches of each kernel of the solve (each launch is
// idempotent), timed per kernel batch with HIP events; launch r solves array
// r % m (its own scratch, input and output).
apsep_status single_time_kernels(apsep_scratch* const* ss, const int32_t* const* ins,
                                 int32_t* const* outs, int m, int n, int incl, int reps,
                                 hipStream_t st, double ms[3]);

}  // namespace apsep
