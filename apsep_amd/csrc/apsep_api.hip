// apsep_api.hip -- the extern "C" boundary (include/apsep/apsep.h).
//
// Argument checks run before any HIP call, so every error path is testable
// on a machine without a GPU.  Nothing here prints or exits.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <new>

#include "apsep_device.hpp"
#include "apsep_internal.hpp"
#include "../../include/apsep/apsep.h"

using namespace apsep;

namespace apsep {

// The calling thread's last recorded failure (apsep_last_error).
static thread_local char g_last_error[192];

void record_hip_error(hipError_t e, const char* func, int line) {
    snprintf(g_last_error, sizeof g_last_error, "%s (%s) in %s, line %d", hipGetErrorName(e),
             hipGetErrorString(e), func, line);
}

static void record_device_error(int code) {
    snprintf(g_last_error, sizeof g_last_error,
             "device-side error word %d: a grid barrier or look-back wait gave up", code);
}

}  // namespace apsep

extern "C" {

const char* apsep_last_error(void) { return g_last_error[0] ? g_last_error : "none"; }

const char* apsep_status_string(apsep_status s) {
    switch (s) {
    case APSEP_OK: return "ok";
    case APSEP_ERR_INVALID_ARG: return "invalid argument";
    case APSEP_ERR_ALIGN: return "device pointer not 16-byte aligned";
    case APSEP_ERR_CAPACITY: return "n exceeds the scratch capacity";
    case APSEP_ERR_ALLOC: return "device allocation failed";
    case APSEP_ERR_LAUNCH: return "kernel launch or runtime call failed";
    case APSEP_ERR_DEVICE: return "no usable HIP device (or not the scratch's device)";
    case APSEP_ERR_TOO_LARGE: return "n exceeds INT32_MAX";
    case APSEP_ERR_UNSUPPORTED: return "operation not supported on this device";
    case APSEP_ERR_BUDGET: return "request exceeds the memory budget";
    }
    return "unknown status";
}

const char* apsep_version(void) { return "0.1.0"; }

static bool valid_algo(apsep_algo a) {
    return a == APSEP_ALGO_WSTL || a == APSEP_ALGO_SPT || a == APSEP_ALGO_LOOKBACK ||
           a == APSEP_ALGO_AUTO;
}

static apsep_algo resolve_algo(apsep_algo a) {
    // the multi-launch solver is the fastest on gfx950: the persistent one
    // pays for agent-scope fences at every grid barrier
    if (a == APSEP_ALGO_AUTO) return APSEP_ALGO_WSTL;
    return a;
}

apsep_status apsep_scratch_bytes(apsep_algo algo, int64_t n_cap, int incl,
                                 int64_t* bytes) {
    if (!bytes || !valid_algo(algo) || n_cap < 0 || (incl != 0 && incl != 1))
        return APSEP_ERR_INVALID_ARG;
    if (n_cap > INT32_MAX) return APSEP_ERR_TOO_LARGE;
    *bytes = n_cap > 0 ? single_scratch_bytes(n_cap, resolve_algo(algo), TILE) : 0;
    return APSEP_OK;
}

// A scratch whose per-tile arrays are sized for tiles of `tile` elements:
// the single-array solvers' tile for apsep_scratch_create, the segmented
// solver's for a batch plan.
static apsep_status scratch_create(apsep_algo algo, int64_t n_cap, int incl, int tile,
                                   apsep_scratch** out) {
    if (!out) return APSEP_ERR_INVALID_ARG;
    *out = nullptr;
    if (!valid_algo(algo) || n_cap < 0 || (incl != 0 && incl != 1))
        return APSEP_ERR_INVALID_ARG;
    if (n_cap > INT32_MAX) return APSEP_ERR_TOO_LARGE;
    int dev = -1, ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 ||
        hipGetDevice(&dev) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_DEVICE;
    }
    apsep_scratch* s = new (std::nothrow) apsep_scratch();  // zeroed
    if (!s) return APSEP_ERR_ALLOC;
    memset(s, 0, sizeof *s);
    s->algo = resolve_algo(algo);
    s->incl = incl;
    s->n_cap = n_cap;
    s->tile = tile;
    s->device = dev;
    if (hipDeviceGetAttribute(&s->num_cus, hipDeviceAttributeMultiprocessorCount, dev) !=
        hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        delete s;
        return APSEP_ERR_DEVICE;
    }
    apsep_status st = single_scratch_init(s);
    if (st != APSEP_OK) {
        apsep_scratch_destroy(s);
        return st;
    }
    *out = s;
    return APSEP_OK;
}

apsep_status apsep_scratch_create(apsep_algo algo, int64_t n_cap, int incl,
                                  apsep_scratch** out) {
    return scratch_create(algo, n_cap, incl, TILE, out);
}

apsep_status apsep_scratch_destroy(apsep_scratch* s) {
    if (!s) return APSEP_OK;
    if (s->base) (void)hipFree(s->base);
    delete s;
    return APSEP_OK;
}

apsep_status apsep_scratch_query(const apsep_scratch* s, apsep_scratch_info* info) {
    if (!s || !info) return APSEP_ERR_INVALID_ARG;
    memset(info, 0, sizeof *info);
    int sz[MAX_LEVELS];
    info->n_cap = s->n_cap;
    info->bytes = s->bytes;
    info->algo = (int32_t)s->algo;
    info->incl = s->incl;
    info->device = s->device;
    info->tile_elems = s->tile;
    info->tiles = s->nt_cap;
    info->levels = s->nt_cap > 0 ? level_sizes(s->nt_cap, sz) : 0;
    info->persistent_blocks = s->spt_grid;
    return APSEP_OK;
}

apsep_status apsep_scratch_set_persistent_blocks(apsep_scratch* s, int32_t blocks) {
    if (!s || blocks < 0) return APSEP_ERR_INVALID_ARG;
    if (s->algo != APSEP_ALGO_SPT) return APSEP_ERR_UNSUPPORTED;
    s->spt_blocks = blocks;
    return APSEP_OK;
}

apsep_status apsep_scratch_set_cooperative(apsep_scratch* s, int32_t enable) {
    if (!s || (enable != 0 && enable != 1)) return APSEP_ERR_INVALID_ARG;
    if (s->algo != APSEP_ALGO_SPT) return APSEP_ERR_UNSUPPORTED;
    s->spt_cooperative = enable;
    return APSEP_OK;
}

apsep_status apsep_scratch_cooperative_refusals(const apsep_scratch* s, int64_t* refused) {
    if (!s || !refused) return APSEP_ERR_INVALID_ARG;
    *refused = s->spt_coop_refused;
    return APSEP_OK;
}

apsep_status apsep_scratch_check(apsep_scratch* s, int32_t* device_error) {
    if (!s || !device_error) return APSEP_ERR_INVALID_ARG;
    int code = 0;
    apsep_status st = single_error(s, &code);
    *device_error = code;
    return st;
}

static apsep_status check_io(const int32_t* in, int32_t* out, int64_t n) {
    if (!in || !out) return APSEP_ERR_INVALID_ARG;
    if (n > INT32_MAX) return APSEP_ERR_TOO_LARGE;
    if (((uintptr_t)in & 15) || ((uintptr_t)out & 15)) return APSEP_ERR_ALIGN;
    const uintptr_t a0 = (uintptr_t)in, a1 = a0 + (uintptr_t)n * 4;
    const uintptr_t b0 = (uintptr_t)out, b1 = b0 + (uintptr_t)n * 4;
    if (a0 < b1 && b0 < a1) return APSEP_ERR_INVALID_ARG;  // overlap
    return APSEP_OK;
}

apsep_status apsep_run(apsep_scratch* s, const int32_t* d_in, int32_t* d_out,
                       int64_t n, void* stream) {
    if (!s) return APSEP_ERR_INVALID_ARG;
    return apsep_run_mode(s, d_in, d_out, n, s->incl, stream);
}

apsep_status apsep_run_mode(apsep_scratch* s, const int32_t* d_in, int32_t* d_out,
                            int64_t n, int incl, void* stream) {
    if (!s || (incl != 0 && incl != 1)) return APSEP_ERR_INVALID_ARG;
    if (n <= 0) return APSEP_OK;
    if (n > INT32_MAX) return APSEP_ERR_TOO_LARGE;
    if (n > s->n_cap) return APSEP_ERR_CAPACITY;
    apsep_status st = check_io(d_in, d_out, n);
    if (st != APSEP_OK) return st;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != s->device) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_DEVICE;
    }
    return single_run(s, d_in, d_out, (int)n, incl, (hipStream_t)stream);
}

apsep_status apsep_run_noalloc(const int32_t* d_in, int32_t* d_out, int64_t n, int incl,
                              void* stream) {
    if (incl != 0 && incl != 1) return APSEP_ERR_INVALID_ARG;
    if (n <= 0) return APSEP_OK;
    if (n > APSEP_NOALLOC_MAX_N) return APSEP_ERR_CAPACITY;
    const apsep_status st = check_io(d_in, d_out, n);
    if (st != APSEP_OK) return st;
    return noalloc_run(d_in, d_out, (int)n, incl, (hipStream_t)stream);
}

apsep_status apsep_time_kernels(apsep_scratch* s, const int32_t* d_in, int32_t* d_out,
                                int64_t n, int incl, int reps, void* stream,
                                apsep_kernel_times* out) {
    if (!s || !out || reps < 1 || (incl != 0 && incl != 1)) return APSEP_ERR_INVALID_ARG;
    memset(out, 0, sizeof *out);
    if (n <= 0) return APSEP_OK;
    apsep_status st = check_io(d_in, d_out, n);
    if (st != APSEP_OK) return st;
    if (n > s->n_cap) return APSEP_ERR_CAPACITY;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != s->device) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_DEVICE;
    }
    double ms[3] = {0, 0, 0};
    st = single_time_kernels(&s, &d_in, &d_out, 1, (int)n, incl, reps, (hipStream_t)stream, ms);
    if (st != APSEP_OK) return st;
    out->runs = reps;
    out->tile_ms = ms[0];
    out->levels_ms = ms[1];
    out->fix_ms = ms[2];
    return APSEP_OK;
}

apsep_status apsep_time_kernels_rotating(apsep_scratch* const* s, const int32_t* const* d_in,
                                         int32_t* const* d_out, int32_t count, int64_t n,
                                         int incl, int reps, void* stream,
                                         apsep_kernel_times* out) {
    if (!s || !d_in || !d_out || !out || count < 1 || count > 8 || reps < 1 ||
        (incl != 0 && incl != 1))
        return APSEP_ERR_INVALID_ARG;
    memset(out, 0, sizeof *out);
    if (n <= 0) return APSEP_OK;
    int dev = -1;
    for (int k = 0; k < count; k++) {
        if (!s[k] || s[k]->algo != s[0]->algo) return APSEP_ERR_INVALID_ARG;
        const apsep_status st = check_io(d_in[k], d_out[k], n);
        if (st != APSEP_OK) return st;
        if (n > s[k]->n_cap) return APSEP_ERR_CAPACITY;
    }
    if (hipGetDevice(&dev) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_DEVICE;
    }
    for (int k = 0; k < count; k++)
        if (s[k]->device != dev) return APSEP_ERR_DEVICE;
    double ms[3] = {0, 0, 0};
    const apsep_status st = single_time_kernels(s, d_in, d_out, count, (int)n, incl, reps,
                                                (hipStream_t)stream, ms);
    if (st != APSEP_OK) return st;
    out->runs = reps;
    out->tile_ms = ms[0];
    out->levels_ms = ms[1];
    out->fix_ms = ms[2];
    return APSEP_OK;
}

// ---- segmented batches ----------------------------------------------------

apsep_status apsep_batch_plan_create(int64_t total_cap, int32_t nseg_cap, int incl,
                                     apsep_batch_plan** out) {
    if (!out) return APSEP_ERR_INVALID_ARG;
    *out = nullptr;
    if (total_cap < 0 || nseg_cap < 0 || (incl != 0 && incl != 1)) return APSEP_ERR_INVALID_ARG;
    if (total_cap > INT32_MAX) return APSEP_ERR_TOO_LARGE;
    apsep_batch_plan* p = new (std::nothrow) apsep_batch_plan();  // zeroed
    if (!p) return APSEP_ERR_ALLOC;
    p->total_cap = total_cap;
    p->nseg_cap = nseg_cap;
    // sized for the smallest tile, which covers every shape: as many tile
    // slots or more, the same chunk-min count (one per 64 keys)
    const int tile = seg128::seg_tile_elems();
    p->tile_lanes = 256;  // until a batch picks its shape by size (or the caller / tuner does)
    apsep_status st = scratch_create(APSEP_ALGO_WSTL, total_cap, incl, tile, &p->s);
    if (st != APSEP_OK) {
        delete p;
        return st;
    }
    const int64_t tiles = (total_cap + tile - 1) / tile;
    if (tiles > 0 && (hipMalloc((void**)&p->tmap, (size_t)tiles * sizeof(int)) != hipSuccess ||
                      hipMalloc((void**)&p->tbound, (size_t)tiles * sizeof(int2)) != hipSuccess ||
                      hipMalloc((void**)&p->pages, (size_t)tiles * sizeof(Page)) != hipSuccess)) {
        APSEP_CLEAR_HIP_ERROR();
        if (p->tmap) (void)hipFree(p->tmap);
        if (p->tbound) (void)hipFree(p->tbound);
        apsep_scratch_destroy(p->s);
        delete p;
        return APSEP_ERR_ALLOC;
    }
    *out = p;
    return APSEP_OK;
}

static void drop_graph(apsep_batch_plan* p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
}

apsep_status apsep_batch_plan_destroy(apsep_batch_plan* p) {
    if (!p) return APSEP_OK;
    drop_graph(p);
    if (p->tmap) (void)hipFree(p->tmap);
    if (p->tbound) (void)hipFree(p->tbound);
    if (p->pages) (void)hipFree(p->pages);
    apsep_scratch_destroy(p->s);
    delete p;
    return APSEP_OK;
}

static apsep_status plan_run(apsep_batch_plan* p, const int32_t* d_in, const int64_t* offs,
                             int32_t* d_out, int nseg, int total, hipStream_t st) {
    return p->tile_lanes == 512   ? seg512::seg_run(p, d_in, offs, d_out, nseg, total, st)
           : p->tile_lanes == 128 ? seg128::seg_run(p, d_in, offs, d_out, nseg, total, st)
                                  : seg256::seg_run(p, d_in, offs, d_out, nseg, total, st);
}

static apsep_status plan_run_jobs(apsep_batch_plan* p, const apsep_job* jobs, int count,
                                  int tiles, hipStream_t st) {
    return p->tile_lanes == 512   ? seg512::jobs_run(p, jobs, count, tiles, st)
           : p->tile_lanes == 128 ? seg128::jobs_run(p, jobs, count, tiles, st)
                                  : seg256::jobs_run(p, jobs, count, tiles, st);
}

// Replay the plan's graph when it was recorded for the same call (kind 0:
// apsep_run_segmented, 1: apsep_run_jobs; same pointers, counts and
// stream), else record it first.
static apsep_status graph_run(apsep_batch_plan* p, int kind, const void* a, const void* b,
                              void* c, int32_t count, int64_t total, void* stream) {
    hipStream_t hs = (hipStream_t)stream;
    const bool same = p->exec && p->g_jobs == kind && p->g_in == a && p->g_offs == b &&
                      p->g_out == c && p->g_nseg == count && p->g_total == total &&
                      p->g_stream == stream;
    if (!same) {
        drop_graph(p);
        if (hipStreamBeginCapture(hs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            APSEP_CLEAR_HIP_ERROR();
            return APSEP_ERR_LAUNCH;
        }
        const apsep_status st =
            kind ? plan_run_jobs(p, (const apsep_job*)a, count, (int)total, hs)
                 : plan_run(p, (const int32_t*)a, (const int64_t*)b, (int32_t*)c, count,
                            (int)total, hs);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(hs, &g);
        if (st != APSEP_OK || ec != hipSuccess || !g) {
            APSEP_CLEAR_HIP_ERROR();
            if (g) (void)hipGraphDestroy(g);
            return APSEP_ERR_LAUNCH;
        }
        p->graph = g;
        if (hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            APSEP_CLEAR_HIP_ERROR();
            drop_graph(p);
            return APSEP_ERR_LAUNCH;
        }
        p->g_jobs = kind;
        p->g_in = a;
        p->g_offs = b;
        p->g_out = c;
        p->g_nseg = count;
        p->g_total = total;
        p->g_stream = stream;
    }
    if (hipGraphLaunch(p->exec, hs) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_LAUNCH;
    }
    return APSEP_OK;
}

static void set_lanes(apsep_batch_plan* p, int lanes) {
    if (lanes != p->tile_lanes) drop_graph(p);  // a recorded graph holds the other shape
    p->tile_lanes = lanes;
}

apsep_status apsep_batch_plan_set_tile_lanes(apsep_batch_plan* p, int32_t lanes) {
    if (!p || (lanes != 128 && lanes != 256 && lanes != 512)) return APSEP_ERR_INVALID_ARG;
    set_lanes(p, lanes);
    p->tuned_auto = 0;  // an explicit shape overrides the tuned table
    p->lanes_explicit = 1;
    return APSEP_OK;
}

apsep_status apsep_batch_plan_tile_lanes(const apsep_batch_plan* p, int32_t* lanes) {
    if (!p || !lanes) return APSEP_ERR_INVALID_ARG;
    *lanes = p->tile_lanes;
    return APSEP_OK;
}

static int plan_tile_elems(const apsep_batch_plan* p) {
    return p->tile_lanes == 512   ? seg512::seg_tile_elems()
           : p->tile_lanes == 128 ? seg128::seg_tile_elems()
                                  : seg256::seg_tile_elems();
}

// Size class of a batch of `total` keys: c with 2^(c-1) < total <= 2^c.
static int size_class(int64_t total) {
    return total <= 1 ? 0 : 64 - __builtin_clzll((unsigned long long)(total - 1));
}

apsep_status apsep_batch_plan_tuned(const apsep_batch_plan* p, int32_t size_class_log2,
                                    int32_t* lanes, int32_t* graph, float* ms) {
    if (!p || !lanes || !graph || size_class_log2 < 0 || size_class_log2 > 31)
        return APSEP_ERR_INVALID_ARG;
    *lanes = p->cls_lanes[size_class_log2];
    *graph = p->cls_graph[size_class_log2];
    if (ms) *ms = p->cls_ms[size_class_log2];
    return APSEP_OK;
}

apsep_status apsep_batch_plan_tuned_persistent(const apsep_batch_plan* p,
                                               int32_t size_class_log2, int32_t* persistent) {
    if (!p || !persistent || size_class_log2 < 0 || size_class_log2 > 31)
        return APSEP_ERR_INVALID_ARG;
    *persistent = p->cls_persist[size_class_log2];
    return APSEP_OK;
}

apsep_status apsep_batch_plan_tile_elems(const apsep_batch_plan* p, int32_t* elems) {
    if (!p || !elems) return APSEP_ERR_INVALID_ARG;
    *elems = plan_tile_elems(p);
    return APSEP_OK;
}

// Tiles a job batch of this plan may span with its current tile shape: the
// scratch covers total_cap keys in tiles of either shape.
static int64_t plan_job_tiles(const apsep_batch_plan* p) {
    const int64_t te = plan_tile_elems(p);
    return (p->total_cap + te - 1) / te;
}

apsep_status apsep_batch_plan_job_tiles(const apsep_batch_plan* p, int64_t* tiles) {
    if (!p || !tiles) return APSEP_ERR_INVALID_ARG;
    *tiles = plan_job_tiles(p);
    return APSEP_OK;
}

apsep_status apsep_batch_plan_use_graph(apsep_batch_plan* p, int enable) {
    if (!p || (enable != 0 && enable != 1)) return APSEP_ERR_INVALID_ARG;
    if (!enable) drop_graph(p);
    p->use_graph = enable;
    p->tuned_auto = 0;  // an explicit mode overrides the tuned table
    return APSEP_OK;
}

static void set_persistent(apsep_batch_plan* p, int enable) {
    if (enable != p->persistent) drop_graph(p);  // a recorded graph holds the other sequence
    p->persistent = enable;
}

apsep_status apsep_batch_plan_set_persistent(apsep_batch_plan* p, int enable) {
    if (!p || (enable != 0 && enable != 1)) return APSEP_ERR_INVALID_ARG;
    set_persistent(p, enable);
    p->tuned_auto = 0;  // an explicit mode overrides the tuned table
    return APSEP_OK;
}

apsep_status apsep_batch_plan_check(apsep_batch_plan* p, int32_t* device_error) {
    if (!p || !device_error) return APSEP_ERR_INVALID_ARG;
    int code = 0;
    const apsep_status st = apsep::single_error(p->s, &code);
    *device_error = code;
    return st;
}

apsep_status apsep_run_segmented(apsep_batch_plan* p, const int32_t* d_in,
                                 const int64_t* d_offsets, int32_t* d_out, int32_t nseg,
                                 int64_t total, void* stream) {
    if (!p || nseg < 0 || total < 0) return APSEP_ERR_INVALID_ARG;
    if (total == 0) return APSEP_OK;
    if (nseg == 0 || !d_offsets) return APSEP_ERR_INVALID_ARG;  // elements outside every segment
    if (total > INT32_MAX) return APSEP_ERR_TOO_LARGE;
    if (nseg > p->nseg_cap || total > p->total_cap) return APSEP_ERR_CAPACITY;
    apsep_status st = check_io(d_in, d_out, total);
    if (st != APSEP_OK) return st;
    if ((uintptr_t)d_offsets & 7) return APSEP_ERR_ALIGN;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != p->s->device) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_DEVICE;
    }
    hipStream_t hs = (hipStream_t)stream;
    const int c = size_class(total);
    if (p->tuned_auto && p->cls_lanes[c]) {
        // the tuned shape and launch mode of this batch's size class
        set_lanes(p, p->cls_lanes[c]);
        set_persistent(p, p->cls_persist[c]);
        if (!p->cls_graph[c]) drop_graph(p);
        p->use_graph = p->cls_graph[c];
    } else if (!p->lanes_explicit) {
        set_lanes(p, default_batch_lanes(total));
    }
    if (!p->use_graph || !stream)
        return plan_run(p, d_in, d_offsets, d_out, nseg, (int)total, hs);
    return graph_run(p, 0, d_in, d_offsets, d_out, nseg, total, stream);
}

apsep_status apsep_run_jobs(apsep_batch_plan* p, const apsep_job* d_jobs, int32_t count,
                            int64_t tiles, void* stream) {
    if (!p || count < 0 || tiles < 0) return APSEP_ERR_INVALID_ARG;
    if (count == 0 || tiles == 0) return APSEP_OK;
    if (!d_jobs || ((uintptr_t)d_jobs & 7)) return !d_jobs ? APSEP_ERR_INVALID_ARG : APSEP_ERR_ALIGN;
    if (count > p->nseg_cap || tiles > plan_job_tiles(p)) return APSEP_ERR_CAPACITY;
    if (tiles * plan_tile_elems(p) > INT32_MAX) return APSEP_ERR_TOO_LARGE;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != p->s->device) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_DEVICE;
    }
    hipStream_t hs = (hipStream_t)stream;
    if (!p->use_graph || !stream) return plan_run_jobs(p, d_jobs, count, (int)tiles, hs);
    return graph_run(p, 1, d_jobs, nullptr, nullptr, count, tiles, stream);
}

apsep_status apsep_batch_plan_tune(apsep_batch_plan* p, const int32_t* d_in,
                                   const int64_t* d_offsets, int32_t* d_out, int32_t nseg,
                                   int64_t total, void* stream, int32_t reps,
                                   float* eager_ms, float* graph_ms) {
    if (!p || reps < 1) return APSEP_ERR_INVALID_ARG;
    if (!stream) return APSEP_ERR_INVALID_ARG;  // graphs need a capturable stream
    hipStream_t hs = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_DEVICE;
    }
    if (hipEventCreate(&e1) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        (void)hipEventDestroy(e0);
        return APSEP_ERR_DEVICE;
    }
    const int explicit_before = p->lanes_explicit;
    // every tile shape x launch sequence (kernel sequence or one persistent
    // launch) x launch mode (direct or graph replay), timed on this batch
    constexpr int NSHAPES = 3;
    const int32_t shapes[NSHAPES] = {128, 256, 512};
    float t[NSHAPES][2][2] = {};
    apsep_status st = APSEP_OK;
    for (int k = 0; k < NSHAPES && st == APSEP_OK; k++) {
        st = apsep_batch_plan_set_tile_lanes(p, shapes[k]);
        for (int q = 0; q < 2 && st == APSEP_OK; q++) {
            st = apsep_batch_plan_set_persistent(p, q);
            for (int g = 0; g < 2 && st == APSEP_OK; g++) {
                st = apsep_batch_plan_use_graph(p, g);
                // one untimed call (records the graph), then the timed ones
                if (st == APSEP_OK)
                    st = apsep_run_segmented(p, d_in, d_offsets, d_out, nseg, total, stream);
                if (st == APSEP_OK && hipEventRecord(e0, hs) != hipSuccess) st = APSEP_ERR_LAUNCH;
                for (int r = 0; r < reps && st == APSEP_OK; r++)
                    st = apsep_run_segmented(p, d_in, d_offsets, d_out, nseg, total, stream);
                if (st == APSEP_OK &&
                    (hipEventRecord(e1, hs) != hipSuccess ||
                     hipEventSynchronize(e1) != hipSuccess ||
                     hipEventElapsedTime(&t[k][q][g], e0, e1) != hipSuccess))
                    st = APSEP_ERR_LAUNCH;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != APSEP_OK) {
        APSEP_CLEAR_HIP_ERROR();
        return st;
    }
    int bk = 0, bq = 0, bg = 0;
    for (int k = 0; k < NSHAPES; k++)
        for (int q = 0; q < 2; q++)
            for (int g = 0; g < 2; g++)
                if (t[k][q][g] < t[bk][bq][bg]) bk = k, bq = q, bg = g;
    if (eager_ms) *eager_ms = t[bk][bq][0] / reps;
    if (graph_ms) *graph_ms = t[bk][bq][1] / reps;
    st = apsep_batch_plan_set_tile_lanes(p, shapes[bk]);
    if (st == APSEP_OK) st = apsep_batch_plan_set_persistent(p, bq);
    // d_out holds the last setting's answer: bit-identical across settings
    if (st == APSEP_OK) st = apsep_batch_plan_use_graph(p, bg);
    if (st != APSEP_OK) return st;
    // remembered for the batch's size class; later batches of a tuned class
    // run with its settings
    const int c = size_class(total);
    p->cls_lanes[c] = shapes[bk];
    p->cls_persist[c] = bq;
    p->cls_graph[c] = bg;
    p->cls_ms[c] = t[bk][bq][bg] / reps;
    p->tuned_auto = 1;
    p->lanes_explicit = explicit_before;
    return APSEP_OK;
}

apsep_status apsep_launch(apsep_algo algo, const int32_t* d_in, int32_t* d_out,
                          int64_t n, int incl) {
    if (!valid_algo(algo) || (incl != 0 && incl != 1)) return APSEP_ERR_INVALID_ARG;
    if (n <= 0) return APSEP_OK;
    apsep_status st = check_io(d_in, d_out, n);
    if (st != APSEP_OK) return st;
    apsep_scratch* s = nullptr;
    st = apsep_scratch_create(algo, n, incl, &s);
    if (st != APSEP_OK) return st;
    st = apsep_run(s, d_in, d_out, n, nullptr);
    if (st == APSEP_OK && hipDeviceSynchronize() != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        st = APSEP_ERR_LAUNCH;
    }
    if (st == APSEP_OK) {
        int32_t derr = 0;
        st = apsep_scratch_check(s, &derr);
        if (st == APSEP_OK && derr) {
            record_device_error(derr);
            st = APSEP_ERR_LAUNCH;
        }
    }
    apsep_scratch_destroy(s);
    return st;
}

}  // extern "C"
