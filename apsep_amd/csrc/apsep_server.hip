// apsep_server.hip -- host-side serving engine: a fixed HBM pool, budget
// admission, and continuous batching of independent solve jobs into
// in-place job batches on several streams.
//
// The reference serves nothing: each launchWSTL / launchSPT call allocates
// its scratch, solves one array and synchronises the device
// (src/apsep.cuh:2598-2605, 3069-3076).  Here:
//   pool       one hipMalloc holding the answers: a job owns a contiguous,
//              16-byte aligned slot of it, allocated in submission order
//              (a ring); its keys stay where the client holds them and are
//              read in place (no copy into the pool)
//   admission  a job is accepted only if its slot fits in the free part of
//              the ring (APSEP_ERR_BUDGET otherwise: nothing is queued, the
//              caller retries after releasing finished jobs)
//   batching   poll() retires finished batches (event queries, never a
//              blocking wait) and, while a stream has a free batch slot,
//              packs the queued jobs into one in-place job batch on it:
//              the batch's job table is uploaded from pinned memory in one
//              copy and solved by apsep_run_jobs with that stream's plan
//   streams    NSTREAMS streams, each with its own plan (scratch) and up to
//              INFLIGHT batches, so consecutive batches overlap on the GPU
// One thread drives a server; servers are independent.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <chrono>
#include <deque>
#include <new>

#include "apsep_internal.hpp"
#include "../../include/apsep/apsep.h"

namespace {

constexpr int NSTREAMS = 2;
constexpr int INFLIGHT = 2;        // batches queued per stream
constexpr int SLOTS = NSTREAMS * INFLIGHT;
constexpr int64_t MAX_TILE = 8192; // the larger tile shape, in keys

enum JobState { JOB_QUEUED = 0, JOB_RUNNING = 1, JOB_DONE = 2 };

struct Job {
    const int32_t* src;
    int64_t n;
    int64_t slot;   // element offset in the answer ring
    int64_t span;   // slot length (n rounded up to 4)
    int state;
    bool released;
    double t_submit, t_done;  // microseconds, steady clock
};

struct Batch {
    bool busy;
    int stream;       // index into the server's streams
    hipEvent_t done;
    apsep_job* h_jobs;  // pinned
    apsep_job* d_jobs;
    int64_t first;    // job id of the batch's first job
    int32_t count;
};

double now_us() {
    using namespace std::chrono;
    return duration<double, std::micro>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct apsep_server {
    int device;
    int incl;
    hipStream_t stream[NSTREAMS];
    apsep_batch_plan* plan[NSTREAMS];
    int inflight[NSTREAMS];   // batches queued per stream
    int next_stream;          // round-robin start
    int32_t* ring_out;
    int64_t cap;          // ring elements
    int64_t max_batch;    // keys per batch
    int32_t max_segs;     // jobs per batch
    std::deque<Job> jobs; // unreleased jobs, in submission (= ring) order
    int64_t front_id;     // id of jobs.front()
    int64_t next_id;
    int64_t next_launch;  // id of the first job not yet in a batch
    Batch batch[SLOTS];
    int busy;             // batches in flight
};

static Job* job_of(apsep_server* s, int64_t id) {
    if (id < s->front_id || id >= s->next_id) return nullptr;
    return &s->jobs[(size_t)(id - s->front_id)];
}

// Slot for `span` elements in the free part of the ring, or -1.
static int64_t ring_alloc(apsep_server* s, int64_t span) {
    if (s->jobs.empty()) return span <= s->cap ? 0 : -1;
    const int64_t head = s->jobs.front().slot;
    const Job& b = s->jobs.back();
    const int64_t tail = b.slot + b.span;
    if (b.slot >= head) {  // not wrapped: [tail, cap) then [0, head)
        if (tail + span <= s->cap) return tail;
        if (span <= head) return 0;
        return -1;
    }
    return (tail + span <= head) ? tail : -1;  // wrapped: [tail, head)
}

static void destroy_server(apsep_server* s) {
    for (int i = 0; i < SLOTS; i++) {
        if (s->batch[i].done) (void)hipEventDestroy(s->batch[i].done);
        if (s->batch[i].h_jobs) (void)hipHostFree(s->batch[i].h_jobs);
        if (s->batch[i].d_jobs) (void)hipFree(s->batch[i].d_jobs);
    }
    for (int k = 0; k < NSTREAMS; k++) {
        if (s->plan[k]) apsep_batch_plan_destroy(s->plan[k]);
        if (s->stream[k]) (void)hipStreamDestroy(s->stream[k]);
    }
    if (s->ring_out) (void)hipFree(s->ring_out);
    delete s;
}

static void finish(apsep_server* s, Batch& f) {
    const double t = now_us();
    for (int64_t id = f.first; id < f.first + f.count; id++) {
        Job* j = job_of(s, id);
        if (j && j->state == JOB_RUNNING) {
            j->state = JOB_DONE;
            j->t_done = t;
        }
    }
    f.busy = false;
    s->inflight[f.stream]--;
    s->busy--;
}

// Retire finished batches, in any order (non-blocking unless `wait`: then
// block for the oldest batch first).
static apsep_status retire(apsep_server* s, bool wait) {
    if (wait && s->busy > 0) {
        int oldest = -1;
        for (int i = 0; i < SLOTS; i++)
            if (s->batch[i].busy && (oldest < 0 || s->batch[i].first < s->batch[oldest].first))
                oldest = i;
        if (hipEventSynchronize(s->batch[oldest].done) != hipSuccess) {
            APSEP_CLEAR_HIP_ERROR();
            return APSEP_ERR_LAUNCH;
        }
        finish(s, s->batch[oldest]);
    }
    for (int i = 0; i < SLOTS; i++) {
        Batch& f = s->batch[i];
        if (!f.busy) continue;
        const hipError_t e = hipEventQuery(f.done);
        if (e == hipErrorNotReady) {
            (void)hipGetLastError();  // not a failure
            continue;
        }
        if (e != hipSuccess) {
            APSEP_CLEAR_HIP_ERROR();
            return APSEP_ERR_LAUNCH;
        }
        finish(s, f);
    }
    return APSEP_OK;
}

// Pack the next queued jobs (in submission order) into one in-place batch on
// the next stream with a free slot.
static apsep_status launch_batch(apsep_server* s) {
    while (s->next_launch < s->next_id && job_of(s, s->next_launch)->n == 0)
        s->next_launch++;  // empty jobs are done at submission
    if (s->next_launch >= s->next_id) return APSEP_OK;
    int k = -1;
    for (int i = 0; i < NSTREAMS && k < 0; i++) {
        const int c = (s->next_stream + i) % NSTREAMS;
        if (s->inflight[c] < INFLIGHT) k = c;
    }
    if (k < 0) return APSEP_OK;
    int slot = 0;
    while (s->batch[slot].busy) slot++;
    Batch& f = s->batch[slot];
    const int64_t first = s->next_launch;
    // the batch's keys, then the tile shape for that size
    int64_t keys = 0;
    int32_t m = 0;
    for (int64_t id = first; id < s->next_id && m < s->max_segs; id++) {
        const Job* j = job_of(s, id);
        if (m > 0 && keys + j->n > s->max_batch) break;
        keys += j->n;
        m++;
    }
    int32_t te = 0;
    int64_t tile_cap = 0;
    if (apsep_batch_plan_set_tile_lanes(s->plan[k], apsep::default_batch_lanes(keys)) != APSEP_OK ||
        apsep_batch_plan_tile_elems(s->plan[k], &te) != APSEP_OK ||
        apsep_batch_plan_job_tiles(s->plan[k], &tile_cap) != APSEP_OK)
        return APSEP_ERR_INVALID_ARG;
    const int32_t m_max = m;
    int64_t tiles = 0;
    keys = 0;
    m = 0;
    for (int64_t id = first; id < s->next_id && m < m_max; id++) {
        const Job* j = job_of(s, id);
        const int64_t jt = (j->n + te - 1) / te;
        if (m > 0 && (keys + j->n > s->max_batch || tiles + jt > tile_cap)) break;
        f.h_jobs[m] = apsep_job{j->src, s->ring_out + j->slot, j->n, tiles};
        keys += j->n;
        tiles += jt;
        m++;
    }
    for (int32_t i = 0; i < m; i++) {
        Job* j = job_of(s, first + i);
        if (j->n > 0) j->state = JOB_RUNNING;
    }
    if (hipMemcpyAsync(f.d_jobs, f.h_jobs, (size_t)m * sizeof(apsep_job), hipMemcpyHostToDevice,
                       s->stream[k]) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_LAUNCH;
    }
    apsep_status st = apsep_run_jobs(s->plan[k], f.d_jobs, m, tiles, s->stream[k]);
    if (st != APSEP_OK) return st;
    if (hipEventRecord(f.done, s->stream[k]) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_LAUNCH;
    }
    f.busy = true;
    f.stream = k;
    f.first = first;
    f.count = m;
    s->inflight[k]++;
    s->busy++;
    s->next_stream = (k + 1) % NSTREAMS;
    s->next_launch = first + m;
    return APSEP_OK;
}

extern "C" {

apsep_status apsep_server_create(int64_t pool_bytes, int64_t max_batch, int32_t max_batch_jobs,
                                 int incl, apsep_server** out) {
    if (!out) return APSEP_ERR_INVALID_ARG;
    *out = nullptr;
    if (pool_bytes < 64 || max_batch < 1 || max_batch_jobs < 1 || (incl != 0 && incl != 1))
        return APSEP_ERR_INVALID_ARG;
    if (max_batch > INT32_MAX) return APSEP_ERR_TOO_LARGE;
    apsep_server* s = new (std::nothrow) apsep_server();
    if (!s) return APSEP_ERR_ALLOC;
    s->incl = incl;
    s->cap = (pool_bytes / 4) & ~(int64_t)3;  // answers, 4 B each
    s->max_batch = max_batch < s->cap ? max_batch : s->cap;
    s->max_segs = max_batch_jobs;
    if (hipGetDevice(&s->device) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        destroy_server(s);
        return APSEP_ERR_DEVICE;
    }
    for (int k = 0; k < NSTREAMS; k++)
        if (hipStreamCreateWithFlags(&s->stream[k], hipStreamNonBlocking) != hipSuccess) {
            APSEP_CLEAR_HIP_ERROR();
            s->stream[k] = nullptr;
            destroy_server(s);
            return APSEP_ERR_DEVICE;
        }
    if (hipMalloc((void**)&s->ring_out, (size_t)s->cap * 4) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        s->ring_out = nullptr;
        destroy_server(s);
        return APSEP_ERR_ALLOC;
    }
    // a batch of max_batch keys in max_batch_jobs jobs spans at most
    // max_batch / tile + max_batch_jobs tiles of either shape
    int64_t plan_cap = s->max_batch + (int64_t)max_batch_jobs * MAX_TILE;
    if (plan_cap > INT32_MAX) plan_cap = INT32_MAX;
    for (int k = 0; k < NSTREAMS; k++) {
        const apsep_status st = apsep_batch_plan_create(plan_cap, max_batch_jobs, incl, &s->plan[k]);
        if (st != APSEP_OK) {
            s->plan[k] = nullptr;
            destroy_server(s);
            return st;
        }
    }
    for (int i = 0; i < SLOTS; i++) {
        Batch& f = s->batch[i];
        if (hipEventCreateWithFlags(&f.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void**)&f.h_jobs, (size_t)max_batch_jobs * sizeof(apsep_job), 0) != hipSuccess ||
            hipMalloc((void**)&f.d_jobs, (size_t)max_batch_jobs * sizeof(apsep_job)) != hipSuccess) {
            APSEP_CLEAR_HIP_ERROR();
            destroy_server(s);
            return APSEP_ERR_ALLOC;
        }
    }
    *out = s;
    return APSEP_OK;
}

apsep_status apsep_server_destroy(apsep_server* s) {
    if (!s) return APSEP_OK;
    for (int k = 0; k < NSTREAMS; k++)
        if (s->stream[k]) (void)hipStreamSynchronize(s->stream[k]);
    destroy_server(s);
    return APSEP_OK;
}

apsep_status apsep_server_submit(apsep_server* s, const int32_t* d_keys, int64_t n,
                                 int64_t* job_id) {
    if (!s || !job_id || n < 0 || (n > 0 && !d_keys)) return APSEP_ERR_INVALID_ARG;
    if (n > 0 && ((uintptr_t)d_keys & 15)) return APSEP_ERR_ALIGN;
    if (n > s->max_batch) return APSEP_ERR_CAPACITY;
    const int64_t span = (n + 3) & ~(int64_t)3;
    int64_t slot = 0;
    if (span > 0) {
        slot = ring_alloc(s, span);
        if (slot < 0) return APSEP_ERR_BUDGET;
    } else if (!s->jobs.empty()) {
        const Job& b = s->jobs.back();
        slot = b.slot + b.span;  // empty job: occupies nothing
    }
    Job j;
    j.src = d_keys;
    j.n = n;
    j.slot = slot;
    j.span = span;
    j.released = false;
    j.t_submit = now_us();
    j.state = n == 0 ? JOB_DONE : JOB_QUEUED;
    j.t_done = n == 0 ? j.t_submit : 0.0;
    if (s->jobs.empty()) s->front_id = s->next_id;
    s->jobs.push_back(j);
    *job_id = s->next_id++;
    return APSEP_OK;
}

apsep_status apsep_server_poll(apsep_server* s, int32_t* batches_launched) {
    if (!s) return APSEP_ERR_INVALID_ARG;
    apsep_status st = retire(s, false);
    if (st != APSEP_OK) return st;
    int32_t launched = 0;
    while (s->busy < SLOTS && s->next_launch < s->next_id) {
        const int before = s->busy;
        st = launch_batch(s);
        if (st != APSEP_OK) return st;
        if (s->busy == before) break;
        launched++;
    }
    if (batches_launched) *batches_launched = launched;
    return APSEP_OK;
}

apsep_status apsep_server_job_info(apsep_server* s, int64_t job_id, apsep_job_info* info) {
    if (!s || !info) return APSEP_ERR_INVALID_ARG;
    const Job* j = job_of(s, job_id);
    if (!j || j->released) return APSEP_ERR_INVALID_ARG;
    info->state = j->state;
    info->reserved = 0;
    info->n = j->n;
    info->submit_us = j->t_submit;
    info->done_us = j->t_done;
    info->d_answers = s->ring_out + j->slot;
    return APSEP_OK;
}

apsep_status apsep_server_release(apsep_server* s, int64_t job_id) {
    if (!s) return APSEP_ERR_INVALID_ARG;
    Job* j = job_of(s, job_id);
    if (!j || j->released || j->state != JOB_DONE) return APSEP_ERR_INVALID_ARG;
    j->released = true;
    while (!s->jobs.empty() && s->jobs.front().released) {
        s->jobs.pop_front();
        s->front_id++;
    }
    return APSEP_OK;
}

apsep_status apsep_server_submit_many(apsep_server* s, const int32_t* const* d_keys,
                                      const int64_t* n, int32_t count, int64_t* first_id,
                                      int32_t* accepted) {
    if (!s || !d_keys || !n || count < 0 || !first_id || !accepted) return APSEP_ERR_INVALID_ARG;
    *first_id = s->next_id;
    *accepted = 0;
    for (int32_t i = 0; i < count; i++) {
        int64_t id = 0;
        const apsep_status st = apsep_server_submit(s, d_keys[i], n[i], &id);
        if (st == APSEP_ERR_BUDGET) return APSEP_OK;  // the rest waits for room
        if (st != APSEP_OK) return st;
        (*accepted)++;
    }
    return APSEP_OK;
}

apsep_status apsep_server_done_upto(apsep_server* s, int64_t* done_upto) {
    if (!s || !done_upto) return APSEP_ERR_INVALID_ARG;
    int64_t id = s->front_id;
    while (id < s->next_id && job_of(s, id)->state == JOB_DONE) id++;
    *done_upto = id;
    return APSEP_OK;
}

apsep_status apsep_server_job_times(apsep_server* s, int64_t first_id, int32_t count,
                                    double* submit_us, double* done_us) {
    if (!s || count < 0 || (count > 0 && (!submit_us || !done_us))) return APSEP_ERR_INVALID_ARG;
    for (int32_t i = 0; i < count; i++) {
        const Job* j = job_of(s, first_id + i);
        if (!j || j->released) return APSEP_ERR_INVALID_ARG;
        submit_us[i] = j->t_submit;
        done_us[i] = j->state == JOB_DONE ? j->t_done : 0.0;
    }
    return APSEP_OK;
}

apsep_status apsep_server_release_upto(apsep_server* s, int64_t id) {
    if (!s) return APSEP_ERR_INVALID_ARG;
    for (int64_t k = s->front_id; k < id && k < s->next_id; k++) {
        Job* j = job_of(s, k);
        if (j->state != JOB_DONE) return APSEP_ERR_INVALID_ARG;
        j->released = true;
    }
    while (!s->jobs.empty() && s->jobs.front().released) {
        s->jobs.pop_front();
        s->front_id++;
    }
    return APSEP_OK;
}

apsep_status apsep_server_drain(apsep_server* s) {
    if (!s) return APSEP_ERR_INVALID_ARG;
    for (;;) {
        apsep_status st = apsep_server_poll(s, nullptr);
        if (st != APSEP_OK) return st;
        if (s->busy == 0 && s->next_launch >= s->next_id) return APSEP_OK;
        st = retire(s, true);
        if (st != APSEP_OK) return st;
    }
}

}  // extern "C"
