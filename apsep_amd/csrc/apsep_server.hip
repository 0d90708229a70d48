// apsep_server.hip -- host-side serving engine: a fixed HBM pool, budget
// admission, and continuous batching of independent solve jobs into
// segmented launches.
//
// The reference serves nothing: each launchWSTL / launchSPT call allocates
// its scratch, solves one array and synchronises the device
// (src/apsep.cuh:2598-2605, 3069-3076).  Here:
//   pool       one hipMalloc split into a key ring and an answer ring with
//              the same indexing; a job owns a contiguous slot in both,
//              allocated in submission order (16-byte aligned)
//   admission  a job is accepted only if its slot fits in the free part of
//              the ring (APSEP_ERR_BUDGET otherwise: nothing is queued, the
//              caller retries after releasing finished jobs)
//   batching   poll() retires finished batches (event queries, never a
//              blocking wait) and, while fewer than INFLIGHT batches are
//              queued on the stream, packs every queued job whose slot
//              continues the previous one into one segmented solve: the
//              batch's offsets and job descriptors are uploaded from pinned
//              memory in one copy, one gather kernel copies every job's keys
//              into its slot, and the batch is solved in place
// Padding between slots (< 4 keys) is absorbed by the segment before it:
// answers depend only on earlier keys, so a job's answers do not see it.
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

constexpr int INFLIGHT = 2;

enum JobState { JOB_QUEUED = 0, JOB_RUNNING = 1, JOB_DONE = 2 };

struct Job {
    const int32_t* src;
    int64_t n;
    int64_t slot;   // element offset in both rings
    int64_t span;   // slot length (n rounded up to 4)
    int state;
    bool released;
    double t_submit, t_done;  // microseconds, steady clock
};

// Copy every job of a batch into its slot: one workgroup per GATHER keys of
// the batch, the jobs overlapping it found by binary search over the slots.
constexpr int GATHER = 4096;

__global__ __launch_bounds__(256) void k_gather(const long long* __restrict__ off,
                                                const unsigned long long* __restrict__ src,
                                                const long long* __restrict__ len, int m,
                                                int* __restrict__ dst, long long total) {
    const long long p0 = (long long)blockIdx.x * GATHER;
    if (p0 >= total) return;
    const long long p1 = p0 + GATHER < total ? p0 + GATHER : total;
    int lo = 0, hi = m - 1;  // last job starting at or before p0
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= p0) lo = mid;
        else hi = mid - 1;
    }
    for (int k = lo; k < m && off[k] < p1; k++) {
        const long long a = off[k] > p0 ? off[k] : p0;
        const long long e = off[k] + len[k];
        const long long b = e < p1 ? e : p1;
        const int* s = reinterpret_cast<const int*>(src[k]) - off[k];
        for (long long p = a + threadIdx.x; p < b; p += blockDim.x) dst[p] = s[p];
    }
}

struct Inflight {
    bool busy;
    hipEvent_t done;
    int64_t* h_off;   // pinned: offsets [m + 1], sources [m], lengths [m]
    int64_t* d_off;
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
    hipStream_t stream;
    int32_t* ring_in;
    int32_t* ring_out;
    int64_t cap;          // ring elements
    int64_t max_batch;    // elements per batch
    int32_t max_segs;     // jobs per batch
    apsep_batch_plan* plan;
    std::deque<Job> jobs; // unreleased jobs, in submission (= ring) order
    int64_t front_id;     // id of jobs.front()
    int64_t next_id;
    int64_t next_launch;  // id of the first job not yet in a batch
    Inflight inf[INFLIGHT];
    int inf_head;         // oldest in-flight batch
    int inf_count;
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
    for (int i = 0; i < INFLIGHT; i++) {
        if (s->inf[i].done) (void)hipEventDestroy(s->inf[i].done);
        if (s->inf[i].h_off) (void)hipHostFree(s->inf[i].h_off);
        if (s->inf[i].d_off) (void)hipFree(s->inf[i].d_off);
    }
    if (s->plan) apsep_batch_plan_destroy(s->plan);
    if (s->ring_in) (void)hipFree(s->ring_in);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

// retire finished batches in order (non-blocking unless `wait`)
static apsep_status retire(apsep_server* s, bool wait) {
    while (s->inf_count > 0) {
        Inflight& f = s->inf[s->inf_head];
        hipError_t e = wait ? hipEventSynchronize(f.done) : hipEventQuery(f.done);
        if (e == hipErrorNotReady) {
            APSEP_CLEAR_HIP_ERROR();
            return APSEP_OK;
        }
        if (e != hipSuccess) {
            APSEP_CLEAR_HIP_ERROR();
            return APSEP_ERR_LAUNCH;
        }
        const double t = now_us();
        for (int64_t id = f.first; id < f.first + f.count; id++) {
            Job* j = job_of(s, id);
            if (j && j->state == JOB_RUNNING) {
                j->state = JOB_DONE;
                j->t_done = t;
            }
        }
        f.busy = false;
        s->inf_head = (s->inf_head + 1) % INFLIGHT;
        s->inf_count--;
        wait = false;  // block for at most one batch
    }
    return APSEP_OK;
}

// pack the next run of queued jobs into one segmented solve
static apsep_status launch_batch(apsep_server* s) {
    while (s->next_launch < s->next_id && job_of(s, s->next_launch)->n == 0)
        s->next_launch++;  // empty jobs are done at submission
    if (s->next_launch >= s->next_id || s->inf_count >= INFLIGHT) return APSEP_OK;
    Inflight& f = s->inf[(s->inf_head + s->inf_count) % INFLIGHT];
    const int64_t first = s->next_launch;
    const Job* j0 = job_of(s, first);
    const int64_t start = j0->slot;
    int64_t end = start;  // end of the last job's slot
    int32_t m = 0;
    for (int64_t id = first; id < s->next_id && m < s->max_segs; id++) {
        const Job* j = job_of(s, id);
        if (j->slot != end && m > 0) break;  // wrapped: next batch
        if (j->slot + j->n - start > s->max_batch) break;  // the first job always fits
        f.h_off[m] = j->slot - start;
        end = j->slot + j->span;
        m++;
    }
    const Job* jl = job_of(s, first + m - 1);
    const int64_t total = jl->slot + jl->n - start;
    f.h_off[m] = total;
    uint64_t* h_src = reinterpret_cast<uint64_t*>(f.h_off + m + 1);
    int64_t* h_len = f.h_off + 2 * m + 1;
    for (int32_t k = 0; k < m; k++) {
        Job* j = job_of(s, first + k);
        h_src[k] = (uint64_t)(uintptr_t)j->src;
        h_len[k] = j->n;
        if (j->n > 0) j->state = JOB_RUNNING;
    }
    if (hipMemcpyAsync(f.d_off, f.h_off, (size_t)(3 * m + 1) * 8, hipMemcpyHostToDevice,
                       s->stream) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_LAUNCH;
    }
    const long long nblk = (total + GATHER - 1) / GATHER;
    hipLaunchKernelGGL(k_gather, dim3((unsigned)nblk), dim3(256), 0, s->stream,
                       (const long long*)f.d_off, (const unsigned long long*)(f.d_off + m + 1),
                       (const long long*)(f.d_off + 2 * m + 1), (int)m, s->ring_in + start,
                       (long long)total);
    if (APSEP_LAUNCH_FAILED()) return APSEP_ERR_LAUNCH;
    apsep_status st = apsep_run_segmented(s->plan, s->ring_in + start, f.d_off,
                                          s->ring_out + start, m, total, s->stream);
    if (st != APSEP_OK) return st;
    if (hipEventRecord(f.done, s->stream) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_LAUNCH;
    }
    f.busy = true;
    f.first = first;
    f.count = m;
    s->inf_count++;
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
    s->cap = (pool_bytes / 8) & ~(int64_t)3;  // keys + answers, 4 B each
    s->max_batch = max_batch < s->cap ? max_batch : s->cap;
    s->max_segs = max_batch_jobs;
    if (hipGetDevice(&s->device) != hipSuccess ||
        hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        destroy_server(s);
        return APSEP_ERR_DEVICE;
    }
    if (hipMalloc((void**)&s->ring_in, (size_t)s->cap * 8) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        s->ring_in = nullptr;
        destroy_server(s);
        return APSEP_ERR_ALLOC;
    }
    s->ring_out = s->ring_in + s->cap;
    apsep_status st = apsep_batch_plan_create(s->max_batch, max_batch_jobs, incl, &s->plan);
    if (st != APSEP_OK) {
        s->plan = nullptr;
        destroy_server(s);
        return st;
    }
    for (int i = 0; i < INFLIGHT; i++) {
        Inflight& f = s->inf[i];
        if (hipEventCreateWithFlags(&f.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void**)&f.h_off, (size_t)(3 * max_batch_jobs + 1) * 8, 0) != hipSuccess ||
            hipMalloc((void**)&f.d_off, (size_t)(3 * max_batch_jobs + 1) * 8) != hipSuccess) {
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
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    destroy_server(s);
    return APSEP_OK;
}

apsep_status apsep_server_submit(apsep_server* s, const int32_t* d_keys, int64_t n,
                                 int64_t* job_id) {
    if (!s || !job_id || n < 0 || (n > 0 && !d_keys)) return APSEP_ERR_INVALID_ARG;
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
    while (s->inf_count < INFLIGHT && s->next_launch < s->next_id) {
        const int before = s->inf_count;
        st = launch_batch(s);
        if (st != APSEP_OK) return st;
        if (s->inf_count == before) break;
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
        if (s->inf_count == 0 && s->next_launch >= s->next_id) return APSEP_OK;
        st = retire(s, true);
        if (st != APSEP_OK) return st;
    }
}

}  // extern "C"
