// apsep_lookback.hip -- single-pass, non-cooperative APSEP for gfx950 with
// decoupled look-back (APSEP_ALGO_LOOKBACK).
//
// The reference's look-back family (src/apsep.cuh:236-254, 1720-1747) chains
// every block behind its predecessor's final prefix.  Here each workgroup
// takes the next tile from a ticket counter, so it only ever waits for
// tiles whose workgroups already run, and:
//   1. solves its tile (the shared tile solver; plain stores, read only by
//      this workgroup);
//   2. publishes the tile min (M), the 64 chunk mins (CM) and, after the
//      look-back, its inclusive prefix-min (S); the last tile of each
//      64- / 4096- / 262144-tile group publishes the group min (G1..G3);
//   3. looks back over its predecessors' S and M words, 64 per wave load,
//      for its exclusive prefix-min P;
//   4. answers its open elements that P beats: the window of the 64
//      previous tile mins, the group mins for older tiles, then the found
//      tile's chunk mins and one 64-key chunk of the input.
// Nothing crosses workgroups through a fence: every published word is a
// 64-bit (run tag << 32 | value) written with one agent-scope atomic store
// and read with agent-scope atomic loads, so a word is either this run's
// complete value or recognisably stale.  A reader that meets a stale word
// polls it (bounded: past the limit the run is flagged as failed and every
// wait ends).  The run tag lives in device memory and is advanced by a
// one-thread kernel before the tile kernel, so the launch sequence holds no
// host-baked state and can be captured into a graph.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "apsep_device.hpp"
#include "apsep_internal.hpp"
#include "../../include/apsep/apsep.h"

namespace apsep {

constexpr unsigned LB_SPIN_LIMIT = 1u << 22;  // polls before a wait gives up (>= 1 s)

__device__ __forceinline__ u64 ld_agent(const u64* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(u64* p, u64 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 tagged(unsigned tag, int v) {
    return ((u64)tag << 32) | (u64)(unsigned)v;
}

struct LbView {
    LbCtl* ctl;
    u64* M;                   // [nt] tile mins
    u64* S;                   // [nt] inclusive prefix-mins
    u64* G[LB_LEVELS];        // G[0] unused; G[l] = min of 64^l tiles
    int gn[LB_LEVELS];        // entries per level (gn[0] = nt)
    u64* CM;                  // [nt * CPT] chunk mins
    unsigned tag;
    int nt;
};

// This run's value of *p, polling while it is stale (bounded).
__device__ __forceinline__ int lb_wait(const LbView& V, const u64* p) {
    u64 w = ld_agent(p);
    unsigned it = 0;
    while ((unsigned)(w >> 32) != V.tag) {
        if (++it > LB_SPIN_LIMIT) {
            __hip_atomic_store(&V.ctl->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return INF;
        }
        if (__hip_atomic_load(&V.ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return INF;
        __builtin_amdgcn_s_sleep(2);
        w = ld_agent(p);
    }
    return (int)(unsigned)w;
}

// (levels picked with selects, not a dynamic index: the view stays in
// registers)
__device__ __forceinline__ const u64* lb_level_ptr(const LbView& V, int lvl) {
    static_assert(LB_LEVELS == 5, "levels");
    return lvl == 0 ? V.M : lvl == 1 ? V.G[1] : lvl == 2 ? V.G[2] : lvl == 3 ? V.G[3] : V.G[4];
}
__device__ __forceinline__ int lb_count(const LbView& V, int lvl) {
    return lvl == 0 ? V.gn[0] : lvl == 1 ? V.gn[1] : lvl == 2 ? V.gn[2] : lvl == 3 ? V.gn[3]
                                                                        : V.gn[4];
}
__device__ __forceinline__ int lb_level(const LbView& V, int lvl, int idx) {
    return lb_wait(V, lb_level_ptr(V, lvl) + idx);
}

// Nearest tile < tb whose min qualifies for v (uniform v; whole wave), or -1.
template <bool INCL>
__device__ int lb_find_tile(const LbView& V, int tb, int v) {
    const int lane = threadIdx.x & (WAVE - 1);
    int x = tb, lvl = 0, found = -1;
    while (found < 0) {
        if (x <= 0) return -1;
        const int gs = ((x - 1) >> LOG_FAN) << LOG_FAN;
        const int idx = gs + lane;
        const bool ok = idx < x && qual<INCL>(lb_level(V, lvl, idx), v);
        const u64 b = __ballot(ok);
        if (b) {
            found = gs + 63 - __clzll(b);
        } else {
            x = (x - 1) >> LOG_FAN;
            if (++lvl >= LB_LEVELS || lb_count(V, lvl) == 0) return -1;
        }
    }
    while (lvl > 0) {
        lvl--;
        const int gs = found << LOG_FAN;
        const int idx = gs + lane;
        const bool ok = idx < lb_count(V, lvl) && qual<INCL>(lb_level(V, lvl, idx), v);
        const u64 b = __ballot(ok);
        found = b ? gs + 63 - __clzll(b) : -1;
        if (found < 0) return -1;
    }
    return found;
}

// Rightmost element < tile s's end qualifying for v, given that s's min
// qualifies (per lane): chunk mins from the published words, then the chunk.
template <bool INCL>
__device__ __forceinline__ int lb_answer_in_tile(const LbView& V, const int* __restrict__ in,
                                                 int s, int v) {
    // 16 words per round, loaded together; a stale word is polled alone
    int c = -1;
    const u64* cm = V.CM + (size_t)s * CPT;
    for (int b = CPT - 16; b >= 0 && c < 0; b -= 16) {
        u64 w[16];
#pragma unroll
        for (int k = 0; k < 16; k++) w[k] = ld_agent(&cm[b + k]);
#pragma unroll
        for (int k = 15; k >= 0; k--) {
            const int x = (unsigned)(w[k] >> 32) == V.tag ? (int)(unsigned)w[k]
                                                           : lb_wait(V, &cm[b + k]);
            if (c < 0 && qual<INCL>(x, v)) c = s * CPT + b + k;
        }
    }
    if (c < 0) return -1;  // unreachable: the tile's min qualifies
    const int e = last_in_32<INCL>(in, c * CHUNK + 32, v);
    return e >= 0 ? e : last_in_32<INCL>(in, c * CHUNK, v);
}

// Publish the min of the 64 entries of group g at level lvl - 1 (or of the
// entries up to the last one) as G[lvl][g].
__device__ __forceinline__ void lb_publish_group(const LbView& V, int lvl, int g) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int idx = (g << LOG_FAN) + lane;
    const int x = idx < lb_count(V, lvl - 1) ? lb_level(V, lvl - 1, idx) : INF;
    const int m = wave_min(x);
    if (lane == 0) st_agent(const_cast<u64*>(lb_level_ptr(V, lvl)) + g, tagged(V.tag, m));
}

template <bool INCL>
__global__ __launch_bounds__(NT) void k_lookback(const int* __restrict__ in,
                                                 int* __restrict__ out, int n, TileOut to,
                                                 Levels L, LbView V) {
    __shared__ TileSmem sm;
    __shared__ int s_t;
    __shared__ int lds[WAVE];
    if (threadIdx.x == 0)
        s_t = (int)__hip_atomic_fetch_add(&V.ctl->ticket, 1u, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int t = s_t;
    if (t >= V.nt) return;
    V.tag = V.ctl->epoch;  // set by the preceding one-thread kernel
    int v[K];
    load_lane(in, n, t, v);
    const int pl = (threadIdx.x == 0 && t > 0) ? in[t * TILE - 1] : 0;
    solve_tile<INCL>(out, n, t, v, pl, sm, to, L);  // ends with a workgroup barrier
    if (threadIdx.x >= WAVE) return;
    const int lane = threadIdx.x;
    const int tile_start = t * TILE;

    // 2. publish the chunk mins and the tile min
    int tmin = sm.wmin[0];
#pragma unroll
    for (int w = 1; w < NWAVE; w++) tmin = min(tmin, sm.wmin[w]);
    for (int c = lane; c < CPT; c += WAVE)
        st_agent(&V.CM[(size_t)t * CPT + c], tagged(V.tag, to.cmin[(size_t)t * CPT + c]));
    if (lane == 0) st_agent(&V.M[t], tagged(V.tag, tmin));

    // 3. look back: exclusive prefix-min over tiles < t
    int P = INF;
    for (int j = t - 1; j >= 0;) {
        const int idx = j - lane;
        u64 ws = 0, wm = 0;
        if (idx >= 0) {
            ws = ld_agent(&V.S[idx]);
            wm = ld_agent(&V.M[idx]);
        }
        const bool vs = idx >= 0 && (unsigned)(ws >> 32) == V.tag;
        const bool vm = idx >= 0 && (unsigned)(wm >> 32) == V.tag;
        const u64 bs = __ballot(vs);
        const int ks = bs ? (int)__builtin_ctzll(bs) : WAVE;  // nearest inclusive prefix
        const u64 need = __ballot(lane < ks && idx >= 0);
        if (need & ~__ballot(vm)) {
            // a predecessor has not published its tile min yet: poll (bounded)
            if (lb_wait(V, &V.M[max(j - (int)__builtin_ctzll(need & ~__ballot(vm)), 0)]) == INF &&
                __hip_atomic_load(&V.ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                break;
            continue;
        }
        int x = INF;
        if (lane < ks && idx >= 0) x = (int)(unsigned)wm;
        if (lane == ks) x = (int)(unsigned)ws;
        P = min(P, wave_min(x));
        if (ks < WAVE) break;
        j -= WAVE;
    }
    if (lane == 0) st_agent(&V.S[t], tagged(V.tag, min(P, tmin)));

    // group mins, by the last tile of each group
    for (int l = 1; l < LB_LEVELS && lb_count(V, l) > 0; l++) {
        const int span_log = LOG_FAN * l;
        const bool last = ((t + 1) & ((1 << span_log) - 1)) == 0 || t == V.nt - 1;
        if (!last) break;
        lb_publish_group(V, l, t >> span_log);
    }

    // 4. open elements that P beats (same rules as the multi-launch fix-up)
    if (t == 0) return;
    const int v1 = __shfl(v[0], 0, WAVE);  // the tile's first element (lane 0's first value)
    if (!qual<INCL>(P, v1)) return;
    const int cnt = to.qcount[t];
    const int tw = t - WAVE + lane;
    lds[lane] = tw >= 0 ? lb_wait(V, &V.M[tw]) : INF;
    wave_lds_sync();
    if (cnt <= QS) {
        int2 e = make_int2(0, INF);
        if (lane < cnt) e = to.qlist[(size_t)t * QS + lane];
        const bool active = lane < cnt && qual<INCL>(P, e.y);
        const int val = e.y;
        int ts = -1;
        if (active) {
            int s = -1;
            for (int i = WAVE - 1; i >= 0 && s < 0; i--)
                if (qual<INCL>(lds[i], val)) s = i;
            if (s >= 0 && t - WAVE + s >= 0) ts = t - WAVE + s;
        }
        u64 deep = __ballot(active && ts < 0);
        while (deep) {
            const int l = __ffsll(deep) - 1;
            deep &= deep - 1;
            const int r = lb_find_tile<INCL>(V, t - WAVE, __shfl(val, l, WAVE));
            if (lane == l) ts = r;
        }
        if (active) out[tile_start + e.x] = ts < 0 ? -1 : lb_answer_in_tile<INCL>(V, in, ts, val);
        return;
    }
    // More open elements than recorded: rescan the tile (its prefix minima
    // are the open elements).  Their answers move monotonically: the answer
    // c of an open element also answers every later open element v with
    // qual(key c, v) (everything between c and the tile failed the earlier,
    // larger element; nothing in the tile before v qualifies for v).  So the
    // answer of the last open element seen is carried, keys it qualifies
    // for take it, and only the others are searched for, one at a time
    // from the first, each new answer serving the keys after it.
    int rm = INF;
    bool have = false;
    int c_pos = -1, c_val = 0;  // carried answer and its key
    bool c_ok = false;
    // (offsets from the tile start: p0 + 64 would overflow near 2^31)
    const int span = min(TILE, n - tile_start);
    for (int o = 0; o < span; o += WAVE) {
        const int p0 = tile_start + o;
        const bool valid = lane < span - o;
        const int idx = valid ? p0 + lane : p0;
        const int xv = valid ? in[idx] : INF;
        int inc = min(xv, dpp<DPP_ROW_SHR + 1>(INF, xv));
        inc = min(inc, dpp<DPP_ROW_SHR + 2>(INF, inc));
        inc = min(inc, dpp<DPP_ROW_SHR + 4>(INF, inc));
        inc = min(inc, dpp<DPP_ROW_SHR + 8>(INF, inc));
        inc = min(inc, dpp<DPP_ROW_BCAST15, 0xa>(INF, inc));
        inc = min(inc, dpp<DPP_ROW_BCAST31, 0xc>(INF, inc));
        const bool ex_real = lane > 0 || have;
        const int ex = min(dpp<DPP_WAVE_SHR1>(INF, inc), rm);
        const bool open = valid && !(ex_real && qual<INCL>(ex, xv));
        bool todo = open && qual<INCL>(P, xv);
        if (todo && c_ok && qual<INCL>(c_val, xv)) {
            out[idx] = c_pos;
            todo = false;
        }
        u64 m = __ballot(todo);
        while (m) {
            const int l = __ffsll(m) - 1;
            const int v = __shfl(xv, l, WAVE);
            const int r = lb_find_tile<INCL>(V, t, v);
            int a = -1, av = 0;
            if (lane == l && r >= 0) {
                a = lb_answer_in_tile<INCL>(V, in, r, v);
                if (a >= 0) av = in[a];
            }
            c_pos = readlane(a, l);
            c_val = readlane(av, l);
            c_ok = c_pos >= 0;
            if (lane == l) {
                out[idx] = c_pos;
                todo = false;
            }
            // the keys after lane l that this answer qualifies for
            if (todo && lane > l && c_ok && qual<INCL>(c_val, xv)) {
                out[idx] = c_pos;
                todo = false;
            }
            m = __ballot(todo);
        }
        rm = min(rm, lane63(inc));
        have = true;
        if (!qual<INCL>(P, rm)) break;  // later open elements are <= rm
    }
}

// next run tag (never 0, which fresh memory holds), ticket counter and error
// word reset
__global__ void k_lookback_begin(LbCtl* c) {
    unsigned e = c->epoch + 1u;
    if (e == 0u) e = 1u;
    c->epoch = e;
    c->ticket = 0u;
    c->err = 0;
}

template <bool INCL>
static apsep_status run_lookback_t(apsep_scratch* s, const int* in, int* out, int n,
                                   hipStream_t st) {
    const int nt = (int)(((long long)n + TILE - 1) / TILE);  // (n + TILE - 1 overflows near 2^31)
    Levels L{};
    L.nt = nt;
    L.nlev = 1;  // no level-1 atomics from the tile solver
    L.mn0 = s->mn0;
    L.n[0] = nt;
    TileOut to{s->cmin, s->qcount, s->qlist, s->scount, s->spine};
    LbView V{};
    V.ctl = s->lb_ctl;
    V.M = s->lb_M;
    V.S = s->lb_S;
    V.CM = s->lb_CM;
    V.nt = nt;
    int sz[MAX_LEVELS];
    const int nl = level_sizes(nt, sz);
    V.gn[0] = nt;
    for (int l = 1; l < LB_LEVELS; l++) {
        V.G[l] = s->lb_G[l];
        V.gn[l] = l < nl ? sz[l] : 0;
    }
    hipLaunchKernelGGL(k_lookback_begin, dim3(1), dim3(1), 0, st, s->lb_ctl);
    hipLaunchKernelGGL(k_lookback<INCL>, dim3(nt), dim3(NT), 0, st, in, out, n, to, L, V);
    return APSEP_LAUNCH_FAILED() ? APSEP_ERR_LAUNCH : APSEP_OK;
}

apsep_status run_lookback(apsep_scratch* s, const int32_t* in, int32_t* out, int n, int incl,
                          hipStream_t st) {
    return incl ? run_lookback_t<true>(s, in, out, n, st)
                : run_lookback_t<false>(s, in, out, n, st);
}

}  // namespace apsep
