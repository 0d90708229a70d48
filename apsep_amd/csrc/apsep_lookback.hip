// apsep_lookback.hip -- single-pass, non-cooperative APSEP for gfx950 with
// decoupled look-back (APSEP_ALGO_LOOKBACK).
//
// The reference's look-back family (src/apsep.cuh:236-254, 1720-1747) has
// each block solve its tile, publish its summary and look back over its
// predecessors' published summaries for the answers its tile leaves open.
// Here one launch holds two kinds of workgroup:
//   tile workgroups (one per tile, index = tile) solve the tile with the
//     shared tile solver, publish its 128 chunk mins (CM) and its min (M),
//     and the last tile of a 64-tile group to publish also publishes the
//     group's min (G1), and so on up the levels (G2, G3: group counters
//     tell which is last).  A tile workgroup never waits for another.
//   answer workgroups (after the tiles, one wave per tile) look back: the
//     exclusive prefix-min P of tile t is the min of the published entries
//     before t in each enclosing group (at most 63 per level, loaded at
//     once); then the open elements that P beats are answered from the 64
//     previous tile mins, the group mins for older tiles, the found tile's
//     chunk mins and one 64-key chunk of the input, and written over the
//     tile's -1.
// Keeping the look-back out of the tile workgroups keeps its round trips
// from holding a tile's LDS (three tiles per CU): the tiles run as in the
// multi-launch tile pass.
// Nothing crosses workgroups through a fence: every published word is a
// 64-bit (run tag << 32 | value) written with one agent-scope atomic store
// and read with agent-scope atomic loads, so a word is either this run's
// complete value or recognisably stale; the open list and the outputs the
// answers overwrite are written device-coherent, all before the tile's M.
// A reader that meets a stale word polls it (bounded: past the limit the
// run is flagged as failed and every wait ends).  An eager run takes its
// tag from the host and clears the group counters of the next eager run
// (two sets in turn); a captured run's tag lives in device memory and is
// advanced, with its own counter set cleared, by a one-workgroup kernel
// recorded before the solve, so a replay holds no host-baked state.
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
    u64* G[LB_LEVELS];        // G[0] unused; G[l] = min of 64^l tiles
    unsigned* GC;             // group counters: entries of group g at level l - 1
                              // published, at GC[gn[1] + .. + gn[l - 1] + g]
    unsigned* GC_next;        // the next eager run's set, cleared by this run (or null)
    int ngc_next;             // counters cleared there: as many as a run at capacity uses
                              // (the next run may be larger than this one)
    int gn[LB_LEVELS];        // entries per level (gn[0] = nt)
    u64* CM;                  // [nt * CPT] chunk mins
    unsigned tag;             // 0: read the captured run's tag from ctl
    int nt;
};

// This run's value of *p, polling while it is stale (bounded).
__device__ __forceinline__ int lb_wait(const LbView& V, const u64* p) {
    u64 w = ld_agent(p);
    unsigned it = 0;
    // the error word holds the tag of the run that gave up: an older run's
    // failure does not end this run's waits
    while ((unsigned)(w >> 32) != V.tag) {
        if (++it > LB_SPIN_LIMIT) {
            __hip_atomic_store(&V.ctl->err, (int)V.tag, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            return INF;
        }
        if (__hip_atomic_load(&V.ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
            (int)V.tag)
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
__device__ __forceinline__ int lb_gc_base(const LbView& V, int lvl) {
    static_assert(LB_LEVELS == 5, "levels");
    return (lvl > 1 ? V.gn[1] : 0) + (lvl > 2 ? V.gn[2] : 0) + (lvl > 3 ? V.gn[3] : 0);
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

// Rightmost element of tile s qualifying for v, given that s's min
// qualifies (per lane): the tile's right spine (its strict suffix minima,
// lowest rank = nearest) with its count, 16 ranks per round; past the
// recorded ranks the published chunk mins, then the chunk.
template <bool INCL>
__device__ int lb_answer_in_tile(const LbView& V, const int* __restrict__ in,
                                 const TileOut& to, int s, int v) {
    const int4* sp = reinterpret_cast<const int4*>(to.spine + (size_t)s * RS);
    const int cnt = ld_i<true>(&to.scount[s]);  // (loaded with the first ranks)
    int pos = -1;
#pragma unroll
    for (int h = 0; h < RS / 16; h++) {
        if (h > 0 && (pos >= 0 || cnt <= 16 * h)) break;
        int4 w[8];
#pragma unroll
        for (int k = 0; k < 8; k++) w[k] = ld_i4<true>(&sp[8 * h + k]);
#pragma unroll
        for (int k = 7; k >= 0; k--) {
            const int r = 16 * h + 2 * k;
            if (r + 1 < cnt && qual<INCL>(w[k].w, v)) pos = w[k].z;
            if (r < cnt && qual<INCL>(w[k].y, v)) pos = w[k].x;
        }
    }
    if (pos >= 0) return s * TILE + pos;
    // 16 chunk mins per round; a stale word is polled alone
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

// Tile role: solve tile t (its outputs written device-coherent: an answer
// workgroup may overwrite the open elements' -1), then publish its chunk
// mins and, once every output and summary of the tile has been written, its
// min.
template <bool INCL>
__device__ __forceinline__ void lb_solve(const int* __restrict__ in, int* __restrict__ out,
                                         int n, int t, TileSmem& sm, const TileOut& to,
                                         const Levels& L, const LbView& V) {
    const int tid = (int)threadIdx.x;
    int v[K];
    load_lane(in, n, t, v);
    const int pl = (tid == 0 && t > 0) ? in[t * TILE - 1] : 0;
    solve_tile<INCL, false, false, false, true>(out, n, t, v, pl, sm, to, L);  // ends with a barrier
    if (tid < CPT)
        st_agent(&V.CM[(size_t)t * CPT + tid], tagged(V.tag, to.cmin[(size_t)t * CPT + tid]));
    wait_own_memory_ops();  // outputs, open list, spine, counts, chunk mins: complete
    __syncthreads();
    if (tid >= WAVE) return;
    int tmin = sm.wmin[0];
#pragma unroll
    for (int w = 1; w < NWAVE; w++) tmin = min(tmin, sm.wmin[w]);
    if (tid == 0) {
        st_agent(&V.M[t], tagged(V.tag, tmin));
        wait_own_memory_ops();
    }
    // the last entry of a group to be published publishes the group's min
    // (every other entry of the group was published before its count)
    int idx = t;
    for (int l = 1; l < LB_LEVELS && lb_count(V, l) > 0; l++) {
        const int g = idx >> LOG_FAN;
        const int members = min(1 << LOG_FAN, lb_count(V, l - 1) - (g << LOG_FAN));
        unsigned c = 0u;
        if (tid == 0)
            c = __hip_atomic_fetch_add(&V.GC[lb_gc_base(V, l) + g], 1u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
        if ((int)__builtin_amdgcn_readfirstlane(c) + 1 != members) break;
        lb_publish_group(V, l, g);
        if (tid == 0) wait_own_memory_ops();
        idx = g;
    }
}

// Answer role, one wave: the open elements of tile t that have an answer
// before the tile, written device-coherent over the tile's -1.
template <bool INCL>
__device__ void lb_answer(const int* __restrict__ in, int* __restrict__ out, int n, int t,
                          const TileOut& to, const LbView& V, int* lds) {
    const int lane = (int)threadIdx.x & (WAVE - 1);
    const int tile_start = t * TILE;
    // Loads in as few dependent rounds as the data allows.  Round 1: the
    // tile's own min (is it solved?), the entries for P (the published mins
    // of the tiles before t in its 64-tile group and of the earlier groups in
    // each enclosing group: at most 63 per level), the 64 previous tile mins
    // (the search window) and the tile's first key.
    const u64 wt = ld_agent(&V.M[t]);
    u64 w[LB_LEVELS];
    const u64* p[LB_LEVELS];
    {
        int idx = t;
#pragma unroll
        for (int l = 0; l < LB_LEVELS; l++) {
            const int j = ((idx >> LOG_FAN) << LOG_FAN) + lane;
            const bool use = j < idx && (l == 0 || lb_count(V, l) > 0);
            p[l] = use ? lb_level_ptr(V, l) + j : nullptr;
            w[l] = use ? ld_agent(p[l]) : tagged(V.tag, INF);
            idx >>= LOG_FAN;
        }
    }
    const int tw = t - WAVE + lane;
    const u64 ww = tw >= 0 ? ld_agent(&V.M[tw]) : tagged(V.tag, INF);
    const int v1 = in[tile_start];
    if ((unsigned)(wt >> 32) != V.tag) {
        lb_wait(V, &V.M[t]);
        // (a failed wait leaves the tile's summaries unknown: nothing is written)
        if ((unsigned)(ld_agent(&V.M[t]) >> 32) != V.tag) return;
    }
    // Round 2 (written before the tile's min): its open count and list.
    const int cnt = ld_i<true>(&to.qcount[t]);
    int2 e = make_int2(0, INF);
    if (lane < QS) e = ld_i2<true>(&to.qlist[(size_t)t * QS + lane]);
    int x = INF;
#pragma unroll
    for (int l = 0; l < LB_LEVELS; l++)
        x = min(x, (unsigned)(w[l] >> 32) == V.tag ? (int)(unsigned)w[l] : lb_wait(V, p[l]));
    const int P = wave_min(x);  // exclusive prefix-min over tiles < t

    // open elements that P beats (same rules as the multi-launch fix-up)
    if (t == 0 || !qual<INCL>(P, v1)) return;  // the first element bounds them all
    if (cnt <= QS) {
        lds[lane] = (unsigned)(ww >> 32) == V.tag ? (int)(unsigned)ww : lb_wait(V, &V.M[tw]);
        wave_lds_sync();
        const bool active = lane < cnt && qual<INCL>(P, e.y) && e.x >= 0 && e.x < n - tile_start;
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
        if (active && ts >= 0)
            st_i<true>(&out[tile_start + e.x], lb_answer_in_tile<INCL>(V, in, to, ts, val));
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
            st_i<true>(&out[idx], c_pos);
            todo = false;
        }
        u64 m = __ballot(todo);
        while (m) {
            const int l = __ffsll(m) - 1;
            const int v = __shfl(xv, l, WAVE);
            const int r = lb_find_tile<INCL>(V, t, v);
            int a = -1, av = 0;
            if (lane == l && r >= 0) {
                a = lb_answer_in_tile<INCL>(V, in, to, r, v);
                if (a >= 0) av = in[a];
            }
            c_pos = readlane(a, l);
            c_val = readlane(av, l);
            c_ok = c_pos >= 0;
            if (lane == l) {
                st_i<true>(&out[idx], c_pos);
                todo = false;
            }
            // the keys after lane l that this answer qualifies for
            if (todo && lane > l && c_ok && qual<INCL>(c_val, xv)) {
                st_i<true>(&out[idx], c_pos);
                todo = false;
            }
            m = __ballot(todo);
        }
        rm = min(rm, lane63(inc));
        have = true;
        if (!qual<INCL>(P, rm)) break;  // later open elements are <= rm
    }
}

// Workgroups 0 .. nt - 1 solve tile blockIdx.x; the nfix after them answer
// the open elements of blocks of NWAVE tiles (block j, j + nfix, ..., one
// wave per tile).  Only answer workgroups wait, and only for words that tile
// workgroups publish; nfix stays below the number of workgroups that fit on
// the GPU at once, so tile workgroups always find room, whatever order the
// workgroups start in.  (Started in index order, as the hardware does, the
// answer workgroups come after every tile.)
template <bool INCL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) void k_lookback(
    const int* __restrict__ in, int* __restrict__ out, int n, TileOut to, Levels L, LbView V,
    int nfix) {
    __shared__ TileSmem sm;
    if (V.tag == 0u) V.tag = V.ctl->epoch;  // a captured run: set by the preceding kernel
    const int b = (int)blockIdx.x;
    if (V.GC_next && b == 0)  // no workgroup of this run uses that set
        for (int i = (int)threadIdx.x; i < V.ngc_next; i += NT)
            __hip_atomic_store(&V.GC_next[i], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (b < V.nt) {
        lb_solve<INCL>(in, out, n, b, sm, to, L, V);
        return;
    }
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    int* lds = reinterpret_cast<int*>(&sm) + wave * WAVE;
    for (int t = (b - V.nt) * NWAVE + wave; t < V.nt; t += nfix * NWAVE)
        lb_answer<INCL>(in, out, n, t, to, V, lds);
}

// Before a captured run: the next captured-run tag (LB_TAG_CAPTURED_MIN and
// up, wrapping within that range) and its group counters cleared (one
// workgroup)
__global__ __launch_bounds__(256) void k_lookback_begin(LbCtl* c, unsigned* gc, int ngc) {
    if (threadIdx.x == 0) {
        unsigned e = c->epoch + 1u;
        if (e < LB_TAG_CAPTURED_MIN) e = LB_TAG_CAPTURED_MIN;
        c->epoch = e;
    }
    for (int i = (int)threadIdx.x; i < ngc; i += (int)blockDim.x) gc[i] = 0u;
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
    V.CM = s->lb_CM;
    V.nt = nt;
    int sz[MAX_LEVELS];
    const int nl = level_sizes(nt, sz);
    V.gn[0] = nt;
    for (int l = 1; l < LB_LEVELS; l++) {
        V.G[l] = s->lb_G[l];
        V.gn[l] = l < nl ? sz[l] : 0;
    }
    int ngc = 0;  // group counters, level after level
    for (int l = 1; l < LB_LEVELS; l++) ngc += V.gn[l];
    V.ngc_next = s->lb_ngc_cap;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_UNSUPPORTED;
    }
    if (cs != hipStreamCaptureStatusNone && cs != hipStreamCaptureStatusActive)
        return APSEP_ERR_UNSUPPORTED;
    const bool cap = cs == hipStreamCaptureStatusActive;
    if (cap) {
        V.GC = s->lb_GC + LB_GC_CAPTURED * s->lb_gc_stride;
        V.GC_next = nullptr;
        V.tag = 0u;
        hipLaunchKernelGGL(k_lookback_begin, dim3(1), dim3(256), 0, st, s->lb_ctl, V.GC, ngc);
    } else {
        unsigned tag = s->lb_tag + 1u;
        if (tag >= LB_TAG_CAPTURED_MIN) {
            // tag wrap (after ~2e9 eager runs): no published word may carry
            // a tag of the new range
            if (hipMemsetAsync(s->lb_M, 0, (size_t)s->lb_m_bytes, st) != hipSuccess ||
                hipMemsetAsync(s->lb_CM, 0, (size_t)s->lb_cm_bytes, st) != hipSuccess) {
                APSEP_CLEAR_HIP_ERROR();
                return APSEP_ERR_LAUNCH;
            }
            for (int l = 1; l < LB_LEVELS; l++)
                if (s->lb_G[l] &&
                    hipMemsetAsync(s->lb_G[l], 0, (size_t)s->lb_g_bytes[l], st) != hipSuccess) {
                    APSEP_CLEAR_HIP_ERROR();
                    return APSEP_ERR_LAUNCH;
                }
            tag = 1u;
        }
        V.tag = tag;
        V.GC = s->lb_GC + s->lb_parity * s->lb_gc_stride;
        V.GC_next = s->lb_GC + (1 - s->lb_parity) * s->lb_gc_stride;
    }
    // answer workgroups: one per block of NWAVE tiles, at most two per CU
    // (the kernel fits three per CU)
    const int nb = (nt + NWAVE - 1) / NWAVE;
    const int nfix = max(1, min(nb, 2 * s->num_cus));
    hipLaunchKernelGGL(k_lookback<INCL>, dim3(nt + nfix), dim3(NT), 0, st, in, out, n, to, L, V,
                       nfix);
    if (APSEP_LAUNCH_FAILED()) return APSEP_ERR_LAUNCH;
    if (!cap) {
        s->lb_tag = V.tag;
        s->lb_parity ^= 1;
    }
    return APSEP_OK;
}

apsep_status run_lookback(apsep_scratch* s, const int32_t* in, int32_t* out, int n, int incl,
                          hipStream_t st) {
    return incl ? run_lookback_t<true>(s, in, out, n, st)
                : run_lookback_t<false>(s, in, out, n, st);
}

}  // namespace apsep
