// apsep_segmented_impl.hpp -- segmented (CSR batch) APSEP for gfx950: many
// independent arrays solved by one launch sequence, with optional hipGraph
// capture and replay of that sequence.
//
// Compiled once per tile shape: apsep_segmented_256.hip and
// apsep_segmented_512.hip define APSEP_TILE_LANES and APSEP_SEG_NS (the
// namespace that keeps the two builds' kernels and entry points apart) and
// include this file.  A batch plan runs one of them; apsep_batch_plan_tune
// times both on the caller's batch.
//
// The reference has no batched entry point; its callers loop over arrays
// with one launchWSTL / launchSPT each (src/apsep.cuh:2598-2605,
// 3069-3076).  Here the segments are concatenated and solved as one array:
//   setup     tile -> segment map and bounds, level-1 keys reset
//   tile pass the single-array tile solver with the segment starts inside
//             each tile marked in LDS; answers rebased to their segments
//   levels    level 1 from the tile mins (no atomics in the tile pass: a
//             batch's 64-tile groups would queue on their keys), [levels
//             >= 2 as for one array]
//   fix-up    prefix-min and searches bounded by the segment that runs into
//             the tile from before
// The level-1 keys are reset by the setup kernel (a kernel, not a memset,
// so the sequence can be captured into a graph), and the run uses a fixed
// epoch: every replay of a captured graph then starts from clean keys.
#if !defined(APSEP_TILE_LANES) || !defined(APSEP_SEG_NS)
#error "include through apsep_segmented_256.hip / apsep_segmented_512.hip"
#endif

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <atomic>
#include <new>

#include "apsep_device.hpp"
#include "apsep_internal.hpp"
#include "../../include/apsep/apsep.h"

namespace apsep {
namespace APSEP_SEG_NS {

int seg_tile_elems() { return TILE; }

constexpr unsigned SEG_EPOCH = 1u;
constexpr int SEG_FIX_WG = 256;  // four tiles per group, up to 8 waves per SIMD (as k_fix_pass)

// tile -> segment holding the tile's first element, and that segment's
// start and end.  A lane per segment: a wave reads 64 segments' offsets in
// one coalesced load, each lane writes the map entries of its segment when
// it holds at most two tile starts (every short segment), and the wave
// writes those of the longer ones together, one segment at a time, its
// lanes over the segment's tiles (a segment of 2^20 keys covers 256 tiles).
// Every tile start lies in exactly one segment, so every map entry is
// written once.  gw / nw: the calling wave's index among the nw waves that
// share the work (uniform per wave).  CO: the entries are read by other
// workgroups of the same launch (device-coherent stores).
template <bool CO>
__device__ __forceinline__ void seg_map_tiles(const long long* __restrict__ offs, int nseg,
                                              int* __restrict__ tmap, int2* __restrict__ tbound,
                                              int nt, int gw, int nw) {
    const int lane = (int)threadIdx.x & (WAVE - 1);
    const long long cap = (long long)nt * TILE;
    for (long long k0 = (long long)gw * WAVE; k0 < nseg; k0 += (long long)nw * WAVE) {
        const int k = (int)k0 + lane;
        long long s = 0, e = 0;
        if (k < nseg) {
            s = max(offs[k], 0ll);
            e = min(offs[k + 1], cap);
        }
        int t0 = 0, t1 = -1;  // the tiles whose first element lies in [s, e)
        if (s < e) {
            t0 = (int)((s + TILE - 1) / TILE);
            t1 = (int)((e - 1) / TILE);
        }
        const int2 b = make_int2((int)min(s, (long long)INT_MAX), (int)min(e, (long long)INT_MAX));
        const bool few = t1 - t0 < 2;
        if (few) {
            for (int t = t0; t <= t1; t++) {
                st_i<CO>(&tmap[t], k);
                st_i2<CO>(&tbound[t], b);
            }
        }
        u64 many = __ballot(!few);
        while (many) {
            const int l = __ffsll(many) - 1;
            many &= many - 1;
            const int u0 = __shfl(t0, l, WAVE), u1 = __shfl(t1, l, WAVE);
            const int2 ub = make_int2(__shfl(b.x, l, WAVE), __shfl(b.y, l, WAVE));
            for (int t = u0 + lane; t <= u1; t += WAVE) {
                st_i<CO>(&tmap[t], (int)k0 + l);
                st_i2<CO>(&tbound[t], ub);
            }
        }
    }
}

// The tile map and bounds (above), and level-1 keys emptied.
__global__ __launch_bounds__(256) void k_seg_setup(const long long* __restrict__ offs,
                                                   int nseg, int* __restrict__ tmap,
                                                   int2* __restrict__ tbound, int nt,
                                                   u64* __restrict__ key1, int nkey1) {
    const int gid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int stride = (int)(gridDim.x * blockDim.x);
    for (int i = gid; i < nkey1; i += stride) key1[key_slot(1, i)] = ~0ull;
    seg_map_tiles<false>(offs, nseg, tmap, tbound, nt, gid >> 6, stride >> 6);
}

// 256-lane tiles: 5 waves per SIMD (88 VGPRs, no spill; 5 tiles per CU)
// beat 6 (80 VGPRs and a 16-byte spill): 1.47 instead of 1.53 ms for the
// 8192-segment batch.  128-lane tiles likewise run at 5 (at 6 they spill 35
// registers).  512-lane tiles keep 6 (3 tiles per CU, not 2).
template <bool INCL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT <= 256 ? 5 : 6))) void k_tile_pass_seg(const int* __restrict__ in,
                                                      int* __restrict__ out, int n,
                                                      TileOut to, Levels L, SegIn sg) {
    __shared__ TileSmem sm;
    tile_step_seg<INCL>(in, out, n, (int)blockIdx.x, sm, to, L, sg);
}

__global__ __launch_bounds__(256) void k_level1_seg(Levels L) { build_level1(L); }
// level lvl >= 2 from the level below, one wave per entry (one launch per
// level: each reads the whole level below)
__global__ __launch_bounds__(256) void k_level_seg(Levels L, int lvl) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int g = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (g >= L.n[lvl]) return;  // uniform per wave
    const int idx = g * FAN + lane;
    const int x = idx < L.n[lvl - 1] ? key_value(key_at(L, lvl - 1, idx), L.epoch) : INF;
    const int m = wave_min(x);
    if (lane == 0) key_at(L, lvl, g) = make_key(L.epoch, m);
}

template <bool INCL>
__global__ __launch_bounds__(SEG_FIX_WG) __attribute__((amdgpu_waves_per_eu(8))) void k_fix_pass_seg(const int* __restrict__ in,
                                                             int* __restrict__ out, int n,
                                                             Levels L, TileOut to, SegIn sg) {
    __shared__ int lds[SEG_FIX_WG / WAVE][FIX_LDS_INTS];
    const int t = (int)(blockIdx.x * (SEG_FIX_WG / WAVE) + (threadIdx.x >> 6));
    if (t >= L.nt) return;
    fix_tile<INCL, true>(FlatMap{in, out}, n, t, L, to, lds[threadIdx.x >> 6], sg);
}

// ---- The whole batch in one persistent launch ----------------------------
// The four steps above in one launch of co-resident workgroups, separated
// by grid barriers (an eager launch clears the barrier set of the next one;
// a captured launch is preceded by k_seg_persist_begin, which clears its
// own, so the pair can be replayed).  Which
// workgroups run together is found at run time (grid_join / grid_count):
// participant `me` of `np` maps segments me*NWAVE.. (a lane each), then
// solves tiles me, me + np, ..., builds level-1 keys with its waves, and
// fixes the tiles it solved, one wave per tile, so each tile's outputs and
// open list stay in the XCD (L2) that wrote them.  Everything another
// workgroup reads after a barrier (map and bounds, tile and chunk mins,
// spines, counts, keys) is written and read device-coherent, so the
// barriers need no cache write-back.  Level 1 is built whole (no reset);
// levels >= 2 one per barrier.
constexpr int PERSIST_WAVES_PER_EU = 4;
static_assert(NWAVE * FIX_LDS_INTS * sizeof(int) <= sizeof(TileSmem), "fix-up LDS");

__global__ __launch_bounds__(BAR_SUB) void k_seg_persist_begin(GridBarrier* bar) {
    grid_barrier_reset(bar);
}

template <bool INCL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(PERSIST_WAVES_PER_EU)))
void k_seg_persist(const int* __restrict__ in, int* __restrict__ out, int n, TileOut to,
                   Levels L, SegIn sg, int* __restrict__ tmap, int2* __restrict__ tbound,
                   GridBarrier* bar, GridBarrier* next) {
    __shared__ TileSmem sm;
    grid_barrier_clear_next(next);
    const int j = grid_join(bar);
    if (j < 0) return;  // uniform per workgroup: not a participant
    const int c = (int)(blockIdx.x % SPT_SLOTS);
    const u64 pub = grid_count(bar, c == 0 && j == 0);
    if (pub == 0ull) return;  // timed out (error word raised)
    int me = j, np = 0;
#pragma unroll
    for (int k = 0; k < SPT_SLOTS; k++) {
        const int nk = (int)((pub >> (8 * k)) & 0xFF);
        if (k < c) me += nk;
        np += nk;
    }
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = lane_id();
    const bool leader = threadIdx.x == 0;
    const int nt = L.nt;
    unsigned gen = grid_barrier_start(bar), jb = 0;  // before this workgroup arrives
    seg_map_tiles<true>(sg.offs, sg.nseg, tmap, tbound, nt, me * NWAVE + wave, np * NWAVE);
    if (!grid_barrier_co(bar, (unsigned)np, (unsigned)me, gen, jb, leader)) return;
    for (int t = me; t < nt; t += np) {
        tile_step_seg<INCL, true>(in, out, n, t, sm, to, L, sg);
        __syncthreads();  // the next tile reuses the LDS
    }
    if (nt <= 1) return;  // uniform across the grid
    if (!grid_barrier_co(bar, (unsigned)np, (unsigned)me, gen, jb, leader)) return;
    if (L.nlev > 1) {
        for (int g = me * NWAVE + wave; g < L.n[1]; g += np * NWAVE) {  // uniform per wave
            const int idx = g * FAN + lane;
            const int m = wave_min(idx < nt ? ld_i<true>(&L.mn0[idx]) : INF);
            if (lane == 0) st_u64<true>(&key_at(L, 1, g), make_key(L.epoch, m));
        }
        if (L.upper) {
            for (int l = 2; l < L.nlev; l++) {
                if (!grid_barrier_co(bar, (unsigned)np, (unsigned)me, gen, jb, leader)) return;
                for (int g = me * NWAVE + wave; g < L.n[l]; g += np * NWAVE) {
                    const int idx = g * FAN + lane;
                    const int x = idx < L.n[l - 1]
                                      ? key_value(ld_u64<true>(&key_at(L, l - 1, idx)), L.epoch)
                                      : INF;
                    const int m = wave_min(x);
                    if (lane == 0) st_u64<true>(&key_at(L, l, g), make_key(L.epoch, m));
                }
            }
        }
        if (!grid_barrier_co(bar, (unsigned)np, (unsigned)me, gen, jb, leader)) return;
    }
    int* lds = reinterpret_cast<int*>(&sm) + wave * FIX_LDS_INTS;
    for (long long t = me + (long long)wave * np; t < nt; t += (long long)NWAVE * np)
        fix_tile<INCL, true, true>(FlatMap{in, out}, n, (int)t, L, to, lds, sg);
}

// Workgroups of k_seg_persist that fit on one CU at once (the smaller of
// the two predicates' occupancy, never above what the waves-per-EU
// attribute guarantees); 0 if the runtime cannot tell.
static int persist_per_cu() {
    static std::atomic<int> cached{-1};
    int v = cached.load(std::memory_order_relaxed);
    if (v >= 0) return v;
    int a = 0, b = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_seg_persist<false>, NT, 0);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_seg_persist<true>, NT, 0);
    if (e != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        return 0;
    }
    v = min(min(a, b), PERSIST_WAVES_PER_EU * 4 / NWAVE);
    cached.store(v < 0 ? 0 : v, std::memory_order_relaxed);
    return v < 0 ? 0 : v;
}

template <bool INCL>
static apsep_status seg_launch_persist(apsep_batch_plan* p, const int* in,
                                       const long long* offs, int* out, int nseg, int n,
                                       hipStream_t st) {
    apsep_scratch* s = p->s;
    const int nt = (int)(((long long)n + TILE - 1) / TILE);
    const int per_cu = persist_per_cu();
    if (per_cu <= 0) return APSEP_ERR_UNSUPPORTED;
    Levels L;
    int sz[MAX_LEVELS];
    L.nt = nt;
    L.nlev = level_sizes(nt, sz);
    L.mn0 = s->mn0;
    L.upper = (L.nlev > 2 && sz[1] > UPPER_MIN) ? 1 : 0;
    for (int l = 0; l < MAX_LEVELS; l++) {
        L.n[l] = l < L.nlev ? sz[l] : 0;
        L.key[l] = (l >= 1 && l < L.nlev) ? s->key[l] : nullptr;
    }
    L.epoch = SEG_EPOCH;
    TileOut to{s->cmin, s->qcount, s->qlist, s->scount, s->spine};
    SegIn sg{offs, p->tmap, nseg, p->tbound};
    // a workgroup per tile, or per NWAVE * 64 segments, up to what runs at
    // once (and what the join's per-slot counts hold)
    long long want = (nseg + (long long)NWAVE * WAVE - 1) / ((long long)NWAVE * WAVE);
    if (want < nt) want = nt;
    long long grid = (long long)per_cu * s->num_cus;
    if (grid > (long long)SPT_SLOTS * SPT_SLOT_MAX) grid = (long long)SPT_SLOTS * SPT_SLOT_MAX;
    if (grid > want) grid = want;
    if (grid < 1) grid = 1;
    // barrier sets as the single-array persistent solve's: a captured run
    // clears its own set with a kernel, an eager run the next run's set
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_UNSUPPORTED;
    }
    if (cs != hipStreamCaptureStatusNone && cs != hipStreamCaptureStatusActive)
        return APSEP_ERR_UNSUPPORTED;
    const bool cap = cs == hipStreamCaptureStatusActive;
    GridBarrier* bar;
    GridBarrier* next = nullptr;
    if (cap) {
        bar = s->bar + BAR_CAPTURED;
        hipLaunchKernelGGL(k_seg_persist_begin, dim3(1), dim3(BAR_SUB), 0, st, bar);
    } else {
        bar = s->bar + s->bar_parity;
        next = s->bar + (1 - s->bar_parity);
    }
    hipLaunchKernelGGL(k_seg_persist<INCL>, dim3((unsigned)grid), dim3(NT), 0, st, in, out, n,
                       to, L, sg, p->tmap, p->tbound, bar, next);
    if (APSEP_LAUNCH_FAILED()) return APSEP_ERR_LAUNCH;
    if (!cap) s->bar_parity ^= 1;
    return APSEP_OK;
}

template <bool INCL>
static apsep_status seg_launch(apsep_batch_plan* p, const int* in, const long long* offs,
                               int* out, int nseg, int n, hipStream_t st) {
    apsep_scratch* s = p->s;
    const int nt = (int)(((long long)n + TILE - 1) / TILE);  // (n + TILE - 1 overflows near 2^31)
    Levels L;
    int sz[MAX_LEVELS];
    L.nt = nt;
    L.nlev = level_sizes(nt, sz);
    L.mn0 = s->mn0;
    L.upper = (L.nlev > 2 && sz[1] > UPPER_MIN) ? 1 : 0;
    for (int l = 0; l < MAX_LEVELS; l++) {
        L.n[l] = l < L.nlev ? sz[l] : 0;
        L.key[l] = (l >= 1 && l < L.nlev) ? s->key[l] : nullptr;
    }
    L.epoch = SEG_EPOCH;
    TileOut to{s->cmin, s->qcount, s->qlist, s->scount, s->spine};
    SegIn sg{offs, p->tmap, nseg, p->tbound};
    const int nkey1 = L.nlev > 1 ? sz[1] : 0;
    // a lane per segment, a thread per level-1 key
    const long long segs = (nseg + 63ll) / 64 * 64;
    const long long work = segs > nkey1 ? segs : nkey1;
    int blocks = (int)((work + 255) / 256);
    if (blocks < 1) blocks = 1;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_seg_setup, dim3(blocks), dim3(256), 0, st, offs, nseg, p->tmap,
                       p->tbound, nt, nkey1 ? s->key[1] : nullptr, nkey1);
    hipLaunchKernelGGL(k_tile_pass_seg<INCL>, dim3(nt), dim3(NT), 0, st, in, out, n, to, L, sg);
    if (nt > 1) {
        if (L.nlev > 1) hipLaunchKernelGGL(k_level1_seg, dim3((L.n[1] + 3) / 4), dim3(256), 0, st, L);
        if (L.upper)
            for (int l = 2; l < L.nlev; l++)
                hipLaunchKernelGGL(k_level_seg, dim3((L.n[l] + 3) / 4), dim3(256), 0, st, L, l);
        constexpr int tpw = SEG_FIX_WG / WAVE;
        hipLaunchKernelGGL(k_fix_pass_seg<INCL>, dim3((nt + tpw - 1) / tpw), dim3(SEG_FIX_WG), 0,
                           st, in, out, n, L, to, sg);
    }
    return APSEP_LAUNCH_FAILED() ? APSEP_ERR_LAUNCH : APSEP_OK;
}

// ---- Batches of arrays read in place (apsep_run_jobs) ----------------------
// Job k's keys and answers stay in the caller's memory.  In the solver's
// virtual index space job k occupies whole tiles from first_tile, so no tile
// holds two jobs: a page per tile gives its keys' and answers' addresses,
// and the tile's segment bound (tbound) is its job's first and last virtual
// index.  Everything else is the segmented solve (answers rebased to the
// job's start, prefix mins and searches bounded by it).
struct JobDesc {          // apsep_job
    const int* keys;
    int* answers;
    long long n;
    long long first_tile;
};

// Page and bound of every tile: tile t belongs to the last job whose
// first_tile <= t (binary search, so each tile is written exactly once); a
// tile no job covers becomes an empty tile (bound [t*TILE, t*TILE)).  Also
// empties level 1.
__global__ __launch_bounds__(256) void k_jobs_setup(const JobDesc* __restrict__ jobs, int njobs,
                                                    Page* __restrict__ pages,
                                                    int2* __restrict__ tbound, int nt,
                                                    u64* __restrict__ key1, int nkey1) {
    const int gid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int stride = (int)(gridDim.x * blockDim.x);
    for (int i = gid; i < nkey1; i += stride) key1[key_slot(1, i)] = ~0ull;
    for (int t = gid; t < nt; t += stride) {
        int lo = 0, hi = njobs - 1;  // last job with first_tile <= t
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (jobs[mid].first_tile <= t) lo = mid;
            else hi = mid - 1;
        }
        const JobDesc j = jobs[lo];
        const long long local = (long long)t - j.first_tile;  // tile within the job
        const long long ntl = j.n > 0 ? (j.n + TILE - 1) / TILE : 0;
        const int ts = t * TILE;
        if (local >= 0 && local < ntl) {
            pages[t] = Page{j.keys + local * TILE, j.answers + local * TILE};
            const long long s = j.first_tile * TILE;
            tbound[t] = make_int2((int)s, (int)min(s + j.n, (long long)INT_MAX));
        } else {
            pages[t] = Page{nullptr, nullptr};
            tbound[t] = make_int2(ts, ts);
        }
    }
}

template <bool INCL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT <= 256 ? 5 : 6))) void k_tile_pass_jobs(const Page* __restrict__ pages,
                                                       const int2* __restrict__ tbound,
                                                       TileOut to, Levels L) {
    __shared__ TileSmem sm;
    const int t = (int)blockIdx.x;
    const Page pg = pages[t];
    const int2 b = tbound[t];
    const int tile_start = t * TILE;
    // the tile's valid keys end at its job's end (tile_start when empty)
    const int n = max(min(b.y, tile_start + TILE), tile_start);
    const int* in = reinterpret_cast<const int*>(PagedMap::virt(pg.src, t));
    int* out = reinterpret_cast<int*>(PagedMap::virt(pg.dst, t));
    int v[K];
    load_lane(in, n, t, v);
    // the key before the tile, when the tile is not its job's first
    const int pl = (threadIdx.x == 0 && b.x < tile_start) ? in[tile_start - 1] : 0;
    solve_tile<INCL, true, false>(out, n, t, v, pl, sm, to, L, SegLane{b.x, 0u});
}

template <bool INCL>
__global__ __launch_bounds__(SEG_FIX_WG) __attribute__((amdgpu_waves_per_eu(8))) void k_fix_pass_jobs(const Page* __restrict__ pages, int n,
                                                              Levels L, TileOut to,
                                                              SegIn sg) {
    __shared__ int lds[SEG_FIX_WG / WAVE][FIX_LDS_INTS];
    const int t = (int)(blockIdx.x * (SEG_FIX_WG / WAVE) + (threadIdx.x >> 6));
    if (t >= L.nt) return;
    fix_tile<INCL, true>(PagedMap{pages}, n, t, L, to, lds[threadIdx.x >> 6], sg);
}

template <bool INCL>
static apsep_status jobs_launch(apsep_batch_plan* p, const JobDesc* jobs, int njobs, int nt,
                                hipStream_t st) {
    apsep_scratch* s = p->s;
    Levels L;
    int sz[MAX_LEVELS];
    L.nt = nt;
    L.nlev = level_sizes(nt, sz);
    L.mn0 = s->mn0;
    L.upper = (L.nlev > 2 && sz[1] > UPPER_MIN) ? 1 : 0;
    for (int l = 0; l < MAX_LEVELS; l++) {
        L.n[l] = l < L.nlev ? sz[l] : 0;
        L.key[l] = (l >= 1 && l < L.nlev) ? s->key[l] : nullptr;
    }
    L.epoch = SEG_EPOCH;
    TileOut to{s->cmin, s->qcount, s->qlist, s->scount, s->spine};
    SegIn sg{nullptr, nullptr, njobs, p->tbound};
    const int nkey1 = L.nlev > 1 ? sz[1] : 0;
    const int work = nt > nkey1 ? nt : nkey1;
    int blocks = (work + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_jobs_setup, dim3(blocks), dim3(256), 0, st, jobs, njobs, p->pages,
                       p->tbound, nt, nkey1 ? s->key[1] : nullptr, nkey1);
    hipLaunchKernelGGL(k_tile_pass_jobs<INCL>, dim3(nt), dim3(NT), 0, st, p->pages, p->tbound,
                       to, L);
    if (nt > 1) {
        if (L.nlev > 1) hipLaunchKernelGGL(k_level1_seg, dim3((L.n[1] + 3) / 4), dim3(256), 0, st, L);
        if (L.upper)
            for (int l = 2; l < L.nlev; l++)
                hipLaunchKernelGGL(k_level_seg, dim3((L.n[l] + 3) / 4), dim3(256), 0, st, L, l);
        constexpr int tpw = SEG_FIX_WG / WAVE;
        hipLaunchKernelGGL(k_fix_pass_jobs<INCL>, dim3((nt + tpw - 1) / tpw), dim3(SEG_FIX_WG), 0,
                           st, p->pages, nt * TILE, L, to, sg);
    }
    return APSEP_LAUNCH_FAILED() ? APSEP_ERR_LAUNCH : APSEP_OK;
}

apsep_status jobs_run(apsep_batch_plan* p, const void* jobs, int njobs, int tiles,
                      hipStream_t st) {
    static_assert(sizeof(JobDesc) == 32, "apsep_job layout");
    const JobDesc* j = reinterpret_cast<const JobDesc*>(jobs);
    return p->s->incl ? jobs_launch<true>(p, j, njobs, tiles, st)
                      : jobs_launch<false>(p, j, njobs, tiles, st);
}

apsep_status seg_run(apsep_batch_plan* p, const int32_t* in, const int64_t* offs,
                     int32_t* out, int nseg, int n, hipStream_t st) {
    const long long* o = reinterpret_cast<const long long*>(offs);
    if (p->persistent)
        return p->s->incl ? seg_launch_persist<true>(p, in, o, out, nseg, n, st)
                          : seg_launch_persist<false>(p, in, o, out, nseg, n, st);
    return p->s->incl ? seg_launch<true>(p, in, o, out, nseg, n, st)
                      : seg_launch<false>(p, in, o, out, nseg, n, st);
}

}  // namespace APSEP_SEG_NS
}  // namespace apsep
