// apsep_single.hip -- single-array APSEP solvers for gfx950 and the C ABI
// that drives them (include/apsep/apsep.h).
//
//   WSTL     tile pass -> level pass (64-ary tile-min hierarchy) -> fix-up
//            pass (multi-launch; replaces runWSTL, src/apsep.cuh:2548-2596)
//   SPT      the same phases in one persistent launch with two grid
//            barriers among the workgroups that run together (replaces
//            runSPT, src/apsep.cuh:3052-3067)
//
// All device work is integer compare/min and index arithmetic; no MFMA.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <mutex>
#include <new>

#include "apsep_device.hpp"
#include "apsep_internal.hpp"
#include "../../include/apsep/apsep.h"

namespace apsep {

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// WSTL pass 1: one workgroup per tile (also accumulates level 1: a separate
// level-1 kernel, as the segmented pass uses, measured slower here).
template <bool INCL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) void k_tile_pass(const int* __restrict__ in,
                                                  int* __restrict__ out, int n,
                                                  TileOut to, Levels L) {
    __shared__ TileSmem sm;
    tile_step<INCL>(in, out, n, (int)blockIdx.x, sm, to, L);
}

// Level-1 keys emptied (above every key of every run).  Launched before the
// tile pass of a run recorded into a graph: such a run uses the fixed epoch
// CAPTURE_EPOCH, so that each replay starts from clean keys instead of
// folding its tile mins into the previous replay's.
__global__ __launch_bounds__(256) void k_reset_level1(u64* __restrict__ key1, int nkey1) {
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < nkey1;
         i += (int)(gridDim.x * blockDim.x))
        key1[key_slot(1, i)] = ~0ull;
}

// Before every captured SPT launch: its grid barrier's counters zeroed (an
// eager launch finds its set cleared by the eager launch before it).
__global__ __launch_bounds__(BAR_SUB) void k_spt_begin(GridBarrier* bar) {
    grid_barrier_reset(bar);
}

// WSTL pass 2 (only when level 1 exceeds UPPER_MIN entries): levels >= 2.
__global__ __launch_bounds__(256) void k_upper_levels(Levels L) {
    build_upper_levels(L, (int)(threadIdx.x >> 6));
}

// WSTL pass 3: one wave per tile.
// (256-lane workgroups of four tiles, held to 64 VGPRs so that 8 waves fit
// per SIMD: the pass is latency-bound, and more tiles in flight beat a few
// spilled registers)
constexpr int FIX_WG = 256;

template <bool INCL>
__global__ __launch_bounds__(FIX_WG) __attribute__((amdgpu_waves_per_eu(8))) void k_fix_pass(const int* __restrict__ in,
                                                     int* __restrict__ out, int n,
                                                     Levels L, TileOut to) {
    __shared__ int lds[FIX_WG / WAVE][FIX_LDS_INTS];
    const int t = (int)(blockIdx.x * (FIX_WG / WAVE) + (threadIdx.x >> 6));
    if (t >= L.nt) return;
    fix_tile<INCL>(FlatMap{in, out}, n, t, L, to, lds[threadIdx.x >> 6]);
}

// SPT: one persistent launch, normally launched with at most as many
// workgroups as fit on the GPU at once (three per CU).  grid_join and
// grid_count find the workgroups that actually run together (normally all
// of them), per slot c = blockIdx.x % SPT_SLOTS.  Slot c's participants
// solve the tiles of residue c (tile c + SPT_SLOTS i for class index i):
// each first its join ticket's, then fixed rounds (participant j takes
// i = j, j + npc, ...), and the last SPT_DYN_ROUNDS rounds' worth of tiles
// or fewer one at a time from the slot's ticket counter, by whichever
// participant is free (tile durations vary with the keys).  A residue whose
// slot has no participant (a grid smaller than SPT_SLOTS, or workgroups kept
// from starting by other work on the GPU) is solved by all participants
// after their own, a tile per ticket.  Then a grid barrier among the
// participants, [upper levels by participant 0, a barrier,] then the fix-up
// of the same tiles, one wave per tile: a tile is fixed by the workgroup
// that solved it, so its outputs and open list stay in one XCD's L2.  A
// workgroup lists the tiles it took by ticket in LDS, the ones past
// SPT_DYN_CAP in a linked list through `link` (one word per tile).  The
// summaries other workgroups read (tile mins, chunk mins, spines, counts,
// upper keys) are written and read with device-coherent accesses, and the
// barrier needs no fence.  The occupancy attribute guarantees
// SPT_WAVES_PER_EU waves per SIMD.  The fix-up's LDS reuses the tile's.
// (A cooperative launch, which has the runtime guarantee co-residency
// instead, costs ~17 us more per launch on MI355X.)
constexpr int SPT_DYN_ROUNDS = 3;
constexpr int SPT_DYN_CAP = 32;  // tiles taken by ticket listed in LDS

template <bool INCL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(SPT_WAVES_PER_EU)))
void k_spt(const int* __restrict__ in, int* __restrict__ out, int n,
           TileOut to, Levels L, GridBarrier* bar, GridBarrier* next, int* link) {
    __shared__ TileSmem sm;
    __shared__ int s_dyn[SPT_DYN_CAP];  // tiles this workgroup took by ticket
    __shared__ int s_head;              // and the last one listed through link
    __shared__ unsigned s_next;         // the next tile's ticket
    static_assert(NWAVE * FIX_LDS_INTS * sizeof(int) <= sizeof(TileSmem), "fix-up LDS");
    if (threadIdx.x == 0) s_head = -1;  // (grid_join's barrier orders it)
    grid_barrier_clear_next(next);
    const int j = grid_join(bar);
    if (j < 0) return;  // uniform per workgroup: not a participant
    // Nothing per-lane stays live across the tile loop (it would be spilled
    // to memory): the wave index is kept in a scalar, and each tile computes
    // the lane's index afresh inside a volatile asm, so the compiler cannot
    // hoist the lane-derived addresses out of the loop.
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nt = L.nt;
    const int c = (int)(blockIdx.x % SPT_SLOTS);
    unsigned* const tickets = &bar->tiles[c * BAR_LINE];
    // tiles of residue c (at least as many as slot c's workgroups: the grid
    // has at most nt workgroups)
    const int nc = (nt - c + SPT_SLOTS - 1) / SPT_SLOTS;
    // schedule, known after the first tile: slot participants, fixed
    // rounds S (the first included), first class index taken by ticket
    int npc = 0, S = 1, tail0 = 0, ntail = 0, ndyn = 0, i = j;
    u64 pub = 0ull;
    // residues with tiles and no participant (known with the schedule): one
    // ticket counter for all of them, ticket o -> the (o mod m)-th such
    // residue's class index o / m
    unsigned empty = 0u;
    bool orphans = false;  // the workgroup's own residue is done
    int t = c + SPT_SLOTS * i;
    while (t >= 0) {
        int lane;
        asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0"
                     : "=v"(lane));
        const int tid = wave * WAVE + lane;
        const bool dyn = orphans || (npc > 0 && i >= tail0);
        // the next tile of the residue comes from the counter once the
        // fixed rounds are done; its ticket is taken inside this tile and
        // lands in s_next before the tile's closing barrier (one slot
        // suffices: the next write comes after the next tile's first
        // barrier, which every wave passes only after reading this one)
        const bool take = !orphans && npc > 0 && i + npc >= tail0;
        int v[K];
        load_lane(in, n, t, v, tid);
        const int pl = (tid == 0 && t > 0) ? in[t * TILE - 1] : 0;
        solve_tile<INCL, false, true, true>(out, n, t, v, pl, sm, to, L, SegLane{0, 0u}, tid,
                                            take ? tickets : nullptr, &s_next);
        if (dyn) {  // a tile taken by ticket, listed for the fix-up
            if (tid == 0) {
                if (ndyn < SPT_DYN_CAP) {
                    s_dyn[ndyn] = t;
                } else {
                    link[t] = s_head;
                    s_head = t;
                }
            }
            ndyn++;
        }
        if (!orphans) {
            if (npc == 0) {  // after the first tile
                pub = grid_count(bar, c == 0 && j == 0);
                npc = (int)((pub >> (8 * c)) & 0xFF);
                if (npc == 0) return;  // timed out (error word raised)
#pragma unroll
                for (int k = 0; k < SPT_SLOTS; k++)
                    if (k < nt && ((pub >> (8 * k)) & 0xFF) == 0) empty |= 1u << k;
                S = max(nc / npc - SPT_DYN_ROUNDS, 1);
                tail0 = S * npc;     // <= nc: npc <= the slot's workgroups <= nc
                ntail = nc - tail0;  // < (SPT_DYN_ROUNDS + 1) * npc
                if (S > 1) {
                    i += npc;
                } else {
                    if (tid == 0)
                        s_next = __hip_atomic_fetch_add(tickets, 1u, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
                    __syncthreads();
                    i = tail0 + (int)min(__builtin_amdgcn_readfirstlane(s_next), (unsigned)ntail);
                }
            } else if (take) {
                i = tail0 + (int)min(__builtin_amdgcn_readfirstlane(s_next), (unsigned)ntail);
            } else {
                i = i + npc < tail0 ? i + npc : nc;
            }
            t = i < nc ? c + SPT_SLOTS * i : -1;
            orphans = t < 0 && empty != 0u;
        }
        if (orphans) {  // (only where some residue has no participant)
            const int m = __builtin_popcount(empty);
            const int most = (nt - __builtin_ctz(empty) + SPT_SLOTS - 1) / SPT_SLOTS;
            for (t = -1;;) {
                __syncthreads();  // every wave has read s_next
                if (tid == 0)
                    s_next = __hip_atomic_fetch_add(&bar->orphans, 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
                __syncthreads();
                const unsigned o = __builtin_amdgcn_readfirstlane(s_next);
                const int oi = (int)(o / (unsigned)m);
                if (oi >= most) break;  // t = -1: done
                unsigned em = empty;
                for (int r = (int)(o % (unsigned)m); r > 0; r--) em &= em - 1;
                const int to_ = __builtin_ctz(em) + SPT_SLOTS * oi;
                if (to_ < nt) {
                    t = to_;
                    break;
                }
            }
        }
    }
    if (nt <= 1) return;  // uniform across the grid
    // participant index (slots in order) and count
    int me = j, np = 0;
#pragma unroll
    for (int k = 0; k < SPT_SLOTS; k++) {
        const int nk = (int)((pub >> (8 * k)) & 0xFF);
        if (k < c) me += nk;
        np += nk;
    }
    const bool leader = wave == 0 && lane_id() == 0;
    unsigned gen = grid_barrier_start(bar), jb = 0;  // before this workgroup arrives
    if (!grid_barrier_co(bar, (unsigned)np, (unsigned)me, gen, jb, leader)) return;
    if (L.upper) {
        if (me == 0) build_upper_levels<true>(L, wave);
        if (!grid_barrier_co(bar, (unsigned)np, (unsigned)me, gen, jb, leader)) return;
    }
    int* lds = reinterpret_cast<int*>(&sm) + wave * FIX_LDS_INTS;
    // this workgroup's tiles: S fixed rounds, then those taken by ticket
    // (listed in LDS, then through link: each wave walks the list and takes
    // every NWAVE-th, a path only a starved grid takes)
    const int nlds = S + min(ndyn, SPT_DYN_CAP);
    for (int k = wave; k < nlds; k += NWAVE) {
        const int t = k < S ? c + SPT_SLOTS * (j + k * npc)
                            : __builtin_amdgcn_readfirstlane(s_dyn[k - S]);
        fix_tile<INCL, false, true>(FlatMap{in, out}, n, t, L, to, lds);
    }
    if (ndyn > SPT_DYN_CAP) {
        int t = __builtin_amdgcn_readfirstlane(s_head);
        for (int k = 0; t >= 0; k++) {
            const int nx = __builtin_amdgcn_readfirstlane(link[t]);
            if (k % NWAVE == wave) fix_tile<INCL, false, true>(FlatMap{in, out}, n, t, L, to, lds);
            t = nx;
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// Run epochs.  Eager runs take 2, 3, ... from the scratch; a run recorded
// into a graph takes CAPTURE_EPOCH and empties level 1 first (its launch
// arguments are replayed unchanged).  A later eager run's keys compare below
// a captured run's, so the two kinds mix on one scratch in any order.
constexpr unsigned CAPTURE_EPOCH = 1u;
constexpr unsigned FIRST_EAGER_EPOCH = 2u;

constexpr int64_t round4(int64_t x) { return (x + 3) & ~(int64_t)3; }

struct LevelSizes {
    int nl;
    int sz[MAX_LEVELS];
};

// Level sizes for nt tiles: entries per level, the top one <= FAN.
constexpr LevelSizes level_sizes_of(int nt) {
    LevelSizes r{0, {0, 0, 0, 0, 0}};
    int m = nt;
    while (true) {
        r.sz[r.nl++] = m;
        if (m <= FAN) break;
        m = (m + FAN - 1) / FAN;
    }
    return r;
}

// Level sizes for nt tiles; returns nlev.
int level_sizes(int nt, int* sizes) {
    const LevelSizes r = level_sizes_of(nt);
    for (int l = 0; l < r.nl; l++) sizes[l] = r.sz[l];
    return r.nl;
}

// Chunk-min slots: whole tiles of the scratch's own shape and of the
// largest shape (this file's tile): a batch plan's scratch, sized for the
// smaller segmented tile, also serves batches run on the larger one, whose
// last partial tile writes all of its chunk mins.
constexpr int64_t chunk_slots(int64_t n_cap, int tile) {
    const int64_t own = (n_cap + tile - 1) / tile * (tile / CHUNK);
    const int64_t big = (n_cap + TILE - 1) / TILE * CPT;
    return own > big ? own : big;
}

constexpr int64_t ALIGN_SLACK = 256 * 12;

// Bytes of the arrays every solver uses (tile mins, keyed levels, chunk
// mins, open lists, spines, grid barrier), without the alignment slack.
constexpr int64_t common_scratch_bytes(int64_t n_cap, int tile) {
    const int nt = (int)((n_cap + tile - 1) / tile);
    const int64_t ncm = chunk_slots(n_cap, tile);
    const LevelSizes ls = level_sizes_of(nt < 1 ? 1 : nt);
    int64_t b = (round4(ls.sz[0]) + 4) * 4;          // tile mins
    for (int l = 1; l < ls.nl; l++) b += ((int64_t)key_slot(l, (int)round4(ls.sz[l]) + 4)) * 8;
    b += (round4(ncm) + 4) * 4;                       // cmin
    b += (round4(nt) + 4) * 4;                        // qcount
    b += ((int64_t)nt * QS + 2) * 8;                  // qlist
    b += (round4(nt) + 4) * 4;                        // scount
    b += ((int64_t)nt * RS + 2) * 8;                  // spine
    b += (int64_t)sizeof(GridBarrier) * BAR_SETS;     // grid barriers
    return b;
}

int64_t single_scratch_bytes(int64_t n_cap, apsep_algo algo, int tile) {
    int64_t b = common_scratch_bytes(n_cap, tile);
    if (algo == APSEP_ALGO_SPT) b += (round4((n_cap + tile - 1) / tile) + 4) * 4;  // link
    if (algo == APSEP_ALGO_LOOKBACK) {
        const int nt = (int)((n_cap + tile - 1) / tile);
        const int64_t ncm = chunk_slots(n_cap, tile);
        const LevelSizes ls = level_sizes_of(nt < 1 ? 1 : nt);
        b += (int64_t)sizeof(LbCtl);
        b += (round4(nt) + 4) * 8;                    // tile mins
        b += (round4(nt) + 4) * 4 * LB_GC_SETS;       // group counter sets (fewer than nt each)
        b += (round4(ncm) + 4) * 8;                   // chunk mins
        for (int l = 1; l < ls.nl; l++) b += (round4(ls.sz[l]) + 4) * 8;  // group mins
        b += 256 * (4 + LB_LEVELS);
    }
    return b + ALIGN_SLACK;                           // alignment slack
}

static void* carve(char*& p, int64_t bytes) {
    uintptr_t a = ((uintptr_t)p + 255) & ~(uintptr_t)255;
    p = (char*)(a + bytes);
    return (void*)a;
}

// Point the scratch's arrays into s->base (s->bytes long), sized for
// s->n_cap keys in tiles of s->tile; false if they do not fit.
static bool carve_arrays(apsep_scratch* s) {
    const int nt = (int)((s->n_cap + s->tile - 1) / s->tile);
    const int64_t ncm = chunk_slots(s->n_cap, s->tile);
    int sz[MAX_LEVELS];
    const int nl = level_sizes(nt, sz);
    char* p = (char*)s->base;
    s->mn0 = (int*)carve(p, (round4(sz[0]) + 4) * 4);
    for (int l = 1; l < nl; l++) s->key[l] = (u64*)carve(p, (int64_t)key_slot(l, round4(sz[l]) + 4) * 8);
    s->cmin = (int*)carve(p, (round4(ncm) + 4) * 4);
    s->qcount = (int*)carve(p, (round4(nt) + 4) * 4);
    s->qlist = (int2*)carve(p, ((int64_t)nt * QS + 2) * 8);
    s->scount = (int*)carve(p, (round4(nt) + 4) * 4);
    s->spine = (int2*)carve(p, ((int64_t)nt * RS + 2) * 8);
    s->bar = (GridBarrier*)carve(p, (int64_t)sizeof(GridBarrier) * BAR_SETS);
    if (s->algo == APSEP_ALGO_SPT) s->spt_link = (int*)carve(p, (round4(nt) + 4) * 4);
    if (s->algo == APSEP_ALGO_LOOKBACK) {
        s->lb_ctl = (LbCtl*)carve(p, (int64_t)sizeof(LbCtl));
        s->lb_m_bytes = (round4(nt) + 4) * 8;
        s->lb_cm_bytes = (round4(ncm) + 4) * 8;
        for (int l = 1; l < nl; l++) s->lb_g_bytes[l] = (round4(sz[l]) + 4) * 8;
        s->lb_M = (u64*)carve(p, (round4(nt) + 4) * 8);
        s->lb_GC = (unsigned*)carve(p, (round4(nt) + 4) * 4 * LB_GC_SETS);
        s->lb_gc_stride = round4(nt) + 4;
        s->lb_ngc_cap = 0;
        for (int l = 1; l < nl; l++) s->lb_ngc_cap += sz[l];
        s->lb_CM = (u64*)carve(p, (round4(ncm) + 4) * 8);
        for (int l = 1; l < nl; l++) s->lb_G[l] = (u64*)carve(p, (round4(sz[l]) + 4) * 8);
    }
    return (char*)p <= (char*)s->base + s->bytes;
}

apsep_status single_scratch_init(apsep_scratch* s) {
    const int nt = (int)((s->n_cap + s->tile - 1) / s->tile);
    s->nt_cap = nt;
    if (nt == 0) return APSEP_OK;
    int sz[MAX_LEVELS];
    const int nl = level_sizes(nt, sz);
    s->bytes = single_scratch_bytes(s->n_cap, s->algo, s->tile);
    if (hipMalloc(&s->base, (size_t)s->bytes) != hipSuccess) {
        s->base = nullptr;
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_ALLOC;
    }
    if (!carve_arrays(s)) return APSEP_ERR_ALLOC;  // sizing bug guard
    s->epoch = FIRST_EAGER_EPOCH;
    if (hipMemset(s->base, 0, (size_t)s->bytes) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_LAUNCH;
    }
    // keys start above every real key (a fresh scratch holds no run)
    for (int l = 1; l < nl; l++)
        if (hipMemset(s->key[l], 0xFF, key_slot(l, round4(sz[l]) + 4) * 8) != hipSuccess) {
            APSEP_CLEAR_HIP_ERROR();
            return APSEP_ERR_LAUNCH;
        }

    if (s->algo == APSEP_ALGO_SPT) {
        // the scratch serves both predicates: size for the tighter one
        int per_cu = 0, per_cu_incl = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_spt<false>, NT, 0);
        if (e == hipSuccess)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_incl, k_spt<true>, NT, 0);
        if (per_cu_incl < per_cu) per_cu = per_cu_incl;
        // never above what the waves-per-EU attribute guarantees
        const int pinned = SPT_WAVES_PER_EU * 4 / NWAVE;
        if (per_cu > pinned) per_cu = pinned;
        if (e != hipSuccess || per_cu <= 0) {
            APSEP_CLEAR_HIP_ERROR();
            return APSEP_ERR_UNSUPPORTED;
        }
        s->spt_grid = per_cu * s->num_cus;
    }
    return APSEP_OK;
}

// 0: the stream is not capturing; 1: it is; -1: it cannot be told (an
// invalidated capture, or the legacy stream under another stream's capture).
static int capture_state(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
        APSEP_CLEAR_HIP_ERROR();
        return -1;
    }
    return cs == hipStreamCaptureStatusNone ? 0 : cs == hipStreamCaptureStatusActive ? 1 : -1;
}

// Levels for a run over nt tiles.  Eager runs advance the scratch's run
// epoch; a captured run (capturing = 1) launches the level-1 reset.
static apsep_status fill_levels(apsep_scratch* s, int nt, Levels& L, hipStream_t st,
                                int capturing = 0) {
    int sz[MAX_LEVELS];
    L.nt = nt;
    L.nlev = level_sizes(nt, sz);
    L.mn0 = s->mn0;
    L.upper = (L.nlev > 2 && sz[1] > UPPER_MIN) ? 1 : 0;
    for (int l = 0; l < MAX_LEVELS; l++) {
        L.n[l] = l < L.nlev ? sz[l] : 0;
        L.key[l] = (l >= 1 && l < L.nlev) ? s->key[l] : nullptr;
    }
    if (capturing) {
        if (L.nlev > 1) {
            const int blocks = (sz[1] + 255) / 256;
            hipLaunchKernelGGL(k_reset_level1, dim3(blocks), dim3(256), 0, st, s->key[1], sz[1]);
            if (APSEP_LAUNCH_FAILED()) return APSEP_ERR_LAUNCH;
        }
        L.epoch = CAPTURE_EPOCH;
        return APSEP_OK;
    }
    if (s->epoch >= 0xFFFFFFF0u) {
        // epoch wrap (after ~4e9 runs): clear every key so none can alias
        int szc[MAX_LEVELS];
        const int nlc = level_sizes(s->nt_cap, szc);
        for (int l = 1; l < nlc; l++)
            if (hipMemsetAsync(s->key[l], 0xFF, key_slot(l, szc[l]) * 8, st) != hipSuccess)
                return APSEP_ERR_LAUNCH;
        s->epoch = FIRST_EAGER_EPOCH;
    }
    L.epoch = s->epoch++;
    return APSEP_OK;
}

// fixed_epoch: every run takes the path of a captured run (level-1 reset,
// CAPTURE_EPOCH) and the scratch keeps no state between runs.
template <bool INCL>
static apsep_status run_wstl(apsep_scratch* s, const int* in, int* out, int n,
                             hipStream_t st, bool fixed_epoch = false) {
    const int nt = (int)(((long long)n + TILE - 1) / TILE);  // (n + TILE - 1 overflows near 2^31)
    const int cap = fixed_epoch ? 1 : capture_state(st);
    if (cap < 0) return APSEP_ERR_UNSUPPORTED;
    Levels L;
    if (fill_levels(s, nt, L, st, cap) != APSEP_OK) return APSEP_ERR_LAUNCH;
    TileOut to{s->cmin, s->qcount, s->qlist, s->scount, s->spine};
    hipLaunchKernelGGL(k_tile_pass<INCL>, dim3(nt), dim3(NT), 0, st, in, out, n, to, L);
    if (APSEP_LAUNCH_FAILED()) return APSEP_ERR_LAUNCH;
    if (nt > 1 && L.upper)
        hipLaunchKernelGGL(k_upper_levels, dim3(1), dim3(256), 0, st, L);
    if (nt > 1) {
        constexpr int tpw = FIX_WG / WAVE;
        hipLaunchKernelGGL(k_fix_pass<INCL>, dim3((nt + tpw - 1) / tpw), dim3(FIX_WG), 0, st,
                           in, out, n, L, to);
        if (APSEP_LAUNCH_FAILED()) return APSEP_ERR_LAUNCH;
    }
    return APSEP_OK;
}

template <bool INCL>
static apsep_status run_spt(apsep_scratch* s, const int* in, int* out, int n,
                            hipStream_t st) {
    const int cap = capture_state(st);
    if (cap < 0) return APSEP_ERR_UNSUPPORTED;
    const int nt = (int)(((long long)n + TILE - 1) / TILE);  // (n + TILE - 1 overflows near 2^31)
    Levels L;
    if (fill_levels(s, nt, L, st, cap) != APSEP_OK) return APSEP_ERR_LAUNCH;
    TileOut to{s->cmin, s->qcount, s->qlist, s->scount, s->spine};
    // no more workgroups than tiles (each participant's first tile follows
    // from its join ticket), and per slot no more than its count can say
    int grid = s->spt_blocks > 0 ? s->spt_blocks : s->spt_grid;
    if (grid > nt) grid = nt;
    if (grid > SPT_SLOTS * SPT_SLOT_MAX) grid = SPT_SLOTS * SPT_SLOT_MAX;
    if (grid < 1) grid = 1;
    GridBarrier* bar;
    GridBarrier* next = nullptr;
    if (cap) {  // a replay clears its own set first
        bar = s->bar + BAR_CAPTURED;
        hipLaunchKernelGGL(k_spt_begin, dim3(1), dim3(BAR_SUB), 0, st, bar);
    } else {    // this launch clears the set the next one uses
        bar = s->bar + s->bar_parity;
        next = s->bar + (1 - s->bar_parity);
    }
    bool launched = false;
    if (s->spt_cooperative && !cap) {
        // the runtime guarantees co-residency (the grid fits at once); the
        // kernel's join then finds every workgroup a participant.  A refused
        // launch runs nothing and is counted; the normal launch follows.
        const int cgrid = grid < s->spt_grid ? grid : s->spt_grid;
        int* link = s->spt_link;
        void* args[] = {(void*)&in, (void*)&out, (void*)&n, (void*)&to, (void*)&L,
                        (void*)&bar, (void*)&next, (void*)&link};
        if (hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_spt<INCL>), dim3(cgrid),
                                       dim3(NT), args, 0, st) == hipSuccess) {
            launched = true;
        } else {
            APSEP_CLEAR_HIP_ERROR();
            s->spt_coop_refused++;
        }
    }
    if (!launched)
        hipLaunchKernelGGL(k_spt<INCL>, dim3(grid), dim3(NT), 0, st, in, out, n, to, L, bar, next,
                           s->spt_link);
    if (APSEP_LAUNCH_FAILED()) return APSEP_ERR_LAUNCH;
    if (!cap) s->bar_parity ^= 1;
    return APSEP_OK;
}

// Launch r of each kernel solves array r % m (its own scratch, input and
// output); with m = 1 the same array every time.
template <bool INCL>
static apsep_status time_wstl(apsep_scratch* const* ss, const int* const* ins, int* const* outs,
                              int m, int n, int reps, hipStream_t st, hipEvent_t* ev) {
    const int nt = (int)(((long long)n + TILE - 1) / TILE);  // (n + TILE - 1 overflows near 2^31)
    constexpr int MAXM = 8;
    Levels L[MAXM];
    TileOut to[MAXM];
    for (int k = 0; k < m; k++) {
        if (fill_levels(ss[k], nt, L[k], st) != APSEP_OK) return APSEP_ERR_LAUNCH;
        to[k] = TileOut{ss[k]->cmin, ss[k]->qcount, ss[k]->qlist, ss[k]->scount, ss[k]->spine};
    }
    constexpr int tpw = FIX_WG / WAVE;
    // re-running a kernel on the same epoch rewrites the same values: the
    // level-1 atomicMin sees the same keys, the fix-up the same summaries
    (void)hipEventRecord(ev[0], st);
    for (int r = 0; r < reps; r++) {
        const int k = r % m;
        hipLaunchKernelGGL(k_tile_pass<INCL>, dim3(nt), dim3(NT), 0, st, ins[k], outs[k], n, to[k],
                           L[k]);
    }
    (void)hipEventRecord(ev[1], st);
    if (nt > 1 && L[0].upper)
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(k_upper_levels, dim3(1), dim3(256), 0, st, L[r % m]);
    (void)hipEventRecord(ev[2], st);
    if (nt > 1)
        for (int r = 0; r < reps; r++) {
            const int k = r % m;
            hipLaunchKernelGGL(k_fix_pass<INCL>, dim3((nt + tpw - 1) / tpw), dim3(FIX_WG), 0,
                               st, ins[k], outs[k], n, L[k], to[k]);
        }
    (void)hipEventRecord(ev[3], st);
    return APSEP_LAUNCH_FAILED() ? APSEP_ERR_LAUNCH : APSEP_OK;
}

apsep_status single_time_kernels(apsep_scratch* const* ss, const int32_t* const* ins,
                                 int32_t* const* outs, int m, int n, int incl, int reps,
                                 hipStream_t st, double ms[3]) {
    if (m < 1 || m > 8) return APSEP_ERR_INVALID_ARG;
    if (capture_state(st) != 0) return APSEP_ERR_UNSUPPORTED;  // it synchronises
    hipEvent_t ev[4];
    int made = 0;
    apsep_status r = APSEP_OK;
    for (; made < 4; made++)
        if (hipEventCreate(&ev[made]) != hipSuccess) {
            APSEP_CLEAR_HIP_ERROR();
            r = APSEP_ERR_DEVICE;
            break;
        }
    if (r == APSEP_OK) {
        if (ss[0]->algo == APSEP_ALGO_WSTL) {
            r = incl ? time_wstl<true>(ss, ins, outs, m, n, reps, st, ev)
                     : time_wstl<false>(ss, ins, outs, m, n, reps, st, ev);
        } else {
            // the persistent kernel is the whole solve
            (void)hipEventRecord(ev[0], st);
            for (int k = 0; k < reps && r == APSEP_OK; k++)
                r = single_run(ss[k % m], ins[k % m], outs[k % m], n, incl, st);
            (void)hipEventRecord(ev[1], st);
            (void)hipEventRecord(ev[2], st);
            (void)hipEventRecord(ev[3], st);
        }
    }
    if (r == APSEP_OK) {
        if (hipEventSynchronize(ev[3]) != hipSuccess) r = APSEP_ERR_LAUNCH;
        for (int k = 0; k < 3 && r == APSEP_OK; k++) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, ev[k], ev[k + 1]) != hipSuccess) r = APSEP_ERR_LAUNCH;
            ms[k] = t;
        }
        if (r != APSEP_OK) APSEP_CLEAR_HIP_ERROR();
    }
    for (int k = 0; k < made; k++) (void)hipEventDestroy(ev[k]);
    return r;
}

apsep_status single_run(apsep_scratch* s, const int32_t* in, int32_t* out, int n,
                        int incl, hipStream_t st) {
    switch (s->algo) {
    case APSEP_ALGO_WSTL:
        return incl ? run_wstl<true>(s, in, out, n, st) : run_wstl<false>(s, in, out, n, st);
    case APSEP_ALGO_SPT:
        return incl ? run_spt<true>(s, in, out, n, st) : run_spt<false>(s, in, out, n, st);
    case APSEP_ALGO_LOOKBACK:
        return run_lookback(s, in, out, n, incl, st);
    default:
        return APSEP_ERR_INVALID_ARG;
    }
}

// ---------------------------------------------------------------------------
// The no-allocation entry (apsep_run_noalloc): the multi-launch solve with
// its scratch in `scratch`, a file-scope device array sized for
// NOALLOC_MAX_N keys, which every GPU the library runs on holds until the
// process ends.  In host code `scratch` names a host-side stand-in: its
// device address comes from hipGetSymbolAddress, once per device with that
// device current, and only that address is used (kernel arguments).  Every
// run empties level 1 and uses the fixed epoch, so no state carries from one
// call to the next and nothing needs clearing at set-up; calls that overlap
// on one GPU share the array and race.
// ---------------------------------------------------------------------------
constexpr int64_t NOALLOC_MAX_N = APSEP_NOALLOC_MAX_N;
constexpr int64_t NOALLOC_BYTES = common_scratch_bytes(NOALLOC_MAX_N, TILE) + ALIGN_SLACK;
static_assert(NOALLOC_MAX_N <= INT32_MAX, "int32 indices");

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static apsep_scratch g_noalloc[NOALLOC_MAX_DEVICES];  // views of each device's array
static bool g_noalloc_ready[NOALLOC_MAX_DEVICES];

// Per-device set-up on the first call: the array's device address, carved
// into the solver's arrays.  Host-side bookkeeping only; no allocation.
static apsep_status noalloc_setup(int dev) {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ready[dev]) return APSEP_OK;
    void* ptr = nullptr;
    if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_DEVICE;
    }
    apsep_scratch& s = g_noalloc[dev];
    memset(&s, 0, sizeof s);
    s.algo = APSEP_ALGO_WSTL;
    s.n_cap = NOALLOC_MAX_N;
    s.device = dev;
    s.tile = TILE;
    s.nt_cap = (int)((NOALLOC_MAX_N + TILE - 1) / TILE);
    s.base = ptr;
    s.bytes = NOALLOC_BYTES;
    if (!carve_arrays(&s)) return APSEP_ERR_ALLOC;  // sizing bug guard
    s.epoch = CAPTURE_EPOCH;
    g_noalloc_ready[dev] = true;
    return APSEP_OK;
}

apsep_status noalloc_run(const int32_t* in, int32_t* out, int n, int incl, hipStream_t st) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) {
        APSEP_CLEAR_HIP_ERROR();
        return APSEP_ERR_DEVICE;
    }
    if (dev >= NOALLOC_MAX_DEVICES) return APSEP_ERR_DEVICE;
    const apsep_status r = noalloc_setup(dev);
    if (r != APSEP_OK) return r;
    apsep_scratch* s = &g_noalloc[dev];
    return incl ? run_wstl<true>(s, in, out, n, st, true) : run_wstl<false>(s, in, out, n, st, true);
}

apsep_status single_error(apsep_scratch* s, int* code) {
    *code = 0;
    if (s->lb_ctl) {
        int lerr = 0;
        if (hipMemcpy(&lerr, &s->lb_ctl->err, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
            APSEP_CLEAR_HIP_ERROR();
            return APSEP_ERR_LAUNCH;
        }
        if (lerr) {
            // a look-back wait gave up (the word holds that run's tag);
            // cleared once read
            const int zero = 0;
            if (hipMemcpy(&s->lb_ctl->err, &zero, sizeof(int), hipMemcpyHostToDevice) !=
                hipSuccess) {
                APSEP_CLEAR_HIP_ERROR();
                return APSEP_ERR_LAUNCH;
            }
            *code = 2;
            return APSEP_OK;
        }
    }
    if (!s->bar) return APSEP_OK;
    int any = 0;
    for (int k = 0; k < BAR_SETS; k++) {
        int err = 0;
        if (hipMemcpy(&err, &s->bar[k].err, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
            APSEP_CLEAR_HIP_ERROR();
            return APSEP_ERR_LAUNCH;
        }
        if (err) {
            // cleared once read (the counters are cleared before every launch)
            const int zero = 0;
            if (hipMemcpy(&s->bar[k].err, &zero, sizeof(int), hipMemcpyHostToDevice) !=
                hipSuccess) {
                APSEP_CLEAR_HIP_ERROR();
                return APSEP_ERR_LAUNCH;
            }
            any = err;
        }
    }
    *code = any;
    return APSEP_OK;
}

}  // namespace apsep
