// apsep_device.hpp -- gfx950 device building blocks shared by every APSEP
// solver variant (WSTL multi-launch, SPT persistent, segmented).
//
// Problem (src/apsep.cuh:1-4): out[i] = largest j < i with in[j] < in[i]
// (strict) or in[j] <= in[i] (INCL), else -1.
//
// Work decomposition (wave64 throughout):
//   tile   = one workgroup of NT lanes x K = 16 consecutive elements per
//            lane: 512 lanes (8 waves, 8192 elements) for single arrays,
//            256 or 512 for segmented batches (APSEP_TILE_LANES), loaded
//            with four dwordx4 loads per lane.
//   chunk  = 64 consecutive elements (4 lanes); its min is published for the
//            inter-tile search.
//   levels = 64-ary min hierarchy over tile mins.  Level 0 (tile mins) is a
//            plain array; levels >= 1 hold 64-bit keys (run epoch in the high
//            word, biased min in the low word).  The tile pass of a plain
//            solve accumulates level 1 itself with device-scope atomicMin
//            (64 tiles per key, keys a cache line apart), with no reset
//            between runs and no memory
//            fence: a key from an older run compares larger and decodes as
//            empty.  Segmented batches build level 1 from the tile mins in
//            a kernel of its own (build_level1).  Levels >= 2 are
//            built by a one-workgroup pass only when level 1 exceeds
//            UPPER_MIN entries; below that, readers fold level 1 directly.
//
// Inside a tile (solve_tile):
//   1. in-lane answers by an unrolled O(K^2) compare/select over registers;
//   2. Hillis-Steele min scan over the 64 lane-mins of each wave; every
//      step's window minimum is kept in LDS (win[s][lane] = min over the 2^s
//      lanes ending at lane), which turns the cross-lane search into a
//      6-step binary lifting per query;
//   3. an element unanswered in its lane is answered inside the tile iff the
//      exclusive prefix-min of the earlier lanes qualifies: the lane's first
//      XQ + 1 such elements are checked against the previous lane's 16
//      values, the others go through a per-wave LDS queue whose entries all
//      64 lanes resolve by lifting, one query per lane per round;
//   4. the rest are the tile's "open" elements: the tile's prefix minima.
//      They are written as -1 and recorded (up to QS (position, value) pairs
//      + count) for the fix-up phase, which only revisits the ones whose
//      value beats the exclusive prefix-min of all earlier tiles;
//   5. the right spine (strict suffix minima, up to RS by rank) is recorded
//      for the fix-up's searches into this tile;
//   6. each wave stages its answers in LDS and stores them coalesced.
//
// Invariants relied on (reference reasoning, docs/spt_design.md:140-151 and
// docs/spt_early_out_regression.md:42-51): open elements of a tile are its
// prefix minima, so their values are non-increasing and the first element
// of a tile is always open; hence prefix_min(t) failing to qualify for the
// first element proves every open element of tile t answers -1.  A found
// earlier tile is complete, so its 64-element chunk can be read unguarded.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <limits.h>

#include <type_traits>

namespace apsep {

#ifndef APSEP_TILE_LANES
#define APSEP_TILE_LANES 512
#endif
constexpr int WAVE = 64;
constexpr int NT = APSEP_TILE_LANES;  // lanes per tile workgroup (256 or 512)
constexpr int K = 16;              // consecutive elements per lane
constexpr int K4 = K / 4;          // dwordx4 per lane
constexpr int TILE = NT * K;       // elements per tile (8192 at 512 lanes)
constexpr int NWAVE = NT / WAVE;   // 8 waves per tile
constexpr int CHUNK = 64;          // elements per published chunk min
constexpr int LPC = CHUNK / K;     // lanes per chunk (4)
constexpr int LOG_LPC = 2;
constexpr int CPT = TILE / CHUNK;  // chunk mins per tile (128 at 512 lanes)
constexpr int QS = 32;             // recorded open elements per tile
constexpr int RS = 32;             // recorded right-spine entries per tile
constexpr int QCAP = 192;          // per-wave LDS queue of cross-lane queries
// further queries per lane tried on the previous lane (measured on 2^25
// random keys: 3 beats 2 and 4), in every solver and tile shape
constexpr int XQ = 3;
constexpr int FAN = 64;            // hierarchy fan-out
constexpr int LOG_FAN = 6;
constexpr int MAX_LEVELS = 5;      // 64^5 tiles >> 2^31 / TILE
constexpr int UPPER_MIN = 512;     // level-1 size above which levels >= 2 are stored
constexpr int K1_LOG = 5;          // level-1 keys 2^5 u64 (256 B) apart (see key_at)
constexpr int INF = INT_MAX;
#define SPT_WAVES_PER_EU 6         // occupancy the persistent kernel is pinned at

static_assert(K == 16 && LPC == 4 && (1 << LOG_LPC) == LPC, "tile shape");
static_assert(NWAVE == 2 || NWAVE == 4 || NWAVE == 8, "the wave searches handle 2, 4 or 8 waves per tile");
static_assert(CPT % 16 == 0, "chunk-min search works in blocks of 16");

typedef unsigned long long u64;

// "x qualifies as the answer for query value v"
template <bool INCL>
__device__ __forceinline__ bool qual(int x, int v) {
    return INCL ? (x <= v) : (x < v);
}

// Hierarchy of tile mins (device pointers + sizes for one solve).
struct Levels {
    int nt;                   // tiles
    int nlev;                 // >= 1; the top level has <= 64 entries
    int upper;                // levels >= 2 stored (level 1 > UPPER_MIN entries)
    int n[MAX_LEVELS];        // entries per level
    int* mn0;                 // level 0: tile mins
    u64* key[MAX_LEVELS];     // levels >= 1: epoch-keyed mins (key[0] unused)
    unsigned epoch;           // this run's tag (never 0 or ~0u)
};

// Entry idx of keyed level lvl.  Level-1 keys take one 256-byte line each:
// the tile pass folds every tile min into its group's key with a device
// atomic, and atomics on keys sharing a line serialise (packed keys cost the
// tile pass ~4 us at N = 2^25 on MI355X).  Upper levels are packed.
__host__ __device__ constexpr size_t key_slot(int lvl, int idx) {
    return (size_t)idx << (lvl == 1 ? K1_LOG : 0);
}
__device__ __forceinline__ u64& key_at(const Levels& L, int lvl, int idx) {
    return L.key[lvl][key_slot(lvl, idx)];
}

__device__ __forceinline__ u64 make_key(unsigned epoch, int v) {
    return ((u64)(~epoch) << 32) | (u64)((unsigned)v ^ 0x80000000u);
}
__device__ __forceinline__ int key_value(u64 k, unsigned epoch) {
    return ((unsigned)(k >> 32) == ~epoch) ? (int)((unsigned)k ^ 0x80000000u) : INF;
}

// Accesses to data that crosses workgroups inside one persistent launch
// (CO = true).  Each XCD of an MI355X has its own L2, and a plain store
// stays in the writer's L2 until it is written back, so another XCD may
// read the old value; a fence that writes the L2 back before a grid barrier
// costs ~30 us at N = 2^25 (it writes back every dirty output line too).
// Instead, the few summary words that other workgroups read are stored and
// loaded as agent-scope relaxed atomics, which the compiler emits as sc1
// (device-coherent) accesses, and the grid barrier only waits for every
// lane's memory operations to complete (grid_barrier_co).  CO = false: plain
// accesses, for data read only after a kernel boundary or by the writer's
// own workgroup.
// Waits for the calling lane's memory operations to complete; the "memory"
// clobber also keeps the compiler from moving any of them past the wait.
__device__ __forceinline__ void wait_own_memory_ops() {
    asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
}

template <bool CO>
__device__ __forceinline__ int ld_i(const int* p) {
    if (CO) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return *p;
}
template <bool CO>
__device__ __forceinline__ u64 ld_u64(const u64* p) {
    if (CO) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return *p;
}
template <bool CO>
__device__ __forceinline__ int2 ld_i2(const int2* p) {
    if (CO) {
        const u64 w = ld_u64<true>(reinterpret_cast<const u64*>(p));
        return make_int2((int)(unsigned)w, (int)(unsigned)(w >> 32));
    }
    return *p;
}
template <bool CO>
__device__ __forceinline__ int4 ld_i4(const int4* p) {
    if (CO) {
        const u64* q = reinterpret_cast<const u64*>(p);
        const u64 a = ld_u64<true>(q), b = ld_u64<true>(q + 1);
        return make_int4((int)(unsigned)a, (int)(unsigned)(a >> 32), (int)(unsigned)b,
                         (int)(unsigned)(b >> 32));
    }
    return *p;
}
template <bool CO>
__device__ __forceinline__ void st_i(int* p, int v) {
    if (CO) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}
template <bool CO>
__device__ __forceinline__ void st_u64(u64* p, u64 v) {
    if (CO) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}
// 16 bytes stored device-coherent: the same cache policy (sc1) as an
// agent-scope atomic store, four dwords wide (no builtin gives it).
__device__ __forceinline__ void st_i4_dev(int* p, int4 x) {
    typedef int v4i __attribute__((ext_vector_type(4)));
    const v4i y = {x.x, x.y, x.z, x.w};
    asm volatile("global_store_dwordx4 %0, %1, off sc1" ::"v"(p), "v"(y) : "memory");
}
template <bool CO>
__device__ __forceinline__ void st_i2(int2* p, int2 v) {
    if (CO) st_u64<true>(reinterpret_cast<u64*>(p), (u64)(unsigned)v.x | ((u64)(unsigned)v.y << 32));
    else *p = v;
}

// Per-tile summaries written by the tile pass.
//   The right spine of a tile is its strict suffix minima (elements smaller
//   than everything after them in the tile).  The nearest qualifying element
//   of a later tile's query inside tile s is always on s's right spine: a
//   later element at most as large would be nearer.  Spine values decrease
//   with rank (rank 0 = the tile's last element), so the answer is the
//   lowest-ranked qualifying entry.
struct TileOut {
    int* cmin;          // [nt * CPT] chunk mins
    int* qcount;        // [nt] number of open elements
    int2* qlist;        // [nt * QS] (tile-relative position, value)
    int* scount;        // [nt] number of right-spine entries
    int2* spine;        // [nt * RS] (tile-relative position, value) by rank
};

// LDS of one tile.  Element e = lane*K + i lives in val4[i >> 2][lane],
// component i & 3: lane-contiguous dwordx4 rows, so every ds_*_b128 of the
// tile pass is bank-conflict free.  Lane-min windows live inside 16-lane
// rows (the reach of a DPP row shift): win[s][l] = min of the lane mins of
// lanes max(l - 2^s + 1, row start) .. l for s = 0..3 (44 KB: three tiles
// per CU).
struct TileSmem {
    int4 val4[K4][NT];          // the tile's values
    int4 gmin[NT];              // each lane's four group-of-4 mins
    int win[4][NT];             // lane-min windows inside rows
    int4 rmin[NWAVE];           // row mins of each wave
    int wmin[NWAVE];            // wave mins
    int wopen[NWAVE];           // open elements per wave
    int wspine[NWAVE];          // right-spine entries per wave
    int wlast[NWAVE];           // segmented solves: last segment start per wave
    union {
        // per-wave queue: query (lane << 4 | slot) -> tile-relative answer
        unsigned short q[NWAVE][QCAP];
        // segmented solves, before the tile is solved: segment starts
        // inside the tile
        unsigned walls[TILE / 32];
    };
};


// Segmented (CSR batch) solves.  Segment k covers [offs[k], offs[k+1]) of
// the concatenated keys; answers are relative to the segment start, -1 when
// the segment holds no qualifying predecessor.  The solve runs on the
// concatenation: the nearest qualifying predecessor j over the whole array
// is the in-segment answer iff j >= the segment start (otherwise no element
// of [start, i) qualifies, since j is the nearest one), so every answer is
// the plain answer filtered and rebased; the fix-up bounds its prefix-min
// and its searches by the segment start.
struct SegIn {
    const long long* offs;  // [nseg + 1], non-decreasing, offs[0] = 0, offs[nseg] = n
    const int* tmap;        // [nt]: segment holding each tile's first element
    int nseg;
    const int2* tbound;     // [nt]: that segment's start and end
};

// Per lane of a segmented tile: the start of the segment holding the lane's
// first element, and the segment starts inside the lane (bit i = element i).
struct SegLane {
    int incoming;
    unsigned wallmask;
};

// Answer `res` (global index or -1) of the lane's element i, rebased to its
// segment (identity for plain solves).
// (res - st is below -1 for every answer before st, -1 included: st >= 0)
template <bool SEG>
__device__ __forceinline__ int seg_rel(int res, int i, int base, const SegLane& sl) {
    if (!SEG) return res;
    const unsigned m = sl.wallmask & ((2u << i) - 1u);
    const int st = m ? base + 31 - __builtin_clz(m) : sl.incoming;
    return max(res - st, -1);
}

__device__ __forceinline__ int tile_value(const TileSmem& sm, int e) {
    // one dword read: val4[(e >> 2) & 3][e >> 4], component e & 3
    const int* w = reinterpret_cast<const int*>(&sm.val4[0][0]);
    return w[((e >> 2) & (K4 - 1)) * (NT * 4) + (e >> 4) * 4 + (e & 3)];
}

// Value i (0 <= i < K) of lane tid, from its own LDS slots.
__device__ __forceinline__ int own_value(const TileSmem& sm, int tid, int i) {
    return reinterpret_cast<const int*>(&sm.val4[(i >> 2) & (K4 - 1)][tid])[i & 3];
}

// The lane's K values back from LDS (four 16-byte reads).
__device__ __forceinline__ void lane_values(const TileSmem& sm, int tid, int* w) {
#pragma unroll
    for (int q = 0; q < K4; q++) {
        const int4 a = sm.val4[q][tid];
        w[4 * q] = a.x; w[4 * q + 1] = a.y; w[4 * q + 2] = a.z; w[4 * q + 3] = a.w;
    }
}

// DPP data-parallel-primitive controls (GFX9 encoding): a VALU operand read
// from another lane of the same row (16 lanes) or wave, at no extra latency.
constexpr int DPP_ROW_SHL = 0x100;      // + n: lane + n inside the row
constexpr int DPP_ROW_SHR = 0x110;      // + n: lane - n inside the row
constexpr int DPP_WAVE_SHR1 = 0x138;    // lane - 1 across the wave
constexpr int DPP_ROW_BCAST15 = 0x142;  // lane 15 of each row, read by the next row
constexpr int DPP_ROW_BCAST31 = 0x143;  // lane 31, read by rows 2 and 3

// The value of x at the lane CTRL names; `ident` where that lane lies outside
// the row / wave or the reading row is not in ROWS.
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ int dpp(int ident, int x) {
    return __builtin_amdgcn_update_dpp(ident, x, CTRL, ROWS, 0xf, false);
}
// Row shift with lanes whose source lies outside the row reading 0 (sums).
template <int CTRL>
__device__ __forceinline__ int dpp0(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xf, 0xf, true);
}
__device__ __forceinline__ int readlane(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
// The lane's index in its wave, computed (bits of an all-ones mask below the
// lane), not read from the work-item id: a persistent kernel that needs the
// id across a long loop would otherwise hold it in a register throughout.
__device__ __forceinline__ int lane_id() {
    return (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ int lane63(int x) { return __builtin_amdgcn_readlane(x, WAVE - 1); }

// Min over the wave (DPP row scans, then the row carries; no LDS round
// trip).  The whole wave must be active.
__device__ __forceinline__ int wave_min(int x) {
    x = min(x, dpp<DPP_ROW_SHR + 1>(INF, x));
    x = min(x, dpp<DPP_ROW_SHR + 2>(INF, x));
    x = min(x, dpp<DPP_ROW_SHR + 4>(INF, x));
    x = min(x, dpp<DPP_ROW_SHR + 8>(INF, x));
    x = min(x, dpp<DPP_ROW_BCAST15, 0xa>(INF, x));
    x = min(x, dpp<DPP_ROW_BCAST31, 0xc>(INF, x));
    return lane63(x);
}

// Intra-wave LDS hand-off: LDS operations of one wave complete in order; this
// keeps the compiler from moving them across and waits for issued ones.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Rightmost qualifying index in the aligned 32-entry block [b, b+32) (b a
// multiple of 4), all loads issued together; -1 if none.
template <bool INCL, bool CO = false>
__device__ __forceinline__ int last_in_32(const int* __restrict__ arr, int b, int v) {
    int4 w[8];
#pragma unroll
    for (int k = 0; k < 8; k++) w[k] = ld_i4<CO>(reinterpret_cast<const int4*>(arr + b + 4 * k));
    int r = -1;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        r = qual<INCL>(w[k].x, v) ? b + 4 * k + 0 : r;
        r = qual<INCL>(w[k].y, v) ? b + 4 * k + 1 : r;
        r = qual<INCL>(w[k].z, v) ? b + 4 * k + 2 : r;
        r = qual<INCL>(w[k].w, v) ? b + 4 * k + 3 : r;
    }
    return r;
}

// Rightmost j in [0, K) with qual(lane lt's value j, v), or -1.
template <bool INCL, int J0>
__device__ __forceinline__ void sel4(int& r, int a0, int a1, int a2, int a3, int x);

// Two steps instead of sixteen selects: the last group of four whose min
// qualifies (the lane's group mins, one read), then the last qualifying
// value in that group (one more read).
template <bool INCL>
__device__ __forceinline__ int last_in_lane(const TileSmem& sm, int lt, int v) {
    const int4 g = sm.gmin[lt];
    int q = -1;
    sel4<INCL, 0>(q, g.x, g.y, g.z, g.w, v);
    // the read is unconditional (group 0 when none qualifies)
    const int4 w = sm.val4[max(q, 0)][lt];
    int j = -1;
    sel4<INCL, 0>(j, w.x, w.y, w.z, w.w, v);
    return q < 0 ? -1 : 4 * q + j;  // j >= 0: the group's min qualifies
}

// Answer (tile-relative element index) of the query value `val` asked by
// tile lane qt, given that some lane before qt holds a qualifying element:
// pick the nearest row holding one (own row, earlier rows of the wave, then
// the nearest qualifying earlier wave's last qualifying row), lift inside
// that row over the windows, then scan the lane.
template <bool INCL>
__device__ __forceinline__ int intile_answer(const TileSmem& sm, int qt, int val) {
    const int wq = qt >> 6, ql = qt & (WAVE - 1), r = ql >> 4;
    // every read is unconditional (clamped index) and every choice a
    // select: the queue resolves one query per lane, and a branch taken by
    // any lane would be paid by all of them
    const int rl = ql & 15;
    const int a1 = sm.win[3][max(qt - 1, 0)];
    const int a9 = sm.win[3][max(qt - 9, 0)];
    const int4 rq = sm.rmin[wq];
    const int4 wa = *reinterpret_cast<const int4*>(&sm.wmin[0]);
    // min of the own row's lanes before qt: two 8-lane windows
    const int rowpre = min(a1, rl > 8 ? a9 : INF);
    // nearest earlier row of the own wave holding a qualifying element
    int rr = -1;
    rr = ((r > 0) & qual<INCL>(rq.x, val)) ? 0 : rr;
    rr = ((r > 1) & qual<INCL>(rq.y, val)) ? 1 : rr;
    rr = ((r > 2) & qual<INCL>(rq.z, val)) ? 2 : rr;
    // nearest earlier wave holding one (wave 0 when none is found: the
    // caller guarantees an answer in the tile)
    int w = 0;
    w = ((wq > 1) & qual<INCL>(wa.y, val)) ? 1 : w;
    w = ((wq > 2) & qual<INCL>(wa.z, val)) ? 2 : w;
    w = ((wq > 3) & qual<INCL>(wa.w, val)) ? 3 : w;
    if constexpr (NWAVE > 4) {
        const int4 wb = *reinterpret_cast<const int4*>(&sm.wmin[4]);
        w = ((wq > 4) & qual<INCL>(wb.x, val)) ? 4 : w;
        w = ((wq > 5) & qual<INCL>(wb.y, val)) ? 5 : w;
        w = ((wq > 6) & qual<INCL>(wb.z, val)) ? 6 : w;
        w = ((wq > 7) & qual<INCL>(wb.w, val)) ? 7 : w;
    }
    const int4 rw = sm.rmin[w];
    const int rw_last = qual<INCL>(rw.w, val) ? 3 : qual<INCL>(rw.z, val) ? 2
                      : qual<INCL>(rw.y, val) ? 1 : 0;
    int p = ((rl > 0) & qual<INCL>(rowpre, val)) ? qt - 1
          : rr >= 0 ? wq * WAVE + rr * 16 + 15
          : w * WAVE + rw_last * 16 + 15;
    // nearest lane <= p of p's row whose min qualifies (one exists)
#pragma unroll
    for (int s = 3; s >= 0; s--)
        if (!qual<INCL>(sm.win[s][p], val)) p -= 1 << s;
    p = max(p, 0);
    const int j = last_in_lane<INCL>(sm, p, val);
    return p * K + (j < 0 ? 0 : j);
}

// ---------------------------------------------------------------------------
// Tile pass: solve tile t inside the workgroup, write every output of the
// tile (open elements as -1), publish chunk mins, open list, tile min and
// the tile's contribution to level 1.  Ends with a workgroup barrier, so
// `sm` may be reused for the next tile right away.
// ---------------------------------------------------------------------------
// This lane's K values of tile t (INF past the end of the array).
__device__ __forceinline__ void load_lane(const int* __restrict__ in, int n, int t, int* v,
                                          int tid = (int)threadIdx.x) {
    const int base = t * TILE + tid * K;
    if (t * TILE <= n - TILE) {  // (not t * TILE + TILE: it overflows near 2^31)
#pragma unroll
        for (int q = 0; q < K4; q++) {
            const int4 a = *reinterpret_cast<const int4*>(in + base + 4 * q);
            v[4 * q] = a.x; v[4 * q + 1] = a.y; v[4 * q + 2] = a.z; v[4 * q + 3] = a.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < K; i++) v[i] = (base + i < n) ? in[base + i] : INF;
    }
}

// r = J0 + k for the last k in 0..3 with qual(a_k, x), else r unchanged.
// The four compares write four separate lane masks before the selects read
// them: a select reading the mask its compare has just written waits two
// cycles (the hardware's VALU-writes-SGPR, VALU-reads-it-as-mask rule), which
// the compiler pays with a nop on every pair when it keeps the mask in VCC.
template <bool INCL, int J0>
__device__ __forceinline__ void sel4(int& r, int a0, int a1, int a2, int a3, int x) {
    u64 m0, m1, m2, m3;
    if (INCL)
        asm volatile(
            "v_cmp_le_i32_e64 %[m0], %[a0], %[x]\n"
            "v_cmp_le_i32_e64 %[m1], %[a1], %[x]\n"
            "v_cmp_le_i32_e64 %[m2], %[a2], %[x]\n"
            "v_cmp_le_i32_e64 %[m3], %[a3], %[x]\n"
            "v_cndmask_b32_e64 %[r], %[r], %[j0], %[m0]\n"
            "v_cndmask_b32_e64 %[r], %[r], %[j1], %[m1]\n"
            "v_cndmask_b32_e64 %[r], %[r], %[j2], %[m2]\n"
            "v_cndmask_b32_e64 %[r], %[r], %[j3], %[m3]\n"
            : [r] "+v"(r), [m0] "=&s"(m0), [m1] "=&s"(m1), [m2] "=&s"(m2), [m3] "=&s"(m3)
            : [a0] "v"(a0), [a1] "v"(a1), [a2] "v"(a2), [a3] "v"(a3), [x] "v"(x),
              [j0] "I"(J0), [j1] "I"(J0 + 1), [j2] "I"(J0 + 2), [j3] "I"(J0 + 3));
    else
        asm volatile(
            "v_cmp_lt_i32_e64 %[m0], %[a0], %[x]\n"
            "v_cmp_lt_i32_e64 %[m1], %[a1], %[x]\n"
            "v_cmp_lt_i32_e64 %[m2], %[a2], %[x]\n"
            "v_cmp_lt_i32_e64 %[m3], %[a3], %[x]\n"
            "v_cndmask_b32_e64 %[r], %[r], %[j0], %[m0]\n"
            "v_cndmask_b32_e64 %[r], %[r], %[j1], %[m1]\n"
            "v_cndmask_b32_e64 %[r], %[r], %[j2], %[m2]\n"
            "v_cndmask_b32_e64 %[r], %[r], %[j3], %[m3]\n"
            : [r] "+v"(r), [m0] "=&s"(m0), [m1] "=&s"(m1), [m2] "=&s"(m2), [m3] "=&s"(m3)
            : [a0] "v"(a0), [a1] "v"(a1), [a2] "v"(a2), [a3] "v"(a3), [x] "v"(x),
              [j0] "I"(J0), [j1] "I"(J0 + 1), [j2] "I"(J0 + 2), [j3] "I"(J0 + 3));
}

// sel4 starting from `seed` instead of r: the first select reads the seed
// and writes r, so a seed shared by many elements is not copied per element.
template <bool INCL, int J0>
__device__ __forceinline__ int sel4_from(int seed, int a0, int a1, int a2, int a3, int x) {
    int r;
    u64 m0, m1, m2, m3;
    if (INCL)
        asm volatile(
            "v_cmp_le_i32_e64 %[m0], %[a0], %[x]\n"
            "v_cmp_le_i32_e64 %[m1], %[a1], %[x]\n"
            "v_cmp_le_i32_e64 %[m2], %[a2], %[x]\n"
            "v_cmp_le_i32_e64 %[m3], %[a3], %[x]\n"
            "v_cndmask_b32_e64 %[r], %[s], %[j0], %[m0]\n"
            "v_cndmask_b32_e64 %[r], %[r], %[j1], %[m1]\n"
            "v_cndmask_b32_e64 %[r], %[r], %[j2], %[m2]\n"
            "v_cndmask_b32_e64 %[r], %[r], %[j3], %[m3]\n"
            : [r] "=v"(r), [m0] "=&s"(m0), [m1] "=&s"(m1), [m2] "=&s"(m2), [m3] "=&s"(m3)
            : [s] "v"(seed), [a0] "v"(a0), [a1] "v"(a1), [a2] "v"(a2), [a3] "v"(a3),
              [x] "v"(x), [j0] "I"(J0), [j1] "I"(J0 + 1), [j2] "I"(J0 + 2), [j3] "I"(J0 + 3));
    else
        asm volatile(
            "v_cmp_lt_i32_e64 %[m0], %[a0], %[x]\n"
            "v_cmp_lt_i32_e64 %[m1], %[a1], %[x]\n"
            "v_cmp_lt_i32_e64 %[m2], %[a2], %[x]\n"
            "v_cmp_lt_i32_e64 %[m3], %[a3], %[x]\n"
            "v_cndmask_b32_e64 %[r], %[s], %[j0], %[m0]\n"
            "v_cndmask_b32_e64 %[r], %[r], %[j1], %[m1]\n"
            "v_cndmask_b32_e64 %[r], %[r], %[j2], %[m2]\n"
            "v_cndmask_b32_e64 %[r], %[r], %[j3], %[m3]\n"
            : [r] "=v"(r), [m0] "=&s"(m0), [m1] "=&s"(m1), [m2] "=&s"(m2), [m3] "=&s"(m3)
            : [s] "v"(seed), [a0] "v"(a0), [a1] "v"(a1), [a2] "v"(a2), [a3] "v"(a3),
              [x] "v"(x), [j0] "I"(J0), [j1] "I"(J0 + 1), [j2] "I"(J0 + 2), [j3] "I"(J0 + 3));
    return r;
}

// r = the last j in [J0, i) with qual(v[j], v[i]), else r (j as an inline
// constant): blocks of four, then the last one to three as plain selects.
template <bool INCL, int J0>
__device__ __forceinline__ void inlane_select(int& r, const int* v, int i) {
    if (J0 >= i) return;
    if (i - J0 >= 4) {
        sel4<INCL, J0>(r, v[J0], v[J0 + 1], v[J0 + 2], v[J0 + 3], v[i]);
        if constexpr (J0 + 4 < K - 1) inlane_select<INCL, J0 + 4>(r, v, i);
    } else {
#pragma unroll
        for (int j = J0; j < J0 + 3; j++)
            if (j < i) r = qual<INCL>(v[j], v[i]) ? j : r;
    }
}

// Bit i = qual(p, v[i]), from the sign of a saturating difference (no lane
// masks): p < v iff sat(p - v) < 0; p <= v iff not sat(v - p) < 0.
template <bool INCL>
__device__ __forceinline__ unsigned qual_bits(int p, const int* v) {
    unsigned m = 0;
#pragma unroll
    for (int i = K - 1; i >= 0; i--) {
        const int d = INCL ? __builtin_elementwise_sub_sat(v[i], p)
                           : __builtin_elementwise_sub_sat(p, v[i]);
        m = __builtin_amdgcn_alignbit(m, (unsigned)d, 31);
    }
    return INCL ? (~m & 0xFFFFu) : m;
}

// OPEN_CO: the outputs are written with device-coherent stores (through
// this XCD's L2), so that another workgroup, on any XCD, can overwrite the
// open elements' -1 with device-coherent stores of its own without a dirty
// copy in this workgroup's L2 being written back over them later; the open
// list and its count are written device-coherent too.
template <bool INCL, bool SEG = false, bool L1 = true, bool CO = false, bool OPEN_CO = false>
__device__ __forceinline__ void solve_tile(int* __restrict__ out, int n, int t, const int* v,
                                           int prev_last, TileSmem& sm, const TileOut& to,
                                           const Levels& L,
                                           const SegLane sl = SegLane{0, 0u},
                                           const int tid = (int)threadIdx.x,
                                           unsigned* tk_ctr = nullptr, unsigned* tk_slot = nullptr) {
    const int lane = tid & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar tests
    const int row = lane >> 4;
    const int tile_start = t * TILE;
    const int base = tile_start + tid * K;
    // elements of this lane inside the array (the last tile is partial; its
    // padding is INF and never becomes a query, an open element or an output)
    const int nvalid = min(max(n - base, 0), K);
    const unsigned vmask = (nvalid >= K) ? 0xFFFFu : ((1u << nvalid) - 1u);

    // 1. in-lane answers (global index, -1 if none in the lane).  The lane
    //    offset of the answer is selected among inline constants, starting
    //    from `none` (base + none == -1), and rebased once.
    int res[K];
    res[0] = -1;
    const int none = -1 - base;
#pragma unroll
    for (int i = 1; i < K; i++) {
        int r = none;
        if (i >= 4) {
            r = sel4_from<INCL, 0>(none, v[0], v[1], v[2], v[3], v[i]);
            inlane_select<INCL, 4>(r, v, i);
        } else {
            inlane_select<INCL, 0>(r, v, i);
        }
        res[i] = base + r;
    }
    int gm[K4];
#pragma unroll
    for (int q = 0; q < K4; q++)
        gm[q] = min(min(v[4 * q], v[4 * q + 1]), min(v[4 * q + 2], v[4 * q + 3]));
    const int lmin = min(min(gm[0], gm[1]), min(gm[2], gm[3]));
    static_assert(K4 == 4, "group mins: one int4 per lane");

#pragma unroll
    for (int q = 0; q < K4; q++)
        sm.val4[q][tid] = make_int4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    sm.gmin[tid] = make_int4(gm[0], gm[1], gm[2], gm[3]);

    // 2. lane-min windows inside rows, wave prefix / suffix mins (DPP)
    const int w2 = min(lmin, dpp<DPP_ROW_SHR + 1>(INF, lmin));
    const int w4 = min(w2, dpp<DPP_ROW_SHR + 2>(INF, w2));
    const int w8 = min(w4, dpp<DPP_ROW_SHR + 4>(INF, w4));
    const int w16 = min(w8, dpp<DPP_ROW_SHR + 8>(INF, w8));  // row start .. lane
    sm.win[0][tid] = lmin;
    sm.win[1][tid] = w2;
    sm.win[2][tid] = w4;
    sm.win[3][tid] = w8;
    // window of LPC lanes = one 64-element chunk
    if ((lane & (LPC - 1)) == LPC - 1) st_i<CO>(&to.cmin[(size_t)t * CPT + (tid >> LOG_LPC)], w4);
    int winc = min(w16, dpp<DPP_ROW_BCAST15, 0xa>(INF, w16));
    winc = min(winc, dpp<DPP_ROW_BCAST31, 0xc>(INF, winc));  // wave start .. lane
    const int wexcl = dpp<DPP_WAVE_SHR1>(INF, winc);          // wave start .. lane - 1
    const int rm0 = readlane(w16, 15), rm1 = readlane(w16, 31);
    const int rm2 = readlane(w16, 47), rm3 = readlane(w16, 63);
    const int r23 = min(rm2, rm3), r123 = min(rm1, r23);
    int sfx = min(lmin, dpp<DPP_ROW_SHL + 1>(INF, lmin));
    sfx = min(sfx, dpp<DPP_ROW_SHL + 2>(INF, sfx));
    sfx = min(sfx, dpp<DPP_ROW_SHL + 4>(INF, sfx));
    sfx = min(sfx, dpp<DPP_ROW_SHL + 8>(INF, sfx));           // lane .. row end
    const int later_rows = row == 0 ? r123 : row == 1 ? r23 : row == 2 ? rm3 : INF;
    const int wafter = min(dpp<DPP_ROW_SHL + 1>(INF, sfx), later_rows);  // lane + 1 .. wave end
    if (lane == 0) {
        sm.rmin[wave] = make_int4(rm0, rm1, rm2, rm3);
        sm.wmin[wave] = min(rm0, r123);
    }
    __syncthreads();
    // a persistent solve's ticket for its next tile (tk_ctr): taken once the
    // tile's keys have arrived and handed over (*tk_slot) before the tile's
    // closing barrier, so its round trip overlaps the tile's work
    unsigned tk = 0u;
    if (tk_ctr && tid == 0)
        tk = __hip_atomic_fetch_add(tk_ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

    // mins of the earlier and of the later waves: uniform, so they are
    // folded in scalar registers
    int cross = INF, after = INF;
    {
        int wm[8];
        // (read in groups of four: with two waves the second half of the
        // group lies past wmin and is not used)
#pragma unroll
        for (int h = 0; h < (NWAVE + 3) / 4; h++) {
            const int4 a = *reinterpret_cast<const int4*>(&sm.wmin[4 * h]);
            wm[4 * h] = __builtin_amdgcn_readfirstlane(a.x);
            wm[4 * h + 1] = __builtin_amdgcn_readfirstlane(a.y);
            wm[4 * h + 2] = __builtin_amdgcn_readfirstlane(a.z);
            wm[4 * h + 3] = __builtin_amdgcn_readfirstlane(a.w);
        }
#pragma unroll
        for (int w = 0; w < NWAVE; w++) {
            if (w < wave) cross = min(cross, wm[w]);
            if (w > wave) after = min(after, wm[w]);
        }
    }
    // exclusive prefix-min of the tile's earlier lanes; meaningful only
    // when has_prev (INCL needs the distinction: INF is a legal key)
    const bool has_prev = tid > 0;
    const int pmx = min(wexcl, cross);

    // classify lane-unanswered elements.  They are the lane's prefix minima
    // (non-increasing, the last one is the lane min), so when pmx beats the
    // lane min it beats them all and every one has an answer in the tile;
    // only lanes holding tile prefix minima need the per-element split.
    // (masks are assembled one bit per op, highest element first:
    // alignbit(m, x, 31) = m << 1 | the sign bit of x)
    unsigned rec = 0;
#pragma unroll
    for (int i = K - 1; i >= 0; i--) rec = __builtin_amdgcn_alignbit(rec, (unsigned)res[i], 31);
    unsigned qmask = rec, omask = 0;
    if (!has_prev || !qual<INCL>(pmx, lmin)) {
        // v[0] is the largest of them: when pmx does not beat it, it beats
        // none (every lane of descending input), and the per-element
        // split is skipped
        unsigned beaten = 0;
        if (has_prev && qual<INCL>(pmx, v[0])) beaten = qual_bits<INCL>(pmx, v);
        qmask = rec & beaten;
        omask = rec & ~beaten;
    }
    qmask &= vmask;
    omask &= vmask;
    if (SEG && sl.wallmask) {
        // an element at or after a segment start in this lane has its
        // answer, if any, inside the lane: the in-lane one is final (seg_rel
        // drops it when it lies before the start), so it is neither a
        // cross-lane query nor open
        const unsigned from_wall = ~((sl.wallmask & (0u - sl.wallmask)) - 1u);
        qmask &= ~from_wall;
        omask &= ~from_wall;
    }
    // The tile's first element is open (nothing precedes it in the tile).
    // When the key just before the tile (prev_last, read by lane 0)
    // qualifies, that key is its answer: settled here, it leaves the
    // fix-up's list, and ascending input leaves the fix-up nothing to do.
    // (The fix-up's early-out stays valid: it tests the first element's
    // value, which bounds every open value.)
    if (tid == 0 && t > 0 && qual<INCL>(prev_last, v[0])) {
        res[0] = tile_start - 1;
        omask &= ~1u;
    }

    // right spine: strict suffix minima of the tile.  A lane holds some iff
    // its min is below everything after it; the tile's last element has
    // nothing after it and is always on the spine (INF stands for "nothing"
    // here, but INF is also a legal key).
    unsigned smask = 0;
    {
        const int A = min(wafter, after);
        if (lmin < A || tid == NT - 1) {
            int m = A;
#pragma unroll
            for (int i = K - 1; i >= 0; i--) {  // bit i = v[i] < m (sign bit, one op)
                smask = __builtin_amdgcn_alignbit(smask, (unsigned)__builtin_elementwise_sub_sat(v[i], m), 31);
                m = min(m, v[i]);
            }
            if (tid == NT - 1) smask |= 1u << (K - 1);
        }
        smask &= vmask;
    }

    // 3a. The lane's first element (always unanswered in-lane; when it is
    //     not a query, no element of the lane is) and up to XQ more of the
    //     lane's queries, in position order, against the previous lane's
    //     values.  The queries are prefix minima of the lane, so each is at
    //     most as large as the one before and its answer lies no later: once
    //     one has no answer in the previous lane, no later one has, and the
    //     answers found are accepted up to the first miss.  All of them are
    //     searched at once, so that their LDS reads (the query values, the
    //     previous lane's group mins, then one group of its values per
    //     query) overlap instead of forming a chain per query.  On uniform
    //     random keys this settles about two thirds of the queue's entries.
    //     The low 16 bits of xa mark the further elements so answered; the
    //     answer of the k-th of them is at offset (xa >> (16 + 4k)) & 15 of
    //     the previous lane.  They are written with the staged or patched
    //     outputs below.
    unsigned xa = 0;
    if (qmask & 1u) {
        const int pt = tid - 1;  // >= 0: the tile's first lane has no queries
        const int4 g = sm.gmin[pt];
        int q0 = -1;
        sel4<INCL, 0>(q0, g.x, g.y, g.z, g.w, v[0]);
        // one group of the previous lane per query (group 0 when none
        // qualifies: the read is unconditional)
        const int4 w0 = sm.val4[max(q0, 0)][pt];
        int j0 = -1;
        sel4<INCL, 0>(j0, w0.x, w0.y, w0.z, w0.w, v[0]);
        if (qmask & ~1u) {  // further queries (none on ascending keys)
            unsigned m = qmask & ~1u;
            int ik[XQ], xv[XQ], qk[XQ];
            int4 wk[XQ];
#pragma unroll
            for (int k = 0; k < XQ; k++) {
                ik[k] = m ? __builtin_ctz(m) : 0;  // (position 0 stands in for none)
                m &= m - 1;
                xv[k] = own_value(sm, tid, ik[k]);
            }
#pragma unroll
            for (int k = 0; k < XQ; k++) {
                qk[k] = -1;
                sel4<INCL, 0>(qk[k], g.x, g.y, g.z, g.w, xv[k]);
                wk[k] = sm.val4[max(qk[k], 0)][pt];
            }
            if (q0 >= 0) {
                unsigned mq = qmask & ~1u;
                bool ok = true;
#pragma unroll
                for (int k = 0; k < XQ; k++) {
                    int j = -1;
                    sel4<INCL, 0>(j, wk[k].x, wk[k].y, wk[k].z, wk[k].w, xv[k]);
                    if (ok && mq && qk[k] >= 0) {
                        xa |= (1u << ik[k]) | ((unsigned)(4 * qk[k] + j) << (16 + 4 * k));
                        mq &= mq - 1;
                    } else {
                        ok = false;
                    }
                }
                qmask = mq | 1u;  // (bit 0 cleared below)
            }
        }
        if (q0 >= 0) {  // j0 >= 0: the group's min qualifies
            res[0] = base - K + 4 * q0 + j0;
            qmask &= ~1u;
        }
    }

    // one scan for both counts: queries (low 16 bits), open (high 16 bits)
    const int cq = __builtin_popcount(qmask), co = __builtin_popcount(omask);
    int packed = cq | (co << 16);
    packed += dpp0<DPP_ROW_SHR + 1>(packed);
    packed += dpp0<DPP_ROW_SHR + 2>(packed);
    packed += dpp0<DPP_ROW_SHR + 4>(packed);
    packed += dpp0<DPP_ROW_SHR + 8>(packed);
    packed += dpp<DPP_ROW_BCAST15, 0xa>(0, packed);
    packed += dpp<DPP_ROW_BCAST31, 0xc>(0, packed);
    const int ptotal = readlane(packed, WAVE - 1);
    const int qoff = (packed & 0xFFFF) - cq, qtotal = ptotal & 0xFFFF;
    const int ooff = (packed >> 16) - co;
    // spine entries after this lane inside the wave (suffix count)
    const int cs = __builtin_popcount(smask);
    int zs = cs;
    zs += dpp0<DPP_ROW_SHL + 1>(zs);
    zs += dpp0<DPP_ROW_SHL + 2>(zs);
    zs += dpp0<DPP_ROW_SHL + 4>(zs);
    zs += dpp0<DPP_ROW_SHL + 8>(zs);  // lane .. row end
    const int rt0 = readlane(zs, 0), rt1 = readlane(zs, 16);
    const int rt2 = readlane(zs, 32), rt3 = readlane(zs, 48);
    zs += row == 0 ? rt1 + rt2 + rt3 : row == 1 ? rt2 + rt3 : row == 2 ? rt3 : 0;
    if (lane == 0) {
        sm.wopen[wave] = ptotal >> 16;
        sm.wspine[wave] = rt0 + rt1 + rt2 + rt3;
    }
    const int safter_w = zs - cs;

    // The outputs of a wave whose queries all fit its queue (nearly every
    // wave) are written after the queue, once, through LDS and coalesced
    // (below).  A wave with more queries writes its own lanes' 16
    // consecutive outputs now and patches the answers in afterwards.
    const bool staged = qtotal <= QCAP;
    if (!staged) {
        if (nvalid == K && OPEN_CO) {
#pragma unroll
            for (int q = 0; q < K4; q++)
                st_i4_dev(out + base + 4 * q, make_int4(res[4 * q], res[4 * q + 1],
                                                        res[4 * q + 2], res[4 * q + 3]));
        } else if (nvalid == K) {
#pragma unroll
            for (int q = 0; q < K4; q++)
                *reinterpret_cast<int4*>(out + base + 4 * q) =
                    make_int4(seg_rel<SEG>(res[4 * q], 4 * q, base, sl),
                              seg_rel<SEG>(res[4 * q + 1], 4 * q + 1, base, sl),
                              seg_rel<SEG>(res[4 * q + 2], 4 * q + 2, base, sl),
                              seg_rel<SEG>(res[4 * q + 3], 4 * q + 3, base, sl));
        } else {
#pragma unroll
            for (int i = 0; i < K; i++)
                if (i < nvalid) st_i<OPEN_CO>(&out[base + i], seg_rel<SEG>(res[i], i, base, sl));
        }
        unsigned xm = xa & 0xFFFFu;
#pragma unroll
        for (int k = 0; k < XQ; k++)
            if (xm) {
                const int i = __builtin_ctz(xm);
                xm &= xm - 1;
                out[base + i] = seg_rel<SEG>(base - K + (int)((xa >> (16 + 4 * k)) & 15u), i,
                                             base, sl);
            }
    }

    // 3b. the rest through the wave's queue: compact, resolve one query per
    //     lane, hand the answers back.  In an unstaged wave the owner
    //     overwrites its placeholder (same-thread store order); entries
    //     beyond QCAP are resolved by their owner.  The lane that resolves
    //     entry `lane` (the first round) keeps the entry and its answer: in
    //     a staged wave it patches the answer into the owner's staged
    //     outputs itself (below).
    unsigned qitem = 0u;
    int qans = 0;
    if (qtotal > 0) {
        unsigned m = qmask;
        int k = qoff;
        if (qtotal <= QCAP) {  // (uniform: nearly every wave) no bound test per entry
            unsigned short* qd = &sm.q[wave][k];
            while (m) {
                *qd++ = (unsigned short)((tid << 4) | __builtin_ctz(m));
                m &= m - 1;
            }
        } else {
            while (m) {
                const int i = __builtin_ctz(m);
                m &= m - 1;
                if (k < QCAP) sm.q[wave][k] = (unsigned short)((tid << 4) | i);
                k++;
            }
        }
        wave_lds_sync();
        for (int r = lane; r < min(qtotal, QCAP); r += WAVE) {
            const int item = sm.q[wave][r];
            const int a = intile_answer<INCL>(sm, item >> 4, tile_value(sm, item));
            sm.q[wave][r] = (unsigned short)a;
            if (!CO && (!SEG || NWAVE <= 4) && r < WAVE) {  // (RESOLVER_PATCH, below)
                qitem = (unsigned)item;
                qans = a;
            }
        }
        wave_lds_sync();
        if (!staged) {
            m = qmask;
            k = qoff;
            while (m) {
                const int i = __builtin_ctz(m);
                m &= m - 1;
                const int r = tile_start + ((k < QCAP) ? (int)sm.q[wave][k]
                                                       : intile_answer<INCL>(sm, tid,
                                                                             tile_value(sm, tid * K + i)));
                out[base + i] = seg_rel<SEG>(r, i, base, sl);
                k++;
            }
        }
    }
    __syncthreads();  // no wave reads the tile's values past this point

    // 4. open elements: (position, value) pairs in position order, up to QS
    int obase = ooff;
#pragma unroll
    for (int w = 0; w < NWAVE - 1; w++)
        if (w < wave) obase += sm.wopen[w];
    // (A lane holds a few at most, except on descending runs, where every
    // key is open: then all sixteen, unrolled; otherwise a loop over the
    // lane's open keys.  The values are read back from the lane's own LDS
    // slots, which only this wave's staging below overwrites, so that the
    // keys' registers need not stay live until here.  On random keys the
    // unrolled form for every lane cost the tile pass ~5 us with the
    // spine's below.)
    if (omask && obase < QS) {
        int2* const ql = to.qlist + (size_t)t * QS;
        if (omask == 0xFFFFu) {
            int w[K];
            lane_values(sm, tid, w);
#pragma unroll
            for (int i = 0; i < K; i++)
                if (obase + i < QS) st_i2<OPEN_CO>(&ql[obase + i], make_int2(tid * K + i, w[i]));
        } else {
            for (unsigned om = omask; om && obase < QS; om &= om - 1, obase++) {
                const int i = __builtin_ctz(om);
                st_i2<OPEN_CO>(&ql[obase], make_int2(tid * K + i, own_value(sm, tid, i)));
            }
        }
    }

    // 5. right spine: ranks from the tile's end, the first RS recorded
    int srank = safter_w;
#pragma unroll
    for (int w = 1; w < NWAVE; w++)
        if (w > wave) srank += sm.wspine[w];
    // (as the open list: every key of an ascending run is on the spine)
    if (smask && srank < RS) {
        int2* const sp = to.spine + (size_t)t * RS;
        if (smask == 0xFFFFu) {
            int w[K];
            lane_values(sm, tid, w);
#pragma unroll
            for (int i = K - 1; i >= 0; i--)
                if (srank + (K - 1 - i) < RS)
                    st_i2<CO || OPEN_CO>(&sp[srank + (K - 1 - i)], make_int2(tid * K + i, w[i]));
        } else {
            for (unsigned m = smask; m && srank < RS; srank++) {
                const int i = 31 - __builtin_clz(m);
                m &= ~(1u << i);
                st_i2<CO || OPEN_CO>(&sp[srank], make_int2(tid * K + i, own_value(sm, tid, i)));
            }
        }
    }
    // staged outputs.  Each wave stages its 1024 outputs in its own lanes'
    // value slots: 16-byte group g of the wave (elements 4g .. 4g+3) goes to
    // val4[g & 3][wave * 64 + ((g >> 2) ^ ((g & 3) << 1))].  The owner writes
    // its groups g = 4 lane + q, the queue answers are patched in, then lane
    // l reads back groups g = 64 q + l and stores them: 1 KB contiguous per
    // store instruction instead of 16 bytes at a 64-byte stride.  (The XOR
    // keeps the writes within aligned 8-slot groups and spreads each 8-lane
    // group of the read-back over 32 banks: bank-conflict cycles per wave
    // 56 -> 24 on descending input, measured, against an XOR by q << 2.)
    if (staged) {
        int* sv = reinterpret_cast<int*>(&sm.val4[0][0]);
        // no segment starts inside this wave's elements (nearly every wave
        // of a batch): every answer is rebased to the same start
        const bool one_start = !SEG || __ballot(sl.wallmask != 0) == 0;
        const int st = SEG ? sl.incoming : 0;  // (that start, when one_start)
        if (one_start) {
#pragma unroll
            for (int q = 0; q < K4; q++)
                sm.val4[q][wave * WAVE + (lane ^ (q << 1))] =
                    SEG ? make_int4(max(res[4 * q] - st, -1), max(res[4 * q + 1] - st, -1),
                                    max(res[4 * q + 2] - st, -1), max(res[4 * q + 3] - st, -1))
                        : make_int4(res[4 * q], res[4 * q + 1], res[4 * q + 2], res[4 * q + 3]);
        } else {
#pragma unroll
            for (int q = 0; q < K4; q++)
                sm.val4[q][wave * WAVE + (lane ^ (q << 1))] =
                    make_int4(seg_rel<SEG>(res[4 * q], 4 * q, base, sl),
                              seg_rel<SEG>(res[4 * q + 1], 4 * q + 1, base, sl),
                              seg_rel<SEG>(res[4 * q + 2], 4 * q + 2, base, sl),
                              seg_rel<SEG>(res[4 * q + 3], 4 * q + 3, base, sl));
        }
        // queue answers: the first round's by the lane that resolved each
        // (after the owners' writes above, in the same wave: LDS order),
        // unless segment starts in the wave make the rebasing per owner;
        // the rest by their owners
        // (not where the two registers held for it would spill: the
        // persistent kernel's loop and 512-lane segmented tiles)
        constexpr bool RESOLVER_PATCH = !CO && (!SEG || NWAVE <= 4);
        int kfrom = 0;  // owners patch entries kfrom ..
        if (RESOLVER_PATCH && one_start) {
            kfrom = WAVE;
            if (lane < qtotal) {
                const int ol = (int)(qitem >> 4) & (WAVE - 1), i = (int)(qitem & 15u), q = i >> 2;
                const int r = tile_start + qans;
                sv[(q * NT + wave * WAVE + (ol ^ (q << 1))) * 4 + (i & 3)] =
                    SEG ? max(r - st, -1) : r;
            }
        }
        if (qtotal > kfrom) {
            unsigned m = qmask;
            int k = qoff;
            while (m) {
                const int i = __builtin_ctz(m);
                m &= m - 1;
                const int q = i >> 2;
                if (k >= kfrom)
                    sv[(q * NT + wave * WAVE + (lane ^ (q << 1))) * 4 + (i & 3)] =
                        seg_rel<SEG>(tile_start + (int)sm.q[wave][k], i, base, sl);
                k++;
            }
        }
        unsigned xm = xa & 0xFFFFu;
#pragma unroll
        for (int x = 0; x < XQ; x++)
            if (xm) {
                const int i = __builtin_ctz(xm), q = i >> 2;
                xm &= xm - 1;
                sv[(q * NT + wave * WAVE + (lane ^ (q << 1))) * 4 + (i & 3)] =
                    seg_rel<SEG>(base - K + (int)((xa >> (16 + 4 * x)) & 15u), i, base, sl);
            }
        wave_lds_sync();
        const int wbase = tile_start + wave * (WAVE * K);
        // the wave's outputs all inside the array (uniform; every wave but
        // the last tile's): a scalar base and 32-bit lane offsets
        int* const wout = out + wbase;
        const bool whole = wbase <= n - WAVE * K;
#pragma unroll
        for (int q = 0; q < K4; q++) {
            const int c = lane & 3, r = 16 * q + (lane >> 2);
            const int4 x = sm.val4[c][wave * WAVE + (r ^ (c << 1))];
            const unsigned o = (unsigned)(q * (WAVE * 4) + lane * 4);
            const int e = wbase + (int)o;
            if (OPEN_CO) {
                if (e < n - 3) {
                    st_i4_dev(out + e, x);
                } else {
                    if (e < n) st_i<true>(&out[e], x.x);
                    if (e < n - 1) st_i<true>(&out[e + 1], x.y);
                    if (e < n - 2) st_i<true>(&out[e + 2], x.z);
                }
            } else if (whole) {
                // streaming (non-temporal) store: the answers are not read
                // back by this solve, and keeping them out of the caches
                // leaves those to the keys and summaries (measured: 1.8%
                // on the alternating-class solves, 4% on the batched
                // configuration; non-temporal key loads were 21% slower)
                typedef int v4i __attribute__((ext_vector_type(4)));
                const v4i y = {x.x, x.y, x.z, x.w};
                __builtin_nontemporal_store(y, reinterpret_cast<v4i*>(wout + o));
            } else if (e + 3 < n) {
                typedef int v4i __attribute__((ext_vector_type(4)));
                const v4i y = {x.x, x.y, x.z, x.w};
                __builtin_nontemporal_store(y, reinterpret_cast<v4i*>(out + e));
            } else {
                if (e < n) out[e] = x.x;
                if (e + 1 < n) out[e + 1] = x.y;
                if (e + 2 < n) out[e + 2] = x.z;
            }
        }
    }

    if (tid == 0) {
        int m = sm.wmin[0], total = sm.wopen[0], stotal = sm.wspine[0];
#pragma unroll
        for (int w = 1; w < NWAVE; w++) {
            m = min(m, sm.wmin[w]);
            total += sm.wopen[w];
            stotal += sm.wspine[w];
        }
        st_i<CO>(&L.mn0[t], m);
        st_i<CO || OPEN_CO>(&to.qcount[t], total);
        st_i<CO || OPEN_CO>(&to.scount[t], stotal);
        if (L1 && L.nlev > 1) atomicMin(&key_at(L, 1, t >> LOG_FAN), make_key(L.epoch, m));
        if (tk_ctr) *tk_slot = tk;
    }
    __syncthreads();  // sm is reused by the next tile
}

// Solve tile t (loads included) in the calling workgroup.  With L1 the tile
// also folds its min into level 1 (an atomic); without, build_level1 makes
// level 1 from the tile mins after the pass.
template <bool INCL, bool L1 = true, bool CO = false>
__device__ __forceinline__ void tile_step(const int* __restrict__ in, int* __restrict__ out,
                                          int n, int t, TileSmem& sm, const TileOut& to,
                                          const Levels& L, const int tid = (int)threadIdx.x) {
    int v[K];
    load_lane(in, n, t, v, tid);
    const int pl = (tid == 0 && t > 0) ? in[t * TILE - 1] : 0;
    solve_tile<INCL, false, L1, CO>(out, n, t, v, pl, sm, to, L, SegLane{0, 0u}, tid);
}

// Segmented tile t: mark the segment starts inside the tile in LDS, give
// every lane the start of the segment holding its first element, then solve
// with answers rebased to their segments.
template <bool INCL, bool CO = false>
__device__ __forceinline__ void tile_step_seg(const int* __restrict__ in, int* __restrict__ out,
                                              int n, int t, TileSmem& sm, const TileOut& to,
                                              const Levels& L, const SegIn& sg) {
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int wave = tid >> 6;
    const int tile_start = t * TILE;
    const int tile_end = tile_start + min(TILE, n - tile_start);  // no overflow near 2^31
    int v[K];
    load_lane(in, n, t, v);
    const int pl = (threadIdx.x == 0 && t > 0) ? in[t * TILE - 1] : 0;
    // offsets are clamped so that malformed ones cannot send a read astray
    const int last = max(sg.nseg - 1, 0);
    const int k0 = min(max(ld_i<CO>(&sg.tmap[t]), 0), last);
    const int k1 = (t + 1 < L.nt) ? min(max(ld_i<CO>(&sg.tmap[t + 1]), k0), last) : last;
    const int s0 = min(max(ld_i2<CO>(&sg.tbound[t]).x, 0), tile_start);
    if (tid < TILE / 32) sm.walls[tid] = 0u;
    __syncthreads();
    for (int k = k0 + 1 + tid; k <= k1; k += NT) {
        const long long w = sg.offs[k];
        if (w > tile_start && w < tile_end) {
            const int r = (int)w - tile_start;
            atomicOr(&sm.walls[r >> 5], 1u << (r & 31));
        }
    }
    __syncthreads();
    const unsigned wm = (sm.walls[tid >> 1] >> ((tid & 1) * 16)) & 0xFFFFu;
    const int base = tile_start + tid * K;
    // last segment start at or before each lane's end: wave max-scan (DPP),
    // then the earlier waves' maxima
    int lw = wm ? base + 31 - __builtin_clz(wm) : s0;
    lw = max(lw, dpp<DPP_ROW_SHR + 1>(s0, lw));
    lw = max(lw, dpp<DPP_ROW_SHR + 2>(s0, lw));
    lw = max(lw, dpp<DPP_ROW_SHR + 4>(s0, lw));
    lw = max(lw, dpp<DPP_ROW_SHR + 8>(s0, lw));
    lw = max(lw, dpp<DPP_ROW_BCAST15, 0xa>(s0, lw));
    lw = max(lw, dpp<DPP_ROW_BCAST31, 0xc>(s0, lw));
    const int lexcl = dpp<DPP_WAVE_SHR1>(s0, lw);
    if (lane == WAVE - 1) sm.wlast[wave] = lw;
    __syncthreads();
    int incoming = lexcl;
#pragma unroll
    for (int w = 0; w < NWAVE - 1; w++)
        if (w < wave) incoming = max(incoming, sm.wlast[w]);
    solve_tile<INCL, true, false, CO>(out, n, t, v, pl, sm, to, L, SegLane{incoming, wm});
}

// ---------------------------------------------------------------------------
// Levels >= 2 (only when level 1 exceeds UPPER_MIN entries): one workgroup
// builds them from the level below after the tile pass.
// ---------------------------------------------------------------------------
// Level 1 from the tile mins, one wave per 64-tile group (after a tile pass
// without level-1 atomics).  Atomics from the tile pass itself queue up:
// the 64 tiles of a group run at the same time and update one key.
__device__ __forceinline__ void build_level1(const Levels& L) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int g = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (g >= L.n[1]) return;  // uniform per wave
    const int idx = g * FAN + lane;
    const int m = wave_min(idx < L.nt ? L.mn0[idx] : INF);
    if (lane == 0) key_at(L, 1, g) = make_key(L.epoch, m);
}

// wv: the calling wave's index in its workgroup
template <bool CO = false>
__device__ __forceinline__ void build_upper_levels(const Levels& L, int wv) {
    const int lane = lane_id();
    const int nwv = blockDim.x >> 6;
    for (int l = 2; l < L.nlev; l++) {
        for (int g = wv; g < L.n[l]; g += nwv) {
            const int idx = g * FAN + lane;
            const int x = idx < L.n[l - 1] ? key_value(ld_u64<CO>(&key_at(L, l - 1, idx)), L.epoch)
                                          : INF;
            const int m = wave_min(x);
            if (lane == 0) st_u64<CO>(&key_at(L, l, g), make_key(L.epoch, m));
        }
        if (CO) wait_own_memory_ops();  // this level's stores complete
        __syncthreads();
    }
}

// find_tile_wave (below): the nearest tile before tb whose min qualifies
// for v (uniform), found by the whole wave.  Every level step is one
// coalesced 64-entry load and a ballot, so an answer arbitrarily far back
// costs a few rounds instead of a per-lane walk.
template <bool CO = false>
__device__ __forceinline__ int level_entry(const Levels& L, int lvl, int idx) {
    return lvl == 0 ? ld_i<CO>(&L.mn0[idx]) : key_value(ld_u64<CO>(&key_at(L, lvl, idx)), L.epoch);
}

template <bool INCL, bool CO = false>
__device__ int find_tile_wave(const Levels& L, int tb, int v) {
    const int lane = lane_id();
    int x = tb, lvl = 0, found = -1;  // x: exclusive bound at level lvl
    while (found < 0) {
        if (x <= 0) return -1;  // unreachable when guaranteed
        const int gs = ((x - 1) >> LOG_FAN) << LOG_FAN;
        const int idx = gs + lane;
        const bool ok = idx < x && qual<INCL>(level_entry<CO>(L, lvl, idx), v);
        const u64 b = __ballot(ok);
        if (b) {
            found = gs + 63 - __clzll(b);
        } else if (lvl == 1 && !L.upper) {
            x = gs;  // no level 2: walk level 1 backwards group by group
        } else {
            x = (x - 1) >> LOG_FAN;
            if (++lvl >= L.nlev) return -1;  // unreachable when guaranteed
        }
    }
    while (lvl > 0) {
        lvl--;
        const int gs = found << LOG_FAN;
        const int idx = gs + lane;
        const bool ok = idx < L.n[lvl] && qual<INCL>(level_entry<CO>(L, lvl, idx), v);
        const u64 b = __ballot(ok);
        found = b ? gs + 63 - __clzll(b) : -1;
        if (found < 0) return -1;  // unreachable: the parent's min qualifies
    }
    return found;
}

// ---------------------------------------------------------------------------
// Fix-up of tile t by one wave: answer the open elements whose value beats
// the exclusive prefix-min of all earlier tiles.  One batch of independent
// loads brings the prefix inputs, the open list, the 64 tile mins before t
// and the recorded right spines of tiles t-1 and t-2 (into the wave's LDS
// window); a query answered in t-1 or t-2 then needs no further global
// round, an older tile one more (its spine), and only a spine longer than
// RS falls back to chunk mins + elements.
// ---------------------------------------------------------------------------
// Control words of a look-back run (device memory): run tag and error word,
// each on its own 128-byte line.
struct LbCtl {
    unsigned epoch;
    unsigned pad0[31];
    int err;
    unsigned pad1[31];
};
constexpr int LB_LEVELS = MAX_LEVELS;
// Look-back group counters: eager runs alternate between sets 0 and 1 and
// clear the other for the next run; captured runs use set 2, cleared by the
// kernel recorded before them.  Eager runs take run tags from the host
// (1 .. LB_TAG_CAPTURED_MIN - 1), captured runs from a device counter
// (LB_TAG_CAPTURED_MIN and up), so the two never share a tag.
constexpr int LB_GC_SETS = 3;
constexpr int LB_GC_CAPTURED = 2;
constexpr unsigned LB_TAG_CAPTURED_MIN = 0x80000000u;

// per fix-up wave: tile-min window (64), its 16 four-tile group mins, two
// spines (2 x RS int2)
constexpr int FIX_LDS_INTS = WAVE + WAVE / 4 + 4 * RS;

// VAL: also return the answer's value in *val.
template <bool INCL, bool VAL = false>
__device__ __forceinline__ int spine_answer_lds(const int2* sp, int cnt, int v, int* val) {
    // lowest rank whose value qualifies (values decrease with rank): batches
    // of 16 ranks read at once and selected, highest rank first
    const int m = min(cnt, RS);
    const int4* p = reinterpret_cast<const int4*>(sp);
    int pos = -1, pv = 0;
    constexpr int B4 = VAL ? RS / 8 : RS / 4;  // dwordx4 reads per batch
#pragma unroll
    for (int h = 0; h < RS / 2 / B4 && pos < 0 && 2 * h * B4 < m; h++) {
        int4 w[B4];
#pragma unroll
        for (int k = 0; k < B4; k++) w[k] = p[h * B4 + k];
#pragma unroll
        for (int k = B4 - 1; k >= 0; k--) {
            const int r = 2 * (h * B4 + k);
            const bool q1 = r + 1 < m && qual<INCL>(w[k].w, v);
            pos = q1 ? w[k].z : pos;
            if (VAL) pv = q1 ? w[k].w : pv;
            const bool q0 = r < m && qual<INCL>(w[k].y, v);
            pos = q0 ? w[k].x : pos;
            if (VAL) pv = q0 ? w[k].y : pv;
        }
    }
    if (VAL) *val = pv;
    return pos;
}

// The nearest qualifying element of tile s for v (global index), read from
// the tile's right spine (from LDS when sp_lds is given), past a truncated
// spine from the chunk mins and the chunk.  VAL: also its value in *val.
template <bool INCL, bool CO = false, bool VAL = false>
__device__ __forceinline__ int answer_in_tile(const int* __restrict__ in, const TileOut& to,
                                              int s, int v, const int2* sp_lds, int cnt,
                                              int* val = nullptr) {
    int pos = -1, pv = 0;
    const int m = min(cnt, RS);
    if (sp_lds) {
        pos = spine_answer_lds<INCL, VAL>(sp_lds, cnt, v, &pv);
    } else {
        // batches of ranks read at once (two of 16 ranks, 8 dwordx4 each;
        // with VAL, for callers that keep more registers live, four of 8),
        // lowest rank wins
        constexpr int B4 = VAL ? RS / 8 : RS / 4;  // dwordx4 loads per batch
        const int4* p = reinterpret_cast<const int4*>(to.spine + (size_t)s * RS);
#pragma unroll
        for (int h = 0; h < RS / 2 / B4 && pos < 0; h++) {
            int4 w[B4];
#pragma unroll
            for (int k = 0; k < B4; k++) w[k] = ld_i4<CO>(&p[h * B4 + k]);
#pragma unroll
            for (int k = B4 - 1; k >= 0; k--) {
                const int r = 2 * (h * B4 + k);
                const bool q1 = r + 1 < m && qual<INCL>(w[k].w, v);
                pos = q1 ? w[k].z : pos;
                if (VAL) pv = q1 ? w[k].w : pv;
                const bool q0 = r < m && qual<INCL>(w[k].y, v);
                pos = q0 ? w[k].x : pos;
                if (VAL) pv = q0 ? w[k].y : pv;
            }
        }
    }
    if (pos >= 0) {
        if (VAL) *val = pv;
        return s * TILE + pos;
    }
    // deeper than the recorded spine: chunk mins, then the chunk
    int c = -1;
    if (CPT % 32 == 0) {
        for (int b = s * CPT + CPT - 32; b >= s * CPT && c < 0; b -= 32)
            c = last_in_32<INCL, CO>(to.cmin, b, v);
    } else {
        for (int b = s * CPT + CPT - 16; b >= s * CPT && c < 0; b -= 16) {
            int4 w[4];
#pragma unroll
            for (int k = 0; k < 4; k++) w[k] = ld_i4<CO>(reinterpret_cast<const int4*>(to.cmin + b + 4 * k));
#pragma unroll
            for (int k = 0; k < 4; k++) {
                c = qual<INCL>(w[k].x, v) ? b + 4 * k : c;
                c = qual<INCL>(w[k].y, v) ? b + 4 * k + 1 : c;
                c = qual<INCL>(w[k].z, v) ? b + 4 * k + 2 : c;
                c = qual<INCL>(w[k].w, v) ? b + 4 * k + 3 : c;
            }
        }
    }
    if (c < 0) return -1;  // unreachable: the tile's min qualifies
    int e = last_in_32<INCL>(in, c * CHUNK + 32, v);
    if (e < 0) e = last_in_32<INCL>(in, c * CHUNK, v);
    if (VAL && e >= 0) *val = in[e];
    return e;
}

// This lane's share of the min over the tiles [a, b) (whole wave, uniform a
// and b; the wave min of the shares is the range min): partial 64-groups at
// both ends on each level, full groups one level up.  The ranges of all
// levels follow from a and b alone, so every level's loads are issued
// before any is used (one round, not one per level).
template <bool CO = false>
__device__ __forceinline__ int range_min_lane(const Levels& L, int a, int b) {
    const int lane = lane_id();
    // raw entries: level 0 as loaded, keys (levels >= 1) converted after
    // every load is issued (0 reads as INF: a key of no run)
    int e0a = INF, e0b = INF;
    u64 ek[MAX_LEVELS][2];
    int mtop = INF;
    bool done = false;
#pragma unroll
    for (int lvl = 0; lvl < MAX_LEVELS; lvl++) {
        ek[lvl][0] = ek[lvl][1] = 0;
        if (done || a >= b) {
            done = true;
            continue;
        }
        const bool up_stored = lvl + 1 < L.nlev && (lvl + 1 < 2 || L.upper);
        if (!up_stored) {
            // the top stored level: the whole rest of the range
            for (int i = a + lane; i < b; i += WAVE) mtop = min(mtop, level_entry<CO>(L, lvl, i));
            done = true;
            continue;
        }
        const int ga = min(((a + FAN - 1) >> LOG_FAN) << LOG_FAN, b);
        const int gb = max((b >> LOG_FAN) << LOG_FAN, ga);
        if (lvl == 0) {
            if (a + lane < ga) e0a = ld_i<CO>(&L.mn0[a + lane]);
            if (gb + lane < b) e0b = ld_i<CO>(&L.mn0[gb + lane]);
        } else {
            if (a + lane < ga) ek[lvl][0] = ld_u64<CO>(&key_at(L, lvl, a + lane));
            if (gb + lane < b) ek[lvl][1] = ld_u64<CO>(&key_at(L, lvl, gb + lane));
        }
        a = ga >> LOG_FAN;
        b = gb >> LOG_FAN;
    }
    int m = min(mtop, min(e0a, e0b));
#pragma unroll
    for (int lvl = 1; lvl < MAX_LEVELS; lvl++)
        m = min(m, min(key_value(ek[lvl][0], L.epoch), key_value(ek[lvl][1], L.epoch)));
    return m;
}

// Where the keys and answers of a tile live, addressed by the tile's global
// (virtual) element indices.  FlatMap: one array for every tile.  PagedMap:
// batches of independent arrays read in place: tile t's keys start at
// pages[t].src and its answers at pages[t].dst (each array starts on a tile
// boundary of the virtual index space), so src(t)[t * TILE + i] is key i of
// the tile.  The returned base is formed by integer arithmetic and only
// dereferenced inside tile t (or, for src, the tile before it in the same
// array).
struct Page {
    const int* src;
    int* dst;
};

struct FlatMap {
    const int* in;
    int* out;
    __device__ __forceinline__ const int* src(int) const { return in; }
    __device__ __forceinline__ int* dst(int) const { return out; }
};

struct PagedMap {
    const Page* pages;
    __device__ __forceinline__ static uintptr_t virt(const void* p, int t) {
        return (uintptr_t)p - (uintptr_t)t * (TILE * sizeof(int));
    }
    __device__ __forceinline__ const int* src(int t) const {
        return reinterpret_cast<const int*>(virt(pages[t].src, t));
    }
    __device__ __forceinline__ int* dst(int t) const {
        return reinterpret_cast<int*>(virt(pages[t].dst, t));
    }
};

// Per-wave state of a tile's fix-up that the search for an open element's
// answer needs: the LDS window of the 64 tile mins before t (win), their
// four-tile group mins (gm), the spines of t-1 and t-2 (sp) and their
// counts, the min of tile t-1, and with two levels the tile group before the
// window (mw2, per lane) and the level-1 values (k1, per lane) from gs on.
struct FixCtx {
    int t;
    const int* win;
    const int* gm;
    const int2* sp;
    int sc1, sc2;
    int m_prev;
    bool two;
    int gs;
    int mw2;
    int k1;
};

// The answer (global index) of the open elements v of tile c.t held by the
// lanes with `active`, all of which have one before the tile; VAL: also its
// value in val.  Called by the whole wave.
// LEAN: deep searches walk the levels from memory (find_tile_wave) instead
// of using the preloaded mw2 / k1, which then need not stay live.
template <bool INCL, bool CO, bool VAL, bool LEAN, class Map>
__device__ __forceinline__ int resolve_open(const FixCtx& c, const Map& io, const Levels& L,
                                            const TileOut& to, bool active, int v, int& val) {
    const int lane = lane_id();
    const int t = c.t;
    // tile: the window first (tiles t-64 .. t-1) ...
    int ts = -1;
    if (active) {
        int s = WAVE - 1;
        if (!qual<INCL>(c.m_prev, v)) {
            // the last qualifying group of four window tiles, then the last
            // qualifying tile in it: two LDS rounds, whatever the distance
            const int4* g4 = reinterpret_cast<const int4*>(c.gm);
            const int4 ga = g4[0], gb = g4[1], gc = g4[2], gd = g4[3];
            int g = -1;
            sel4<INCL, 0>(g, ga.x, ga.y, ga.z, ga.w, v);
            sel4<INCL, 4>(g, gb.x, gb.y, gb.z, gb.w, v);
            sel4<INCL, 8>(g, gc.x, gc.y, gc.z, gc.w, v);
            sel4<INCL, 12>(g, gd.x, gd.y, gd.z, gd.w, v);
            s = -1;
            if (g >= 0) {
                const int4 w = reinterpret_cast<const int4*>(c.win)[g];
                int k = -1;
                sel4<INCL, 0>(k, w.x, w.y, w.z, w.w, v);
                s = 4 * g + k;  // k >= 0: the group's min qualifies
            }
        }
        // window entries of tiles < 0 hold INF, which never qualifies
        // except under INCL with v == INF: exclude them explicitly
        if (s >= 0 && t - WAVE + s >= 0) ts = t - WAVE + s;
    }
    // ... then, for the rare lanes whose tile is older, the whole wave walks
    // the hierarchy for one of them at a time
    u64 deep = __ballot(active && ts < 0);
    while (deep) {
        const int l = __ffsll(deep) - 1;
        deep &= deep - 1;
        const int vv = __shfl(v, l, WAVE);
        int r = -1;
        if (LEAN || !c.two) {
            r = find_tile_wave<INCL, CO>(L, t - WAVE, vv);
        } else if (t > WAVE) {
            const u64 b0 = __ballot(c.gs + lane < t - WAVE && qual<INCL>(c.mw2, vv));
            if (b0) {
                r = c.gs + 63 - __clzll(b0);
            } else {
                const u64 b1 = __ballot(lane < (c.gs >> LOG_FAN) && qual<INCL>(c.k1, vv));
                if (b1) {
                    const int g = 63 - __clzll(b1);
                    const int idx = (g << LOG_FAN) + lane;
                    const u64 b2 = __ballot(idx < L.n[0] && qual<INCL>(ld_i<CO>(&L.mn0[idx]), vv));
                    r = b2 ? (g << LOG_FAN) + 63 - __clzll(b2) : -1;
                }
            }
        }
        if (lane == l) ts = r;
    }
    int r = -1;
    if (active && ts >= 0) {  // (ts < 0 is unreachable: P qualifies, so some tile does)
        // spines of t-1 and t-2 are in LDS, older ones are read
        const int2* spl = ts == t - 1 ? c.sp : ts == t - 2 ? c.sp + RS : nullptr;
        const int cnt = ts == t - 1 ? c.sc1 : ts == t - 2 ? c.sc2 : ld_i<CO>(&to.scount[ts]);
        r = answer_in_tile<INCL, CO, VAL>(io.src(ts), to, ts, v, spl, cnt, &val);
    }
    return r;
}

// Fix-up of a tile with more open elements than its recorded list holds.
// P: the prefix-min before the tile (segmented: before it within the
// segment running into it, which starts at s0); lim: tile-relative end of
// that segment.
template <bool INCL, bool SEG, bool CO, class Map>
__device__ __forceinline__ void fix_rescan(const FixCtx& c, const Map& io, int n,
                                           const Levels& L, const TileOut& to, int P, int s0,
                                           int lim) {
    // Rescan the tile, 4 keys per lane and 256 per round.  Open
    // elements are the tile's prefix minima: an element is open iff no
    // earlier element of the tile qualifies for it.  Their answers move
    // monotonically: the answer c of an open element u also answers every
    // later open element v with qual(key c, v), since every element between
    // c and the tile failed u, which is at least as large as v, and no
    // element of the tile before v qualifies for v.  So an element first
    // tries the answer of the previous open element of its lane (for the
    // lane's first one, of the last open element of the previous round, or
    // in the first round of the tile's first open element), and only a miss
    // is searched for.
    const int lane = lane_id();
    const int t = c.t;
    const int tile_start = t * TILE;
    const int* src = io.src(t) + tile_start;
    int* dst = io.dst(t) + tile_start;
    const int span = min(lim, n - tile_start);
    int rm = INF;                       // min of the rounds before
    int c_pos = -1, c_val = 0;          // carried answer (global index) and its value
    bool c_ok = false;
    for (int o = 0; o < span; o += 4 * WAVE) {
        const int q = o + 4 * lane;  // tile-relative index of this lane's first key
        int x[4];
        if (q + 4 <= span) {
            const int4 w = *reinterpret_cast<const int4*>(src + q);
            x[0] = w.x; x[1] = w.y; x[2] = w.z; x[3] = w.w;
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) x[k] = q + k < span ? src[q + k] : INF;
        }
        // prefix-min before each key: the rounds before, the earlier lanes
        // (DPP scan: rows, then across rows), the lane's earlier keys.
        // Every key of the range is real, so for any key but the tile's
        // first the min is over real keys.
        const int l1 = min(x[0], x[1]), l2 = min(l1, x[2]), l3 = min(l2, x[3]);
        int inc = min(l3, dpp<DPP_ROW_SHR + 1>(INF, l3));
        inc = min(inc, dpp<DPP_ROW_SHR + 2>(INF, inc));
        inc = min(inc, dpp<DPP_ROW_SHR + 4>(INF, inc));
        inc = min(inc, dpp<DPP_ROW_SHR + 8>(INF, inc));
        inc = min(inc, dpp<DPP_ROW_BCAST15, 0xa>(INF, inc));
        inc = min(inc, dpp<DPP_ROW_BCAST31, 0xc>(INF, inc));
        const int rm_next = min(rm, lane63(inc));
        const int ex = min(dpp<DPP_WAVE_SHR1>(INF, inc), rm);
        unsigned need = 0;  // open keys with an answer before the tile
        {
            const int pm[4] = {ex, min(ex, x[0]), min(ex, l1), min(ex, l2)};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const bool open = q + k == 0 || !qual<INCL>(pm[k], x[k]);
                if (q + k < span && open && qual<INCL>(P, x[k])) need |= 1u << k;
            }
        }
        // Each lane answers its keys in order from a candidate while it
        // qualifies: first the carried answer (usually for all four keys:
        // one 16-byte store), after a miss the miss's own answer.  A miss is
        // searched for; before any answer is carried (the tile's first
        // round), only the round's first such key is, and its answer then
        // serves every lane.  One search call site keeps the code small.
        unsigned todo = need;
        // the lane's leading keys that candidate (pos, val) qualifies for
        auto take = [&](int pos, int val) {
            const int a = SEG ? (pos >= s0 ? pos - s0 : -1) : pos;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (todo && __ffs(todo) - 1 == k && qual<INCL>(val, x[k])) {
                    dst[q + k] = a;
                    todo &= todo - 1;
                }
            }
        };
        if (c_ok) {
            if (need == 0xFu && qual<INCL>(c_val, x[3])) {
                // x[3] is the lane's smallest open key: all four qualify
                const int a = SEG ? (c_pos >= s0 ? c_pos - s0 : -1) : c_pos;
                *reinterpret_cast<int4*>(dst + q) = make_int4(a, a, a, a);
                todo = 0;
            } else {
                take(c_pos, c_val);
            }
        }
        int cp = c_pos, cv = c_val;
        while (true) {
            const u64 b = __ballot(todo != 0);
            if (!b) break;
            const int l0 = __ffsll(b) - 1;
            const bool act = todo != 0 && (c_ok || lane == l0);
            const int k = act ? __ffs(todo) - 1 : 0;
            const int v = k == 0 ? x[0] : k == 1 ? x[1] : k == 2 ? x[2] : x[3];
            int val = 0;
            const int r = resolve_open<INCL, CO, true, true>(c, io, L, to, act, v, val);
            if (act) {
                dst[q + k] = SEG ? (r >= s0 ? r - s0 : -1) : r;
                todo &= todo - 1;
                cp = r;
                cv = val;
                if (r >= 0) take(r, val);
            }
            if (!c_ok) {
                c_pos = readlane(r, l0);
                c_val = readlane(val, l0);
                c_ok = true;
                if (c_pos < 0) break;  // unreachable: P qualifies for the key
                if (lane != l0) {
                    cp = c_pos;
                    cv = c_val;
                    take(c_pos, c_val);
                }
            }
        }
        // carry the answer of the round's last open key
        const u64 b = __ballot(need != 0);
        if (b) {
            const int l = 63 - __clzll(b);
            c_pos = readlane(cp, l);
            c_val = readlane(cv, l);
            c_ok = c_pos >= 0;
        }
        rm = rm_next;
        if (!qual<INCL>(P, rm)) break;  // later open elements are <= rm
    }
}

// Min of in[s0 .. t * TILE) for s0 < t * TILE (whole wave, uniform result):
// the rest of s0's chunk, the rest of its tile's chunk mins, then whole
// tiles through the hierarchy.  Tile s0 / TILE precedes t, so it is complete.
static_assert(CPT <= 2 * WAVE, "the rest of a tile's chunk mins takes two loads per lane");
template <bool CO = false>
__device__ __forceinline__ int seg_prefix_min(const int* __restrict__ in, const TileOut& to,
                                              const Levels& L, int s0, int t) {
    const int lane = lane_id();
    const int u = s0 / TILE;
    const bool part = s0 % TILE != 0;
    // all loads first, one wave min at the end
    const int m3 = range_min_lane<CO>(L, part ? u + 1 : u, t);
    int m0 = INF, m1 = INF, m2 = INF;
    if (part) {
        const int c0 = s0 / CHUNK;
        const int cend = (u + 1) * CPT;
        if (s0 + lane < (c0 + 1) * CHUNK) m0 = in[s0 + lane];
        if (c0 + 1 + lane < cend) m1 = ld_i<CO>(&to.cmin[c0 + 1 + lane]);
        if (c0 + 1 + WAVE + lane < cend) m2 = ld_i<CO>(&to.cmin[c0 + 1 + WAVE + lane]);
    }
    return wave_min(min(min(m0, m1), min(m2, m3)));
}

// CO: the fix-up runs in the persistent launch that wrote the summaries
// (tile mins, keys, spines, chunk mins are read coherently; the tile's own
// open list and outputs were written by this workgroup).  io: where each
// tile's keys and answers live (FlatMap or PagedMap).
template <bool INCL, bool SEG = false, bool CO = false, class Map = FlatMap>
__device__ __forceinline__ void fix_tile(const Map& io, int n, int t, const Levels& L,
                                         const TileOut& to, int* lds,
                                         const SegIn sg = SegIn{nullptr, nullptr, 0, nullptr}) {
    if (t == 0) return;
    const int lane = lane_id();
    const int tile_start = t * TILE;
    const int2 tb = SEG ? ld_i2<CO>(&sg.tbound[t]) : make_int2(0, 0);
    const int tw = t - WAVE + lane;  // tile whose min this lane holds
    const int mw = tw >= 0 ? ld_i<CO>(&L.mn0[tw]) : INF;
    // spines of t-1 (lanes 0..RS-1) and t-2 (lanes RS..2RS-1)
    const int st = (lane < RS) ? t - 1 : t - 2;
    int2 sw = make_int2(-1, INF);
    if (lane < 2 * RS && st >= 0) sw = ld_i2<CO>(&to.spine[(size_t)st * RS + (lane & (RS - 1))]);
    const int sc1 = ld_i<CO>(&to.scount[t - 1]);
    const int sc2 = (t >= 2) ? ld_i<CO>(&to.scount[t - 2]) : 0;
    const int cnt = ld_i<CO>(&to.qcount[t]);
    // the tile's first key; a tile of a paged batch that no job covers has
    // no keys (and no page: nothing of it is read)
    const int v1 = (!SEG || tb.y > tile_start) ? io.src(t)[tile_start] : INF;
    // the list slots are read with the rest of the first batch, before the
    // count is known; slots past the count are ignored below
    int2 e = to.qlist[(size_t)t * QS + (lane & (QS - 1))];
    if (lane >= min(cnt, QS)) e = make_int2(0, INF);
    // two levels (at most 64 groups of 64 tiles): every level-1 value and
    // the tile group just before the window come with the first batch, so
    // a search older than the window costs one more round, not three
    const bool two = L.nlev == 2;
    const int gs = ((t - WAVE - 1) >> LOG_FAN) << LOG_FAN;  // group of tile t - 65
    u64 k1raw = 0;  // (a key of no run: reads as INF)
    int mw2 = INF;
    if (two) {
        // only groups before the tile's own are used (prefix, deep search)
        if (lane < (t >> LOG_FAN)) k1raw = ld_u64<CO>(&key_at(L, 1, lane));
        if (t > WAVE && gs + lane < t - WAVE) mw2 = ld_i<CO>(&L.mn0[gs + lane]);
    }
    const int k1 = key_value(k1raw, L.epoch);
    // segmented: only the open elements of the segment that runs into the
    // tile from before (tile-relative positions < lim) can have an answer
    int s0 = 0, lim = TILE;
    if (SEG) {
        s0 = min(max(tb.x, 0), tile_start);
        lim = (int)(min(max((long long)tb.y, (long long)tile_start + 1),
                        (long long)tile_start + TILE) - tile_start);
        if (s0 >= tile_start) return;
    }
    // no open element (every tile of an ascending run): nothing to fix, and
    // a segmented tile skips its prefix-min's round of loads
    if (cnt == 0) return;
    // exclusive prefix-min of tile t: level 0 from the window, upper levels
    // from the keys
    int p = (tw >= ((t >> LOG_FAN) << LOG_FAN)) ? mw : INF;
    if (two) {
        if (lane < (t >> LOG_FAN)) p = min(p, k1);
    } else if (L.nlev > 1) {
        const int a1 = t >> LOG_FAN;
        const int b1 = (a1 >> LOG_FAN) << LOG_FAN;
        if (b1 + lane < a1) p = min(p, key_value(ld_u64<CO>(&key_at(L, 1, b1 + lane)), L.epoch));
        if (L.upper) {
#pragma unroll
            for (int l = 2; l < MAX_LEVELS; l++) {
                if (l < L.nlev) {
                    const int a = t >> (LOG_FAN * l);
                    const int idx = ((a >> LOG_FAN) << LOG_FAN) + lane;
                    if (idx < a) p = min(p, key_value(ld_u64<CO>(&key_at(L, l, idx)), L.epoch));
                }
            }
        } else {
            for (int i = lane; i < b1; i += WAVE) p = min(p, key_value(ld_u64<CO>(&key_at(L, 1, i)), L.epoch));
        }
    }
    const int P = SEG ? seg_prefix_min<CO>(io.src(s0 / TILE), to, L, s0, t) : wave_min(p);
    if (!qual<INCL>(P, v1)) return;  // every open element answers -1
    int* win = lds;
    int* gm = lds + WAVE;  // min of window tiles 4g .. 4g+3
    int2* sp = reinterpret_cast<int2*>(lds + WAVE + WAVE / 4);
    win[lane] = mw;
    {
        int g4 = min(mw, dpp<DPP_ROW_SHR + 1>(INF, mw));
        g4 = min(g4, dpp<DPP_ROW_SHR + 2>(INF, g4));
        if ((lane & 3) == 3) gm[lane >> 2] = g4;
    }
    if (lane < 2 * RS) sp[lane] = sw;
    wave_lds_sync();
    const FixCtx fc{t, win, gm, sp, sc1, sc2, readlane(mw, WAVE - 1), two, gs, mw2, k1};
    if (cnt <= QS) {
        const bool active = lane < cnt && qual<INCL>(P, e.y) && (!SEG || e.x < lim);
        int unused;
        int r = resolve_open<INCL, CO, false, false>(fc, io, L, to, active, e.y, unused);
        if (active) {
            if (SEG) r = r >= s0 ? r - s0 : -1;
            io.dst(t)[tile_start + e.x] = r;
        }
        return;
    }
    fix_rescan<INCL, SEG, CO>(fc, io, n, L, to, P, s0, lim);
}

// ---------------------------------------------------------------------------
// Grid barrier for persistent launches.  Arrivals are spread
// over 64 sub-counters (one 128-byte line each); the last arriver of a
// sub-group arrives at the root; the last root arriver bumps the generation
// word (its own line).  The counters only grow inside a launch (barrier j
// of a launch is complete for a sub-group at (j + 1) x its size), and they
// are zeroed (grid_barrier_reset) before every launch, by the previous eager
// launch or by a one-workgroup kernel (see BAR_SETS), so a launch cut short
// by a timeout leaves nothing behind for the next.
// Only the workgroups that grid_join found running take part, so every
// participant reaches every barrier.  The spin is still bounded: any stall
// raises `err` and lets every workgroup leave instead of hanging the
// device; `err` stays set until the host reads it (apsep_scratch_check).
// ---------------------------------------------------------------------------
constexpr int BAR_SUB = 64;
constexpr int BAR_LINE = 32;  // uints per 128-byte line

// Workgroups of a persistent launch join through SPT_SLOTS counters, one
// per residue of blockIdx.x, each on its own line: a device atomic on one
// address serialises (~13 ns each on MI355X), so 768 joins on one counter
// kept the last workgroup waiting ~10 us.  Slot c's participants work on
// the tiles of residue c (see k_spt).
constexpr int SPT_SLOTS = 8;
constexpr int SPT_SLOT_MAX = 255;  // participants per slot (a byte of the published word)

struct GridBarrier {
    unsigned gen;
    int err;
    unsigned pad0[BAR_LINE - 2];
    unsigned root;
    unsigned pad1[BAR_LINE - 1];
    u64 pub;          // participants per slot, one byte each (0: not yet published)
    unsigned pad2[BAR_LINE - 2];
    unsigned orphans; // tickets of the tiles of residues no participant's slot has
    unsigned pad3[BAR_LINE - 1];
    unsigned disc[SPT_SLOTS * BAR_LINE];   // per slot: joins | DISC_CLOSED
    unsigned tiles[SPT_SLOTS * BAR_LINE];  // per slot: tickets of dynamically taken tiles
    unsigned sub[BAR_SUB * BAR_LINE];
};

constexpr unsigned BARRIER_SPIN_LIMIT = 1u << 22;  // >= 1 s at >= 250 ns per poll
constexpr unsigned DISC_CLOSED = 1u << 31;
// how long the closing workgroup waits for the others to start: 50 us of
// the 100 MHz real-time clock (all of them start within a few us on an
// idle GPU)
constexpr unsigned long long DISC_WAIT_TICKS = 5000;

// Clears a barrier's counters (the first BAR_SUB threads of a workgroup;
// device-coherent stores, like the atomics that use them).
__device__ __forceinline__ void grid_barrier_reset(GridBarrier* b) {
    const int i = (int)threadIdx.x;
    constexpr int sc = __HIP_MEMORY_SCOPE_AGENT;
    if (i < BAR_SUB) __hip_atomic_store(&b->sub[i * BAR_LINE], 0u, __ATOMIC_RELAXED, sc);
    if (i < SPT_SLOTS) {
        __hip_atomic_store(&b->disc[i * BAR_LINE], 0u, __ATOMIC_RELAXED, sc);
        __hip_atomic_store(&b->tiles[i * BAR_LINE], 0u, __ATOMIC_RELAXED, sc);
    }
    if (i == 0) {
        __hip_atomic_store(&b->root, 0u, __ATOMIC_RELAXED, sc);
        __hip_atomic_store(&b->pub, 0ull, __ATOMIC_RELAXED, sc);
        __hip_atomic_store(&b->orphans, 0u, __ATOMIC_RELAXED, sc);
    }
}

// A scratch holds BAR_SETS barriers.  Eager persistent launches alternate
// between sets 0 and 1 (the host keeps the parity): each launch's workgroup
// 0 clears the other set for the next one (grid_barrier_clear_next), so no
// launch of its own is needed.  That is safe because no workgroup of a
// launch touches the other set, every workgroup of a launch runs (late ones
// join, see the close and leave), and the launch that used the other set
// last ended before this one started.  Captured launches use set
// BAR_CAPTURED, cleared by a one-workgroup kernel recorded before them
// (a replay has no host to flip a parity).
constexpr int BAR_SETS = 3;
constexpr int BAR_CAPTURED = 2;

__device__ __forceinline__ void grid_barrier_clear_next(GridBarrier* next) {
    if (next && blockIdx.x == 0) grid_barrier_reset(next);
}

// Which workgroups of a normally launched persistent grid run at the same
// time.  Each workgroup takes a ticket from its slot's counter (grid_join);
// one that arrives after its slot was closed (it could not start while the
// others held the GPU) leaves at once.  A workgroup that took its ticket
// before the close is running, so it is a participant and will reach every
// barrier; slot c's tickets 0 .. npc-1 go to its participants.  The slots
// are closed by the workgroup holding slot 0's ticket 0 (grid_count),
// which first waits, bounded in time, until every workgroup of the grid has
// taken a ticket, then publishes the per-slot counts in one word.
// Participants may work before they need the counts: the first tile of a
// persistent solve follows from the ticket alone.  Uniform per workgroup.
__device__ __forceinline__ int grid_join(GridBarrier* b) {
    __shared__ int s_id;
    if (threadIdx.x == 0) {
        const unsigned c = blockIdx.x % SPT_SLOTS;
        const unsigned old = __hip_atomic_fetch_add(&b->disc[c * BAR_LINE], 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        s_id = (old & DISC_CLOSED) ? -1 : (int)old;
    }
    __syncthreads();
    return s_id;
}

// The published per-slot participant counts (0 after a timeout: the caller
// leaves).  `closer`: slot 0's ticket 0.
__device__ __forceinline__ u64 grid_count(GridBarrier* b, bool closer) {
    __shared__ u64 s_pub;
    if (threadIdx.x == 0) {
        u64 pub = 0ull;
        if (closer) {
            const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
            for (;;) {
                unsigned joined = 0u;
                for (int c = 0; c < SPT_SLOTS; c++)
                    joined += __hip_atomic_load(&b->disc[c * BAR_LINE], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT) & ~DISC_CLOSED;
                if (joined >= gridDim.x || __builtin_amdgcn_s_memrealtime() - t0 >= DISC_WAIT_TICKS)
                    break;
                __builtin_amdgcn_s_sleep(1);
            }
            for (int c = 0; c < SPT_SLOTS; c++) {
                const unsigned cnt = __hip_atomic_fetch_or(&b->disc[c * BAR_LINE], DISC_CLOSED,
                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
                                     ~DISC_CLOSED;
                pub |= (u64)min(cnt, (unsigned)SPT_SLOT_MAX) << (8 * c);
            }
            __hip_atomic_store(&b->pub, pub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned spins = 0;
            while ((pub = __hip_atomic_load(&b->pub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ==
                   0ull) {
                if (++spins > BARRIER_SPIN_LIMIT) {
                    __hip_atomic_store(&b->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        s_pub = pub;
    }
    __syncthreads();
    return s_pub;
}

// Generation a launch starts from.  The counter persists across launches;
// it cannot advance before every workgroup has arrived at the first
// barrier, and each workgroup reads it (at any point) before arriving, so
// all workgroups of a launch read the same value.
__device__ __forceinline__ unsigned grid_barrier_start(GridBarrier* b) {
    return __hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Barrier number `j` (0, 1, ...) of a launch whose cross-workgroup data is
// written and read with device-coherent (CO) accesses, so no fence is
// needed: every lane waits for its own memory operations to complete (the
// "memory" clobber also keeps the compiler from moving any of them past the
// wait), then the workgroup arrives; after the generation changes, the CO
// loads read the other workgroups' words from the coherent level.  `leader`
// is true for exactly one lane of the workgroup.  Returns false (uniformly
// per workgroup) if the barrier timed out.
// `me`: the workgroup's participant index, `nblocks`: the participant count.
__device__ __forceinline__ bool grid_barrier_co(GridBarrier* b, unsigned nblocks, unsigned me,
                                                unsigned& gen, unsigned& j, bool leader) {
    __shared__ int s_ok;
    wait_own_memory_ops();
    __syncthreads();
    if (leader) {
        int ok = 1;
        const unsigned g = me % BAR_SUB;
        const unsigned nsub = nblocks < (unsigned)BAR_SUB ? nblocks : (unsigned)BAR_SUB;
        const unsigned gsz = nblocks / BAR_SUB + (g < nblocks % BAR_SUB ? 1u : 0u);
        bool release = false;
        if (atomicAdd(&b->sub[g * BAR_LINE], 1u) == (j + 1) * gsz - 1 &&
            atomicAdd(&b->root, 1u) == (j + 1) * nsub - 1)
            release = true;
        if (release) {
            __hip_atomic_store(&b->gen, gen + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned spins = 0;
            while (__hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (++spins > BARRIER_SPIN_LIMIT) {
                    __hip_atomic_store(&b->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        s_ok = ok;
    }
    __syncthreads();
    gen++;
    j++;
    return s_ok != 0;
}

}  // namespace apsep
