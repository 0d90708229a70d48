// apsep_device.hpp -- gfx950 device building blocks shared by every APSEP
// solver variant (WSTL multi-launch, SPT cooperative, look-back, segmented).
//
// Problem (src/apsep.cuh:1-4): out[i] = largest j < i with in[j] < in[i]
// (strict) or in[j] <= in[i] (INCL), else -1.
//
// Work decomposition (wave64 throughout):
//   tile   = one workgroup of NT = 256 lanes (4 waves) x K = 8 consecutive
//            elements per lane = 2048 elements, loaded with two dwordx4
//            loads per lane.
//   chunk  = 64 consecutive elements (8 lanes); its min is published for the
//            inter-tile search.
//   levels = 64-ary min hierarchy over tile mins (level 0 = tile mins); each
//            level also stores the exclusive prefix-min of every entry within
//            its 64-entry group, so the exclusive prefix-min of tile t is the
//            min over levels of pre[l][t >> 6l].
//
// Inside a tile (solve_tile):
//   1. in-lane answers by an unrolled O(K^2) compare/select over registers;
//   2. Hillis-Steele min scan over the 64 lane-mins of each wave; every
//      step's window minimum is kept in LDS (win[s][lane] = min over the
//      2^s lanes ending at lane), which turns the cross-lane search into a
//      6-step binary lifting per query;
//   3. elements still unanswered whose value beats the exclusive prefix-min
//      of earlier lanes are answered inside the tile (lifting, then the
//      rightmost qualifying element of the found lane, read from LDS);
//   4. the rest are the tile's "open" elements: the tile's prefix minima.
//      They are written as -1 and recorded (first QS positions + count) for
//      the fix-up phase, which only revisits the ones whose value beats the
//      exclusive prefix-min of all earlier tiles.
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

namespace apsep {

constexpr int WAVE = 64;
constexpr int NT = 256;            // lanes per tile workgroup
constexpr int K = 8;               // consecutive elements per lane
constexpr int TILE = NT * K;       // 2048 elements per tile
constexpr int NWAVE = NT / WAVE;   // 4 waves per tile
constexpr int CHUNK = 64;          // elements per published chunk min
constexpr int CPT = TILE / CHUNK;  // 32 chunk mins per tile
constexpr int QS = 32;             // recorded open positions per tile
constexpr int FAN = 64;            // hierarchy fan-out
constexpr int LOG_FAN = 6;
constexpr int MAX_LEVELS = 5;      // 64^5 tiles >> 2^31 / TILE
constexpr int INF = INT_MAX;

static_assert(K == 8, "tile loader assumes two dwordx4 per lane");
static_assert(CHUNK % K == 0 && (CHUNK / K) == 8, "chunk = 8 lanes");

// "x qualifies as the answer for query value v"
template <bool INCL>
__device__ __forceinline__ bool qual(int x, int v) {
    return INCL ? (x <= v) : (x < v);
}

// Hierarchy of tile mins (device pointers + sizes for one solve).
struct Levels {
    int nlev;                 // >= 1; the top level has <= 64 entries
    int n[MAX_LEVELS];        // entries per level
    int* mn[MAX_LEVELS];      // mins (mn[0] = tile mins)
    int* pre[MAX_LEVELS];     // exclusive prefix-min within the 64-group
};

// Per-tile summaries written by the tile pass.
struct TileOut {
    int* tmin;          // [nt] tile mins (== Levels::mn[0])
    int* cmin;          // [nt * CPT] chunk mins
    int* qcount;        // [nt] number of open elements
    uint16_t* qlist;    // [nt * QS] tile-relative positions of the first QS
};

struct TileSmem {
    int val[TILE];              // the tile's values
    int win[6][NT];             // window mins of lane mins, sizes 1..32
    int wmin[NWAVE];            // wave mins
    int wopen[NWAVE];           // open-element counts per wave
};

__device__ __forceinline__ int wave_shfl_up(int x, int d) {
    return __shfl_up(x, d, WAVE);
}

// Inclusive min scan over the wave (Hillis-Steele), lanes below the shift
// keep their value, so every step is a window clipped at lane 0.
__device__ __forceinline__ int wave_incl_min(int x) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) x = min(x, wave_shfl_up(x, d));
    return x;
}

__device__ __forceinline__ int wave_incl_add(int x) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        int y = wave_shfl_up(x, d);
        if (lane >= d) x += y;
    }
    return x;
}

// Rightmost index in [lo, hi] with qual(arr[idx], v), or -1.  Reads aligned
// dwordx4 groups; arr must be readable (padded) up to (hi | 3).
template <bool INCL>
__device__ __forceinline__ int scan_back(const int* __restrict__ arr, int lo,
                                         int hi, int v) {
    int i = hi;
    while (i >= lo) {
        const int a4 = i & ~3;
        const int4 w = *reinterpret_cast<const int4*>(arr + a4);
        if (a4 + 3 <= i && a4 + 3 >= lo && qual<INCL>(w.w, v)) return a4 + 3;
        if (a4 + 2 <= i && a4 + 2 >= lo && qual<INCL>(w.z, v)) return a4 + 2;
        if (a4 + 1 <= i && a4 + 1 >= lo && qual<INCL>(w.y, v)) return a4 + 1;
        if (a4 >= lo && qual<INCL>(w.x, v)) return a4;
        i = a4 - 1;
    }
    return -1;
}

// Exclusive prefix-min of tile t over tiles [0, t) (t >= 1).
__device__ __forceinline__ int tile_prefix(const Levels& L, int t) {
    int p = INF;
#pragma unroll
    for (int l = 0; l < MAX_LEVELS; l++)
        if (l < L.nlev) p = min(p, L.pre[l][t >> (LOG_FAN * l)]);
    return p;
}

// Global index of the nearest j < tile_start(t) with qual(in[j], v).  The
// caller guarantees one exists (qual(tile_prefix(t), v)).  Ascends the
// 64-ary hierarchy until an earlier entry qualifies, descends to the tile,
// then to the chunk, then to the element.
template <bool INCL>
__device__ int search_before_tile(const int* __restrict__ in, const Levels& L,
                                  const int* __restrict__ cmin, int t, int v) {
    int lvl = 0, x = t, found = -1;
    while (true) {
        const int gstart = ((x - 1) >> LOG_FAN) << LOG_FAN;
        found = scan_back<INCL>(L.mn[lvl], gstart, x - 1, v);
        if (found >= 0) break;
        x = (x - 1) >> LOG_FAN;
        lvl++;
        if (lvl >= L.nlev || x <= 0) return -1;  // unreachable when guaranteed
    }
    while (lvl > 0) {
        lvl--;
        const int lo = found << LOG_FAN;
        const int hi = min(lo + FAN - 1, L.n[lvl] - 1);
        found = scan_back<INCL>(L.mn[lvl], lo, hi, v);
    }
    const int s = found;
    const int c = scan_back<INCL>(cmin, s * CPT, s * CPT + CPT - 1, v);
    const int e0 = c * CHUNK;  // == s*TILE + (c - s*CPT)*CHUNK
    return scan_back<INCL>(in, e0, e0 + CHUNK - 1, v);
}

// ---------------------------------------------------------------------------
// Tile pass: solve tile t inside the workgroup, write every output of the
// tile (open elements as -1), publish tile min, chunk mins and open list.
// The caller must __syncthreads() before reusing `sm` for another tile.
// ---------------------------------------------------------------------------
template <bool INCL>
__device__ __forceinline__ void solve_tile(const int* __restrict__ in,
                                           int* __restrict__ out, int n, int t,
                                           TileSmem& sm, const TileOut& to) {
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int wave = tid >> 6;
    const int tile_start = t * TILE;
    const int base = tile_start + tid * K;
    const bool full = tile_start + TILE <= n;

    int v[K];
    if (full) {
        const int4 a = *reinterpret_cast<const int4*>(in + base);
        const int4 b = *reinterpret_cast<const int4*>(in + base + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
        for (int i = 0; i < K; i++) v[i] = (base + i < n) ? in[base + i] : INF;
    }

    // 1. in-lane answers (lane-relative index or -1)
    int loc[K];
#pragma unroll
    for (int i = 0; i < K; i++) {
        loc[i] = -1;
#pragma unroll
        for (int j = 0; j < i; j++)
            if (qual<INCL>(v[j], v[i])) loc[i] = j;
    }
    int lmin = v[0];
#pragma unroll
    for (int i = 1; i < K; i++) lmin = min(lmin, v[i]);

    *reinterpret_cast<int4*>(&sm.val[tid * K]) = make_int4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<int4*>(&sm.val[tid * K + 4]) = make_int4(v[4], v[5], v[6], v[7]);

    // 2. window mins of lane mins (Hillis-Steele within the wave)
    int x = lmin, chunk_min = 0;
    sm.win[0][tid] = x;
#pragma unroll
    for (int s = 1; s <= 6; s++) {
        x = min(x, wave_shfl_up(x, 1 << (s - 1)));
        if (s < 6) sm.win[s][tid] = x;
        if (s == 3) chunk_min = x;  // window of 8 lanes = 64 elements
    }
    const int wexcl = wave_shfl_up(x, 1);  // valid for lane > 0
    if (lane == WAVE - 1) sm.wmin[wave] = x;
    if ((lane & 7) == 7) to.cmin[(size_t)t * CPT + (tid >> 3)] = chunk_min;
    __syncthreads();

    int cross = INF;
#pragma unroll
    for (int w = 0; w < NWAVE - 1; w++)
        if (w < wave) cross = min(cross, sm.wmin[w]);

    // classify lane-unanswered elements
    unsigned qmask = 0, omask = 0;
    int res[K];
#pragma unroll
    for (int i = 0; i < K; i++) {
        res[i] = (loc[i] >= 0) ? base + loc[i] : -1;
        const bool valid = full || (base + i < n);
        if (valid && loc[i] < 0) {
            const bool in_tile = (lane > 0 && qual<INCL>(wexcl, v[i])) ||
                                 (wave > 0 && qual<INCL>(cross, v[i]));
            if (in_tile) qmask |= 1u << i;
            else omask |= 1u << i;
        }
    }

    // write the tile's outputs (answers found so far; -1 elsewhere)
    if (full) {
        *reinterpret_cast<int4*>(out + base) = make_int4(res[0], res[1], res[2], res[3]);
        *reinterpret_cast<int4*>(out + base + 4) = make_int4(res[4], res[5], res[6], res[7]);
    } else {
#pragma unroll
        for (int i = 0; i < K; i++)
            if (base + i < n) out[base + i] = res[i];
    }

    // 3. cross-lane answers inside the tile (same-thread overwrite of the
    //    placeholder written above)
    while (qmask) {
        const int i = __builtin_ctz(qmask);
        qmask &= qmask - 1;
        const int val = sm.val[tid * K + i];
        int wb, p;
        if (lane > 0 && qual<INCL>(wexcl, val)) {
            wb = wave * WAVE;
            p = lane - 1;
        } else {
            // an earlier wave qualifies (qual(cross, val)); bounded anyway
            int w = wave - 1;
            while (w > 0 && !qual<INCL>(sm.wmin[w], val)) w--;
            wb = w * WAVE;
            p = WAVE - 1;
        }
#pragma unroll
        for (int s = 5; s >= 0; s--)
            if (!qual<INCL>(sm.win[s][wb + p], val)) p -= 1 << s;
        const int lt = wb + p;
        const int4 e0 = *reinterpret_cast<const int4*>(&sm.val[lt * K]);
        const int4 e1 = *reinterpret_cast<const int4*>(&sm.val[lt * K + 4]);
        const int j = qual<INCL>(e1.w, val) ? 7 : qual<INCL>(e1.z, val) ? 6
                    : qual<INCL>(e1.y, val) ? 5 : qual<INCL>(e1.x, val) ? 4
                    : qual<INCL>(e0.w, val) ? 3 : qual<INCL>(e0.z, val) ? 2
                    : qual<INCL>(e0.y, val) ? 1 : 0;
        out[base + i] = tile_start + lt * K + j;
    }

    // 4. open elements: count, scan, record the first QS positions
    const int mine = __builtin_popcount(omask);
    const int incl_cnt = wave_incl_add(mine);
    if (lane == WAVE - 1) sm.wopen[wave] = incl_cnt;
    __syncthreads();
    int before = incl_cnt - mine, total = 0;
#pragma unroll
    for (int w = 0; w < NWAVE; w++) {
        const int c = sm.wopen[w];
        if (w < wave) before += c;
        total += c;
    }
    unsigned om = omask;
    while (om && before < QS) {
        const int i = __builtin_ctz(om);
        om &= om - 1;
        to.qlist[(size_t)t * QS + before] = (uint16_t)(tid * K + i);
        before++;
    }
    if (tid == 0) {
        int m = sm.wmin[0];
#pragma unroll
        for (int w = 1; w < NWAVE; w++) m = min(m, sm.wmin[w]);
        to.tmin[t] = m;
        to.qcount[t] = total;
    }
}

// ---------------------------------------------------------------------------
// Hierarchy: one wave builds one 64-entry group of level l (exclusive
// prefix within the group, group min into level l+1).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void build_group(const int* __restrict__ mn, int* __restrict__ pre,
                                            int* __restrict__ up, int n, int g) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int idx = g * FAN + lane;
    const int x = (idx < n) ? mn[idx] : INF;
    const int inc = wave_incl_min(x);
    int ex = wave_shfl_up(inc, 1);
    if (lane == 0) ex = INF;
    if (idx < n) pre[idx] = ex;
    if (up && lane == WAVE - 1) up[g] = inc;
}

// ---------------------------------------------------------------------------
// Fix-up of tile t by one wave: answer the open elements whose value beats
// the exclusive prefix-min of all earlier tiles.
// ---------------------------------------------------------------------------
template <bool INCL>
__device__ __forceinline__ void fix_tile(const int* __restrict__ in, int* __restrict__ out,
                                         int n, int t, const Levels& L,
                                         const int* __restrict__ cmin,
                                         const int* __restrict__ qcount,
                                         const uint16_t* __restrict__ qlist) {
    if (t == 0) return;
    const int lane = threadIdx.x & (WAVE - 1);
    const int tile_start = t * TILE;
    const int P = tile_prefix(L, t);
    const int v1 = in[tile_start];
    if (!qual<INCL>(P, v1)) return;  // every open element answers -1
    const int cnt = qcount[t];
    const int nq = min(cnt, QS);
    if (lane < nq) {
        const int pos = tile_start + qlist[(size_t)t * QS + lane];
        const int v = in[pos];
        if (qual<INCL>(P, v)) out[pos] = search_before_tile<INCL>(in, L, cmin, t, v);
    }
    if (cnt > QS) {
        // open elements beyond the recorded ones: rescan the tile's tail.
        // The prefix-min of the tile up to the last recorded open element is
        // that element's value (open elements are prefix minima).
        const int lastpos = tile_start + qlist[(size_t)t * QS + QS - 1];
        int rm = in[lastpos];
        const int tile_end = min(tile_start + TILE, n);
        for (int p0 = lastpos + 1; p0 < tile_end && qual<INCL>(P, rm); p0 += WAVE) {
            const int idx = p0 + lane;
            const bool valid = idx < tile_end;
            const int xv = valid ? in[idx] : INF;
            const int inc = wave_incl_min(xv);
            int ex = wave_shfl_up(inc, 1);
            if (lane == 0) ex = INF;
            ex = min(ex, rm);
            if (valid && !qual<INCL>(ex, xv) && qual<INCL>(P, xv))
                out[idx] = search_before_tile<INCL>(in, L, cmin, t, xv);
            rm = min(rm, __shfl(inc, WAVE - 1, WAVE));
        }
    }
}

}  // namespace apsep
