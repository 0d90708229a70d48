// apsep_single.hip -- single-array APSEP solvers for gfx950 and the C ABI
// that drives them (include/apsep/apsep.h).
//
//   WSTL     tile pass -> one 64-ary level build per hierarchy level ->
//            fix-up pass (multi-launch; replaces runWSTL,
//            src/apsep.cuh:2548-2596)
//   SPT      the same three phases in one cooperative persistent launch,
//            separated by grid barriers (replaces runSPT,
//            src/apsep.cuh:3052-3067)
//
// All device work is integer compare/min and index arithmetic; no MFMA.
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <new>

#include "apsep_device.hpp"
#include "apsep_internal.hpp"
#include "../../include/apsep/apsep.h"

namespace cg = cooperative_groups;

namespace apsep {

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

template <bool INCL>
__global__ __launch_bounds__(NT) void k_tile_pass(const int* __restrict__ in,
                                                  int* __restrict__ out, int n,
                                                  TileOut to) {
    __shared__ TileSmem sm;
    solve_tile<INCL>(in, out, n, (int)blockIdx.x, sm, to);
}

__global__ __launch_bounds__(256) void k_build_level(const int* __restrict__ mn,
                                                     int* __restrict__ pre,
                                                     int* __restrict__ up, int n) {
    const int g = (int)(blockIdx.x * (blockDim.x / WAVE) + (threadIdx.x >> 6));
    if (g * FAN >= n) return;
    build_group(mn, pre, up, n, g);
}

template <bool INCL>
__global__ __launch_bounds__(256) void k_fix_pass(const int* __restrict__ in,
                                                  int* __restrict__ out, int n,
                                                  int nt, Levels L,
                                                  const int* __restrict__ cmin,
                                                  const int* __restrict__ qcount,
                                                  const uint16_t* __restrict__ qlist) {
    const int t = (int)(blockIdx.x * (blockDim.x / WAVE) + (threadIdx.x >> 6));
    if (t >= nt) return;
    fix_tile<INCL>(in, out, n, t, L, cmin, qcount, qlist);
}

// SPT: one cooperative persistent launch.  Phase 1 grid-strides over tiles,
// each level of the hierarchy is one grid-strided wave pass, phase 3
// grid-strides waves over tiles.  Grid barriers: 1 + nlev.
template <bool INCL>
__global__ __launch_bounds__(NT) void k_spt(const int* __restrict__ in,
                                            int* __restrict__ out, int n, int nt,
                                            TileOut to, Levels L) {
    __shared__ TileSmem sm;
    cg::grid_group grid = cg::this_grid();
    for (int t = (int)blockIdx.x; t < nt; t += (int)gridDim.x) {
        solve_tile<INCL>(in, out, n, t, sm, to);
        __syncthreads();
    }
    if (nt <= 1) return;  // uniform across the grid
    grid.sync();
    const int gw = (int)(blockIdx.x * NWAVE + (threadIdx.x >> 6));
    const int nw = (int)(gridDim.x * NWAVE);
    for (int l = 0; l < L.nlev; l++) {
        const int groups = (L.n[l] + FAN - 1) / FAN;
        int* up = (l + 1 < L.nlev) ? L.mn[l + 1] : nullptr;
        for (int g = gw; g < groups; g += nw) build_group(L.mn[l], L.pre[l], up, L.n[l], g);
        grid.sync();
    }
    for (int t = gw; t < nt; t += nw)
        fix_tile<INCL>(in, out, n, t, L, to.cmin, to.qcount, to.qlist);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static inline int64_t round4(int64_t x) { return (x + 3) & ~(int64_t)3; }

// Level sizes for nt tiles; returns nlev.
int level_sizes(int nt, int* sizes) {
    int nl = 0, m = nt;
    while (true) {
        sizes[nl++] = m;
        if (m <= FAN) break;
        m = (m + FAN - 1) / FAN;
    }
    return nl;
}

int64_t single_scratch_bytes(int64_t n_cap) {
    const int nt = (int)((n_cap + TILE - 1) / TILE);
    int sz[MAX_LEVELS];
    const int nl = level_sizes(nt < 1 ? 1 : nt, sz);
    int64_t b = 0;
    for (int l = 0; l < nl; l++) b += 2 * (round4(sz[l]) + 4) * 4;
    b += (round4((int64_t)nt * CPT) + 4) * 4;       // cmin
    b += (round4(nt) + 4) * 4;                        // qcount
    b += ((int64_t)nt * QS + 8) * 2;                  // qlist
    return b + 256 * 8;                               // alignment slack
}

static void* carve(char*& p, int64_t bytes) {
    uintptr_t a = ((uintptr_t)p + 255) & ~(uintptr_t)255;
    p = (char*)(a + bytes);
    return (void*)a;
}

apsep_status single_scratch_init(apsep_scratch* s) {
    const int nt = (int)((s->n_cap + TILE - 1) / TILE);
    s->nt_cap = nt;
    if (nt == 0) return APSEP_OK;
    int sz[MAX_LEVELS];
    const int nl = level_sizes(nt, sz);
    s->bytes = single_scratch_bytes(s->n_cap);
    if (hipMalloc(&s->base, (size_t)s->bytes) != hipSuccess) {
        s->base = nullptr;
        (void)hipGetLastError();
        return APSEP_ERR_ALLOC;
    }
    char* p = (char*)s->base;
    for (int l = 0; l < nl; l++) {
        s->lv_mn[l] = (int*)carve(p, (round4(sz[l]) + 4) * 4);
        s->lv_pre[l] = (int*)carve(p, (round4(sz[l]) + 4) * 4);
    }
    s->cmin = (int*)carve(p, (round4((int64_t)nt * CPT) + 4) * 4);
    s->qcount = (int*)carve(p, (round4(nt) + 4) * 4);
    s->qlist = (uint16_t*)carve(p, ((int64_t)nt * QS + 8) * 2);
    if (hipMemset(s->base, 0, (size_t)s->bytes) != hipSuccess) {
        (void)hipGetLastError();
        return APSEP_ERR_LAUNCH;
    }

    if (s->algo == APSEP_ALGO_SPT) {
        // the scratch serves both predicates: size for the tighter one
        int per_cu = 0, per_cu_incl = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_spt<false>, NT, 0);
        if (e == hipSuccess)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_incl, k_spt<true>, NT, 0);
        if (per_cu_incl < per_cu) per_cu = per_cu_incl;
        if (e != hipSuccess || per_cu <= 0) {
            (void)hipGetLastError();
            return APSEP_ERR_UNSUPPORTED;
        }
        int coop = 0;
        if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, s->device) != hipSuccess || !coop) {
            (void)hipGetLastError();
            return APSEP_ERR_UNSUPPORTED;
        }
        s->coop_grid = per_cu * s->num_cus;
    }
    return APSEP_OK;
}

static void fill_levels(const apsep_scratch* s, int nt, Levels& L) {
    int sz[MAX_LEVELS];
    L.nlev = level_sizes(nt, sz);
    for (int l = 0; l < MAX_LEVELS; l++) {
        L.n[l] = l < L.nlev ? sz[l] : 0;
        L.mn[l] = l < L.nlev ? s->lv_mn[l] : nullptr;
        L.pre[l] = l < L.nlev ? s->lv_pre[l] : nullptr;
    }
}

template <bool INCL>
static apsep_status run_wstl(apsep_scratch* s, const int* in, int* out, int n,
                             hipStream_t st) {
    const int nt = (n + TILE - 1) / TILE;
    Levels L;
    fill_levels(s, nt, L);
    TileOut to{s->lv_mn[0], s->cmin, s->qcount, s->qlist};
    hipLaunchKernelGGL(k_tile_pass<INCL>, dim3(nt), dim3(NT), 0, st, in, out, n, to);
    if (hipGetLastError() != hipSuccess) return APSEP_ERR_LAUNCH;
    if (nt <= 1) return APSEP_OK;
    for (int l = 0; l < L.nlev; l++) {
        const int groups = (L.n[l] + FAN - 1) / FAN;
        int* up = (l + 1 < L.nlev) ? L.mn[l + 1] : nullptr;
        hipLaunchKernelGGL(k_build_level, dim3((groups + 3) / 4), dim3(256), 0, st,
                           L.mn[l], L.pre[l], up, L.n[l]);
    }
    hipLaunchKernelGGL(k_fix_pass<INCL>, dim3((nt + 3) / 4), dim3(256), 0, st, in, out,
                       n, nt, L, s->cmin, s->qcount, s->qlist);
    if (hipGetLastError() != hipSuccess) return APSEP_ERR_LAUNCH;
    return APSEP_OK;
}

template <bool INCL>
static apsep_status run_spt(apsep_scratch* s, const int* in, int* out, int n,
                            hipStream_t st) {
    int nt = (n + TILE - 1) / TILE;
    Levels L;
    fill_levels(s, nt, L);
    TileOut to{s->lv_mn[0], s->cmin, s->qcount, s->qlist};
    int grid = s->coop_grid < nt ? s->coop_grid : nt;
    if (grid < 1) grid = 1;
    void* args[] = {(void*)&in, (void*)&out, (void*)&n, (void*)&nt, (void*)&to, (void*)&L};
    hipError_t e = hipLaunchCooperativeKernel((const void*)k_spt<INCL>, dim3(grid), dim3(NT),
                                              args, 0, st);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return APSEP_ERR_LAUNCH;
    }
    return APSEP_OK;
}

apsep_status single_run(apsep_scratch* s, const int32_t* in, int32_t* out, int n,
                        int incl, hipStream_t st) {
    switch (s->algo) {
    case APSEP_ALGO_WSTL:
        return incl ? run_wstl<true>(s, in, out, n, st) : run_wstl<false>(s, in, out, n, st);
    case APSEP_ALGO_SPT:
        return incl ? run_spt<true>(s, in, out, n, st) : run_spt<false>(s, in, out, n, st);
    default:
        return APSEP_ERR_INVALID_ARG;
    }
}

}  // namespace apsep
