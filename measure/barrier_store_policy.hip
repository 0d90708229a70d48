// barrier_store_policy.hip -- what it costs to make one workgroup's stores
// visible to another XCD's workgroups inside one cooperative launch.  Each
// workgroup writes its share of 128 MB with one store cache policy (plain,
// nt, sc1, sc0 sc1, nt sc1), passes a grid barrier (relaxed atomics after
// s_waitcnt, or with agent-scope release/acquire fences), then reads the
// next workgroup's share and counts stale words ("bad").  On MI355X plain
// stores need the fence (+~34 us: the L2 write-back of every dirty line),
// while sc1 stores cost nothing extra and are visible with the relaxed
// barrier -- the scheme the persistent solver's summaries use.
//
//   hipcc --offload-arch=gfx950 -O3 measure/barrier_store_policy.hip -o /tmp/bsp
//   /tmp/bsp
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <algorithm>

#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP %s line %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while (0)
typedef unsigned v4u __attribute__((vector_size(16)));

struct Bar {
    unsigned gen; int err; unsigned pad0[30];
    unsigned root; unsigned pad1[31];
    unsigned sub[64 * 32];
};
constexpr unsigned SPIN = 1u << 22;

template <int AUX>
__device__ __forceinline__ void st4(__amdgpu_buffer_rsrc_t r, unsigned off, v4u v) {
    __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, AUX);
}

// V: 1 relaxed barrier + waitcnt, 2 agent release fence before arrival + acquire fence after
template <int V, int AUX>
__global__ __launch_bounds__(512) void k(int* buf, int per, Bar* b, int* bad, int salt) {
    const unsigned nb = gridDim.x;
    const unsigned gen = __hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int* my = buf + (size_t)blockIdx.x * per;
    __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(my, 0, 0x7fffffff, 0x00020000);
    for (int i = threadIdx.x * 4; i < per; i += 512 * 4) {
        v4u v = {(unsigned)(i + salt), (unsigned)(i + 1 + salt), (unsigned)(i + 2 + salt), (unsigned)(i + 3 + salt)};
        st4<AUX>(r, (unsigned)i * 4u, v);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (V == 2) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned g = blockIdx.x % 64;
        const unsigned nsub = nb < 64 ? nb : 64;
        const unsigned gsz = nb / 64 + (g < nb % 64 ? 1u : 0u);
        if (atomicAdd(&b->sub[g * 32], 1u) == gsz - 1 && atomicAdd(&b->root, 1u) == nsub - 1) {
            __hip_atomic_store(&b->gen, gen + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned s = 0;
            while (__hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (++s > SPIN) { __hip_atomic_store(&b->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); break; }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        if (V == 2) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    const unsigned nx = (blockIdx.x + 1) % nb;
    const int* ot = buf + (size_t)nx * per;
    int errs = 0;
    for (int i = threadIdx.x * 4; i < per; i += 512 * 4) {
        const int4 v = *reinterpret_cast<const int4*>(ot + i);
        errs += (v.x != i + salt) + (v.w != i + 3 + salt);
    }
    if (errs) atomicAdd(bad, errs);
}

template <int V, int AUX>
void run(const char* name, int* buf, Bar* b, int* bad, int G, int per, hipEvent_t e0, hipEvent_t e1) {
    std::vector<float> ts;
    int bad_h = 0, err_h = 0;
    for (int r = 0; r < 14; r++) {
        CK(hipMemset(b, 0, sizeof(Bar)));
        CK(hipMemset(bad, 0, 4));
        CK(hipDeviceSynchronize());
        int salt = r * 7 + V * 3 + AUX;
        void* args[] = {&buf, (void*)&per, &b, &bad, &salt};
        CK(hipEventRecord(e0, 0));
        CK(hipLaunchCooperativeKernel((const void*)k<V, AUX>, dim3(G), dim3(512), args, 0, 0));
        CK(hipEventRecord(e1, 0));
        CK(hipEventSynchronize(e1));
        float ms; CK(hipEventElapsedTime(&ms, e0, e1));
        if (r >= 2) ts.push_back(ms);
        int bh; CK(hipMemcpy(&bh, bad, 4, hipMemcpyDeviceToHost)); bad_h += bh;
        int eh; CK(hipMemcpy(&eh, &b->err, 4, hipMemcpyDeviceToHost)); err_h += eh;
    }
    std::sort(ts.begin(), ts.end());
    printf("  %-28s median %.2f us  min %.2f us  bad=%d err=%d\n", name, ts[ts.size() / 2] * 1e3, ts[0] * 1e3, bad_h, err_h);
    fflush(stdout);
}

int main() {
    int cus = 0;
    CK(hipSetDevice(0));
    CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    const size_t N = 1 << 25;
    int* buf; Bar* b; int* bad;
    CK(hipMalloc(&buf, N * 4));
    CK(hipMalloc(&b, sizeof(Bar)));
    CK(hipMalloc(&bad, 4));
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    int occ = 1 << 30;
    const void* ks[] = {(const void*)k<1, 0>, (const void*)k<2, 0>, (const void*)k<1, 2>, (const void*)k<2, 2>,
                        (const void*)k<1, 16>, (const void*)k<2, 16>, (const void*)k<1, 17>, (const void*)k<1, 18>};
    for (const void* kf : ks) { int o = 0; CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, kf, 512, 0)); occ = std::min(occ, o); }
    const int G = std::min(occ, 3) * cus;
    const int per = (int)(N / G) & ~3;
    printf("CUs=%d occ=%d grid=%d per=%d ints\n", cus, occ, G, per);
    run<1, 0>("plain  relaxed", buf, b, bad, G, per, e0, e1);
    run<2, 0>("plain  fence", buf, b, bad, G, per, e0, e1);
    run<1, 2>("nt     relaxed", buf, b, bad, G, per, e0, e1);
    run<2, 2>("nt     fence", buf, b, bad, G, per, e0, e1);
    run<1, 16>("sc1    relaxed", buf, b, bad, G, per, e0, e1);
    run<2, 16>("sc1    fence", buf, b, bad, G, per, e0, e1);
    run<1, 17>("sc0sc1 relaxed", buf, b, bad, G, per, e0, e1);
    run<1, 18>("nt+sc1 relaxed", buf, b, bad, G, per, e0, e1);
    run<1, 0>("plain  relaxed (again)", buf, b, bad, G, per, e0, e1);
    printf("done\n");
    return 0;
}
