// measure/copy_probe.hip -- build: hipcc -O3 --offload-arch=gfx950 -Xclang -target-feature
//   -Xclang -packed-fp32-ops -o copy_probe measure/copy_probe.hip; run: ./copy_probe
// Copy-rate probe: 3 input/output pairs of 2^25 int32 used in turn, one
// 8192-int tile per 512-thread workgroup (grid = n / 8192, exact).
//  mode 0: lane loads its 16 consecutive ints (4 x dwordx4 at 64 B lane stride),
//          stores them back the same way (plain)
//  mode 1: same loads, stores coalesced (lane l writes 16 B at 16 B stride, 4 x 1 KB per wave)
//  mode 2: mode 1 with non-temporal stores
//  mode 3: coalesced loads and non-temporal coalesced stores
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>
typedef int v4i __attribute__((ext_vector_type(4)));
template <int MODE>
__global__ __launch_bounds__(512) void k_copy(const int* __restrict__ in, int* __restrict__ out) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long tile = (long long)blockIdx.x * 8192;
    v4i v[4];
    if (MODE == 3) {
        const long long wb = tile + wave * 1024;
        for (int q = 0; q < 4; q++) v[q] = *reinterpret_cast<const v4i*>(in + wb + q * 256 + lane * 4);
    } else {
        const long long b = tile + tid * 16;
        for (int q = 0; q < 4; q++) v[q] = *reinterpret_cast<const v4i*>(in + b + q * 4);
    }
    if (MODE == 0) {
        const long long b = tile + tid * 16;
        for (int q = 0; q < 4; q++) *reinterpret_cast<v4i*>(out + b + q * 4) = v[q];
    } else {
        // (the values are not transposed: a timing probe of the access pattern only)
        const long long wb = tile + wave * 1024;
        for (int q = 0; q < 4; q++) {
            int* p = out + wb + q * 256 + lane * 4;
            if (MODE == 1) *reinterpret_cast<v4i*>(p) = v[q];
            else __builtin_nontemporal_store(v[q], reinterpret_cast<v4i*>(p));
        }
    }
}
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP %s line %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)
int main() {
    const long long n = 1ll << 25;
    int* in[3]; int* out[3];
    for (int i = 0; i < 3; i++) { CK(hipMalloc(&in[i], n * 4)); CK(hipMalloc(&out[i], n * 4)); CK(hipMemset(in[i], i, n * 4)); }
    hipEvent_t a, b; CK(hipEventCreate(&a)); CK(hipEventCreate(&b));
    const int grid = (int)(n / 8192);
    for (int mode = 0; mode < 4; mode++) {
        for (int rot = 1; rot <= 3; rot += 2) {
            auto launch = [&](int r) {
                const int k = r % rot;
                if (mode == 0) hipLaunchKernelGGL(k_copy<0>, dim3(grid), dim3(512), 0, 0, in[k], out[k]);
                if (mode == 1) hipLaunchKernelGGL(k_copy<1>, dim3(grid), dim3(512), 0, 0, in[k], out[k]);
                if (mode == 2) hipLaunchKernelGGL(k_copy<2>, dim3(grid), dim3(512), 0, 0, in[k], out[k]);
                if (mode == 3) hipLaunchKernelGGL(k_copy<3>, dim3(grid), dim3(512), 0, 0, in[k], out[k]);
            };
            for (int r = 0; r < 300; r++) launch(r);  // settle
            CK(hipDeviceSynchronize());
            CK(hipEventRecord(a, 0));
            const int reps = 300;
            for (int r = 0; r < reps; r++) launch(r);
            CK(hipEventRecord(b, 0));
            CK(hipEventSynchronize(b));
            float ms = 0; CK(hipEventElapsedTime(&ms, a, b));
            const double us = ms * 1e3 / reps;
            printf("mode %d %s: %.2f us per copy, %.1f GB/s\n", mode, rot == 3 ? "rotating" : "warm", us, 8.0 * n / (us * 1e-6) / 1e9);
            fflush(stdout);
        }
    }
    return 0;
}
