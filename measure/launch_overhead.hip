// launch_overhead.hip -- per-launch cost of a normal and of a cooperative
// launch (hipLaunchCooperativeKernel) of a near-empty kernel, back to back
// on one stream, alone and each preceded by a one-workgroup kernel (as the
// persistent solve is preceded by its barrier reset).  This is the fixed cost
// the cooperative solver pays per call.
//
//   hipcc --offload-arch=gfx950 -O3 measure/launch_overhead.hip -o /tmp/launch_overhead
//   /tmp/launch_overhead
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP %s line %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while (0)
__global__ __launch_bounds__(512) void k_empty(int* p) { if (threadIdx.x == 0 && blockIdx.x == 0 && p[0] == 12345) p[1] = 1; }
__global__ __launch_bounds__(64) void k_tiny(int* p) { if (threadIdx.x == 0) p[2] = 0; }
int main() {
    int* p; CK(hipMalloc(&p, 64)); CK(hipMemset(p, 0, 64));
    hipEvent_t a, b; CK(hipEventCreate(&a)); CK(hipEventCreate(&b));
    const int R = 2000;
    for (int grid : {1, 768}) {
        for (int mode = 0; mode < 4; mode++) {
            void* args[] = {&p};
            for (int w = 0; w < 2; w++) {
                CK(hipDeviceSynchronize());
                CK(hipEventRecord(a, 0));
                for (int r = 0; r < R; r++) {
                    if (mode == 2 || mode == 3) hipLaunchKernelGGL(k_tiny, dim3(1), dim3(64), 0, 0, p);
                    if (mode == 0 || mode == 2) hipLaunchKernelGGL(k_empty, dim3(grid), dim3(512), 0, 0, p);
                    else CK(hipLaunchCooperativeKernel((const void*)k_empty, dim3(grid), dim3(512), args, 0, 0));
                }
                CK(hipEventRecord(b, 0));
                CK(hipEventSynchronize(b));
                float ms; CK(hipEventElapsedTime(&ms, a, b));
                if (w == 1) printf("grid %d mode %s: %.2f us per launch\n", grid,
                                   mode == 0 ? "normal" : mode == 1 ? "cooperative" : mode == 2 ? "tiny+normal" : "tiny+cooperative",
                                   ms * 1e3 / R);
            }
        }
    }
    return 0;
}
