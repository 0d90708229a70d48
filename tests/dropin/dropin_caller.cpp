// dropin_caller.cpp -- a reference-style C++ caller of the drop-in shims.
//
// It includes only apsep/apsep.hpp (the reference-named templates) and the
// HIP runtime API for device memory, links libapsep.so, and calls the eight
// template entry points the way the reference's own drivers do:
//   * launchWSTL / launchSPT through lambdas, a device allocation per case,
//     as src/main.cu:228-252 (smallCases, stressKernel) does;
//   * allocSPTScratch<int,128,4,INCL> once at the largest size, then
//     runSPT<int,128,4,INCL> per case on 0xAB-poisoned output, as
//     src/bench_spt_le.cu:60-91 does; the same for allocWSTLScratch/runWSTL;
//   * freeSPTScratch / freeWSTLScratch at the end.
// The answers are written to a file for the test (tests/test_dropin_gpu.py)
// to compare with the oracle.
//
// Input file:  int32 count, then per case: int32 n, int32 entry mask, n keys.
// Output file: per case, per set bit of its mask (in bit order): n answers.
// Entry bits:
//   0 launchWSTL<int,128,4>               strict
//   1 launchSPT<int,128,4>                strict
//   2 launchSPT<int,128,4,true>           <=
//   3 runWSTL<int,128,4>                  strict, one scratch for every case
//   4 runSPT<int,128,4,false>             strict, one scratch for every case
//   5 runSPT<int,128,4,true>              <=,     one scratch for every case
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "apsep/apsep.hpp"

#define HIP_OK(x)                                                                    \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            std::fprintf(stderr, "HIP error %s:%d: %s\n", __FILE__, __LINE__,        \
                         hipGetErrorString(e_));                                     \
            std::exit(2);                                                            \
        }                                                                            \
    } while (0)

struct Case {
    int n;
    int mask;
    std::vector<int> keys;
};

static std::vector<Case> read_cases(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        std::perror(path);
        std::exit(2);
    }
    int32_t count = 0;
    if (std::fread(&count, 4, 1, f) != 1) std::exit(2);
    std::vector<Case> cs(count);
    for (auto& c : cs) {
        int32_t hdr[2];
        if (std::fread(hdr, 4, 2, f) != 2) std::exit(2);
        c.n = hdr[0];
        c.mask = hdr[1];
        c.keys.resize(c.n);
        if (c.n && std::fread(c.keys.data(), 4, c.n, f) != (size_t)c.n) std::exit(2);
    }
    std::fclose(f);
    return cs;
}

// The reference's per-case pattern: allocate, copy in, launch, synchronise,
// copy out, free.
template <typename LaunchFn>
static std::vector<int> via_launch(LaunchFn launch, const std::vector<int>& h) {
    const int n = (int)h.size();
    int *d_in = nullptr, *d_out = nullptr;
    HIP_OK(hipMalloc(&d_in, n * sizeof(int)));
    HIP_OK(hipMalloc(&d_out, n * sizeof(int)));
    HIP_OK(hipMemcpy(d_in, h.data(), n * sizeof(int), hipMemcpyHostToDevice));
    launch(d_in, d_out, n);
    HIP_OK(hipDeviceSynchronize());
    std::vector<int> got(n);
    HIP_OK(hipMemcpy(got.data(), d_out, n * sizeof(int), hipMemcpyDeviceToHost));
    HIP_OK(hipFree(d_in));
    HIP_OK(hipFree(d_out));
    return got;
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s cases.bin answers.bin\n", argv[0]);
        return 2;
    }
    const std::vector<Case> cases = read_cases(argv[1]);
    int nmax = 1;
    for (const auto& c : cases) nmax = c.n > nmax ? c.n : nmax;

    auto wstlLaunch = [](int* di, int* dou, int n) { launchWSTL<int, 128, 4>(di, dou, n); };
    auto sptLaunch = [](int* di, int* dou, int n) { launchSPT<int, 128, 4>(di, dou, n); };
    auto sptLaunchLe = [](int* di, int* dou, int n) { launchSPT<int, 128, 4, true>(di, dou, n); };

    // persistent scratches at the largest size, as bench_spt_le.cu allocates
    auto s_w = allocWSTLScratch<int, 128, 4>(nmax);
    auto s_lt = allocSPTScratch<int, 128, 4, false>(nmax);
    auto s_le = allocSPTScratch<int, 128, 4, true>(nmax);
    std::printf("dropin: %zu cases, nmax=%d, num_blocks=%d, num_phys strict=%d incl=%d\n",
                cases.size(), nmax, s_w.num_blocks, s_lt.num_phys, s_le.num_phys);
    int *d_in = nullptr, *d_out = nullptr;
    HIP_OK(hipMalloc(&d_in, (size_t)nmax * sizeof(int)));
    HIP_OK(hipMalloc(&d_out, (size_t)nmax * sizeof(int)));

    FILE* fo = std::fopen(argv[2], "wb");
    if (!fo) {
        std::perror(argv[2]);
        return 2;
    }
    std::vector<int> got((size_t)nmax);
    for (size_t k = 0; k < cases.size(); k++) {
        const Case& c = cases[k];
        const int n = c.n;
        for (int bit = 0; bit < 6; bit++) {
            if (!(c.mask & (1 << bit))) continue;
            std::vector<int> r;
            if (bit == 0) r = via_launch(wstlLaunch, c.keys);
            if (bit == 1) r = via_launch(sptLaunch, c.keys);
            if (bit == 2) r = via_launch(sptLaunchLe, c.keys);
            if (bit >= 3) {
                HIP_OK(hipMemcpy(d_in, c.keys.data(), (size_t)n * sizeof(int),
                                 hipMemcpyHostToDevice));
                HIP_OK(hipMemset(d_out, 0xAB, (size_t)n * sizeof(int)));
                if (bit == 3) runWSTL<int, 128, 4>(d_in, d_out, n, s_w);
                if (bit == 4) runSPT<int, 128, 4, false>(d_in, d_out, n, s_lt);
                if (bit == 5) runSPT<int, 128, 4, true>(d_in, d_out, n, s_le);
                HIP_OK(hipDeviceSynchronize());
                HIP_OK(hipMemcpy(got.data(), d_out, (size_t)n * sizeof(int),
                                 hipMemcpyDeviceToHost));
                r.assign(got.begin(), got.begin() + n);
            }
            if (n && std::fwrite(r.data(), 4, (size_t)n, fo) != (size_t)n) return 2;
        }
        if ((k + 1) % 100 == 0 || k + 1 == cases.size())
            std::printf("dropin: case %zu/%zu done (n=%d)\n", k + 1, cases.size(), n);
        std::fflush(stdout);
    }
    std::fclose(fo);
    HIP_OK(hipFree(d_in));
    HIP_OK(hipFree(d_out));
    freeWSTLScratch<int, 128, 4>(s_w);
    freeSPTScratch<int, 128, 4>(s_lt);
    freeSPTScratch<int, 128, 4>(s_le);
    std::printf("dropin: ok\n");
    return 0;
}
