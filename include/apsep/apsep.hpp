// apsep.hpp -- drop-in C++ shims with the reference's host API names.
//
// A caller of the reference (src/apsep.cuh) swaps
//     #include "apsep.cuh"   for   #include "apsep/apsep.hpp"
// links libapsep.so, and keeps its calls unchanged:
//
//   allocWSTLScratch / freeWSTLScratch / runWSTL / launchWSTL
//       (src/apsep.cuh:2520-2605)
//   allocSPTScratch / freeSPTScratch / runSPT / launchSPT
//       (src/apsep.cuh:3013-3076)
//
// Differences a caller can observe:
//   * T must be int (the reference's SPT is int-only in practice: its int4
//     loads value-convert lanes, src/apsep.cuh:2689-2691).
//   * BLOCK_SIZE and IPT are accepted and ignored: the gfx950 engine uses its
//     own wave64 tile shape (512 lanes x 16 elements).
//   * runSPT / launchSPT run the fastest solver on the device unless
//     APSEP_SHIM_SPT_ALGO selects another (the SPT section below).
//   * run* launch on the default stream, asynchronously, like the reference;
//     an optional trailing stream argument (hipStream_t as void*) is added.
//   * Errors print "APSEP error <file>:<line>: <text> (<HIP error>)" and
//     exit, as the reference's gpuAssert does (src/apsep.cuh:39-47).  The C ABI underneath
//     (apsep.h) returns status codes instead.
#ifndef APSEP_APSEP_HPP
#define APSEP_APSEP_HPP

#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "apsep.h"

#define APSEP_SHIM_CHECK(x) ::apsep_shim::check((x), __FILE__, __LINE__)

namespace apsep_shim {
inline void check(apsep_status st, const char* file, int line) {
    if (st != APSEP_OK) {
        std::fprintf(stderr, "APSEP error %s:%d: %s (%s)\n", file, line,
                     apsep_status_string(st), apsep_last_error());
        std::exit(EXIT_FAILURE);
    }
}

// Scratch handle plus the reference's informational fields.
struct ScratchBase {
    apsep_scratch* handle = nullptr;
    int num_blocks = 0;   // tiles (the engine's tile = 8192 elements)
    int num_phys = 0;     // SPT persistent workgroups (occupancy x CUs)
    int n_cap = 0;
};

inline void fill(ScratchBase& s) {
    apsep_scratch_info info;
    APSEP_SHIM_CHECK(apsep_scratch_query(s.handle, &info));
    s.num_blocks = info.tiles;
    s.num_phys = info.persistent_blocks;
    s.n_cap = (int)info.n_cap;
}
}  // namespace apsep_shim

template <typename T, int BLOCK_SIZE = 128, int IPT = 4>
struct WSTLScratch : apsep_shim::ScratchBase {};

template <typename T, int BLOCK_SIZE = 128, int IPT = 4>
struct SPTScratch : apsep_shim::ScratchBase {};

// ----------------------------------------------------------------- WSTL

template <typename T, int BLOCK_SIZE = 128, int IPT = 4>
WSTLScratch<T, BLOCK_SIZE, IPT> allocWSTLScratch(int n) {
    static_assert(std::is_same<T, int>::value, "apsep: int32 keys only");
    WSTLScratch<T, BLOCK_SIZE, IPT> s;
    APSEP_SHIM_CHECK(apsep_scratch_create(APSEP_ALGO_WSTL, n > 0 ? n : 0, 0, &s.handle));
    apsep_shim::fill(s);
    return s;
}

template <typename T, int BLOCK_SIZE = 128, int IPT = 4>
void freeWSTLScratch(WSTLScratch<T, BLOCK_SIZE, IPT>& s) {
    APSEP_SHIM_CHECK(apsep_scratch_destroy(s.handle));
    s = WSTLScratch<T, BLOCK_SIZE, IPT>{};
}

template <typename T, int BLOCK_SIZE = 128, int IPT = 4>
void runWSTL(const T* d_in, int* d_out, int n, WSTLScratch<T, BLOCK_SIZE, IPT>& s,
             void* stream = nullptr) {
    static_assert(std::is_same<T, int>::value, "apsep: int32 keys only");
    if (n <= 0) return;
    APSEP_SHIM_CHECK(apsep_run_mode(s.handle, d_in, d_out, n, 0, stream));
}

template <typename T, int BLOCK_SIZE = 128, int IPT = 4>
void launchWSTL(const T* d_in, int* d_out, int n) {
    static_assert(std::is_same<T, int>::value, "apsep: int32 keys only");
    if (n <= 0) return;
    APSEP_SHIM_CHECK(apsep_launch(APSEP_ALGO_WSTL, d_in, d_out, n, 0));
}

// ----------------------------------------------------------------- SPT
//
// The SPT names solve with APSEP_SHIM_SPT_ALGO, by default APSEP_ALGO_AUTO:
// the fastest solver on the device, which on MI355X is the multi-launch
// one (the single persistent kernel gives each workgroup a fixed share of
// the tiles, where separate launches let the hardware balance them, and
// adds a grid barrier: 93 against 63 us on 2^25 random keys).  Every
// algorithm gives the same answers.  Define APSEP_SHIM_SPT_ALGO as
// APSEP_ALGO_SPT before including this header to run the single persistent
// kernel instead.
#ifndef APSEP_SHIM_SPT_ALGO
#define APSEP_SHIM_SPT_ALGO APSEP_ALGO_AUTO
#endif

template <typename T, int BLOCK_SIZE = 128, int IPT = 4, bool INCL = false>
SPTScratch<T, BLOCK_SIZE, IPT> allocSPTScratch(int n) {
    static_assert(std::is_same<T, int>::value, "apsep: int32 keys only");
    SPTScratch<T, BLOCK_SIZE, IPT> s;
    APSEP_SHIM_CHECK(apsep_scratch_create(APSEP_SHIM_SPT_ALGO, n > 0 ? n : 0, INCL ? 1 : 0,
                                          &s.handle));
    apsep_shim::fill(s);
    return s;
}

template <typename T, int BLOCK_SIZE = 128, int IPT = 4>
void freeSPTScratch(SPTScratch<T, BLOCK_SIZE, IPT>& s) {
    APSEP_SHIM_CHECK(apsep_scratch_destroy(s.handle));
    s = SPTScratch<T, BLOCK_SIZE, IPT>{};
}

template <typename T, int BLOCK_SIZE = 128, int IPT = 4, bool INCL = false>
void runSPT(const T* d_in, int* d_out, int n, SPTScratch<T, BLOCK_SIZE, IPT>& s,
            void* stream = nullptr) {
    static_assert(std::is_same<T, int>::value, "apsep: int32 keys only");
    if (n <= 0) return;
    APSEP_SHIM_CHECK(apsep_run_mode(s.handle, d_in, d_out, n, INCL ? 1 : 0, stream));
}

template <typename T, int BLOCK_SIZE = 128, int IPT = 4, bool INCL = false>
void launchSPT(const T* d_in, int* d_out, int n) {
    static_assert(std::is_same<T, int>::value, "apsep: int32 keys only");
    if (n <= 0) return;
    APSEP_SHIM_CHECK(apsep_launch(APSEP_SHIM_SPT_ALGO, d_in, d_out, n, INCL ? 1 : 0));
}

#endif  // APSEP_APSEP_HPP
