// apsep_internal.hpp -- host-side state shared by the solver translation
// units (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/apsep/apsep.h"

struct apsep_scratch {
    apsep_algo algo;      // resolved (never AUTO)
    int incl;
    int64_t n_cap;
    int device;
    int num_cus;
    int nt_cap;
    void* base;
    int64_t bytes;
    int* lv_mn[5];
    int* lv_pre[5];
    int* cmin;
    int* qcount;
    uint16_t* qlist;
    int coop_grid;        // SPT persistent grid size (occupancy x CUs)
};

namespace apsep {

int level_sizes(int nt, int* sizes);
int64_t single_scratch_bytes(int64_t n_cap);
apsep_status single_scratch_init(apsep_scratch* s);
apsep_status single_run(apsep_scratch* s, const int32_t* in, int32_t* out, int n,
                        int incl, hipStream_t st);

}  // namespace apsep
