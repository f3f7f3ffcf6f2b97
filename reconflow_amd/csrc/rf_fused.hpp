// Fused solve schedule: one kernel per solver iteration (data + dual + projection +
// triangulation + XYZ median/cleanup + reprojection + Sk update), computed tile by tile with
// an LDS halo, with the same per-pixel arithmetic as the per-step kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "rf_solver.hpp"

namespace rf {

struct FusedCtx {
    const Geometry& g;
    const rf_params& p;
    hipStream_t st;
    DevParams* dparams;
    bool prior, sparse;
    float flow_scale;
    Profiler* prof;
};

hipError_t fused_flow(const FusedCtx& c, Fields& f);
hipError_t fused_recon(const FusedCtx& c, Fields& f);

}  // namespace rf
