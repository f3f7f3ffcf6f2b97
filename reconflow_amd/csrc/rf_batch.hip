// Batch API (rfhip.h rf_batch_*): placeholder entry points until the batched schedule lands.

#include "../../include/rfhip.h"
#include "rf_solver.hpp"

extern "C" {

int rf_batch_create(rf_batch** out, const rf_params* params, int n, const int* widths, const int* heights) {
    (void)params;
    (void)n;
    (void)widths;
    (void)heights;
    if (out) *out = nullptr;
    rf::set_error("rf_batch_create: not implemented");
    return RF_ERR_UNSUPPORTED;
}
int rf_batch_destroy(rf_batch* b) {
    (void)b;
    return RF_OK;
}
int rf_batch_stride(const rf_batch* b, int i) {
    (void)b;
    (void)i;
    return RF_ERR_UNSUPPORTED;
}
int rf_batch_set_images_device(rf_batch* b, int i, const float* d_i0, const float* d_i1) {
    (void)b;
    (void)i;
    (void)d_i0;
    (void)d_i1;
    return RF_ERR_UNSUPPORTED;
}
int rf_batch_set_camera(rf_batch* b, int i, const double K0[9], const double K1[9], const double R1[9],
                        const double t1[3]) {
    (void)b;
    (void)i;
    (void)K0;
    (void)K1;
    (void)R1;
    (void)t1;
    return RF_ERR_UNSUPPORTED;
}
int rf_batch_solve(rf_batch* b) {
    (void)b;
    return RF_ERR_UNSUPPORTED;
}
int rf_batch_synchronize(rf_batch* b) {
    (void)b;
    return RF_ERR_UNSUPPORTED;
}
int rf_batch_get_result_device(const rf_batch* b, int i, const float** u, const float** v, const float** X,
                               const float** Y, const float** Z) {
    (void)b;
    (void)i;
    (void)u;
    (void)v;
    (void)X;
    (void)Y;
    (void)Z;
    return RF_ERR_UNSUPPORTED;
}
void* rf_batch_get_stream(const rf_batch* b) {
    (void)b;
    return nullptr;
}

}  // extern "C"
