// Launch helpers of the per-step gfx950 kernels (one kernel per reference step). All take
// device pointers and enqueue on `st`; they return hipGetLastError() of the launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rf_device.hpp"

namespace rf {

// Zero up to 16 device buffers of n floats each in one launch (fill kernel instead of
// hipMemsetAsync so the schedule can be captured in a graph).
hipError_t launch_fill_zero(float* const* bufs, int nbuf, long n, hipStream_t st);
// Warm start: clears *bad, checks the level-0 flow (W, H, S) against the rule (finite, at most
// max(W, H)) and, when it holds, writes it downscaled to level (w, h, s) into u, v.
hipError_t launch_warm_start(const float* wu, const float* wv, int W, int H, int S, int w, int h, int s, int* bad,
                             float* u, float* v, hipStream_t st);

hipError_t launch_gray_rgb(const uint8_t* rgb, float* out, int w, int h, int s, hipStream_t st);
hipError_t launch_gray_u8(const uint8_t* in, float* out, int w, int h, int s, hipStream_t st);
hipError_t launch_gray_u16(const uint16_t* in, float* out, int w, int h, int s, hipStream_t st);
hipError_t launch_gray_f32(const float* in, float* out, int w, int h, int s, hipStream_t st);
hipError_t launch_downscale(const float* src, int W, int H, int S, int w, int h, int s, bool scaled, float scale,
                            float* out, hipStream_t st);
hipError_t launch_upscale(const float* src, int w, int h, int s, int W, int H, int S, float scale, float* out,
                          hipStream_t st);
hipError_t launch_warp(const float* src, int w, int h, int s, const float* u, const float* v, float* out,
                       hipStream_t st);
hipError_t launch_derivatives(const float* I0, const float* I1, int w, int h, int s, float* Ix, float* Iy,
                              float* Iz, hipStream_t st);
hipError_t launch_data_l1(const float* du, const float* dv, const float* pu1, const float* pu2, const float* pv1,
                          const float* pv2, const float* Ix, const float* Iy, const float* Iz, int w, int h, int s,
                          float lambda, float theta, float* du1, float* dv1, hipStream_t st);
hipError_t launch_dual_tv(const float* du, const float* dv, const float* pu1, const float* pu2, const float* pv1,
                          const float* pv2, int w, int h, int s, float tau, float theta, float* pu1s, float* pu2s,
                          float* pv1s, float* pv2s, hipStream_t st);
hipError_t launch_median5(const float* u, const float* v, int w, int h, int s, float* ou, float* ov,
                          hipStream_t st);
hipError_t launch_median3d(const float* X, const float* Y, const float* Z, int w, int h, int s, float* X1,
                           float* Y1, float* Z1, hipStream_t st);
hipError_t launch_add(const float* a, const float* b, long count, float* out, hipStream_t st);
hipError_t launch_udiv(const float* n, long count, float d, float* q, hipStream_t st);
hipError_t launch_recon_data_l1(const float* u, const float* v, const float* du, const float* dv, const float* up,
                                const float* vp, const float* sku, const float* skv, const float* pu1,
                                const float* pu2, const float* pv1, const float* pv2, const float* Ix,
                                const float* Iy, const float* Iz, int w, int h, int s, float lambda, float aTv,
                                float aP, const float* ufn, const float* vfn, float aFn, const float* mask,
                                float* du1, float* dv1, bool wide, hipStream_t st);
hipError_t launch_uvproj(const float* u, const float* v, const float* du, const float* dv, const float* uc,
                         const float* vc, const float* sku, const float* skv, int w, int h, int s, float lf,
                         float aP, float* up, float* vp, bool wide, hipStream_t st);
// dpq / dc (device pointers, may be null) override the by-value camera data.
hipError_t launch_solve3d(const float* up, const float* vp, const float* X0, const float* Y0, const float* Z0,
                          const PQ& pq, const PQ* dpq, float lf, float lms, int w, int h, int s, float* X, float* Y,
                          float* Z, hipStream_t st);
// The triangulation evaluated in FP64 with the double camera matrices.
hipError_t launch_solve3d_f64(const float* up, const float* vp, const float* X0, const float* Y0, const float* Z0,
                              const PQd& pq, const PQd* dpq, float lf, float lms, int w, int h, int s, float* X,
                              float* Y, float* Z, hipStream_t st);
hipError_t launch_cleanup3d(const float* X0, const float* Y0, const float* Z0, const float* Xm, const float* Ym,
                            const float* Zm, int w, int h, int s, float* X1, float* Y1, float* Z1, hipStream_t st);
hipError_t launch_solve_uc(const float* X, const float* Y, const float* Z, const UcParams& c, const UcParams* dc,
                           int w, int h, int s, float* uc, float* vc, hipStream_t st);
hipError_t launch_update_sk(const float* sku, const float* skv, const float* u, const float* v, const float* du,
                            const float* dv, const float* up, const float* vp, int w, int h, int s, float* sku1,
                            float* skv1, hipStream_t st);
hipError_t launch_flow_to_hsv(const float* u, const float* v, int w, int h, int s, float* uvrgb, float flowscale,
                              hipStream_t st);
hipError_t launch_remove_edge(const float* X0, const float* Y0, const float* Z0, int range, float maxarea, int w,
                              int h, int s, float* X1, float* Y1, float* Z1, hipStream_t st);

hipError_t launch_deriv_single(const float* I, int w, int h, int s, float* Ix, float* Iy, hipStream_t st);
hipError_t launch_derivatives_patch(const float* I0, const float* I1, int w, int h, int s, float* Ix, float* Iy,
                                    float* Iz, hipStream_t st);
// Ixx == nullptr: the Charbonnier form; otherwise the gradient-constancy form.
hipError_t launch_data_charb(const float* du0, const float* dv0, const float* duhat0, const float* dvhat0,
                             const float* pu1, const float* pu2, const float* pv1, const float* pv2, const float* Ix,
                             const float* Iy, const float* It, const float* Ixx, const float* Ixy, const float* Ixt,
                             const float* Iyx, const float* Iyy, const float* Iyt, int w, int h, int s, float lambda,
                             float lambdagrad, float theta, float* du1, float* dv1, float* duhat1, float* dvhat1,
                             hipStream_t st);
hipError_t launch_data_l1_inpaint(const float* du0, const float* dv0, const float* mask0, const float* pu1,
                                  const float* pu2, const float* pv1, const float* pv2, const float* Ix,
                                  const float* Iy, const float* Iz, int w, int h, int s, float lambda, float theta,
                                  float* du1, float* dv1, hipStream_t st);
hipError_t launch_cv16u_to_32f(const uint16_t* in, float* out, int w, int h, int s, hipStream_t st);
hipError_t launch_scene_flow(const float* u, const float* v, const float* depth0, const float* depth1, float fx,
                             float fy, float cx, float cy, int w, int h, int s, float* pc0, float* pc1, float* sf,
                             hipStream_t st);
hipError_t launch_scene_flow_to_hsv(const float* sf, int w, int h, int s, float* rgb, float flowscale,
                                    hipStream_t st);

PQ make_pq(const double* P, const double* Q);
PQd make_pqd(const double* P, const double* Q);
UcParams make_uc(const double* R, const double* t, const double* K1);

}  // namespace rf
