// Per-kernel C entry points (rfhip.h rf_k_*): the public kernel wrappers of the reference
// (Cudaflow.h:101-193, Reconflow.h:86-158) on device pointers, enqueued on a caller stream.

#include <cmath>
#include <string>

#include "../../include/rfhip.h"
#include "rf_kernels.hpp"
#include "rf_solver.hpp"

using namespace rf;

static int done(hipError_t e, const char* what) {
    if (e == hipSuccess) return RF_OK;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return RF_ERR_HIP;
}

static bool bad_geom(int w, int h, int s, const char* what) {
    if (w < 0 || h < 0 || s < w) {
        set_error(std::string(what) + ": invalid width/height/stride");
        return true;
    }
    return false;
}

#define ST ((hipStream_t)stream)
#define GEOM(name)                                    \
    if (bad_geom(w, h, s, name)) return RF_ERR_ARG;

extern "C" {

int rf_k_rgb_to_gray(const uint8_t* rgb, float* gray, int w, int h, int s, void* stream) {
    GEOM("rf_k_rgb_to_gray");
    return done(launch_gray_rgb(rgb, gray, w, h, s, ST), "rf_k_rgb_to_gray");
}
int rf_k_cv8u_to_gray(const uint8_t* in, float* gray, int w, int h, int s, void* stream) {
    GEOM("rf_k_cv8u_to_gray");
    return done(launch_gray_u8(in, gray, w, h, s, ST), "rf_k_cv8u_to_gray");
}
int rf_k_cv16u_to_gray(const uint16_t* in, float* gray, int w, int h, int s, void* stream) {
    GEOM("rf_k_cv16u_to_gray");
    return done(launch_gray_u16(in, gray, w, h, s, ST), "rf_k_cv16u_to_gray");
}
int rf_k_cv32f_to_gray(const float* in, float* gray, int w, int h, int s, void* stream) {
    GEOM("rf_k_cv32f_to_gray");
    return done(launch_gray_f32(in, gray, w, h, s, ST), "rf_k_cv32f_to_gray");
}
int rf_k_downscale(const float* src, int W, int H, int S, int w, int h, int s, float* out, void* stream) {
    GEOM("rf_k_downscale");
    if (W <= 0 || H <= 0 || S < W) return RF_ERR_ARG;
    return done(launch_downscale(src, W, H, S, w, h, s, false, 1.0f, out, ST), "rf_k_downscale");
}
int rf_k_downscale_scaled(const float* src, int W, int H, int S, int w, int h, int s, float scale, float* out,
                          void* stream) {
    GEOM("rf_k_downscale_scaled");
    if (W <= 0 || H <= 0 || S < W) return RF_ERR_ARG;
    return done(launch_downscale(src, W, H, S, w, h, s, true, scale, out, ST), "rf_k_downscale_scaled");
}
int rf_k_upscale(const float* src, int w, int h, int s, int W, int H, int S, float scale, float* out, void* stream) {
    GEOM("rf_k_upscale");
    if (W < 0 || H < 0 || S < W) return RF_ERR_ARG;
    return done(launch_upscale(src, w, h, s, W, H, S, scale, out, ST), "rf_k_upscale");
}
int rf_k_warp_image(const float* src, int w, int h, int s, const float* u, const float* v, float* out, void* stream) {
    GEOM("rf_k_warp_image");
    return done(launch_warp(src, w, h, s, u, v, out, ST), "rf_k_warp_image");
}
int rf_k_compute_derivatives(const float* I0, const float* I1, int w, int h, int s, float* Ix, float* Iy, float* Iz,
                             void* stream) {
    GEOM("rf_k_compute_derivatives");
    return done(launch_derivatives(I0, I1, w, h, s, Ix, Iy, Iz, ST), "rf_k_compute_derivatives");
}
int rf_k_solve_data_l1(const float* duhat0, const float* dvhat0, const float* pu1, const float* pu2, const float* pv1,
                       const float* pv2, const float* Ix, const float* Iy, const float* Iz, int w, int h, int s,
                       float lambda, float theta, float* duhat1, float* dvhat1, void* stream) {
    GEOM("rf_k_solve_data_l1");
    return done(launch_data_l1(duhat0, dvhat0, pu1, pu2, pv1, pv2, Ix, Iy, Iz, w, h, s, lambda, theta, duhat1,
                               dvhat1, ST),
                "rf_k_solve_data_l1");
}
int rf_k_solve_smooth_dual_tv_global(const float* duhat, const float* dvhat, const float* pu1, const float* pu2,
                                     const float* pv1, const float* pv2, int w, int h, int s, float tau, float theta,
                                     float* pu1s, float* pu2s, float* pv1s, float* pv2s, void* stream) {
    GEOM("rf_k_solve_smooth_dual_tv_global");
    return done(launch_dual_tv(duhat, dvhat, pu1, pu2, pv1, pv2, w, h, s, tau, theta, pu1s, pu2s, pv1s, pv2s, ST),
                "rf_k_solve_smooth_dual_tv_global");
}
int rf_k_median_filter(const float* u, const float* v, int w, int h, int s, float* ou, float* ov, int kernel_size,
                       void* stream) {
    GEOM("rf_k_median_filter");
    if (kernel_size != 5) {
        set_error("rf_k_median_filter: only kernel size 5 is on the solve path");
        return RF_ERR_UNSUPPORTED;
    }
    return done(launch_median5(u, v, w, h, s, ou, ov, ST), "rf_k_median_filter");
}
int rf_k_median_filter_3d(const float* X, const float* Y, const float* Z, int w, int h, int s, float* X1, float* Y1,
                          float* Z1, int kernel_size, void* stream) {
    GEOM("rf_k_median_filter_3d");
    if (kernel_size != 3) {
        set_error("rf_k_median_filter_3d: only kernel size 3 exists in the reference");
        return RF_ERR_UNSUPPORTED;
    }
    return done(launch_median3d(X, Y, Z, w, h, s, X1, Y1, Z1, ST), "rf_k_median_filter_3d");
}
int rf_k_add(const float* a, const float* b, int count, float* sum, void* stream) {
    if (count < 0) return RF_ERR_ARG;
    return done(launch_add(a, b, count, sum, ST), "rf_k_add");
}
int rf_k_uniform_divide(const float* n, int count, float d, float* q, void* stream) {
    if (count < 0 || !std::isfinite(d) || !(d > 0.0f)) {
        set_error("rf_k_uniform_divide: count must be >= 0 and d finite and positive");
        return RF_ERR_ARG;
    }
    return done(launch_udiv(n, count, d, q, ST), "rf_k_uniform_divide");
}
int rf_k_solve_recon_data_l1(const float* u, const float* v, const float* duhat0, const float* dvhat0,
                             const float* uproj, const float* vproj, const float* sku, const float* skv,
                             const float* pu1, const float* pu2, const float* pv1, const float* pv2, const float* Ix,
                             const float* Iy, const float* Iz, int w, int h, int s, float lambda, float alpha_tv,
                             float alpha_proj, float* duhat1, float* dvhat1, void* stream) {
    GEOM("rf_k_solve_recon_data_l1");
    return done(launch_recon_data_l1(u, v, duhat0, dvhat0, uproj, vproj, sku, skv, pu1, pu2, pv1, pv2, Ix, Iy, Iz, w,
                                     h, s, lambda, alpha_tv, alpha_proj, nullptr, nullptr, 0.0f, nullptr, duhat1,
                                     dvhat1, false, ST),
                "rf_k_solve_recon_data_l1");
}
int rf_k_solve_recon_data_l1_fn(const float* u, const float* v, const float* duhat0, const float* dvhat0,
                                const float* uproj, const float* vproj, const float* ufn, const float* vfn,
                                const float* sku, const float* skv, const float* pu1, const float* pu2,
                                const float* pv1, const float* pv2, const float* Ix, const float* Iy, const float* Iz,
                                int w, int h, int s, float lambda, float alpha_tv, float alpha_proj, float alpha_fn,
                                const float* fnmask, float* duhat1, float* dvhat1, void* stream) {
    GEOM("rf_k_solve_recon_data_l1_fn");
    if (!ufn || !vfn) return RF_ERR_ARG;
    return done(launch_recon_data_l1(u, v, duhat0, dvhat0, uproj, vproj, sku, skv, pu1, pu2, pv1, pv2, Ix, Iy, Iz, w,
                                     h, s, lambda, alpha_tv, alpha_proj, ufn, vfn, alpha_fn, fnmask, duhat1, dvhat1,
                                     false, ST),
                "rf_k_solve_recon_data_l1_fn");
}
int rf_k_solve_uvproj(const float* u, const float* v, const float* du, const float* dv, const float* uc,
                      const float* vc, const float* sku, const float* skv, int w, int h, int s, float lambdaf,
                      float alpha_proj, float* uproj, float* vproj, void* stream) {
    GEOM("rf_k_solve_uvproj");
    return done(launch_uvproj(u, v, du, dv, uc, vc, sku, skv, w, h, s, lambdaf, alpha_proj, uproj, vproj, false, ST),
                "rf_k_solve_uvproj");
}
int rf_k_solve_3d_minimal_area(const float* uproj, const float* vproj, const float* X0, const float* Y0,
                               const float* Z0, const double P[12], const double Q[12], float lambdaf, float lambdams,
                               int w, int h, int s, float* X, float* Y, float* Z, void* stream) {
    GEOM("rf_k_solve_3d_minimal_area");
    if (!P || !Q) return RF_ERR_ARG;
    return done(launch_solve3d(uproj, vproj, X0, Y0, Z0, make_pq(P, Q), nullptr, lambdaf, lambdams, w, h, s, X, Y, Z,
                               ST),
                "rf_k_solve_3d_minimal_area");
}
int rf_k_cleanup_3d(const float* X0, const float* Y0, const float* Z0, const float* Xm, const float* Ym,
                    const float* Zm, int w, int h, int s, float* X1, float* Y1, float* Z1, void* stream) {
    GEOM("rf_k_cleanup_3d");
    return done(launch_cleanup3d(X0, Y0, Z0, Xm, Ym, Zm, w, h, s, X1, Y1, Z1, ST), "rf_k_cleanup_3d");
}
int rf_k_solve_uc(const float* X, const float* Y, const float* Z, const double R[9], const double t[3],
                  const double K1[9], int w, int h, int s, float* uc, float* vc, void* stream) {
    GEOM("rf_k_solve_uc");
    if (!R || !t || !K1) return RF_ERR_ARG;
    return done(launch_solve_uc(X, Y, Z, make_uc(R, t, K1), nullptr, w, h, s, uc, vc, ST), "rf_k_solve_uc");
}
int rf_k_update_sk(const float* sku0, const float* skv0, const float* u, const float* v, const float* du,
                   const float* dv, const float* uproj, const float* vproj, int w, int h, int s, float* sku1,
                   float* skv1, void* stream) {
    GEOM("rf_k_update_sk");
    return done(launch_update_sk(sku0, skv0, u, v, du, dv, uproj, vproj, w, h, s, sku1, skv1, ST), "rf_k_update_sk");
}
int rf_k_flow_to_hsv(const float* u, const float* v, int w, int h, int s, float* uvrgb, float flow_scale,
                     void* stream) {
    GEOM("rf_k_flow_to_hsv");
    return done(launch_flow_to_hsv(u, v, w, h, s, uvrgb, flow_scale, ST), "rf_k_flow_to_hsv");
}
int rf_k_remove_edge_artifacts_3d(const float* X0, const float* Y0, const float* Z0, int surface_width,
                                  float max_surface_area, int w, int h, int s, float* X1, float* Y1, float* Z1,
                                  void* stream) {
    GEOM("rf_k_remove_edge_artifacts_3d");
    if (surface_width < 1 || surface_width >= w || surface_width >= h) return RF_ERR_ARG;
    return done(launch_remove_edge(X0, Y0, Z0, surface_width, max_surface_area, w, h, s, X1, Y1, Z1, ST),
                "rf_k_remove_edge_artifacts_3d");
}

int rf_k_compute_deriv(const float* I, int w, int h, int s, float* Ix, float* Iy, void* stream) {
    GEOM("rf_k_compute_deriv");
    return done(launch_deriv_single(I, w, h, s, Ix, Iy, ST), "rf_k_compute_deriv");
}
int rf_k_compute_derivatives_patch(const float* I0, const float* I1, int w, int h, int s, float* Ix, float* Iy,
                                   float* Iz, void* stream) {
    GEOM("rf_k_compute_derivatives_patch");
    return done(launch_derivatives_patch(I0, I1, w, h, s, Ix, Iy, Iz, ST), "rf_k_compute_derivatives_patch");
}
int rf_k_solve_data_charb_for_tv(const float* du0, const float* dv0, const float* duhat0, const float* dvhat0,
                                 const float* pu1, const float* pu2, const float* pv1, const float* pv2,
                                 const float* Ix, const float* Iy, const float* It, const float* Ixx,
                                 const float* Ixy, const float* Ixt, const float* Iyx, const float* Iyy,
                                 const float* Iyt, int w, int h, int s, float lambda, float lambdagrad, float theta,
                                 float* du1, float* dv1, float* duhat1, float* dvhat1, void* stream) {
    GEOM("rf_k_solve_data_charb_for_tv");
    const bool any = Ixx || Ixy || Ixt || Iyx || Iyy || Iyt, all = Ixx && Ixy && Ixt && Iyx && Iyy && Iyt;
    if (any && !all) {
        set_error("rf_k_solve_data_charb_for_tv: pass all six gradient-image derivatives or none");
        return RF_ERR_ARG;
    }
    return done(launch_data_charb(du0, dv0, duhat0, dvhat0, pu1, pu2, pv1, pv2, Ix, Iy, It, Ixx, Ixy, Ixt, Iyx, Iyy,
                                  Iyt, w, h, s, lambda, lambdagrad, theta, du1, dv1, duhat1, dvhat1, ST),
                "rf_k_solve_data_charb_for_tv");
}
int rf_k_solve_data_l1_inpaint(const float* duhat0, const float* dvhat0, const float* mask0, const float* mask1,
                               const float* pu1, const float* pu2, const float* pv1, const float* pv2,
                               const float* Ix, const float* Iy, const float* It, int w, int h, int s, float lambda,
                               float theta, float* duhat1, float* dvhat1, void* stream) {
    GEOM("rf_k_solve_data_l1_inpaint");
    (void)mask1;
    return done(launch_data_l1_inpaint(duhat0, dvhat0, mask0, pu1, pu2, pv1, pv2, Ix, Iy, It, w, h, s, lambda, theta,
                                       duhat1, dvhat1, ST),
                "rf_k_solve_data_l1_inpaint");
}
int rf_k_cv16u_to_32f(const uint16_t* in, float* out, int w, int h, int s, void* stream) {
    GEOM("rf_k_cv16u_to_32f");
    return done(launch_cv16u_to_32f(in, out, w, h, s, ST), "rf_k_cv16u_to_32f");
}
int rf_k_solve_scene_flow(const float* u, const float* v, const float* depth0, const float* depth1, float fx,
                          float fy, float cx, float cy, int w, int h, int s, float* pcloud0, float* pcloud1,
                          float* sceneflow, void* stream) {
    GEOM("rf_k_solve_scene_flow");
    return done(launch_scene_flow(u, v, depth0, depth1, fx, fy, cx, cy, w, h, s, pcloud0, pcloud1, sceneflow, ST),
                "rf_k_solve_scene_flow");
}
int rf_k_scene_flow_to_hsv(const float* sceneflow, int w, int h, int s, float* rgb, float flow_scale, void* stream) {
    GEOM("rf_k_scene_flow_to_hsv");
    return done(launch_scene_flow_to_hsv(sceneflow, w, h, s, rgb, flow_scale, ST), "rf_k_scene_flow_to_hsv");
}

}  // extern "C"
