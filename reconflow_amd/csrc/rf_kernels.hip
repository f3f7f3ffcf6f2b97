// Per-step gfx950 kernels: one kernel per reference kernel, one thread per pixel.
//
// Workgroups are 64 x 4 threads: one wave64 spans 64 consecutive pixels of a row, so every
// field access of a wave is one contiguous 256-byte segment. These kernels are the plain
// path that the fused path is checked against; each evaluates the per-pixel functions of
// rf_device.hpp, so both paths share one arithmetic definition.

#include <math.h>

#include "rf_kernels.hpp"

namespace rf {
namespace {

constexpr int BX = 64, BY = 4;

inline dim3 grid2(int w, int h) { return dim3((w + BX - 1) / BX, (h + BY - 1) / BY); }

#define RF_PIX                                          \
    const int ix = blockIdx.x * BX + threadIdx.x;       \
    const int iy = blockIdx.y * BY + threadIdx.y;       \
    if (ix >= w || iy >= h) return;                     \
    const long pos = ix + (long)iy * s;

struct FillArgs {
    float* bufs[16];
};

__global__ void __launch_bounds__(256) k_fill_zero(FillArgs a, int nbuf, long n) {
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    long stride = (long)gridDim.x * 256;
    for (int b = 0; b < nbuf; ++b)
        for (long j = i; j < n; j += stride) a.bufs[b][j] = 0.0f;
}

// rgbToGray.cu:3-19
__global__ void __launch_bounds__(256) k_gray_rgb(const uint8_t* rgb, float* out, int w, int h, int s) {
    RF_PIX
    float g = ((float)rgb[3 * pos] + (float)rgb[3 * pos + 1] + (float)rgb[3 * pos + 2]) / 3.0f;
    out[pos] = g / 256.0f;
}
// rgbToGray.cu:21-34
__global__ void __launch_bounds__(256) k_gray_u8(const uint8_t* in, float* out, int w, int h, int s) {
    RF_PIX
    out[pos] = (float)in[pos] / 256.0f;
}
// rgbToGray.cu:36-49
__global__ void __launch_bounds__(256) k_gray_u16(const uint16_t* in, float* out, int w, int h, int s) {
    RF_PIX
    out[pos] = (float)in[pos] / 65536.0f;
}
// rgbToGray.cu:66-77
__global__ void __launch_bounds__(256) k_gray_f32(const float* in, float* out, int w, int h, int s) {
    RF_PIX
    out[pos] = in[pos];
}

// downscale.cu:6-48. (w, h, s) is the coarse output; (W, H, S) the fine source.
__global__ void __launch_bounds__(256) k_downscale(const float* src, int W, int H, int S, int w, int h, int s,
                                                    int scaled, float scale, float* out) {
    RF_PIX
    int dxn = 4 * w, dyn = 4 * h;
    int xc = (4 * ix + 2) * W - 2 * w;
    int yc = (4 * iy + 2) * H - 2 * h;
    float a = sample_ratio(src, W, H, S, xc - W, dxn, yc, dyn);
    float b = sample_ratio(src, W, H, S, xc + W, dxn, yc, dyn);
    float c = sample_ratio(src, W, H, S, xc, dxn, yc - H, dyn);
    float d = sample_ratio(src, W, H, S, xc, dxn, yc + H, dyn);
    float sum = ((a + b) + c) + d;
    out[pos] = scaled ? (scale * 0.25f) * sum : 0.25f * sum;
}

// Warm start (no counterpart in the reference, which starts every solve from zero): the rule
// over the w x h level-0 pixels is that every component is finite and at most `bound` in
// magnitude; a violation sets *bad, which k_set_int clears first.
__global__ void __launch_bounds__(256) k_ws_check(const float* u, const float* v, int w, int h, int s, float bound,
                                                   int* bad) {
    RF_PIX
    if (!(fabsf(u[pos]) <= bound) || !(fabsf(v[pos]) <= bound)) atomicOr(bad, 1);
}

__global__ void k_set_int(int* p, int value) {
    if (threadIdx.x == 0) *p = value;
}

// The warm start at the coarsest level (w, h, s) from level 0 (W, H, S), by the prior flow's
// DownscaleScalingKernel (downscale.cu:28-48, as k_downscale with scaling): u, v keep their zeros
// when the rule failed.
__global__ void __launch_bounds__(256) k_ws_apply(const float* wu, const float* wv, int W, int H, int S, int w, int h,
                                                   int s, float scale, const int* bad, float* u, float* v) {
    RF_PIX
    if (*bad) return;
    const int dxn = 4 * w, dyn = 4 * h;
    const int xc = (4 * ix + 2) * W - 2 * w, yc = (4 * iy + 2) * H - 2 * h;
    auto down = [&](const float* src) {
        const float a = sample_ratio(src, W, H, S, xc - W, dxn, yc, dyn);
        const float b = sample_ratio(src, W, H, S, xc + W, dxn, yc, dyn);
        const float c = sample_ratio(src, W, H, S, xc, dxn, yc - H, dyn);
        const float d = sample_ratio(src, W, H, S, xc, dxn, yc + H, dyn);
        return (scale * 0.25f) * (((a + b) + c) + d);
    };
    u[pos] = down(wu);
    v[pos] = down(wv);
}

// upscale.cu:14-27. (w, h, s) is the fine output; (cw, ch, cs) the coarse source.
__global__ void __launch_bounds__(256) k_upscale(const float* src, int cw, int ch, int cs, int w, int h, int s,
                                                  float scale, float* out) {
    RF_PIX
    float v = sample_ratio(src, cw, ch, cs, (2 * ix + 1) * cw - w, 2 * w, (2 * iy + 1) * ch - h, 2 * h);
    out[pos] = v * scale;
}

// warpImage.cu:15-29
__global__ void __launch_bounds__(256) k_warp(const float* src, int w, int h, int s, const float* u,
                                               const float* v, float* out) {
    RF_PIX
    out[pos] = warp_sample(src, w, h, s, ix, iy, u[pos], v[pos]);
}

// computeDerivatives.cu:8-57
__global__ void __launch_bounds__(256) k_derivatives(const float* I0, const float* I1, int w, int h, int s,
                                                      float* Ix, float* Iy, float* Iz) {
    RF_PIX
    const long r = (long)iy * s;
    int xm2 = refl(ix - 2, w), xm1 = refl(ix - 1, w), xp1 = refl(ix + 1, w), xp2 = refl(ix + 2, w);
    int ym2 = refl(iy - 2, h), ym1 = refl(iy - 1, h), yp1 = refl(iy + 1, h), yp2 = refl(iy + 2, h);
    float t0 = d5(I0[r + xm2], I0[r + xm1], I0[r + xp1], I0[r + xp2]);
    float t1 = d5(I1[r + xm2], I1[r + xm1], I1[r + xp1], I1[r + xp2]);
    Ix[pos] = (t0 + t1) * 0.5f;
    Iz[pos] = I1[pos] - I0[pos];
    t0 = d5(I0[ix + (long)ym2 * s], I0[ix + (long)ym1 * s], I0[ix + (long)yp1 * s], I0[ix + (long)yp2 * s]);
    t1 = d5(I1[ix + (long)ym2 * s], I1[ix + (long)ym1 * s], I1[ix + (long)yp1 * s], I1[ix + (long)yp2 * s]);
    Iy[pos] = (t0 + t1) * 0.5f;
}

// solveDataL1.cu:4-91
__global__ void __launch_bounds__(256) k_data_l1(const float* du0, const float* dv0, const float* pu1,
                                                  const float* pu2, const float* pv1, const float* pv2,
                                                  const float* Ix, const float* Iy, const float* Iz, int w, int h,
                                                  int s, float lambda, float theta, float* du1, float* dv1) {
    RF_PIX
    float du, dv;
    l1_threshold(Ix[pos], Iy[pos], Iz[pos], du0[pos], dv0[pos], lambda, theta, du, dv);
    bool hl = ix > 0, hu = iy > 0;
    float divu = div_p(pu1[pos], hl ? pu1[pos - 1] : 0.0f, pu2[pos], hu ? pu2[pos - s] : 0.0f, hl, hu);
    float divv = div_p(pv1[pos], hl ? pv1[pos - 1] : 0.0f, pv2[pos], hu ? pv2[pos - s] : 0.0f, hl, hu);
    du1[pos] = du + theta * divu;
    dv1[pos] = dv + theta * divv;
}

// solveSmoothDualTVGlobal.cu:4-65
__global__ void __launch_bounds__(256) k_dual_tv(const float* du, const float* dv, const float* pu1,
                                                  const float* pu2, const float* pv1, const float* pv2, int w, int h,
                                                  int s, float tau, float theta, float* pu1s, float* pu2s,
                                                  float* pv1s, float* pv2s) {
    RF_PIX
    float dux = 0.0f, dvx = 0.0f, duy = 0.0f, dvy = 0.0f;
    if (ix + 1 < w) {
        dux = du[pos + 1] - du[pos];
        dvx = dv[pos + 1] - dv[pos];
    }
    if (iy + 1 < h) {
        duy = du[pos + s] - du[pos];
        dvy = dv[pos + s] - dv[pos];
    }
    float fac = tau / theta;
    float a, b;
    dual_px(dux, duy, pu1[pos], pu2[pos], fac, a, b);
    pu1s[pos] = a;
    pu2s[pos] = b;
    dual_px(dvx, dvy, pv1[pos], pv2[pos], fac, a, b);
    pv1s[pos] = a;
    pv2s[pos] = b;
}

// medianFilter.cu:3-75
__global__ void __launch_bounds__(256) k_median5(const float* u, const float* v, int w, int h, int s, float* ou,
                                                  float* ov) {
    RF_PIX
    float mu[25], mv[25];
    gather5x5(u, v, ix, iy, w, h, s, mu, mv);
    ou[pos] = median25(mu);
    ov[pos] = median25(mv);
}

// medianFilter.cu:166-243
__global__ void __launch_bounds__(256) k_median3d(const float* X, const float* Y, const float* Z, int w, int h,
                                                   int s, float* X1, float* Y1, float* Z1) {
    RF_PIX
    float mX[9], mY[9], mZ[9];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            long q = median3_tap(ix + i - 1, iy + j - 1, w, h, s);
            mX[j * 3 + i] = q >= 0 ? X[q] : 0.0f;
            mY[j * 3 + i] = q >= 0 ? Y[q] : 0.0f;
            mZ[j * 3 + i] = q >= 0 ? Z[q] : 0.0f;
        }
    X1[pos] = median9(mX);
    Y1[pos] = median9(mY);
    Z1[pos] = median9(mZ);
}

// addKernel.cu:3-11
__global__ void __launch_bounds__(256) k_add(const float* a, const float* b, long count, float* out) {
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < count) out[i] = a[i] + b[i];
}

// solveReconDataL1.cu:4-102 (prior == false) and :139-378 (prior == true; mask may be null).
// WIDE: the blend in binary64 (the FP64 mode, recon_blend_wide / recon_blend_fn_wide).
template <bool PRIOR, bool WIDE>
__global__ void __launch_bounds__(256) k_recon_data_l1(const float* u, const float* v, const float* du0,
                                                        const float* dv0, const float* up, const float* vp,
                                                        const float* sku, const float* skv, const float* pu1,
                                                        const float* pu2, const float* pv1, const float* pv2,
                                                        const float* Ix, const float* Iy, const float* Iz, int w,
                                                        int h, int s, float lambda, float aTv, float aP,
                                                        const float* ufn, const float* vfn, float aFn0,
                                                        const float* mask, float* du1, float* dv1) {
    RF_PIX
    float A, bu, bv;
    if (PRIOR) {
        float aFn = mask ? aFn0 * mask[pos] : aFn0;
        A = (aTv + aP) + aFn;
        if (WIDE) {
            const double AAd = ((double)aTv + (double)aP) + (double)aFn;
            bu = recon_blend_fn_wide(aTv, aP, aFn, AAd, du0[pos], up[pos], u[pos], sku[pos], ufn[pos]);
            bv = recon_blend_fn_wide(aTv, aP, aFn, AAd, dv0[pos], vp[pos], v[pos], skv[pos], vfn[pos]);
        } else {
            bu = recon_blend_fn(aTv, aP, aFn, A, du0[pos], up[pos], u[pos], sku[pos], ufn[pos]);
            bv = recon_blend_fn(aTv, aP, aFn, A, dv0[pos], vp[pos], v[pos], skv[pos], vfn[pos]);
        }
    } else {
        A = aTv + aP;
        if (WIDE) {
            const double rA = 1.0 / ((double)aTv + (double)aP);
            bu = recon_blend_wide(aTv, aP, rA, du0[pos], up[pos], u[pos], sku[pos]);
            bv = recon_blend_wide(aTv, aP, rA, dv0[pos], vp[pos], v[pos], skv[pos]);
        } else {
            bu = recon_blend(aTv, aP, A, du0[pos], up[pos], u[pos], sku[pos]);
            bv = recon_blend(aTv, aP, A, dv0[pos], vp[pos], v[pos], skv[pos]);
        }
    }
    float du, dv;
    l1_threshold_div(Ix[pos], Iy[pos], Iz[pos], bu, bv, lambda, A, du, dv);
    bool hl = ix > 0, hu = iy > 0;
    float divu = div_p(pu1[pos], hl ? pu1[pos - 1] : 0.0f, pu2[pos], hu ? pu2[pos - s] : 0.0f, hl, hu);
    float divv = div_p(pv1[pos], hl ? pv1[pos - 1] : 0.0f, pv2[pos], hu ? pv2[pos - s] : 0.0f, hl, hu);
    du1[pos] = du + divu / A;
    dv1[pos] = dv + divv / A;
}

// solveUVProj.cu:4-22. WIDE: the blend in binary64 (the FP64 mode, uvproj_px_wide).
template <bool WIDE>
__global__ void __launch_bounds__(256) k_uvproj(const float* u, const float* v, const float* du, const float* dv,
                                                 const float* uc, const float* vc, const float* sku,
                                                 const float* skv, int w, int h, int s, float lf, float aP,
                                                 float* up, float* vp) {
    RF_PIX
    if (WIDE) {
        const double rP = 1.0 / ((double)lf + (double)aP);
        up[pos] = uvproj_px_wide(lf, aP, rP, uc[pos], u[pos], du[pos], sku[pos]);
        vp[pos] = uvproj_px_wide(lf, aP, rP, vc[pos], v[pos], dv[pos], skv[pos]);
    } else {
        up[pos] = uvproj_px(lf, aP, uc[pos], u[pos], du[pos], sku[pos]);
        vp[pos] = uvproj_px(lf, aP, vc[pos], v[pos], dv[pos], skv[pos]);
    }
}

// solve3dMinimalArea.cu:3-193. Reads column w and row h of X0 (see solve3d_px).
// R = double evaluates the triangulation in FP64 (inputs widened, results rounded once).
template <typename R>
__global__ void __launch_bounds__(256) k_solve3d(const float* up, const float* vp, const float* X0,
                                                  const float* Y0, const float* Z0, PQT<R> pqv, const PQT<R>* pqp,
                                                  float lf, float lms, int w, int h, int s, float* X, float* Y,
                                                  float* Z) {
    RF_PIX
    const PQT<R> pq = pqp ? *pqp : pqv;
    bool hl = ix > 0, hu = iy > 0;
    long pl = hl ? pos - 1 : pos, pu = hu ? pos - s : pos;
    float oX, oY, oZ;
    if constexpr (sizeof(R) == 8)  // the FP64 triangulation mode
        solve3d_px_f64(pq, (double)lf, (double)lms, ix, iy, up[pos], vp[pos], X0[pos], Y0[pos], Z0[pos],
                       X0[pos + 1], Y0[pos + 1], Z0[pos + 1], X0[pos + s], Y0[pos + s], Z0[pos + s], X0[pl], Y0[pl],
                       Z0[pl], X0[pu], Y0[pu], Z0[pu], hl, hu, oX, oY, oZ);
    else
        solve3d_px<R>(pq, (R)lf, (R)lms, ix, iy, up[pos], vp[pos], X0[pos], Y0[pos], Z0[pos], X0[pos + 1],
                      Y0[pos + 1], Z0[pos + 1], X0[pos + s], Y0[pos + s], Z0[pos + s], X0[pl], Y0[pl], Z0[pl], X0[pu],
                      Y0[pu], Z0[pu], hl, hu, oX, oY, oZ);
    X[pos] = oX;
    Y[pos] = oY;
    Z[pos] = oZ;
}

// solve3dMinimalArea.cu:248-297
__global__ void __launch_bounds__(256) k_cleanup3d(const float* X0, const float* Y0, const float* Z0,
                                                    const float* Xm, const float* Ym, const float* Zm, int w, int h,
                                                    int s, float* X1, float* Y1, float* Z1) {
    RF_PIX
    float a, b, c;
    cleanup_px(X0[pos], Y0[pos], Z0[pos], Xm[pos], Ym[pos], Zm[pos], a, b, c);
    X1[pos] = a;
    Y1[pos] = b;
    Z1[pos] = c;
}

// solveUcSk.cu:3-35
__global__ void __launch_bounds__(256) k_solve_uc(const float* X, const float* Y, const float* Z, UcParams cv,
                                                   const UcParams* cp, int w, int h, int s, float* uc, float* vc) {
    RF_PIX
    const UcParams c = cp ? *cp : cv;
    float a, b;
    uc_px(c, X[pos], Y[pos], Z[pos], ix, iy, a, b);
    uc[pos] = a;
    vc[pos] = b;
}

// solveUcSk.cu:80-97
__global__ void __launch_bounds__(256) k_update_sk(const float* sku, const float* skv, const float* u,
                                                    const float* v, const float* du, const float* dv,
                                                    const float* up, const float* vp, int w, int h, int s,
                                                    float* sku1, float* skv1) {
    RF_PIX
    sku1[pos] = sk_px(sku[pos], u[pos], du[pos], up[pos]);
    skv1[pos] = sk_px(skv[pos], v[pos], dv[pos], vp[pos]);
}

// flowToHSV.cu:194-302 (Middlebury colour wheel). The reference indexes its 55-entry tables
// with k0, k1 in [1, 55]; index 55 lies past the table there and wraps to 0 here.
__constant__ float c_wheel[3][55] = {
    {255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 213, 170, 128,
     85,  43,  0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   19,
     39,  58,  78,  98,  117, 137, 156, 176, 196, 215, 235, 255, 255, 255, 255, 255, 255},
    {0,   17,  34,  51,  68,  85,  102, 119, 136, 153, 170, 187, 204, 221, 238, 255, 255, 255, 255,
     255, 255, 255, 255, 255, 255, 255, 232, 209, 186, 163, 140, 116, 93,  70,  47,  24,  0,   0,
     0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0},
    {0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,
     0,   0,   0,   63,  127, 191, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255,
     255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 213, 170, 128, 85,  43}};

__global__ void __launch_bounds__(256) k_flow_to_hsv(const float* u, const float* v, int w, int h, int s,
                                                      float* rgb, float flowscale) {
    RF_PIX
    float du = u[pos] / flowscale, dv = v[pos] / flowscale;
    const int ncols = 55;
    float rad = sqrtf(du * du + dv * dv);
    float a = atan2f(-dv, -du) / 3.14159f;
    float fk = (a + 1.0f) / 2.0f * ((float)ncols - 1.0f) + 1.0f;
    int k0 = (int)floorf(fk);
    int k1 = k0 + 1;
    if (k1 == ncols + 1) k1 = 1;
    float f = fk - (float)k0;
    k0 = ((k0 % ncols) + ncols) % ncols;
    k1 = ((k1 % ncols) + ncols) % ncols;
    float col[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float cc = (1.0f - f) * (c_wheel[c][k0] / 255.0f) + f * (c_wheel[c][k1] / 255.0f);
        col[c] = rad <= 1.0f ? 1.0f - rad * (1.0f - cc) : cc * 0.75f;
    }
    rgb[3 * pos + 0] = col[2];  // B
    rgb[3 * pos + 1] = col[1];  // G
    rgb[3 * pos + 2] = col[0];  // R
}

// ComputeDerivKernel (computeDeriv.cu:6-40): 5-point gradient of one image, mirror border.
__global__ void __launch_bounds__(256) k_deriv_single(const float* I, int w, int h, int s, float* Ix, float* Iy) {
    RF_PIX
    const long r = (long)iy * s;
    Ix[pos] = d5(I[r + refl(ix - 2, w)], I[r + refl(ix - 1, w)], I[r + refl(ix + 1, w)], I[r + refl(ix + 2, w)]);
    Iy[pos] = d5(I[ix + (long)refl(iy - 2, h) * s], I[ix + (long)refl(iy - 1, h) * s],
                 I[ix + (long)refl(iy + 1, h) * s], I[ix + (long)refl(iy + 2, h) * s]);
}

// ComputeDerivativesPatchKernel (computeDerivativesPatch.cu:8-67): the derivatives of
// ComputeDerivatives at every tap of the 5x5 patch (each tap's stencil mirror-indexed on its
// own), summed in row-major tap order.
__global__ void __launch_bounds__(256) k_derivatives_patch(const float* I0, const float* I1, int w, int h, int s,
                                                            float* Ix, float* Iy, float* Iz) {
    RF_PIX
    auto at = [&](const float* I, int x, int y) { return I[refl(x, w) + (long)refl(y, h) * s]; };
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
    for (int j = 0; j < 5; ++j)
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const int col = ix + i - 2, row = iy + j - 2;
            float t0 = d5(at(I0, col - 2, row), at(I0, col - 1, row), at(I0, col + 1, row), at(I0, col + 2, row));
            float t1 = d5(at(I1, col - 2, row), at(I1, col - 1, row), at(I1, col + 1, row), at(I1, col + 2, row));
            sx += (t0 + t1) * 0.5f;
            sz += at(I1, col, row) - at(I0, col, row);
            t0 = d5(at(I0, col, row - 2), at(I0, col, row - 1), at(I0, col, row + 1), at(I0, col, row + 2));
            t1 = d5(at(I1, col, row - 2), at(I1, col, row - 1), at(I1, col, row + 1), at(I1, col, row + 2));
            sy += (t0 + t1) * 0.5f;
        }
    Ix[pos] = sx;
    Iy[pos] = sy;
    Iz[pos] = sz;
}

// SolveDataCharbForTVKernel / SolveDataCharbForTVGradKernel (solveDataCharbForTV.cu:3-85,
// solveDataCharbForTVGrad.cu:3-103): reweighted solve, then the primal step with the
// divergence of SolveDataL1's border rule.
template <bool GRAD>
__global__ void __launch_bounds__(256) k_data_charb(const float* du0, const float* dv0, const float* duhat0,
                                                     const float* dvhat0, const float* pu1, const float* pu2,
                                                     const float* pv1, const float* pv2, const float* Ix,
                                                     const float* Iy, const float* It, const float* Ixx,
                                                     const float* Ixy, const float* Ixt, const float* Iyx,
                                                     const float* Iyy, const float* Iyt, int w, int h, int s,
                                                     float lambda, float lambdagrad, float theta, float* du1,
                                                     float* dv1, float* duhat1, float* dvhat1) {
    RF_PIX
    float du, dv;
    if (GRAD)
        charb_px<true>(du0[pos], dv0[pos], Ix[pos], Iy[pos], It[pos], Ixx[pos], Ixy[pos], Ixt[pos], Iyx[pos],
                       Iyy[pos], Iyt[pos], duhat0[pos], dvhat0[pos], lambda, lambdagrad, theta, du, dv);
    else
        charb_px<false>(du0[pos], dv0[pos], Ix[pos], Iy[pos], It[pos], 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f,
                        duhat0[pos], dvhat0[pos], lambda, lambdagrad, theta, du, dv);
    du1[pos] = du;
    dv1[pos] = dv;
    bool hl = ix > 0, hu = iy > 0;
    float divu = div_p(pu1[pos], hl ? pu1[pos - 1] : 0.0f, pu2[pos], hu ? pu2[pos - s] : 0.0f, hl, hu);
    float divv = div_p(pv1[pos], hl ? pv1[pos - 1] : 0.0f, pv2[pos], hu ? pv2[pos - s] : 0.0f, hl, hu);
    duhat1[pos] = du + theta * divu;
    dvhat1[pos] = dv + theta * divv;
}

// SolveDataL1InpaintKernel (solveDataL1Inpaint.cu:4-98): the L1 step, kept at uhat where
// mask0 == 0.
__global__ void __launch_bounds__(256) k_data_l1_inpaint(const float* du0, const float* dv0, const float* mask0,
                                                          const float* pu1, const float* pu2, const float* pv1,
                                                          const float* pv2, const float* Ix, const float* Iy,
                                                          const float* Iz, int w, int h, int s, float lambda,
                                                          float theta, float* du1, float* dv1) {
    RF_PIX
    float du, dv;
    const float duhat = du0[pos], dvhat = dv0[pos];
    l1_threshold(Ix[pos], Iy[pos], Iz[pos], duhat, dvhat, lambda, theta, du, dv);
    if (mask0[pos] == 0.0f) {
        du = duhat;
        dv = dvhat;
    }
    bool hl = ix > 0, hu = iy > 0;
    float divu = div_p(pu1[pos], hl ? pu1[pos - 1] : 0.0f, pu2[pos], hu ? pu2[pos - s] : 0.0f, hl, hu);
    float divv = div_p(pv1[pos], hl ? pv1[pos - 1] : 0.0f, pv2[pos], hu ? pv2[pos - s] : 0.0f, hl, hu);
    du1[pos] = du + theta * divu;
    dv1[pos] = dv + theta * divv;
}

// Cv16uTo32fKernel (rgbToGray.cu:51-64): depth in sensor units, unscaled.
__global__ void __launch_bounds__(256) k_cv16u_to_32f(const uint16_t* in, float* out, int w, int h, int s) {
    RF_PIX
    out[pos] = (float)in[pos];
}

// SolveSceneFlowAndPointCloudKernel (solveSceneFlow.cu:49-91): depth at the pixel and, bilinear
// with mirror border, at the flow target; both back-projected; scene flow zeroed outside
// 500 <= d0 <= 2500. Outputs are (x, y, z) triples per pixel; pc0 / pc1 may be null.
__global__ void __launch_bounds__(256) k_scene_flow(const float* u, const float* v, const float* depth0,
                                                     const float* depth1, float fx, float fy, float cx, float cy,
                                                     int w, int h, int s, float* pc0, float* pc1, float* sf) {
    RF_PIX
    const float d0 = depth0[pos];
    const float d1 = warp_sample(depth1, w, h, s, ix, iy, u[pos], v[pos]);
    const float p0z = d0, p0x = (((float)ix - cx) * p0z) / fx, p0y = (((float)iy - cy) * p0z) / fy;
    const float p1z = d1, p1x = (((float)ix - cx) * p1z) / fx, p1y = (((float)iy - cy) * p1z) / fy;
    float sx = p0x - p1x, sy = p0y - p1y, sz = p0z - p1z;
    if ((d0 < 500.0f) || (d0 > 2500.0f)) sx = sy = sz = 0.0f;
    if (pc0) {
        pc0[3 * pos] = p0x;
        pc0[3 * pos + 1] = p0y;
        pc0[3 * pos + 2] = p0z;
    }
    if (pc1) {
        pc1[3 * pos] = p1x;
        pc1[3 * pos + 1] = p1y;
        pc1[3 * pos + 2] = p1z;
    }
    sf[3 * pos] = sx;
    sf[3 * pos + 1] = sy;
    sf[3 * pos + 2] = sz;
}

// SceneFlowToHSVKernel (flowToHSV.cu:132-190): hue from the (x, y) direction, brightness
// offset by z; the sector bounds are double literals in the reference, compared in double.
__global__ void __launch_bounds__(256) k_scene_flow_to_hsv(const float* sf, int w, int h, int s, float* rgb,
                                                            float flowscale) {
    RF_PIX
    const float um = sf[3 * pos] / flowscale, vm = sf[3 * pos + 1] / flowscale, wm = sf[3 * pos + 2] / flowscale;
    const float H = atan2f(vm, um) + 3.1416f;
    const float C = sqrtf(um * um + vm * vm);
    const float Hp = H / 1.0472f;
    const float X = C * (1.0f - fabsf(fmodf(Hp, 2.0f) - 1.0f));
    const float m = wm - C;
    const double Hd = H;
    float r, g, b;
    if (Hd < 6.2832 && Hd >= 5.236) r = C, g = 0.0f, b = X;
    else if (Hd < 5.236 && Hd >= 4.1888) r = X, g = 0.0f, b = C;
    else if (Hd < 4.1888 && Hd >= 3.1416) r = 0.0f, g = X, b = C;
    else if (Hd < 3.1416 && Hd >= 2.0944) r = 0.0f, g = C, b = X;
    else if (Hd < 2.0944 && Hd >= 1.0472) r = X, g = C, b = 0.0f;
    else r = C, g = X, b = 0.0f;
    rgb[3 * pos] = r + m;
    rgb[3 * pos + 1] = g + m;
    rgb[3 * pos + 2] = b + m;
}

// solve3dMinimalArea.cu:320-378. The forward difference at ix + range == w (or row h) reads
// the padding column (or the next row / guard row), as in the reference.
__global__ void __launch_bounds__(256) k_remove_edge(const float* X0, const float* Y0, const float* Z0, int range,
                                                      float maxarea, int w, int h, int s, float* X1, float* Y1,
                                                      float* Z1) {
    RF_PIX
    float X = X0[pos], Y = Y0[pos], Z = Z0[pos];
    float Xx, Yx, Zx, Xy, Yy, Zy;
    if (ix + range > w) {
        long q = pos - range;
        Xx = -X0[q] + X;
        Yx = -Y0[q] + Y;
        Zx = -Z0[q] + Z;
    } else {
        long q = pos + range;
        Xx = X0[q] - X;
        Yx = Y0[q] - Y;
        Zx = Z0[q] - Z;
    }
    if (iy + range > h) {
        long q = pos - (long)range * s;
        Xy = -X0[q] + X;
        Yy = -Y0[q] + Y;
        Zy = -Z0[q] + Z;
    } else {
        long q = pos + (long)range * s;
        Xy = X0[q] - X;
        Yy = Y0[q] - Y;
        Zy = Z0[q] - Z;
    }
    float c1 = Yx * Zy - Zx * Yy, c2 = Zx * Xy - Xx * Zy, c3 = Xx * Yy - Yx * Xy;
    float area = sqrtf((c1 * c1 + c2 * c2) + c3 * c3);
    bool drop = area > maxarea;
    X1[pos] = drop ? 0.0f : X;
    Y1[pos] = drop ? 0.0f : Y;
    Z1[pos] = drop ? 0.0f : Z;
}

}  // namespace

PQ make_pq(const double* P, const double* Q) {
    PQ m;
    for (int i = 0; i < 12; ++i) {
        m.P[i] = (float)P[i];
        m.Q[i] = (float)Q[i];
    }
    return m;
}

PQd make_pqd(const double* P, const double* Q) {
    PQd m;
    for (int i = 0; i < 12; ++i) {
        m.P[i] = P[i];
        m.Q[i] = Q[i];
    }
    return m;
}

UcParams make_uc(const double* R, const double* t, const double* K1) {
    UcParams c;
    c.R11 = (float)R[0];
    c.R12 = (float)R[1];
    c.R13 = (float)R[2];
    c.R21 = (float)R[3];
    c.R22 = (float)R[4];
    c.R23 = (float)R[5];
    c.R31 = (float)R[6];
    c.R32 = (float)R[7];
    c.R33 = (float)R[8];
    c.t1 = (float)t[0];
    c.t2 = (float)t[1];
    c.t3 = (float)t[2];
    c.f = (float)K1[0];
    c.cx = (float)K1[2];
    c.cy = (float)K1[5];
    return c;
}

#define RF_L2(K, ...)                                                       \
    do {                                                                    \
        if (w <= 0 || h <= 0) return hipSuccess;                            \
        hipLaunchKernelGGL(K, grid2(w, h), dim3(BX, BY), 0, st, __VA_ARGS__); \
        return hipGetLastError();                                           \
    } while (0)

hipError_t launch_warm_start(const float* wu, const float* wv, int W, int H, int S, int w, int h, int s, int* bad,
                             float* u, float* v, hipStream_t st) {
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, bad, 0);
    if (W > 0 && H > 0)
        hipLaunchKernelGGL(k_ws_check, grid2(W, H), dim3(BX, BY), 0, st, wu, wv, W, H, S, (float)std::max(W, H), bad);
    const float scale = (float)w / (float)W;
    RF_L2(k_ws_apply, wu, wv, W, H, S, w, h, s, scale, bad, u, v);
}

hipError_t launch_fill_zero(float* const* bufs, int nbuf, long n, hipStream_t st) {
    FillArgs a{};
    if (nbuf > 16 || nbuf < 0) return hipErrorInvalidValue;
    for (int i = 0; i < nbuf; ++i) a.bufs[i] = bufs[i];
    long blocks = (n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, st, a, nbuf, n);
    return hipGetLastError();
}

hipError_t launch_gray_rgb(const uint8_t* rgb, float* out, int w, int h, int s, hipStream_t st) {
    RF_L2(k_gray_rgb, rgb, out, w, h, s);
}
hipError_t launch_gray_u8(const uint8_t* in, float* out, int w, int h, int s, hipStream_t st) {
    RF_L2(k_gray_u8, in, out, w, h, s);
}
hipError_t launch_gray_u16(const uint16_t* in, float* out, int w, int h, int s, hipStream_t st) {
    RF_L2(k_gray_u16, in, out, w, h, s);
}
hipError_t launch_gray_f32(const float* in, float* out, int w, int h, int s, hipStream_t st) {
    RF_L2(k_gray_f32, in, out, w, h, s);
}
hipError_t launch_downscale(const float* src, int W, int H, int S, int w, int h, int s, bool scaled, float scale,
                            float* out, hipStream_t st) {
    RF_L2(k_downscale, src, W, H, S, w, h, s, scaled ? 1 : 0, scale, out);
}
hipError_t launch_upscale(const float* src, int cw, int ch, int cs, int w, int h, int s, float scale, float* out,
                          hipStream_t st) {
    RF_L2(k_upscale, src, cw, ch, cs, w, h, s, scale, out);
}
hipError_t launch_warp(const float* src, int w, int h, int s, const float* u, const float* v, float* out,
                       hipStream_t st) {
    RF_L2(k_warp, src, w, h, s, u, v, out);
}
hipError_t launch_derivatives(const float* I0, const float* I1, int w, int h, int s, float* Ix, float* Iy,
                              float* Iz, hipStream_t st) {
    RF_L2(k_derivatives, I0, I1, w, h, s, Ix, Iy, Iz);
}
hipError_t launch_data_l1(const float* du, const float* dv, const float* pu1, const float* pu2, const float* pv1,
                          const float* pv2, const float* Ix, const float* Iy, const float* Iz, int w, int h, int s,
                          float lambda, float theta, float* du1, float* dv1, hipStream_t st) {
    RF_L2(k_data_l1, du, dv, pu1, pu2, pv1, pv2, Ix, Iy, Iz, w, h, s, lambda, theta, du1, dv1);
}
hipError_t launch_dual_tv(const float* du, const float* dv, const float* pu1, const float* pu2, const float* pv1,
                          const float* pv2, int w, int h, int s, float tau, float theta, float* pu1s, float* pu2s,
                          float* pv1s, float* pv2s, hipStream_t st) {
    RF_L2(k_dual_tv, du, dv, pu1, pu2, pv1, pv2, w, h, s, tau, theta, pu1s, pu2s, pv1s, pv2s);
}
hipError_t launch_median5(const float* u, const float* v, int w, int h, int s, float* ou, float* ov,
                          hipStream_t st) {
    RF_L2(k_median5, u, v, w, h, s, ou, ov);
}
hipError_t launch_median3d(const float* X, const float* Y, const float* Z, int w, int h, int s, float* X1,
                           float* Y1, float* Z1, hipStream_t st) {
    RF_L2(k_median3d, X, Y, Z, w, h, s, X1, Y1, Z1);
}
__global__ void __launch_bounds__(256) k_udiv(const float* n, long count, UDiv d, float* q) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i < count) q[i] = udiv(n[i], d);
}

hipError_t launch_udiv(const float* n, long count, float d, float* q, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_udiv, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, n, count, make_udiv(d), q);
    return hipGetLastError();
}

hipError_t launch_add(const float* a, const float* b, long count, float* out, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_add, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, a, b, count, out);
    return hipGetLastError();
}
hipError_t launch_recon_data_l1(const float* u, const float* v, const float* du, const float* dv, const float* up,
                                const float* vp, const float* sku, const float* skv, const float* pu1,
                                const float* pu2, const float* pv1, const float* pv2, const float* Ix,
                                const float* Iy, const float* Iz, int w, int h, int s, float lambda, float aTv,
                                float aP, const float* ufn, const float* vfn, float aFn, const float* mask,
                                float* du1, float* dv1, bool wide, hipStream_t st) {
    if (ufn && vfn) {
        if (wide)
            RF_L2((k_recon_data_l1<true, true>), u, v, du, dv, up, vp, sku, skv, pu1, pu2, pv1, pv2, Ix, Iy, Iz, w, h,
                  s, lambda, aTv, aP, ufn, vfn, aFn, mask, du1, dv1);
        RF_L2((k_recon_data_l1<true, false>), u, v, du, dv, up, vp, sku, skv, pu1, pu2, pv1, pv2, Ix, Iy, Iz, w, h, s,
              lambda, aTv, aP, ufn, vfn, aFn, mask, du1, dv1);
    }
    if (wide)
        RF_L2((k_recon_data_l1<false, true>), u, v, du, dv, up, vp, sku, skv, pu1, pu2, pv1, pv2, Ix, Iy, Iz, w, h, s,
              lambda, aTv, aP, ufn, vfn, aFn, mask, du1, dv1);
    RF_L2((k_recon_data_l1<false, false>), u, v, du, dv, up, vp, sku, skv, pu1, pu2, pv1, pv2, Ix, Iy, Iz, w, h, s,
          lambda, aTv, aP, ufn, vfn, aFn, mask, du1, dv1);
}
hipError_t launch_uvproj(const float* u, const float* v, const float* du, const float* dv, const float* uc,
                         const float* vc, const float* sku, const float* skv, int w, int h, int s, float lf,
                         float aP, float* up, float* vp, bool wide, hipStream_t st) {
    if (wide) RF_L2(k_uvproj<true>, u, v, du, dv, uc, vc, sku, skv, w, h, s, lf, aP, up, vp);
    RF_L2(k_uvproj<false>, u, v, du, dv, uc, vc, sku, skv, w, h, s, lf, aP, up, vp);
}
hipError_t launch_solve3d(const float* up, const float* vp, const float* X0, const float* Y0, const float* Z0,
                          const PQ& pq, const PQ* dpq, float lf, float lms, int w, int h, int s, float* X, float* Y,
                          float* Z, hipStream_t st) {
    RF_L2(k_solve3d<float>, up, vp, X0, Y0, Z0, pq, dpq, lf, lms, w, h, s, X, Y, Z);
}
hipError_t launch_solve3d_f64(const float* up, const float* vp, const float* X0, const float* Y0, const float* Z0,
                              const PQd& pq, const PQd* dpq, float lf, float lms, int w, int h, int s, float* X,
                              float* Y, float* Z, hipStream_t st) {
    RF_L2(k_solve3d<double>, up, vp, X0, Y0, Z0, pq, dpq, lf, lms, w, h, s, X, Y, Z);
}
hipError_t launch_cleanup3d(const float* X0, const float* Y0, const float* Z0, const float* Xm, const float* Ym,
                            const float* Zm, int w, int h, int s, float* X1, float* Y1, float* Z1, hipStream_t st) {
    RF_L2(k_cleanup3d, X0, Y0, Z0, Xm, Ym, Zm, w, h, s, X1, Y1, Z1);
}
hipError_t launch_solve_uc(const float* X, const float* Y, const float* Z, const UcParams& c, const UcParams* dc,
                           int w, int h, int s, float* uc, float* vc, hipStream_t st) {
    RF_L2(k_solve_uc, X, Y, Z, c, dc, w, h, s, uc, vc);
}
hipError_t launch_update_sk(const float* sku, const float* skv, const float* u, const float* v, const float* du,
                            const float* dv, const float* up, const float* vp, int w, int h, int s, float* sku1,
                            float* skv1, hipStream_t st) {
    RF_L2(k_update_sk, sku, skv, u, v, du, dv, up, vp, w, h, s, sku1, skv1);
}
hipError_t launch_flow_to_hsv(const float* u, const float* v, int w, int h, int s, float* uvrgb, float flowscale,
                              hipStream_t st) {
    RF_L2(k_flow_to_hsv, u, v, w, h, s, uvrgb, flowscale);
}
hipError_t launch_deriv_single(const float* I, int w, int h, int s, float* Ix, float* Iy, hipStream_t st) {
    hipLaunchKernelGGL(k_deriv_single, grid2(w, h), dim3(BX, BY), 0, st, I, w, h, s, Ix, Iy);
    return hipGetLastError();
}
hipError_t launch_derivatives_patch(const float* I0, const float* I1, int w, int h, int s, float* Ix, float* Iy,
                                    float* Iz, hipStream_t st) {
    hipLaunchKernelGGL(k_derivatives_patch, grid2(w, h), dim3(BX, BY), 0, st, I0, I1, w, h, s, Ix, Iy, Iz);
    return hipGetLastError();
}
hipError_t launch_data_charb(const float* du0, const float* dv0, const float* duhat0, const float* dvhat0,
                             const float* pu1, const float* pu2, const float* pv1, const float* pv2, const float* Ix,
                             const float* Iy, const float* It, const float* Ixx, const float* Ixy, const float* Ixt,
                             const float* Iyx, const float* Iyy, const float* Iyt, int w, int h, int s, float lambda,
                             float lambdagrad, float theta, float* du1, float* dv1, float* duhat1, float* dvhat1,
                             hipStream_t st) {
    if (Ixx)
        hipLaunchKernelGGL(k_data_charb<true>, grid2(w, h), dim3(BX, BY), 0, st, du0, dv0, duhat0, dvhat0, pu1, pu2,
                           pv1, pv2, Ix, Iy, It, Ixx, Ixy, Ixt, Iyx, Iyy, Iyt, w, h, s, lambda, lambdagrad, theta,
                           du1, dv1, duhat1, dvhat1);
    else
        hipLaunchKernelGGL(k_data_charb<false>, grid2(w, h), dim3(BX, BY), 0, st, du0, dv0, duhat0, dvhat0, pu1,
                           pu2, pv1, pv2, Ix, Iy, It, Ixx, Ixy, Ixt, Iyx, Iyy, Iyt, w, h, s, lambda, lambdagrad,
                           theta, du1, dv1, duhat1, dvhat1);
    return hipGetLastError();
}
hipError_t launch_data_l1_inpaint(const float* du0, const float* dv0, const float* mask0, const float* pu1,
                                  const float* pu2, const float* pv1, const float* pv2, const float* Ix,
                                  const float* Iy, const float* Iz, int w, int h, int s, float lambda, float theta,
                                  float* du1, float* dv1, hipStream_t st) {
    hipLaunchKernelGGL(k_data_l1_inpaint, grid2(w, h), dim3(BX, BY), 0, st, du0, dv0, mask0, pu1, pu2, pv1, pv2, Ix,
                       Iy, Iz, w, h, s, lambda, theta, du1, dv1);
    return hipGetLastError();
}
hipError_t launch_cv16u_to_32f(const uint16_t* in, float* out, int w, int h, int s, hipStream_t st) {
    hipLaunchKernelGGL(k_cv16u_to_32f, grid2(w, h), dim3(BX, BY), 0, st, in, out, w, h, s);
    return hipGetLastError();
}
hipError_t launch_scene_flow(const float* u, const float* v, const float* depth0, const float* depth1, float fx,
                             float fy, float cx, float cy, int w, int h, int s, float* pc0, float* pc1, float* sf,
                             hipStream_t st) {
    hipLaunchKernelGGL(k_scene_flow, grid2(w, h), dim3(BX, BY), 0, st, u, v, depth0, depth1, fx, fy, cx, cy, w, h, s,
                       pc0, pc1, sf);
    return hipGetLastError();
}
hipError_t launch_scene_flow_to_hsv(const float* sf, int w, int h, int s, float* rgb, float flowscale,
                                    hipStream_t st) {
    hipLaunchKernelGGL(k_scene_flow_to_hsv, grid2(w, h), dim3(BX, BY), 0, st, sf, w, h, s, rgb, flowscale);
    return hipGetLastError();
}
hipError_t launch_remove_edge(const float* X0, const float* Y0, const float* Z0, int range, float maxarea, int w,
                              int h, int s, float* X1, float* Y1, float* Z1, hipStream_t st) {
    RF_L2(k_remove_edge, X0, Y0, Z0, range, maxarea, w, h, s, X1, Y1, Z1);
}

}  // namespace rf
