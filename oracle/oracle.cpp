// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the Reconflow coarse-to-fine TV-L1 optical-flow path and its
// epipolar-constrained joint flow/depth path, written from the text of the reference
// kernels. Every routine is a template over the arithmetic type T:
//   * T = double  -> the FP64 oracle the GPU path is measured against (tolerances);
//   * T = float   -> an FP32 restatement with the reference's exact operation order,
//                    built with -ffp-contract=off, used to show bit-level agreement.
// Only tests, the smoke check and the bench's CPU-baseline leg load this library.
// Nothing in the product links or calls it.
//
// Semantics fixed here where the reference leaves them to hardware or to undefined
// behaviour (each also stated where it is used):
//   * texture sampling S(I, tx, ty): bilinear with exact weights, texel-centre
//     coordinates, edge-duplicating mirror addressing over the bound w x h region;
//   * fresh device buffers are zero; one zero guard row follows every level-0-sized
//     buffer (it stands for the read past the allocation in Solve3dMinimalArea);
//   * a NaN rho in the L1 data steps gives NaN du/dv (the reference leaves them
//     unassigned);
//   * a warp coordinate that is not finite or exceeds 2^24 in magnitude samples NaN.
//
// All file:line citations are relative to the reference repository root.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#include <omp.h>

namespace orc {

// ---------------------------------------------------------------------------------
// Texture emulation. Mirror addressing on normalized coordinates with linear
// filtering (e.g. warpImage.cu:40-47): texel index i maps into [0, n) with period 2n,
// duplicating the edge texel (-1 -> 0, n -> n-1).
static inline long refl(long i, long n) {
    long p = 2 * n;
    long m = i % p;
    if (m < 0) m += p;
    return m < n ? m : p - 1 - m;
}

static inline long floor_div(long a, long b) {  // b > 0
    long q = a / b;
    if ((a % b != 0) && (a < 0)) --q;
    return q;
}

// Bilinear sample at texel coordinate (x0 + fx, y0 + fy), weights as in the public
// CUDA texture-fetch formula (1-a)(1-b)T00 + a(1-b)T10 + (1-a)bT01 + abT11.
template <typename T>
static inline T bilin(const T* img, long w, long h, long s, long x0, T fx, long y0, T fy) {
    long xa = refl(x0, w), xb = refl(x0 + 1, w);
    long ya = refl(y0, h), yb = refl(y0 + 1, h);
    T one = T(1);
    T w00 = (one - fx) * (one - fy);
    T w10 = fx * (one - fy);
    T w01 = (one - fx) * fy;
    T w11 = fx * fy;
    T r = w00 * img[xa + ya * s];
    r = r + w10 * img[xb + ya * s];
    r = r + w01 * img[xa + yb * s];
    r = r + w11 * img[xb + yb * s];
    return r;
}

// Sample at the rational texel coordinate (nx/dx, ny/dy) with integer numerators.
template <typename T>
static inline T sample_ratio(const T* img, long w, long h, long s, long nx, long dx, long ny, long dy) {
    long x0 = floor_div(nx, dx), y0 = floor_div(ny, dy);
    T fx = T(nx - x0 * dx) / T(dx);
    T fy = T(ny - y0 * dy) / T(dy);
    return bilin(img, w, h, s, x0, fx, y0, fy);
}

// ---------------------------------------------------------------------------------
// A2: rgbToGray (rgbToGray.cu:3-19): g = ((float)r + (float)g + (float)b)/3; g /= 256.
template <typename T>
void gray_rgb(const uint8_t* rgb, T* out, int w, int h, int s) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            T g = ((T)rgb[3 * p] + (T)rgb[3 * p + 1] + (T)rgb[3 * p + 2]) / T(3);
            out[p] = g / T(256);
        }
}

// A2: Cv32fToGray (rgbToGray.cu:66-77): copy of the valid region.
template <typename T>
void gray_f32(const float* in, T* out, int w, int h, int s) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) out[ix + (long)iy * s] = (T)in[ix + (long)iy * s];
}

// f2: Cv8uToGray / Cv16uToGray (rgbToGray.cu:21-49).
template <typename T>
void gray_u8(const uint8_t* in, T* out, int w, int h, int s) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) out[ix + (long)iy * s] = (T)in[ix + (long)iy * s] / T(256);
}
template <typename T>
void gray_u16(const uint16_t* in, T* out, int w, int h, int s) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) out[ix + (long)iy * s] = (T)in[ix + (long)iy * s] / T(65536);
}

// A3: DownscaleKernel (downscale.cu:6-26). Coarse pixel centre in fine texels is
// (ix+1/2)W/w - 1/2; the four taps sit a quarter coarse pixel to each side.
// scale != 1 gives DownscaleScalingKernel (downscale.cu:28-48).
template <typename T>
void downscale(const T* src, int W, int H, int S, int w, int h, int s, T scale, bool scaled, T* out) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long dxn = 4L * w, dyn = 4L * h;
            long xc = (4L * ix + 2) * W - 2L * w;
            long yc = (4L * iy + 2) * H - 2L * h;
            T a = sample_ratio(src, W, H, S, xc - W, dxn, yc, dyn);
            T b = sample_ratio(src, W, H, S, xc + W, dxn, yc, dyn);
            T c = sample_ratio(src, W, H, S, xc, dxn, yc - H, dyn);
            T d = sample_ratio(src, W, H, S, xc, dxn, yc + H, dyn);
            T sum = ((a + b) + c) + d;
            out[ix + (long)iy * s] = scaled ? (scale * T(0.25)) * sum : T(0.25) * sum;
        }
}

// A10: UpscaleKernel (upscale.cu:14-27): fine pixel centre in coarse texels is
// (ix+1/2)w/W - 1/2; result times scale.
template <typename T>
void upscale(const T* src, int w, int h, int s, int W, int H, int S, T scale, T* out) {
#pragma omp parallel for
    for (int iy = 0; iy < H; ++iy)
        for (int ix = 0; ix < W; ++ix) {
            T v = sample_ratio(src, w, h, s, (2L * ix + 1) * w - W, 2L * W, (2L * iy + 1) * h - H, 2L * H);
            out[ix + (long)iy * S] = v * scale;
        }
}

// A4: WarpingKernel (warpImage.cu:15-29): out = S(I1, ix + u, iy + v).
template <typename T>
void warp(const T* src, int w, int h, int s, const T* u, const T* v, T* out) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            T uu = u[p], vv = v[p];
            if (!(std::fabs(uu) <= T(16777216)) || !(std::fabs(vv) <= T(16777216))) {
                out[p] = std::numeric_limits<T>::quiet_NaN();
                continue;
            }
            T fu = std::floor(uu), fv = std::floor(vv);
            out[p] = bilin(src, w, h, s, ix + (long)fu, uu - fu, iy + (long)fv, vv - fv);
        }
}

// A5: ComputeDerivativesKernel (computeDerivatives.cu:8-57). Taps sit on exact texel
// centres, so each is a mirror-indexed read; operation order as at :26-56.
template <typename T>
void derivatives(const T* I0, const T* I1, int w, int h, int s, T* Ix, T* Iy, T* Iz) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            auto at = [&](const T* I, long x, long y) { return I[refl(x, w) + refl(y, h) * s]; };
            T t0, t1;
            t0 = at(I0, ix - 2, iy);
            t0 -= at(I0, ix - 1, iy) * T(8);
            t0 += at(I0, ix + 1, iy) * T(8);
            t0 -= at(I0, ix + 2, iy);
            t0 /= T(12);
            t1 = at(I1, ix - 2, iy);
            t1 -= at(I1, ix - 1, iy) * T(8);
            t1 += at(I1, ix + 1, iy) * T(8);
            t1 -= at(I1, ix + 2, iy);
            t1 /= T(12);
            Ix[p] = (t0 + t1) * T(0.5);
            Iz[p] = at(I1, ix, iy) - at(I0, ix, iy);
            t0 = at(I0, ix, iy - 2);
            t0 -= at(I0, ix, iy - 1) * T(8);
            t0 += at(I0, ix, iy + 1) * T(8);
            t0 -= at(I0, ix, iy + 2);
            t0 /= T(12);
            t1 = at(I1, ix, iy - 2);
            t1 -= at(I1, ix, iy - 1) * T(8);
            t1 += at(I1, ix, iy + 1) * T(8);
            t1 -= at(I1, ix, iy + 2);
            t1 /= T(12);
            Iy[p] = (t0 + t1) * T(0.5);
        }
}

// Divergence of the dual field with the reference's left/top border rule
// (solveDataL1.cu:64-88): backward differences, missing neighbours dropped.
template <typename T>
static inline T div_p(const T* p1, const T* p2, int ix, int iy, int s) {
    long p = ix + (long)iy * s;
    if (ix - 1 < 0) {
        if (iy - 1 < 0) return p1[p] + p2[p];
        return (p1[p] + p2[p]) - p2[p - s];
    }
    if (iy - 1 < 0) return (p1[p] - p1[p - 1]) + p2[p];
    return ((p1[p] - p1[p - 1]) + p2[p]) - p2[p - s];
}

// A6: SolveDataL1Kernel (solveDataL1.cu:4-91): pointwise L1 threshold + primal step.
template <typename T>
void data_l1(const T* duhat0, const T* dvhat0, const T* pu1, const T* pu2, const T* pv1, const T* pv2,
             const T* Ix, const T* Iy, const T* It, int w, int h, int s, T lambda, T theta, T* duhat1,
             T* dvhat1) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            T dix = Ix[p], diy = Iy[p], dit = It[p];
            T duhat = duhat0[p], dvhat = dvhat0[p];
            T rho = (dix * duhat + diy * dvhat) + dit;
            T lt = lambda * theta;
            T g2 = dix * dix + diy * diy;
            T upper = lt * g2;
            T lower = ((-lambda) * theta) * g2;
            T du = std::numeric_limits<T>::quiet_NaN(), dv = du;
            if ((rho <= upper) && (rho >= lower)) {
                T magi = dix * dix + diy * diy;
                if (magi != T(0)) {
                    du = duhat - (rho * dix) / magi;
                    dv = dvhat - (rho * diy) / magi;
                } else {
                    du = duhat;
                    dv = dvhat;
                }
            } else if (rho < lower) {
                du = duhat + lt * dix;
                dv = dvhat + lt * diy;
            } else if (rho > upper) {
                du = duhat - lt * dix;
                dv = dvhat - lt * diy;
            }
            duhat1[p] = du + theta * div_p(pu1, pu2, ix, iy, s);
            dvhat1[p] = dv + theta * div_p(pv1, pv2, ix, iy, s);
        }
}

// A7: SolveSmoothDualTVGlobalKernel (solveSmoothDualTVGlobal.cu:4-65).
template <typename T>
void dual_tv(const T* duhat, const T* dvhat, const T* pu1, const T* pu2, const T* pv1, const T* pv2, int w,
             int h, int s, T tau, T theta, T* pu1s, T* pu2s, T* pv1s, T* pv2s) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            T dux, dvx, duy, dvy;
            if (ix + 1 >= w) {
                dux = 0;
                dvx = 0;
            } else {
                dux = duhat[p + 1] - duhat[p];
                dvx = dvhat[p + 1] - dvhat[p];
            }
            if (iy + 1 >= h) {
                duy = 0;
                dvy = 0;
            } else {
                duy = duhat[p + s] - duhat[p];
                dvy = dvhat[p + s] - dvhat[p];
            }
            T magdu = std::sqrt(dux * dux + duy * duy);
            T magdv = std::sqrt(dvx * dvx + dvy * dvy);
            T fac = tau / theta;
            pu1s[p] = (pu1[p] + fac * dux) / (T(1) + fac * magdu);
            pu2s[p] = (pu2[p] + fac * duy) / (T(1) + fac * magdu);
            pv1s[p] = (pv1[p] + fac * dvx) / (T(1) + fac * magdv);
            pv2s[p] = (pv2[p] + fac * dvy) / (T(1) + fac * magdv);
        }
}

// Partial exchange sort of the reference median kernels: for j < nsel, for i > j,
// swap when m[j] > m[i]; the answer is m[nsel-1].
template <typename T, int N>
static inline T exchange_select(T (&m)[N], int nsel) {
    for (int j = 0; j < nsel; ++j)
        for (int i = j + 1; i < N; ++i)
            if (m[j] > m[i]) std::swap(m[j], m[i]);
    return m[nsel - 1];
}

// A8: MedianFilterKernel5 (medianFilter.cu:3-75). Taps start at 0; fill rule at :31-50
// (a tap outside in both row and column keeps its 0).
template <typename T>
void median5(const T* u, const T* v, int w, int h, int s, T* ou, T* ov) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            T mu[25], mv[25];
            for (int k = 0; k < 25; ++k) mu[k] = mv[k] = T(0);
            for (int j = 0; j < 5; ++j)
                for (int i = 0; i < 5; ++i) {
                    int col = ix + i - 2, row = iy + j - 2;
                    long q = -1;
                    if (col >= 0 && col < w && row >= 0 && row < h) q = col + (long)s * row;
                    else if (col < 0 && row >= 0 && row < h) q = (long)s * row;
                    else if (col >= w && row >= 0 && row < h) q = w - 1 + (long)s * row;
                    else if (col >= 0 && col < w && row < 0) q = col;
                    else if (col >= 0 && col < w && row >= h) q = col + (long)s * (h - 1);
                    if (q >= 0) {
                        mu[j * 5 + i] = u[q];
                        mv[j * 5 + i] = v[q];
                    }
                }
            long p = ix + (long)iy * s;
            ou[p] = exchange_select(mu, 13);
            ov[p] = exchange_select(mv, 13);
        }
}

// A15: MedianFilter3DKernel3 (medianFilter.cu:166-243). The right and bottom
// replacements test col > width / row > height, so taps at col == w or row == h stay 0.
template <typename T>
void median3d(const T* X, const T* Y, const T* Z, int w, int h, int s, T* X1, T* Y1, T* Z1) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            T mX[9], mY[9], mZ[9];
            for (int k = 0; k < 9; ++k) mX[k] = mY[k] = mZ[k] = T(0);
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) {
                    int col = ix + i - 1, row = iy + j - 1;
                    long q = -1;
                    if (col >= 0 && col < w && row >= 0 && row < h) q = col + (long)s * row;
                    else if (col < 0 && row >= 0 && row < h) q = (long)s * row;
                    else if (col > w && row >= 0 && row < h) q = w - 1 + (long)s * row;
                    else if (col >= 0 && col < w && row < 0) q = col;
                    else if (col >= 0 && col < w && row > h) q = col + (long)s * (h - 1);
                    if (q >= 0) {
                        mX[j * 3 + i] = X[q];
                        mY[j * 3 + i] = Y[q];
                        mZ[j * 3 + i] = Z[q];
                    }
                }
            long p = ix + (long)iy * s;
            X1[p] = exchange_select(mX, 5);
            Y1[p] = exchange_select(mY, 5);
            Z1[p] = exchange_select(mZ, 5);
        }
}

// A9: AddKernel (addKernel.cu:3-11) over count elements (padding included).
template <typename T>
void add(const T* a, const T* b, long count, T* out) {
#pragma omp parallel for
    for (long i = 0; i < count; ++i) out[i] = a[i] + b[i];
}

// A12: SolveReconDataL1Kernel (solveReconDataL1.cu:4-102).
template <typename T>
void recon_data_l1(const T* u0, const T* v0, const T* duhat0, const T* dvhat0, const T* uproj0,
                   const T* vproj0, const T* sku0, const T* skv0, const T* pu1, const T* pu2, const T* pv1,
                   const T* pv2, const T* Ix, const T* Iy, const T* It, int w, int h, int s, T lambda,
                   T alphaTv, T alphaProj, T* duhat1, T* dvhat1) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            T dix = Ix[p], diy = Iy[p], dit = It[p];
            T A = alphaTv + alphaProj;
            T dusub = (alphaTv * duhat0[p] + alphaProj * ((uproj0[p] - u0[p]) - sku0[p])) / A;
            T dvsub = (alphaTv * dvhat0[p] + alphaProj * ((vproj0[p] - v0[p]) - skv0[p])) / A;
            T rho = (dix * dusub + diy * dvsub) + dit;
            T g2 = dix * dix + diy * diy;
            T upper = (lambda * g2) / A;
            T lower = ((-lambda) * g2) / A;
            T du = std::numeric_limits<T>::quiet_NaN(), dv = du;
            if ((rho <= upper) && (rho >= lower)) {
                T magi = dix * dix + diy * diy;
                if (magi != T(0)) {
                    du = dusub - (rho * dix) / magi;
                    dv = dvsub - (rho * diy) / magi;
                } else {
                    du = dusub;
                    dv = dvsub;
                }
            } else if (rho < lower) {
                du = dusub + (lambda * dix) / A;
                dv = dvsub + (lambda * diy) / A;
            } else if (rho > upper) {
                du = dusub - (lambda * dix) / A;
                dv = dvsub - (lambda * diy) / A;
            }
            duhat1[p] = du + div_p(pu1, pu2, ix, iy, s) / A;
            dvhat1[p] = dv + div_p(pv1, pv2, ix, iy, s) / A;
        }
}

// f1: SolveReconDataL1FlownetKernel, dense and sparse (solveReconDataL1.cu:139-240,
// 276-378): the data blend gains an alphaFn pull towards the prior flow; the sparse
// form weights alphaFn by the mask.
template <typename T>
void recon_data_l1_fn(const T* u0, const T* v0, const T* duhat0, const T* dvhat0, const T* uproj0,
                      const T* vproj0, const T* ufn0, const T* vfn0, const T* sku0, const T* skv0,
                      const T* pu1, const T* pu2, const T* pv1, const T* pv2, const T* Ix, const T* Iy,
                      const T* It, int w, int h, int s, T lambda, T alphaTv, T alphaProj, T alphaFn0,
                      const T* mask, T* duhat1, T* dvhat1) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            T alphaFn = mask ? alphaFn0 * mask[p] : alphaFn0;
            T dix = Ix[p], diy = Iy[p], dit = It[p];
            T A = (alphaTv + alphaProj) + alphaFn;
            T dusub = ((alphaFn * (ufn0[p] - u0[p]) + alphaTv * duhat0[p]) +
                       alphaProj * ((uproj0[p] - u0[p]) - sku0[p])) / A;
            T dvsub = ((alphaFn * (vfn0[p] - v0[p]) + alphaTv * dvhat0[p]) +
                       alphaProj * ((vproj0[p] - v0[p]) - skv0[p])) / A;
            T rho = (dix * dusub + diy * dvsub) + dit;
            T g2 = dix * dix + diy * diy;
            T upper = (lambda * g2) / A;
            T lower = ((-lambda) * g2) / A;
            T du = std::numeric_limits<T>::quiet_NaN(), dv = du;
            if ((rho <= upper) && (rho >= lower)) {
                T magi = dix * dix + diy * diy;
                if (magi != T(0)) {
                    du = dusub - (rho * dix) / magi;
                    dv = dvsub - (rho * diy) / magi;
                } else {
                    du = dusub;
                    dv = dvsub;
                }
            } else if (rho < lower) {
                du = dusub + (lambda * dix) / A;
                dv = dvsub + (lambda * diy) / A;
            } else if (rho > upper) {
                du = dusub - (lambda * dix) / A;
                dv = dvsub - (lambda * diy) / A;
            }
            duhat1[p] = du + div_p(pu1, pu2, ix, iy, s) / A;
            dvhat1[p] = dv + div_p(pv1, pv2, ix, iy, s) / A;
        }
}

// A13: SolveUVProjKernel (solveUVProj.cu:4-22).
template <typename T>
void uvproj(const T* u0, const T* v0, const T* du0, const T* dv0, const T* uc0, const T* vc0, const T* sku0,
            const T* skv0, int w, int h, int s, T lambdaf, T alphaProj, T* uproj1, T* vproj1) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            T den = lambdaf + alphaProj;
            uproj1[p] = (lambdaf * uc0[p] + alphaProj * ((u0[p] + du0[p]) + sku0[p])) / den;
            vproj1[p] = (lambdaf * vc0[p] + alphaProj * ((v0[p] + dv0[p]) + skv0[p])) / den;
        }
}

// A14: Solve3dMinimalAreaKernel (solve3dMinimalArea.cu:3-193). The guards ix+1 > width
// and iy+1 > height (:65, :76, :96, :112) are never true, so column w and row h are read;
// the caller's buffers carry the flat level-0 layout plus a zero guard row.
template <typename T>
void solve3d(const T* uproj0, const T* vproj0, const T* X0, const T* Y0, const T* Z0, const double* P,
             const double* Q, T lambdaf, T lambdams, int w, int h, int s, T* X, T* Y, T* Z) {
    const T P11 = (T)P[0], P12 = (T)P[1], P13 = (T)P[2], P14 = (T)P[3];
    const T P21 = (T)P[4], P22 = (T)P[5], P23 = (T)P[6], P24 = (T)P[7];
    const T P31 = (T)P[8], P32 = (T)P[9], P33 = (T)P[10], P34 = (T)P[11];
    const T Q11 = (T)Q[0], Q12 = (T)Q[1], Q13 = (T)Q[2], Q14 = (T)Q[3];
    const T Q21 = (T)Q[4], Q22 = (T)Q[5], Q23 = (T)Q[6], Q24 = (T)Q[7];
    const T Q31 = (T)Q[8], Q32 = (T)Q[9], Q33 = (T)Q[10], Q34 = (T)Q[11];
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long pos = ix + (long)iy * s;
            T x1 = (T)ix, y1 = (T)iy;
            T x2 = (T)ix + uproj0[pos];
            T y2 = (T)iy + vproj0[pos];
            T p1 = x1 * P31 - P11, p2 = x1 * P32 - P12, p3 = x1 * P33 - P13, p4 = x1 * P34 - P14;
            T q1 = y1 * P31 - P21, q2 = y1 * P32 - P22, q3 = y1 * P33 - P23, q4 = y1 * P34 - P24;
            T r1 = x2 * Q31 - Q11, r2 = x2 * Q32 - Q12, r3 = x2 * Q33 - Q13, r4 = x2 * Q34 - Q14;
            T s1 = y2 * Q31 - Q21, s2 = y2 * Q32 - Q22, s3 = y2 * Q33 - Q23, s4 = y2 * Q34 - Q24;
            auto dot4 = [](T a1, T b1, T a2, T b2, T a3, T b3, T a4, T b4) {
                return ((a1 * b1 + a2 * b2) + a3 * b3) + a4 * b4;
            };
            T ax = lambdaf * dot4(p1, p1, q1, q1, r1, r1, s1, s1);
            T bx = lambdaf * dot4(p1, p2, q1, q2, r1, r2, s1, s2);
            T cx = lambdaf * dot4(p1, p3, q1, q3, r1, r3, s1, s3);
            T dx = lambdaf * dot4(p1, p4, q1, q4, r1, r4, s1, s4);
            T ay = lambdaf * dot4(p2, p1, q2, q1, r2, r1, s2, s1);
            T by = lambdaf * dot4(p2, p2, q2, q2, r2, r2, s2, s2);
            T cy = lambdaf * dot4(p2, p3, q2, q3, r2, r3, s2, s3);
            T dy = lambdaf * dot4(p2, p4, q2, q4, r2, r4, s2, s4);
            T az = lambdaf * dot4(p3, p1, q3, q1, r3, r1, s3, s1);
            T bz = lambdaf * dot4(p3, p2, q3, q2, r3, r2, s3, s2);
            T cz = lambdaf * dot4(p3, p3, q3, q3, r3, r3, s3, s3);
            T dz = lambdaf * dot4(p3, p4, q3, q4, r3, r4, s3, s4);

            // forward differences; the border branches of the reference never fire
            T Xx = X0[pos + 1] - X0[pos], Yx = Y0[pos + 1] - Y0[pos], Zx = Z0[pos + 1] - Z0[pos];
            T Xy = X0[pos + s] - X0[pos], Yy = Y0[pos + s] - Y0[pos], Zy = Z0[pos + s] - Z0[pos];

            T c1 = Yx * Zy - Zx * Yy;
            T c2 = Zx * Xy - Xx * Zy;
            T c3 = Xx * Yy - Yx * Xy;
            T lambetams = lambdams / std::sqrt(((c1 * c1 + c2 * c2) + c3 * c3) + T(0.001f));

            T Xibar, Yibar, Zibar, Xjbar, Yjbar, Zjbar;
            if (ix - 1 < 0) {
                Xibar = X0[pos + 1];
                Yibar = Y0[pos + 1];
                Zibar = Z0[pos + 1];
            } else {
                Xibar = T(0.5) * X0[pos - 1] + T(0.5) * X0[pos + 1];
                Yibar = T(0.5) * Y0[pos - 1] + T(0.5) * Y0[pos + 1];
                Zibar = T(0.5) * Z0[pos - 1] + T(0.5) * Z0[pos + 1];
            }
            if (iy - 1 < 0) {
                Xjbar = X0[pos + s];
                Yjbar = Y0[pos + s];
                Zjbar = Z0[pos + s];
            } else {
                Xjbar = T(0.5) * X0[pos - s] + T(0.5) * X0[pos + s];
                Yjbar = T(0.5) * Y0[pos - s] + T(0.5) * Y0[pos + s];
                Zjbar = T(0.5) * Z0[pos - s] + T(0.5) * Z0[pos + s];
            }

            T a11 = ax + lambetams * (((Yx * Yx + Zx * Zx) + Yy * Yy) + Zy * Zy);
            T a12 = bx - lambetams * (Xx * Yx + Xy * Yy);
            T a13 = cx - lambetams * (Xx * Zx + Xy * Zy);
            T a21 = ay - lambetams * (Xx * Yx + Xy * Yy);
            T a22 = by + lambetams * (((Xx * Xx + Zx * Zx) + Xy * Xy) + Zy * Zy);
            T a23 = cy - lambetams * (Yx * Zx + Yy * Zy);
            T a31 = az - lambetams * (Xx * Zx + Xy * Zy);
            T a32 = bz - lambetams * (Yx * Zx + Yy * Zy);
            T a33 = cz + lambetams * (((Xx * Xx + Yx * Yx) + Xy * Xy) + Yy * Yy);

            T detA = (a11 * (a22 * a33 - a23 * a32) - a12 * (a21 * a33 - a23 * a31)) +
                     a13 * (a21 * a32 - a22 * a31);

            T E11 = a22 * a33 - a23 * a32, E12 = a13 * a32 - a12 * a33, E13 = a12 * a23 - a13 * a22;
            T E21 = a23 * a31 - a21 * a33, E22 = a11 * a33 - a13 * a31, E23 = a13 * a21 - a11 * a23;
            T E31 = a21 * a32 - a22 * a31, E32 = a12 * a31 - a11 * a32, E33 = a11 * a22 - a12 * a21;

            T b1 = lambetams * ((((((Zy * Zy + Yy * Yy) * Xibar + (Zx * Zx + Yx * Yx) * Xjbar) -
                                  (Xy * Yy) * Yibar) - (Xx * Yx) * Yjbar) - (Xy * Zy) * Zibar) -
                                (Xx * Zx) * Zjbar) - dx;
            T b2 = lambetams * (((((-(Xy * Yy) * Xibar - (Xx * Yx) * Xjbar) + (Xy * Xy + Zy * Zy) * Yibar) +
                                  (Xx * Xx + Zx * Zx) * Yjbar) - (Yy * Zy) * Zibar) -
                                (Yx * Zx) * Zjbar) - dy;
            T b3 = lambetams * (((((-(Xy * Zy) * Xibar - (Xx * Zx) * Xjbar) - (Yy * Zy) * Yibar) -
                                  (Yx * Zx) * Yjbar) + (Xy * Xy + Yy * Yy) * Zibar) +
                                (Xx * Xx + Yx * Yx) * Zjbar) - dz;

            T Xsub = ((E11 * b1 + E12 * b2) + E13 * b3) / detA;
            T Ysub = ((E21 * b1 + E22 * b2) + E23 * b3) / detA;
            T Zsub = ((E31 * b1 + E32 * b2) + E33 * b3) / detA;

            if (std::isfinite(Xsub) && std::isfinite(Ysub) && std::isfinite(Zsub)) {
                X[pos] = Xsub;
                Y[pos] = Ysub;
                Z[pos] = Zsub;
            } else {
                X[pos] = X0[pos];
                Y[pos] = Y0[pos];
                Z[pos] = Z0[pos];
            }
        }
}

// One pixel of the FP64 FMA triangulation (see solve3d_fma). r, d, l, u: X0 at ix+1, iy+1, ix-1
// (used when hl), iy-1 (used when hu). On a non-finite solution the pixel keeps X0.
static inline void tri_fma(const double* P, const double* Q, double lambdaf, double lambdams, int ix, int iy,
                           double up, double vp, double X0, double Y0, double Z0, double Xr, double Yr, double Zr,
                           double Xd, double Yd, double Zd, double Xl, double Yl, double Zl, double Xu, double Yu,
                           double Zu, bool hl, bool hu, double& oX, double& oY, double& oZ) {
    using std::fma;
    const double x1 = (double)ix, y1 = (double)iy, x2 = (double)ix + up, y2 = (double)iy + vp;
    double pp[4], qq[4], rr[4], ss[4];
    for (int k = 0; k < 4; ++k) {
        pp[k] = fma(x1, P[8 + k], -P[k]);
        qq[k] = fma(y1, P[8 + k], -P[4 + k]);
        rr[k] = fma(x2, Q[8 + k], -Q[k]);
        ss[k] = fma(y2, Q[8 + k], -Q[4 + k]);
    }
    const double Xx = Xr - X0, Yx = Yr - Y0, Zx = Zr - Z0;
    const double Xy = Xd - X0, Yy = Yd - Y0, Zy = Zd - Z0;
    const double c1 = fma(Yx, Zy, -(Zx * Yy)), c2 = fma(Zx, Xy, -(Xx * Zy)), c3 = fma(Xx, Yy, -(Yx * Xy));
    // the system scaled by sq = sqrt(|Xx x Xy|^2 + 0.001): data terms weighted lambdaf * sq, the
    // regulariser lambdams (the same solution, no division)
    const double sq = std::sqrt(fma(c3, c3, fma(c2, c2, c1 * c1)) + (double)0.001f);
    const double ls = lambdaf * sq, lm = lambdams;
    auto dot = [&](int i, int j) {
        return ls * fma(ss[i], ss[j], fma(rr[i], rr[j], fma(qq[i], qq[j], pp[i] * pp[j])));
    };
    const double ax = dot(0, 0), bx = dot(0, 1), cx = dot(0, 2), dx = dot(0, 3);
    const double by = dot(1, 1), cy = dot(1, 2), dy = dot(1, 3);
    const double cz = dot(2, 2), dz = dot(2, 3);
    const double Xib = hl ? fma(0.5, Xl, 0.5 * Xr) : Xr, Yib = hl ? fma(0.5, Yl, 0.5 * Yr) : Yr,
                 Zib = hl ? fma(0.5, Zl, 0.5 * Zr) : Zr;
    const double Xjb = hu ? fma(0.5, Xu, 0.5 * Xd) : Xd, Yjb = hu ? fma(0.5, Yu, 0.5 * Yd) : Yd,
                 Zjb = hu ? fma(0.5, Zu, 0.5 * Zd) : Zd;
    // products and squares of the forward differences
    const double xxy = Xx * Yx, xyy = Xy * Yy, xxz = Xx * Zx, xyz = Xy * Zy, yxz = Yx * Zx, yyz = Yy * Zy;
    const double syx = fma(Zx, Zx, Yx * Yx), syy = fma(Zy, Zy, Yy * Yy);  // Yx^2 + Zx^2, Yy^2 + Zy^2
    const double sxx = fma(Zx, Zx, Xx * Xx), sxy = fma(Zy, Zy, Xy * Xy);  // Xx^2 + Zx^2, Xy^2 + Zy^2
    const double txx = fma(Yx, Yx, Xx * Xx), txy = fma(Yy, Yy, Xy * Xy);  // Xx^2 + Yx^2, Xy^2 + Yy^2
    const double a11 = fma(lm, syx + syy, ax), a22 = fma(lm, sxx + sxy, by), a33 = fma(lm, txx + txy, cz);
    const double a12 = fma(-lm, fma(Xy, Yy, xxy), bx), a13 = fma(-lm, fma(Xy, Zy, xxz), cx),
                 a23 = fma(-lm, fma(Yy, Zy, yxz), cy);
    // adjugate (symmetric) and determinant
    const double E11 = fma(a22, a33, -(a23 * a23)), E12 = fma(a13, a23, -(a12 * a33)),
                 E13 = fma(a12, a23, -(a13 * a22));
    const double E22 = fma(a11, a33, -(a13 * a13)), E23 = fma(a13, a12, -(a11 * a23)),
                 E33 = fma(a11, a22, -(a12 * a12));
    const double det = fma(a13, E13, fma(a12, E12, a11 * E11));
    double b1 = syy * Xib;
    b1 = fma(syx, Xjb, b1);
    b1 = fma(-xyy, Yib, b1);
    b1 = fma(-xxy, Yjb, b1);
    b1 = fma(-xyz, Zib, b1);
    b1 = fma(-xxz, Zjb, b1);
    b1 = fma(lm, b1, -dx);
    double b2 = -xyy * Xib;
    b2 = fma(-xxy, Xjb, b2);
    b2 = fma(sxy, Yib, b2);
    b2 = fma(sxx, Yjb, b2);
    b2 = fma(-yyz, Zib, b2);
    b2 = fma(-yxz, Zjb, b2);
    b2 = fma(lm, b2, -dy);
    double b3 = -xyz * Xib;
    b3 = fma(-xxz, Xjb, b3);
    b3 = fma(-yyz, Yib, b3);
    b3 = fma(-yxz, Yjb, b3);
    b3 = fma(txy, Zib, b3);
    b3 = fma(txx, Zjb, b3);
    b3 = fma(lm, b3, -dz);
    const double inv = 1.0 / det;
    const double Xs = fma(E13, b3, fma(E12, b2, E11 * b1)) * inv;
    const double Ys = fma(E23, b3, fma(E22, b2, E12 * b1)) * inv;
    const double Zs = fma(E33, b3, fma(E23, b2, E13 * b1)) * inv;
    if (std::isfinite(Xs) && std::isfinite(Ys) && std::isfinite(Zs)) {
        oX = Xs;
        oY = Ys;
        oZ = Zs;
    } else {
        oX = X0;
        oY = Y0;
        oZ = Z0;
    }
}

// The FP64 triangulation mode: the normal equations of Solve3dMinimalAreaKernel
// (solve3dMinimalArea.cu:21-185) built and solved in FP64 with fused multiply-adds, on T-typed
// fields (inputs widened, outputs rounded). Same terms as solve3d; the evaluation differs: the
// 4x3 products and all sums of products are FMA chains, the system is scaled by
// sq = sqrt(|Xx x Xy|^2 + 0.001) (data terms weighted lambdaf * sq, the regulariser lambdams, no
// division), the symmetric normal matrix and its adjugate are formed once (a21 = a12, a31 = a13,
// a32 = a23 hold bit for bit in solve3d too),
// the determinant is the cofactor expansion along the first row, and the solution is the
// adjugate times b scaled by one reciprocal of the determinant. Differs from solve3d<double> by
// FP64 rounding only (pinned against it in the tests).
template <typename T>
void solve3d_fma(const T* uproj0, const T* vproj0, const T* X0, const T* Y0, const T* Z0, const double* P,
                 const double* Q, T lambdaf_t, T lambdams_t, int w, int h, int s, T* X, T* Y, T* Z) {
    const double lambdaf = (double)lambdaf_t, lambdams = (double)lambdams_t;
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            const long pos = ix + (long)iy * s;
            const bool hl = ix > 0, hu = iy > 0;
            const long pl = hl ? pos - 1 : pos, pu = hu ? pos - s : pos;
            double oX, oY, oZ;
            tri_fma(P, Q, lambdaf, lambdams, ix, iy, (double)uproj0[pos], (double)vproj0[pos], (double)X0[pos],
                    (double)Y0[pos], (double)Z0[pos], (double)X0[pos + 1], (double)Y0[pos + 1], (double)Z0[pos + 1],
                    (double)X0[pos + s], (double)Y0[pos + s], (double)Z0[pos + s], (double)X0[pl], (double)Y0[pl],
                    (double)Z0[pl], (double)X0[pu], (double)Y0[pu], (double)Z0[pu], hl, hu, oX, oY, oZ);
            X[pos] = (T)oX;
            Y[pos] = (T)oY;
            Z[pos] = (T)oZ;
        }
}

// A16: CleanUp3DKernel (solve3dMinimalArea.cu:248-297).
template <typename T>
void cleanup3d(const T* X0, const T* Y0, const T* Z0, const T* Xm, const T* Ym, const T* Zm, int w, int h,
               int s, T* X1, T* Y1, T* Z1) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            T X = X0[p], Y = Y0[p], Z = Z0[p], xm = Xm[p], ym = Ym[p], zm = Zm[p];
            if (std::isfinite(X) && std::isfinite(Y) && std::isfinite(Z) && std::isfinite(xm) &&
                std::isfinite(ym) && std::isfinite(zm)) {
                X1[p] = (std::fabs(xm - X) > T(5)) ? xm : X;
                Y1[p] = (std::fabs(ym - Y) > T(5)) ? ym : Y;
                Z1[p] = (std::fabs(zm - Z) > T(5)) ? zm : Z;
            } else {
                X1[p] = 0;
                Y1[p] = 0;
                Z1[p] = 0;
            }
        }
}

// A17: SolveUcKernel (solveUcSk.cu:3-35): the focal length K1[0] serves both axes.
template <typename T>
void solve_uc(const T* X0, const T* Y0, const T* Z0, const double* R, const double* t, const double* K1,
              int w, int h, int s, T* uc, T* vc) {
    const T R11 = (T)R[0], R12 = (T)R[1], R13 = (T)R[2];
    const T R21 = (T)R[3], R22 = (T)R[4], R23 = (T)R[5];
    const T R31 = (T)R[6], R32 = (T)R[7], R33 = (T)R[8];
    const T t1 = (T)t[0], t2 = (T)t[1], t3 = (T)t[2];
    const T f = (T)K1[0], cx = (T)K1[2], cy = (T)K1[5];
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            T X = X0[p], Y = Y0[p], Z = Z0[p];
            T X2 = ((R11 * X + R12 * Y) + R13 * Z) + t1;
            T Y2 = ((R21 * X + R22 * Y) + R23 * Z) + t2;
            T Z2 = ((R31 * X + R32 * Y) + R33 * Z) + t3;
            T xproj = (f * X2) / Z2 + cx;
            T yproj = (f * Y2) / Z2 + cy;
            uc[p] = xproj - (T)ix;
            vc[p] = yproj - (T)iy;
        }
}

// A18: UpdateSkKernel (solveUcSk.cu:80-97).
template <typename T>
void update_sk(const T* sku0, const T* skv0, const T* u0, const T* v0, const T* du0, const T* dv0,
               const T* uproj0, const T* vproj0, int w, int h, int s, T* sku1, T* skv1) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            sku1[p] = ((sku0[p] + u0[p]) + du0[p]) - uproj0[p];
            skv1[p] = ((skv0[p] + v0[p]) + dv0[p]) - vproj0[p];
        }
}

// ---------------------------------------------------------------------------------
// Precise-projection mode (Params::proj_f64; this implementation's arithmetic, the reference
// evaluates these steps in FP32). The projection blend (solveUVProj.cu:19-20), the Sk update
// (solveUcSk.cu:93-94) and the joint data step's blend (solveReconDataL1.cu:35-36) add and subtract
// quantities of the size of the flow (u, uproj, sk: 20 px and more at 640x480) to form differences
// of ~1e-3 px, so each FP32 evaluation loses ~1e-6 px, and the Sk sum carries the loss through the
// 1250 iterations of a frame. The selected steps are evaluated in binary64 from the stored FP32
// values (sums of the launch constants in binary64 too, divisions as products with their binary64
// reciprocals) and rounded once to the field type.
enum { PREC_PROJ = 1, PREC_BLEND = 2, PREC_SK = 4, PREC_DATA = 8 };

template <typename T>
void uvproj_wide(const T* u0, const T* v0, const T* du0, const T* dv0, const T* uc0, const T* vc0, const T* sku0,
                 const T* skv0, int w, int h, int s, T lambdaf, T alphaProj, T* uproj1, T* vproj1) {
    const double lf = lambdaf, aP = alphaProj, r = 1.0 / (lf + aP);
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            uproj1[p] = (T)((lf * (double)uc0[p] + aP * (((double)u0[p] + (double)du0[p]) + (double)sku0[p])) * r);
            vproj1[p] = (T)((lf * (double)vc0[p] + aP * (((double)v0[p] + (double)dv0[p]) + (double)skv0[p])) * r);
        }
}

template <typename T>
void update_sk_wide(const T* sku0, const T* skv0, const T* u0, const T* v0, const T* du0, const T* dv0,
                    const T* uproj0, const T* vproj0, int w, int h, int s, T* sku1, T* skv1) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            sku1[p] = (T)((((double)sku0[p] + (double)u0[p]) + (double)du0[p]) - (double)uproj0[p]);
            skv1[p] = (T)((((double)skv0[p] + (double)v0[p]) + (double)dv0[p]) - (double)vproj0[p]);
        }
}

// The joint data step (solveReconDataL1.cu:4-102) with its blend in binary64 (PREC_BLEND: the
// threshold and the divergence then in T, as recon_data_l1) or the whole step in binary64
// (PREC_DATA).
template <typename T, typename C>
static void recon_data_px(C dix, C diy, C dit, C dusub, C dvsub, C lambda, C A, C divu, C divv, T& o1, T& o2) {
    C rho = (dix * dusub + diy * dvsub) + dit;
    C g2 = dix * dix + diy * diy;
    C upper = (lambda * g2) / A;
    C lower = ((-lambda) * g2) / A;
    C du = std::numeric_limits<C>::quiet_NaN(), dv = du;
    if ((rho <= upper) && (rho >= lower)) {
        if (g2 != C(0)) {
            du = dusub - (rho * dix) / g2;
            dv = dvsub - (rho * diy) / g2;
        } else {
            du = dusub;
            dv = dvsub;
        }
    } else if (rho < lower) {
        du = dusub + (lambda * dix) / A;
        dv = dvsub + (lambda * diy) / A;
    } else if (rho > upper) {
        du = dusub - (lambda * dix) / A;
        dv = dvsub - (lambda * diy) / A;
    }
    o1 = (T)(du + divu / A);
    o2 = (T)(dv + divv / A);
}

template <typename T>
void recon_data_l1_wide(const T* u0, const T* v0, const T* duhat0, const T* dvhat0, const T* uproj0,
                        const T* vproj0, const T* sku0, const T* skv0, const T* pu1, const T* pu2, const T* pv1,
                        const T* pv2, const T* Ix, const T* Iy, const T* It, int w, int h, int s, T lambda,
                        T alphaTv, T alphaProj, bool whole, T* duhat1, T* dvhat1) {
    const double aT = alphaTv, aP = alphaProj, Ad = aT + aP, rA = 1.0 / Ad;
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            double bu = (aT * (double)duhat0[p] + aP * (((double)uproj0[p] - (double)u0[p]) - (double)sku0[p])) * rA;
            double bv = (aT * (double)dvhat0[p] + aP * (((double)vproj0[p] - (double)v0[p]) - (double)skv0[p])) * rA;
            if (whole)
                recon_data_px<T, double>(Ix[p], Iy[p], It[p], bu, bv, lambda, Ad, div_p(pu1, pu2, ix, iy, s),
                                         div_p(pv1, pv2, ix, iy, s), duhat1[p], dvhat1[p]);
            else
                recon_data_px<T, T>(Ix[p], Iy[p], It[p], (T)bu, (T)bv, lambda, alphaTv + alphaProj,
                                    div_p(pu1, pu2, ix, iy, s), div_p(pv1, pv2, ix, iy, s), duhat1[p], dvhat1[p]);
        }
}

// The prior-flow data step (solveReconDataL1.cu:139-240, 276-378) with its blend in binary64:
// AA = (aTv + aP) + aFn formed in binary64 (aFn = alphaFn times the mask in T), the blend divided
// by it and rounded once; the threshold and the divergence in T with the T-typed AA.
template <typename T>
void recon_data_l1_fn_wide(const T* u0, const T* v0, const T* duhat0, const T* dvhat0, const T* uproj0,
                           const T* vproj0, const T* ufn0, const T* vfn0, const T* sku0, const T* skv0, const T* pu1,
                           const T* pu2, const T* pv1, const T* pv2, const T* Ix, const T* Iy, const T* It, int w,
                           int h, int s, T lambda, T alphaTv, T alphaProj, T alphaFn0, const T* mask, T* duhat1,
                           T* dvhat1) {
    const double aT = alphaTv, aP = alphaProj;
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            T alphaFn = mask ? alphaFn0 * mask[p] : alphaFn0;
            const double aF = alphaFn, AAd = (aT + aP) + aF;
            T bu = (T)(((aF * ((double)ufn0[p] - (double)u0[p]) + aT * (double)duhat0[p]) +
                        aP * (((double)uproj0[p] - (double)u0[p]) - (double)sku0[p])) / AAd);
            T bv = (T)(((aF * ((double)vfn0[p] - (double)v0[p]) + aT * (double)dvhat0[p]) +
                        aP * (((double)vproj0[p] - (double)v0[p]) - (double)skv0[p])) / AAd);
            recon_data_px<T, T>(Ix[p], Iy[p], It[p], bu, bv, lambda, (alphaTv + alphaProj) + alphaFn,
                                div_p(pu1, pu2, ix, iy, s), div_p(pv1, pv2, ix, iy, s), duhat1[p], dvhat1[p]);
        }
}

// ---------------------------------------------------------------------------------
// f2 / f3 / f4: display, post-filter and alternate data terms.

// ComputeColorKernel (flowToHSV.cu:194-302): Middlebury colour wheel, BGR output. The
// reference indexes its 55-entry tables with k0, k1 in [1, 55]; index 55 lies past the table
// and is taken modulo 55 here (parity unpinned for that entry).
static const double kWheel[3][55] = {
    {255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 213, 170, 128,
     85,  43,  0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   19,
     39,  58,  78,  98,  117, 137, 156, 176, 196, 215, 235, 255, 255, 255, 255, 255, 255},
    {0,   17,  34,  51,  68,  85,  102, 119, 136, 153, 170, 187, 204, 221, 238, 255, 255, 255, 255,
     255, 255, 255, 255, 255, 255, 255, 232, 209, 186, 163, 140, 116, 93,  70,  47,  24,  0,   0,
     0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0},
    {0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,
     0,   0,   0,   63,  127, 191, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255,
     255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 213, 170, 128, 85,  43}};

template <typename T>
void flow_to_hsv(const T* u, const T* v, int w, int h, int s, T* rgb, T flowscale) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            T du = u[p] / flowscale, dv = v[p] / flowscale;
            const int ncols = 55;
            T rad = std::sqrt(du * du + dv * dv);
            T a = std::atan2(-dv, -du) / T(3.14159f);
            T fk = (a + T(1)) / T(2) * (T(ncols) - T(1)) + T(1);
            int k0 = (int)std::floor(fk);
            int k1 = k0 + 1;
            if (k1 == ncols + 1) k1 = 1;
            T f = fk - (T)k0;
            k0 = ((k0 % ncols) + ncols) % ncols;
            k1 = ((k1 % ncols) + ncols) % ncols;
            T col[3];
            for (int c = 0; c < 3; ++c) {
                T cc = (T(1) - f) * (T(kWheel[c][k0]) / T(255)) + f * (T(kWheel[c][k1]) / T(255));
                col[c] = rad <= T(1) ? T(1) - rad * (T(1) - cc) : cc * T(0.75);
            }
            rgb[3 * p + 0] = col[2];
            rgb[3 * p + 1] = col[1];
            rgb[3 * p + 2] = col[0];
        }
}

// RemoveEdgeArtifacts3DKernel (solve3dMinimalArea.cu:320-378): zero points whose surface-area
// proxy over a `range`-pixel neighbourhood exceeds maxarea. Reads column w / row h at the border.
template <typename T>
void remove_edge(const T* X0, const T* Y0, const T* Z0, int range, T maxarea, int w, int h, int s, T* X1, T* Y1,
                 T* Z1) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            T X = X0[p], Y = Y0[p], Z = Z0[p];
            T Xx, Yx, Zx, Xy, Yy, Zy;
            if (ix + range > w) {
                long q = p - range;
                Xx = -X0[q] + X;
                Yx = -Y0[q] + Y;
                Zx = -Z0[q] + Z;
            } else {
                long q = p + range;
                Xx = X0[q] - X;
                Yx = Y0[q] - Y;
                Zx = Z0[q] - Z;
            }
            if (iy + range > h) {
                long q = p - (long)range * s;
                Xy = -X0[q] + X;
                Yy = -Y0[q] + Y;
                Zy = -Z0[q] + Z;
            } else {
                long q = p + (long)range * s;
                Xy = X0[q] - X;
                Yy = Y0[q] - Y;
                Zy = Z0[q] - Z;
            }
            T c1 = Yx * Zy - Zx * Yy, c2 = Zx * Xy - Xx * Zy, c3 = Xx * Yy - Yx * Xy;
            T area = std::sqrt((c1 * c1 + c2 * c2) + c3 * c3);
            bool drop = area > maxarea;
            X1[p] = drop ? T(0) : X;
            Y1[p] = drop ? T(0) : Y;
            Z1[p] = drop ? T(0) : Z;
        }
}

// ComputeDerivKernel (computeDeriv.cu:6-40): single-image 5-point gradient.
template <typename T>
void deriv_single(const T* I, int w, int h, int s, T* Ix, T* Iy) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            auto at = [&](long x, long y) { return I[refl(x, w) + refl(y, h) * s]; };
            T t0 = at(ix - 2, iy);
            t0 -= at(ix - 1, iy) * T(8);
            t0 += at(ix + 1, iy) * T(8);
            t0 -= at(ix + 2, iy);
            t0 /= T(12);
            Ix[p] = t0;
            t0 = at(ix, iy - 2);
            t0 -= at(ix, iy - 1) * T(8);
            t0 += at(ix, iy + 1) * T(8);
            t0 -= at(ix, iy + 2);
            t0 /= T(12);
            Iy[p] = t0;
        }
}

// ComputeDerivativesPatchKernel (computeDerivativesPatch.cu:8-67): the derivatives of
// ComputeDerivatives summed over the 5x5 patch; each tap's own stencil is mirror-indexed.
template <typename T>
void derivatives_patch(const T* I0, const T* I1, int w, int h, int s, T* Ix, T* Iy, T* Iz) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            auto at = [&](const T* I, long x, long y) { return I[refl(x, w) + refl(y, h) * s]; };
            T sx = 0, sy = 0, sz = 0;
            for (int j = 0; j < 5; ++j)
                for (int i = 0; i < 5; ++i) {
                    const long col = ix + i - 2, row = iy + j - 2;
                    T t0, t1;
                    t0 = at(I0, col - 2, row);
                    t0 -= at(I0, col - 1, row) * T(8);
                    t0 += at(I0, col + 1, row) * T(8);
                    t0 -= at(I0, col + 2, row);
                    t0 /= T(12);
                    t1 = at(I1, col - 2, row);
                    t1 -= at(I1, col - 1, row) * T(8);
                    t1 += at(I1, col + 1, row) * T(8);
                    t1 -= at(I1, col + 2, row);
                    t1 /= T(12);
                    sx += (t0 + t1) * T(0.5);
                    sz += at(I1, col, row) - at(I0, col, row);
                    t0 = at(I0, col, row - 2);
                    t0 -= at(I0, col, row - 1) * T(8);
                    t0 += at(I0, col, row + 1) * T(8);
                    t0 -= at(I0, col, row + 2);
                    t0 /= T(12);
                    t1 = at(I1, col, row - 2);
                    t1 -= at(I1, col, row - 1) * T(8);
                    t1 += at(I1, col, row + 1) * T(8);
                    t1 -= at(I1, col, row + 2);
                    t1 /= T(12);
                    sy += (t0 + t1) * T(0.5);
                }
            long p = ix + (long)iy * s;
            Ix[p] = sx;
            Iy[p] = sy;
            Iz[p] = sz;
        }
}

// sqrtf(x*x + 0.001) of the Charbonnier kernels: the literal 0.001 is double there, so the sum
// is formed in double and rounded to float before the float square root.
static inline float charb_root(float x) { return std::sqrt((float)((double)(x * x) + 0.001)); }
static inline double charb_root(double x) { return std::sqrt(x * x + 0.001); }

// SolveDataCharbForTVKernel (solveDataCharbForTV.cu:3-85) and SolveDataCharbForTVGradKernel
// (solveDataCharbForTVGrad.cu:3-103): IRLS 2x2 solves (1 and 2 passes), then the primal step.
// With gradient terms null this is the plain Charbonnier form.
template <typename T>
void data_charb(const T* du0, const T* dv0, const T* duhat0, const T* dvhat0, const T* pu1, const T* pu2,
                const T* pv1, const T* pv2, const T* Ix, const T* Iy, const T* It, const T* Ixx, const T* Ixy,
                const T* Ixt, const T* Iyx, const T* Iyy, const T* Iyt, int w, int h, int s, T lambda, T lambdagrad,
                T theta, T* du1, T* dv1, T* duhat1, T* dvhat1) {
    const bool grad = Ixx != nullptr;
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            T du = du0[p], dv = dv0[p];
            const T dix = Ix[p], diy = Iy[p], dit = It[p], duhat = duhat0[p], dvhat = dvhat0[p];
            const T it = T(1) / theta;
            if (!grad) {
                T innerSum = (dix * du + diy * dv) + dit;
                T beta = lambda / charb_root(innerSum);
                T a11 = (dix * dix) * beta + it;
                T a12 = (dix * diy) * beta;
                T a21 = a12;
                T a22 = (diy * diy) * beta + it;
                T detA = a11 * a22 - a12 * a21;
                T b1 = it * duhat - (dix * dit) * beta;
                T b2 = it * dvhat - (diy * dit) * beta;
                du = (a22 * b1 - a12 * b2) / detA;
                dv = ((-a21) * b1 + a11 * b2) / detA;
            } else {
                const T dixx = Ixx[p], dixy = Ixy[p], dixt = Ixt[p], diyx = Iyx[p], diyy = Iyy[p], diyt = Iyt[p];
                for (int k = 0; k < 2; ++k) {
                    T innerSum = (dix * du + diy * dv) + dit;
                    T innerSumx = (dixx * du + dixy * dv) + dixt;
                    T innerSumy = (diyx * du + diyy * dv) + diyt;
                    T beta = lambda / charb_root(innerSum);
                    T betax = lambdagrad / charb_root(innerSumx);
                    T betay = lambdagrad / charb_root(innerSumy);
                    T a11 = (((dix * dix) * beta + (dixx * dixx) * betax) + (diyx * diyx) * betay) + it;
                    T a12 = ((dix * diy) * beta + (dixx * dixy) * betax) + (diyx * diyy) * betay;
                    T a21 = a12;
                    T a22 = (((diy * diy) * beta + (dixy * dixy) * betax) + (diyy * diyy) * betay) + it;
                    T detA = a11 * a22 - a12 * a21;
                    T b1 = ((it * duhat - (dix * dit) * beta) - (dixx * dixt) * betax) - (diyx * diyt) * betay;
                    T b2 = ((it * dvhat - (diy * dit) * beta) - (dixy * dixt) * betax) - (diyy * diyt) * betay;
                    du = (a22 * b1 - a12 * b2) / detA;
                    dv = ((-a21) * b1 + a11 * b2) / detA;
                }
            }
            du1[p] = du;
            dv1[p] = dv;
            duhat1[p] = du + theta * div_p(pu1, pu2, ix, iy, s);
            dvhat1[p] = dv + theta * div_p(pv1, pv2, ix, iy, s);
        }
}

// SolveDataL1InpaintKernel (solveDataL1Inpaint.cu:4-98): the L1 step, skipped where mask0 == 0.
template <typename T>
void data_l1_inpaint(const T* duhat0, const T* dvhat0, const T* mask0, const T* pu1, const T* pu2, const T* pv1,
                     const T* pv2, const T* Ix, const T* Iy, const T* It, int w, int h, int s, T lambda, T theta,
                     T* duhat1, T* dvhat1) {
    data_l1(duhat0, dvhat0, pu1, pu2, pv1, pv2, Ix, Iy, It, w, h, s, lambda, theta, duhat1, dvhat1);
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            if (mask0[p] == T(0)) {
                duhat1[p] = duhat0[p] + theta * div_p(pu1, pu2, ix, iy, s);
                dvhat1[p] = dvhat0[p] + theta * div_p(pv1, pv2, ix, iy, s);
            }
        }
}

// SolveSceneFlowAndPointCloudKernel (solveSceneFlow.cu:49-91): back-projection of the depth at
// the pixel and at the flow target; the host passes depth0 as both maps (Cudaflow.cu:843).
template <typename T>
void scene_flow(const T* u, const T* v, const T* depth0, const T* depth1, T fx, T fy, T cx, T cy, int w, int h,
                int s, T* pc0, T* pc1, T* sf) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            T d0 = depth0[p];
            T uu = u[p], vv = v[p];
            T d1;
            if (!(std::fabs(uu) <= T(16777216)) || !(std::fabs(vv) <= T(16777216))) {
                d1 = std::numeric_limits<T>::quiet_NaN();
            } else {
                T fu = std::floor(uu), fv = std::floor(vv);
                d1 = bilin(depth1, w, h, s, ix + (long)fu, uu - fu, iy + (long)fv, vv - fv);
            }
            T p0z = d0, p0x = (((T)ix - cx) * p0z) / fx, p0y = (((T)iy - cy) * p0z) / fy;
            T p1z = d1, p1x = (((T)ix - cx) * p1z) / fx, p1y = (((T)iy - cy) * p1z) / fy;
            T sx = p0x - p1x, sy = p0y - p1y, sz = p0z - p1z;
            if ((d0 < T(500)) || (d0 > T(2500))) sx = sy = sz = T(0);
            pc0[3 * p] = p0x;
            pc0[3 * p + 1] = p0y;
            pc0[3 * p + 2] = p0z;
            pc1[3 * p] = p1x;
            pc1[3 * p + 1] = p1y;
            pc1[3 * p + 2] = p1z;
            sf[3 * p] = sx;
            sf[3 * p + 1] = sy;
            sf[3 * p + 2] = sz;
        }
}

// SceneFlowToHSVKernel (flowToHSV.cu:132-190): sf holds (x, y, z) per pixel. The sector bounds
// are double literals in the reference, so H is compared in double.
template <typename T>
void scene_flow_to_hsv(const T* sf, int w, int h, int s, T* rgb, T flowscale) {
#pragma omp parallel for
    for (int iy = 0; iy < h; ++iy)
        for (int ix = 0; ix < w; ++ix) {
            long p = ix + (long)iy * s;
            T um = sf[3 * p] / flowscale, vm = sf[3 * p + 1] / flowscale, wm = sf[3 * p + 2] / flowscale;
            T H = std::atan2(vm, um) + T(3.1416f);
            T C = std::sqrt(um * um + vm * vm);
            T Hp = H / T(1.0472f);
            T X = C * (T(1) - std::fabs(std::fmod(Hp, T(2)) - T(1)));
            T m = wm - C;
            double Hd = (double)H;
            T r, g, b;
            if (Hd < 6.2832 && Hd >= 5.236) r = C, g = 0, b = X;
            else if (Hd < 5.236 && Hd >= 4.1888) r = X, g = 0, b = C;
            else if (Hd < 4.1888 && Hd >= 3.1416) r = 0, g = X, b = C;
            else if (Hd < 3.1416 && Hd >= 2.0944) r = 0, g = C, b = X;
            else if (Hd < 2.0944 && Hd >= 1.0472) r = X, g = C, b = 0;
            else r = C, g = X, b = 0;
            rgb[3 * p] = r + m;
            rgb[3 * p + 1] = g + m;
            rgb[3 * p + 2] = b + m;
        }
}

// ---------------------------------------------------------------------------------
// Host orchestration.

// Parameters of CudaFlow::initialize / ReconFlow::_initializeR (Cudaflow.cu:99-127;
// Reconflow.cu:27-41). Layout mirrored by the Python wrapper.
struct Params {
    int width, height, n_levels;
    float scale;
    int method;
    float lambda, lambdagrad, theta, tau;
    int n_warp_iters, n_solver_iters;
    float lambdaf, lambdams, alpha_tv, alpha_proj, alpha_fn;
    int stride_alignment;
    int input_type;    // 0: float gray, 1: 8-bit RGB (uchar3), 2: 8-bit gray, 3: 16-bit gray
    int fix_coarse_k;  // 0: reference intrinsics pyramid (K[2][2] = 0 on levels >= 1); 1: K[2][2] = 1
    int tri_f64;       // 1: the FP64 mode: the triangulation step in FP64 whatever T is, and in the FP32
                       //    solve the projection and data blends in binary64 (PREC_PROJ | PREC_BLEND)
    // Test-only knobs (0 = off); neither is reference semantics.
    //   mutation:   plants one known schedule mistake (MUT_* below) so that tests can show their
    //               checks detect it;
    //   round_mask: rounds the state groups whose RND_* bit is set to FP32 after every write
    //               (FP64 arithmetic, FP32 storage of that group), to find which stored field
    //               carries the FP32 solve's distance from the FP64 one.
    int mutation;
    int round_mask;
    //   f32_steps:  in the FP64 joint solve, runs the steps whose STEP_* bit is set in FP32
    //               arithmetic (their fields converted to float and back), to find which step's
    //               arithmetic carries the FP32 solve's distance from the FP64 one.
    int f32_steps;
    int proj_f64;  // PREC_* bits: the precise-projection mode (for FP32 fields)
    // FP16 storage of the fused GPU path (rfhip.h RF_STORAGE_*): 1 = images, derivatives and the
    // TV-L1 duals in binary16; 2 = derivatives and the TV-L1 duals; 3 = derivatives and the duals of
    // both solvers; 4 (a measurement mode with no GPU counterpart) = images, derivatives and the
    // duals of both solvers. Every store of such a field is rounded to nearest-even binary16 and
    // widened exactly on load; updates stay in T.
    int storage;
};

// Round to the nearest binary16 value (ties to even), returned as a float: quantum q of the value's
// binade (2^-24 below the smallest normal 2^-14), rounding by the float addition of C = q 2^23 (whose
// float ulp is q) and its exact subtraction; 65520 and above (the tie above the largest finite
// binary16, 65504) go to infinity; NaN and infinities pass through.
static inline float round_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return f;
    const float a = std::fabs(f);
    if (a >= 65520.0f) return std::copysign(std::numeric_limits<float>::infinity(), f);
    const int e = (int)(ax >> 23) - 127;
    const float C = std::ldexp(1.0f, (e < -14 ? -24 : e - 10) + 23);
    volatile float t = a + C;  // one float rounding, to a multiple of q
    const float r = t - C;
    return std::copysign(r, f);
}

template <typename T>
static void round_f16(int n_or_zero, long n, std::initializer_list<T*> bufs) {
    if (!n_or_zero) return;
    for (T* b : bufs)
        for (long i = 0; i < n; ++i) b[i] = (T)round_half((float)b[i]);
}

// Steps of Params::f32_steps (joint solve).
enum {
    STEP_WARP = 1, STEP_DERIV = 2, STEP_DATA = 4, STEP_DUAL = 8, STEP_PROJ = 16, STEP_TRI = 32,
    STEP_CLEAN = 64, STEP_UC = 128, STEP_SK = 256, STEP_MED = 512, STEP_UPS = 1024,
};

// Float copies of a step's fields for Params::f32_steps: in(slot, src) converts a field to float
// (keeping what the step does not overwrite), out(slot, dst) converts a step output back.
template <typename T>
struct F32Fields {
    long m = 0;
    std::vector<std::vector<float>> b{16};
    float* in(int slot, const T* src) {
        b[slot].resize(m);
        for (long i = 0; i < m; ++i) b[slot][i] = (float)src[i];
        return b[slot].data();
    }
    void out(int slot, T* dst) {
        for (long i = 0; i < m; ++i) dst[i] = (T)b[slot][i];
    }
};

// Planted schedule mistakes (Params::mutation). TV-L1 schedule (Cudaflow.cu:403-621):
enum {
    MUT_NONE = 0,
    MUT_FLOW_NO_WARP_RESET = 1,   // uhat and duals not zeroed at the start of each warp (:474-487)
    MUT_FLOW_NO_MEDIAN = 2,       // the 5x5 median before the add skipped (:580-583)
    MUT_FLOW_NO_DUAL_SWAP = 3,    // the duals' swap after the dual step dropped (:568-571)
    MUT_FLOW_UPSCALE_NO_SCALE = 4,// upscale without the fScale factor (:602-611)
    MUT_FLOW_WARP_ZERO_FLOW = 5,  // frame 1 warped with zero flow (:490)
    MUT_FLOW_NO_UHAT_SWAP = 6,    // the uhat swap after the data step dropped (:548-549)
    // joint schedule (Reconflow.cu:147-466):
    MUT_RECON_NO_UV_MEDIAN = 11,  // the extra MedianFilter(u, v) per warp skipped (:418-420)
    MUT_RECON_SK_RESET = 12,      // sk zeroed at every warp (the reference keeps it: :223-224, 260-261 are commented out)
    MUT_RECON_NO_UC_UPSCALE = 13, // uc, vc not upscaled to the next level (:438-443)
    MUT_RECON_NO_SK_SWAP = 14,    // the sk swap after UpdateSk dropped (:405)
    MUT_RECON_NO_CLEANUP = 15,    // CleanUp3D skipped (:383-391)
    MUT_RECON_NO_PROJ_UPSCALE = 16, // uproj, vproj not upscaled (:445-450)
    MUT_RECON_NO_WARP_RESET = 17, // uhat and duals not zeroed at the start of each warp (:245-258)
    MUT_RECON_XYZ_LEVEL_RESET = 18, // X, Y, Z zeroed at every level (the reference zeroes them once, :216-218)
    MUT_RECON_UC_LEVEL0_K = 19,   // SolveUc with the level-0 intrinsics on every level (:397)
};

// State groups of Params::round_mask.
enum {
    RND_IMG = 1,    // image pyramids and the warped image
    RND_DER = 2,    // Ix, Iy, Iz
    RND_UHAT = 4,   // uhat (dumed, dvmed)
    RND_DUAL = 8,   // p (pu1, pu2, pv1, pv2)
    RND_PROJ = 16,  // uproj, vproj
    RND_XYZ = 32,   // X, Y, Z and their 3x3 medians
    RND_UC = 64,    // uc, vc
    RND_SK = 128,   // sku, skv
    RND_UV = 256,   // u, v
};

template <typename T>
static void round_group(const Params& P, int bit, long n, std::initializer_list<T*> bufs) {
    if (sizeof(T) == sizeof(float) || !(P.round_mask & bit)) return;
    for (T* b : bufs)
        for (long i = 0; i < n; ++i) b[i] = (T)(float)b[i];
}

// The triangulation step of the FP64-triangulation mode on T-typed fields: solve3d_fma.
template <typename T>
void solve3d_wide(const T* up, const T* vp, const T* X0, const T* Y0, const T* Z0, const double* P,
                  const double* Q, T lambdaf, T lambdams, int w, int h, int s, T* X, T* Y, T* Z) {
    solve3d_fma<T>(up, vp, X0, Y0, Z0, P, Q, lambdaf, lambdams, w, h, s, X, Y, Z);
}

// iAlignUp (Cudaflow.cu:985-994).
static int align_up(int n, int m) {
    int mod = n % m;
    return mod ? n + m - mod : n;
}

// Pyramid geometry (Cudaflow.cu:146-171): newW = (int)(W / fScale) in float.
static void pyramid(const Params& P, std::vector<int>& W, std::vector<int>& H, std::vector<int>& S) {
    W.resize(P.n_levels);
    H.resize(P.n_levels);
    S.resize(P.n_levels);
    int nh = P.height, nw = P.width, ns = align_up(P.width, P.stride_alignment);
    for (int l = 0; l < P.n_levels; ++l) {
        W[l] = nw;
        H[l] = nh;
        S[l] = ns;
        nh = (int)((float)nh / P.scale);
        nw = (int)((float)nw / P.scale);
        ns = align_up(nw, P.stride_alignment);
    }
}

// Flat device-buffer model: every buffer holds s0*h0 elements plus one zero guard row.
template <typename T>
struct Pool {
    long n, guard;
    std::vector<std::vector<T>> store;
    Pool(long n_, long guard_) : n(n_), guard(guard_) {}
    T* make() {
        store.emplace_back(n + guard, T(0));
        return store.back().data();
    }
};

template <typename T>
static void zero(T* p, long n) {
    std::fill(p, p + n, T(0));
}

template <typename T>
static void gray_input(const Params& P, const void* in, T* out, int s0) {
    if (P.input_type == 1) gray_rgb((const uint8_t*)in, out, P.width, P.height, s0);
    else if (P.input_type == 2) gray_u8((const uint8_t*)in, out, P.width, P.height, s0);
    else if (P.input_type == 3) gray_u16((const uint16_t*)in, out, P.width, P.height, s0);
    else gray_f32((const float*)in, out, P.width, P.height, s0);
}

// Warm start (this implementation's addition; the reference always starts from zero,
// Cudaflow.cu:467-468, Reconflow.cu:209-220): level-0 flow ws (stride s0) replaces the zero
// initial u, v at the coarsest level, downscaled there by the prior flow's operator
// (DownscaleScalingKernel straight from level 0, as Reconflow.cu:232-238 does for the prior). A
// warm start with a non-finite component, or one larger than max(w0, h0), over the w0 x h0
// pixels is not used.
inline bool warm_start_usable(const float* wsu, const float* wsv, int w0, int h0, int s0) {
    if (!wsu || !wsv) return false;
    const float bound = (float)std::max(w0, h0);
    for (int y = 0; y < h0; ++y)
        for (int x = 0; x < w0; ++x) {
            const long i = x + (long)y * s0;
            if (!(std::fabs(wsu[i]) <= bound) || !(std::fabs(wsv[i]) <= bound)) return false;
        }
    return true;
}

template <typename T>
void apply_warm_start(const float* wsu, const float* wsv, const std::vector<int>& pW, const std::vector<int>& pH,
                      const std::vector<int>& pS, Pool<T>& pool, T* u, T* v) {
    const int L = (int)pW.size() - 1;
    if (!warm_start_usable(wsu, wsv, pW[0], pH[0], pS[0])) return;
    const long n = (long)pS[0] * pH[0];
    T *a = pool.make(), *b = pool.make();
    for (long i = 0; i < n; ++i) {
        a[i] = (T)wsu[i];
        b[i] = (T)wsv[i];
    }
    const float sc = (float)pW[L] / (float)pW[0];
    downscale(a, pW[0], pH[0], pS[0], pW[L], pH[L], pS[L], (T)sc, true, u);
    downscale(b, pW[0], pH[0], pS[0], pW[L], pH[L], pS[L], (T)sc, true, v);
}

// A11: CudaFlow::_solveOpticalFlow for METHOD_TVL1 (Cudaflow.cu:403-621), with the
// pointer swaps mirrored. Images are stride x height, outputs u, v likewise.
// Method branches follow Cudaflow.cu:426-578: Charbonnier-gradient for methods 2 and 7 (with
// gradient pyramids, :426-461, and three warps / derivative sets per warp, :489-504),
// Charbonnier for method 1, patch derivatives for method 3, the L1 step otherwise. With mask0_in
// it is CudaFlow::_solveInpainting (Inpainting.cu:13-152): the masked L1 step, plain derivatives.
template <typename T>
int solve_flow(const Params& P, const void* i0, const void* i1, float* u_out, float* v_out, const float* mask0_in = nullptr,
               const float* wsu = nullptr, const float* wsv = nullptr) {
    std::vector<int> pW, pH, pS;
    pyramid(P, pW, pH, pS);
    const int w0 = pW[0], h0 = pH[0], s0 = pS[0];
    const long n = (long)s0 * h0;
    const int m = P.method;
    const bool cg = (m == 2 || m == 7), inpaint = (mask0_in != nullptr);
    Pool<T> pool(n, s0);
    std::vector<T*> pI0(P.n_levels), pI1(P.n_levels), pIx0(P.n_levels), pIx1(P.n_levels), pIy0(P.n_levels),
        pIy1(P.n_levels), pM0(P.n_levels);
    for (int l = 0; l < P.n_levels; ++l) {
        pI0[l] = pool.make();
        pI1[l] = pool.make();
        if (cg) {
            pIx0[l] = pool.make();
            pIx1[l] = pool.make();
            pIy0[l] = pool.make();
            pIy1[l] = pool.make();
        }
        if (inpaint) pM0[l] = pool.make();
    }
    T *i1warp = pool.make(), *dumed = pool.make(), *dvmed = pool.make(), *dumeds = pool.make(),
      *dvmeds = pool.make();
    T *du = pool.make(), *dv = pool.make(), *dus = pool.make(), *dvs = pool.make();
    T *pu1 = pool.make(), *pu2 = pool.make(), *pv1 = pool.make(), *pv2 = pool.make();
    T *pu1s = pool.make(), *pu2s = pool.make(), *pv1s = pool.make(), *pv2s = pool.make();
    T *Ix = pool.make(), *Iy = pool.make(), *Iz = pool.make();
    T *ix1warp = nullptr, *iy1warp = nullptr, *Ixx = nullptr, *Ixy = nullptr, *Ixz = nullptr, *Iyx = nullptr,
      *Iyy = nullptr, *Iyz = nullptr;
    if (cg) {
        ix1warp = pool.make();
        iy1warp = pool.make();
        Ixx = pool.make();
        Ixy = pool.make();
        Ixz = pool.make();
        Iyx = pool.make();
        Iyy = pool.make();
        Iyz = pool.make();
    }
    T *u = pool.make(), *v = pool.make(), *us = pool.make(), *vs = pool.make();
    (void)w0;

    gray_input(P, i0, pI0[0], s0);
    gray_input(P, i1, pI1[0], s0);
    round_group<T>(P, RND_IMG, n, {pI0[0], pI1[0]});
    const int mut = P.mutation;
    const int img16 = P.storage == 1 || P.storage == 4, der16 = P.storage >= 1;  // the TV-L1 duals: der16
    round_f16<T>(img16, n, {pI0[0], pI1[0]});
    if (cg) {
        deriv_single(pI0[0], pW[0], pH[0], s0, pIx0[0], pIy0[0]);
        deriv_single(pI1[0], pW[0], pH[0], s0, pIx1[0], pIy1[0]);
    }
    if (inpaint) gray_f32(mask0_in, pM0[0], pW[0], pH[0], s0);
    for (int l = 1; l < P.n_levels; ++l) {
        auto down = [&](std::vector<T*>& py) {
            downscale(py[l - 1], pW[l - 1], pH[l - 1], pS[l - 1], pW[l], pH[l], pS[l], T(1), false, py[l]);
        };
        down(pI0);
        down(pI1);
        round_group<T>(P, RND_IMG, n, {pI0[l], pI1[l]});
        round_f16<T>(img16, n, {pI0[l], pI1[l]});
        if (cg) {
            down(pIx0);
            down(pIx1);
            down(pIy0);
            down(pIy1);
        }
        if (inpaint) down(pM0);
    }
    zero(u, n);
    zero(v, n);
    apply_warm_start(wsu, wsv, pW, pH, pS, pool, u, v);
    const T lambda = P.lambda, lambdagrad = P.lambdagrad, theta = P.theta, tau = P.tau;
    for (int level = P.n_levels - 1; level >= 0; --level) {
        const int w = pW[level], h = pH[level], s = pS[level];
        for (int wi = 0; wi < P.n_warp_iters; ++wi) {
            if (!(mut == MUT_FLOW_NO_WARP_RESET && (wi > 0 || level < P.n_levels - 1)))
                for (T* b : {du, dv, dus, dvs, dumed, dvmed, dumeds, dvmeds, pu1, pu2, pv1, pv2}) zero(b, n);
            if (mut == MUT_FLOW_WARP_ZERO_FLOW)
                std::copy(pI1[level], pI1[level] + n, i1warp);
            else
                warp(pI1[level], w, h, s, u, v, i1warp);
            round_group<T>(P, RND_IMG, n, {i1warp});
            round_f16<T>(img16, n, {i1warp});
            if (cg) {
                warp(pIx1[level], w, h, s, u, v, ix1warp);
                warp(pIy1[level], w, h, s, u, v, iy1warp);
            }
            if (m == 3 && !inpaint)
                derivatives_patch(pI0[level], i1warp, w, h, s, Ix, Iy, Iz);
            else
                derivatives(pI0[level], i1warp, w, h, s, Ix, Iy, Iz);
            round_group<T>(P, RND_DER, n, {Ix, Iy, Iz});
            round_f16<T>(der16, n, {Ix, Iy, Iz});
            if (cg) {
                derivatives(pIx0[level], ix1warp, w, h, s, Ixx, Ixy, Ixz);
                derivatives(pIy0[level], iy1warp, w, h, s, Iyx, Iyy, Iyz);
            }
            for (int it = 0; it < P.n_solver_iters; ++it) {
                if (inpaint) {
                    data_l1_inpaint(dumed, dvmed, pM0[level], pu1, pu2, pv1, pv2, Ix, Iy, Iz, w, h, s, lambda, theta,
                                    dumeds, dvmeds);
                } else if (cg || m == 1) {
                    data_charb(du, dv, dumed, dvmed, pu1, pu2, pv1, pv2, Ix, Iy, Iz, Ixx, Ixy, Ixz, Iyx, Iyy, Iyz, w,
                               h, s, lambda, lambdagrad, theta, dus, dvs, dumeds, dvmeds);
                    std::swap(du, dus);
                    std::swap(dv, dvs);
                } else {
                    data_l1(dumed, dvmed, pu1, pu2, pv1, pv2, Ix, Iy, Iz, w, h, s, lambda, theta, dumeds, dvmeds);
                }
                if (mut != MUT_FLOW_NO_UHAT_SWAP) {
                    std::swap(dumed, dumeds);
                    std::swap(dvmed, dvmeds);
                }
                round_group<T>(P, RND_UHAT, n, {dumed, dvmed});
                dual_tv(dumed, dvmed, pu1, pu2, pv1, pv2, w, h, s, tau, theta, pu1s, pu2s, pv1s, pv2s);
                if (mut != MUT_FLOW_NO_DUAL_SWAP) {
                    std::swap(pu1, pu1s);
                    std::swap(pu2, pu2s);
                    std::swap(pv1, pv1s);
                    std::swap(pv2, pv2s);
                }
                round_group<T>(P, RND_DUAL, n, {pu1, pu2, pv1, pv2});
                round_f16<T>(der16, n, {pu1, pu2, pv1, pv2});
            }
            if (mut != MUT_FLOW_NO_MEDIAN) {
                median5(dumed, dvmed, w, h, s, dumeds, dvmeds);
                std::swap(dumed, dumeds);
                std::swap(dvmed, dvmeds);
            }
            add(u, dumed, (long)h * s, u);
            add(v, dvmed, (long)h * s, v);
            round_group<T>(P, RND_UV, n, {u, v});
        }
        if (level > 0) {
            T sc = mut == MUT_FLOW_UPSCALE_NO_SCALE ? T(1) : (T)P.scale;
            upscale(u, w, h, s, pW[level - 1], pH[level - 1], pS[level - 1], sc, us);
            upscale(v, w, h, s, pW[level - 1], pH[level - 1], pS[level - 1], sc, vs);
            std::swap(u, us);
            std::swap(v, vs);
            round_group<T>(P, RND_UV, n, {u, v});
        }
    }
    for (long i = 0; i < n; ++i) {
        u_out[i] = (float)u[i];
        v_out[i] = (float)v[i];
    }
    return 0;
}

// 3x3 (row-major) times [R|t] (3x4): the K*[R|t] product of Reconflow.cu:225-226.
static void kmul(const double* K, const double* R, const double* t, double* out) {
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            double acc = 0;
            for (int k = 0; k < 3; ++k) acc += K[i * 3 + k] * R[k * 3 + j];
            out[i * 4 + j] = acc;
        }
        double acc = 0;
        for (int k = 0; k < 3; ++k) acc += K[i * 3 + k] * t[k];
        out[i * 4 + 3] = acc;
    }
}

// Intrinsics pyramid of ReconFlow::setCameraMatrices (Reconflow.cu:78-116): levels >= 1
// keep only fx, fy, cx, cy (divided by fScale); K[2][2] stays 0 there.
// fix_coarse_k = 1 sets K[2][2] = 1 on every level (an option, not the reference).
static void k_pyramid(const double* K, int n_levels, float scale, int fix_coarse_k,
                      std::vector<std::vector<double>>& out) {
    out.assign(n_levels, std::vector<double>(9, 0.0));
    for (int i = 0; i < 9; ++i) out[0][i] = K[i];
    for (int l = 1; l < n_levels; ++l) {
        out[l][0] = out[l - 1][0] / (double)scale;
        out[l][2] = out[l - 1][2] / (double)scale;
        out[l][4] = out[l - 1][4] / (double)scale;
        out[l][5] = out[l - 1][5] / (double)scale;
        if (fix_coarse_k) out[l][8] = 1.0;
    }
}

// A19: ReconFlow::solveReconFlow for METHODR_TVL1_MS (Reconflow.cu:147-466) and, with a
// prior flow, METHODR_TVL1_MS_FN / _FNSPARSE (Reconflow.cu:232-241, 301-329). Pointer
// swaps mirrored, so the buffer Solve3d reads is always the one zeroed at solve start.
template <typename T>
int solve_recon(const Params& P, const double* K0, const double* K1, const double* R0, const double* t0,
                const double* R1, const double* t1, const void* i0, const void* i1, const float* ufn_in,
                const float* vfn_in, const float* mask_in, float* u_out, float* v_out, float* X_out,
                float* Y_out, float* Z_out, const float* wsu = nullptr, const float* wsv = nullptr) {
    std::vector<int> pW, pH, pS;
    pyramid(P, pW, pH, pS);
    const int h0 = pH[0], s0 = pS[0];
    const long n = (long)s0 * h0;
    Pool<T> pool(n, s0);
    std::vector<T*> pI0(P.n_levels), pI1(P.n_levels);
    for (int l = 0; l < P.n_levels; ++l) {
        pI0[l] = pool.make();
        pI1[l] = pool.make();
    }
    T *i1warp = pool.make(), *dumed = pool.make(), *dvmed = pool.make(), *dumeds = pool.make(),
      *dvmeds = pool.make();
    T *pu1 = pool.make(), *pu2 = pool.make(), *pv1 = pool.make(), *pv2 = pool.make();
    T *pu1s = pool.make(), *pu2s = pool.make(), *pv1s = pool.make(), *pv2s = pool.make();
    T *Ix = pool.make(), *Iy = pool.make(), *Iz = pool.make();
    T *u = pool.make(), *v = pool.make(), *us = pool.make(), *vs = pool.make();
    T *X = pool.make(), *Y = pool.make(), *Z = pool.make(), *Xs = pool.make(), *Ys = pool.make(),
      *Zs = pool.make(), *Xm = pool.make(), *Ym = pool.make(), *Zm = pool.make();
    T *uproj = pool.make(), *vproj = pool.make(), *uc = pool.make(), *vc = pool.make();
    T *sku = pool.make(), *skv = pool.make(), *skus = pool.make(), *skvs = pool.make();
    const bool fn = (P.method == 17 || P.method == 20) && ufn_in && vfn_in;
    T *ufn = nullptr, *vfn = nullptr, *ufn_l = nullptr, *vfn_l = nullptr, *mask = nullptr, *mask_l = nullptr;
    if (fn) {
        ufn = pool.make();
        vfn = pool.make();
        ufn_l = pool.make();
        vfn_l = pool.make();
        mask = pool.make();
        mask_l = pool.make();
        for (long i = 0; i < n; ++i) {
            ufn[i] = ufn_in[i];
            vfn[i] = vfn_in[i];
            mask[i] = mask_in ? mask_in[i] : 1.0f;
        }
    }

    std::vector<std::vector<double>> pK0, pK1;
    k_pyramid(K0, P.n_levels, P.scale, P.fix_coarse_k, pK0);
    k_pyramid(K1, P.n_levels, P.scale, P.fix_coarse_k, pK1);

    const int mut = P.mutation;
    const int img16 = P.storage == 1 || P.storage == 4, der16 = P.storage >= 1;
    const int dual16 = P.storage >= 3;  // the joint duals in binary16: storage 3 and 4 (measurement modes)
    // binary64 blends: the FP64 mode (tri_f64) takes the projection and data blends along with the
    // triangulation; proj_f64 selects them one by one. The FP64 solve computes in binary64 anyway.
    const int prec = sizeof(T) == sizeof(float) ? (P.proj_f64 | (P.tri_f64 ? PREC_PROJ | PREC_BLEND : 0)) : 0;
    gray_input(P, i0, pI0[0], s0);
    gray_input(P, i1, pI1[0], s0);
    round_group<T>(P, RND_IMG, n, {pI0[0], pI1[0]});
    round_f16<T>(img16, n, {pI0[0], pI1[0]});
    for (int l = 1; l < P.n_levels; ++l) {
        downscale(pI0[l - 1], pW[l - 1], pH[l - 1], pS[l - 1], pW[l], pH[l], pS[l], T(1), false, pI0[l]);
        downscale(pI1[l - 1], pW[l - 1], pH[l - 1], pS[l - 1], pW[l], pH[l], pS[l], T(1), false, pI1[l]);
        round_group<T>(P, RND_IMG, n, {pI0[l], pI1[l]});
        round_f16<T>(img16, n, {pI0[l], pI1[l]});
    }
    for (T* b : {u, v, uproj, vproj, uc, vc, X, Y, Z, sku, skv}) zero(b, n);
    apply_warm_start(wsu, wsv, pW, pH, pS, pool, u, v);

    const T lambda = P.lambda, tau = P.tau, alphaTv = P.alpha_tv, alphaProj = P.alpha_proj;
    const T thetaTv = (T)(1.0f / P.alpha_tv);  // Reconflow.cu:32
    const T lambdaf = P.lambdaf, lambdams = P.lambdams, alphaFn = P.alpha_fn;
    for (int level = P.n_levels - 1; level >= 0; --level) {
        const int w = pW[level], h = pH[level], s = pS[level];
        double Pm[12], Qm[12];
        kmul(pK0[level].data(), R0, t0, Pm);
        kmul(pK1[level].data(), R1, t1, Qm);
        const double* Kuc = mut == MUT_RECON_UC_LEVEL0_K ? pK1[0].data() : pK1[level].data();
        if (mut == MUT_RECON_XYZ_LEVEL_RESET)
            for (T* b : {X, Y, Z, Xs, Ys, Zs}) zero(b, n);
        if (fn) {
            float sc = (float)pW[level] / (float)pW[0];
            downscale(ufn, pW[0], pH[0], pS[0], w, h, s, (T)sc, true, ufn_l);
            downscale(vfn, pW[0], pH[0], pS[0], w, h, s, (T)sc, true, vfn_l);
            downscale(mask, pW[0], pH[0], pS[0], w, h, s, T(1), false, mask_l);
        }
        for (int wi = 0; wi < P.n_warp_iters; ++wi) {
            if (!(mut == MUT_RECON_NO_WARP_RESET && (wi > 0 || level < P.n_levels - 1)))
                for (T* b : {dumed, dvmed, dumeds, dvmeds, pu1, pu2, pv1, pv2}) zero(b, n);
            if (mut == MUT_RECON_SK_RESET)
                for (T* b : {sku, skv}) zero(b, n);
            F32Fields<T> q;
            q.m = (long)(h + 1) * s;
            auto a32 = [&](int bit) { return sizeof(T) == sizeof(double) && (P.f32_steps & bit); };
            if (a32(STEP_WARP)) {
                warp<float>(q.in(0, pI1[level]), w, h, s, q.in(1, u), q.in(2, v), q.in(3, i1warp));
                q.out(3, i1warp);
            } else {
                warp(pI1[level], w, h, s, u, v, i1warp);
            }
            round_group<T>(P, RND_IMG, n, {i1warp});
            round_f16<T>(img16, n, {i1warp});
            if (a32(STEP_DERIV)) {
                derivatives<float>(q.in(0, pI0[level]), q.in(1, i1warp), w, h, s, q.in(2, Ix), q.in(3, Iy),
                                   q.in(4, Iz));
                q.out(2, Ix);
                q.out(3, Iy);
                q.out(4, Iz);
            } else {
                derivatives(pI0[level], i1warp, w, h, s, Ix, Iy, Iz);
            }
            round_group<T>(P, RND_DER, n, {Ix, Iy, Iz});
            round_f16<T>(der16, n, {Ix, Iy, Iz});
            for (int it = 0; it < P.n_solver_iters; ++it) {
                if (fn && (prec & PREC_BLEND))
                    recon_data_l1_fn_wide(u, v, dumed, dvmed, uproj, vproj, ufn_l, vfn_l, sku, skv, pu1, pu2, pv1,
                                          pv2, Ix, Iy, Iz, w, h, s, lambda, alphaTv, alphaProj, alphaFn,
                                          P.method == 20 ? mask_l : nullptr, dumeds, dvmeds);
                else if (fn)
                    recon_data_l1_fn(u, v, dumed, dvmed, uproj, vproj, ufn_l, vfn_l, sku, skv, pu1, pu2, pv1,
                                     pv2, Ix, Iy, Iz, w, h, s, lambda, alphaTv, alphaProj, alphaFn,
                                     P.method == 20 ? mask_l : nullptr, dumeds, dvmeds);
                else if (a32(STEP_DATA)) {
                    recon_data_l1<float>(q.in(0, u), q.in(1, v), q.in(2, dumed), q.in(3, dvmed), q.in(4, uproj),
                                         q.in(5, vproj), q.in(6, sku), q.in(7, skv), q.in(8, pu1), q.in(9, pu2),
                                         q.in(10, pv1), q.in(11, pv2), q.in(12, Ix), q.in(13, Iy), q.in(14, Iz), w,
                                         h, s, (float)lambda, (float)alphaTv, (float)alphaProj, q.in(0, dumeds),
                                         q.in(1, dvmeds));
                    q.out(0, dumeds);
                    q.out(1, dvmeds);
                } else if (prec & (PREC_BLEND | PREC_DATA))
                    recon_data_l1_wide(u, v, dumed, dvmed, uproj, vproj, sku, skv, pu1, pu2, pv1, pv2, Ix, Iy, Iz,
                                       w, h, s, lambda, alphaTv, alphaProj, (prec & PREC_DATA) != 0, dumeds,
                                       dvmeds);
                else
                    recon_data_l1(u, v, dumed, dvmed, uproj, vproj, sku, skv, pu1, pu2, pv1, pv2, Ix, Iy, Iz,
                                  w, h, s, lambda, alphaTv, alphaProj, dumeds, dvmeds);
                std::swap(dumed, dumeds);
                std::swap(dvmed, dvmeds);
                round_group<T>(P, RND_UHAT, n, {dumed, dvmed});
                if (a32(STEP_DUAL)) {
                    dual_tv<float>(q.in(0, dumed), q.in(1, dvmed), q.in(2, pu1), q.in(3, pu2), q.in(4, pv1),
                                   q.in(5, pv2), w, h, s, (float)tau, (float)thetaTv, q.in(6, pu1s), q.in(7, pu2s),
                                   q.in(8, pv1s), q.in(9, pv2s));
                    q.out(6, pu1s);
                    q.out(7, pu2s);
                    q.out(8, pv1s);
                    q.out(9, pv2s);
                } else
                dual_tv(dumed, dvmed, pu1, pu2, pv1, pv2, w, h, s, tau, thetaTv, pu1s, pu2s, pv1s, pv2s);
                std::swap(pu1, pu1s);
                std::swap(pu2, pu2s);
                std::swap(pv1, pv1s);
                std::swap(pv2, pv2s);
                round_group<T>(P, RND_DUAL, n, {pu1, pu2, pv1, pv2});
                round_f16<T>(dual16, n, {pu1, pu2, pv1, pv2});
                if (a32(STEP_PROJ)) {
                    uvproj<float>(q.in(0, u), q.in(1, v), q.in(2, dumed), q.in(3, dvmed), q.in(4, uc), q.in(5, vc),
                                  q.in(6, sku), q.in(7, skv), w, h, s, (float)lambdaf, (float)alphaProj,
                                  q.in(8, uproj), q.in(9, vproj));
                    q.out(8, uproj);
                    q.out(9, vproj);
                } else if (prec & PREC_PROJ)
                    uvproj_wide(u, v, dumed, dvmed, uc, vc, sku, skv, w, h, s, lambdaf, alphaProj, uproj, vproj);
                else
                    uvproj(u, v, dumed, dvmed, uc, vc, sku, skv, w, h, s, lambdaf, alphaProj, uproj, vproj);
                round_group<T>(P, RND_PROJ, n, {uproj, vproj});
                if (a32(STEP_TRI)) {  // float fields; FP64 arithmetic inside with tri_f64, as on the GPU
                    float *a = q.in(0, uproj), *b = q.in(1, vproj), *x0 = q.in(2, X), *y0 = q.in(3, Y),
                          *z0 = q.in(4, Z), *x1 = q.in(5, Xs), *y1 = q.in(6, Ys), *z1 = q.in(7, Zs);
                    if (P.tri_f64)
                        solve3d_wide<float>(a, b, x0, y0, z0, Pm, Qm, (float)lambdaf, (float)lambdams, w, h, s, x1,
                                            y1, z1);
                    else
                        solve3d<float>(a, b, x0, y0, z0, Pm, Qm, (float)lambdaf, (float)lambdams, w, h, s, x1, y1,
                                       z1);
                    q.out(5, Xs);
                    q.out(6, Ys);
                    q.out(7, Zs);
                } else if (P.tri_f64)
                    solve3d_wide(uproj, vproj, X, Y, Z, Pm, Qm, lambdaf, lambdams, w, h, s, Xs, Ys, Zs);
                else
                    solve3d(uproj, vproj, X, Y, Z, Pm, Qm, lambdaf, lambdams, w, h, s, Xs, Ys, Zs);
                std::swap(X, Xs);
                std::swap(Y, Ys);
                std::swap(Z, Zs);
                round_group<T>(P, RND_XYZ, n, {X, Y, Z});
                if (mut != MUT_RECON_NO_CLEANUP && a32(STEP_CLEAN)) {
                    float *x = q.in(0, X), *y = q.in(1, Y), *z = q.in(2, Z), *xm = q.in(3, Xm), *ym = q.in(4, Ym),
                          *zm = q.in(5, Zm);
                    median3d<float>(x, y, z, w, h, s, xm, ym, zm);
                    cleanup3d<float>(x, y, z, xm, ym, zm, w, h, s, q.in(6, Xs), q.in(7, Ys), q.in(8, Zs));
                    q.out(3, Xm);
                    q.out(4, Ym);
                    q.out(5, Zm);
                    q.out(6, Xs);
                    q.out(7, Ys);
                    q.out(8, Zs);
                    std::swap(X, Xs);
                    std::swap(Y, Ys);
                    std::swap(Z, Zs);
                    round_group<T>(P, RND_XYZ, n, {X, Y, Z});
                } else if (mut != MUT_RECON_NO_CLEANUP) {
                    median3d(X, Y, Z, w, h, s, Xm, Ym, Zm);
                    cleanup3d(X, Y, Z, Xm, Ym, Zm, w, h, s, Xs, Ys, Zs);
                    std::swap(X, Xs);
                    std::swap(Y, Ys);
                    std::swap(Z, Zs);
                    round_group<T>(P, RND_XYZ, n, {X, Y, Z});
                }
                if (a32(STEP_UC)) {
                    solve_uc<float>(q.in(0, X), q.in(1, Y), q.in(2, Z), R1, t1, Kuc, w, h, s, q.in(3, uc),
                                    q.in(4, vc));
                    q.out(3, uc);
                    q.out(4, vc);
                } else
                solve_uc(X, Y, Z, R1, t1, Kuc, w, h, s, uc, vc);
                round_group<T>(P, RND_UC, n, {uc, vc});
                if (a32(STEP_SK)) {
                    update_sk<float>(q.in(0, sku), q.in(1, skv), q.in(2, u), q.in(3, v), q.in(4, dumed),
                                     q.in(5, dvmed), q.in(6, uproj), q.in(7, vproj), w, h, s, q.in(8, skus),
                                     q.in(9, skvs));
                    q.out(8, skus);
                    q.out(9, skvs);
                } else if (prec & PREC_SK)
                    update_sk_wide(sku, skv, u, v, dumed, dvmed, uproj, vproj, w, h, s, skus, skvs);
                else
                    update_sk(sku, skv, u, v, dumed, dvmed, uproj, vproj, w, h, s, skus, skvs);
                if (mut != MUT_RECON_NO_SK_SWAP) {
                    std::swap(sku, skus);
                    std::swap(skv, skvs);
                }
                round_group<T>(P, RND_SK, n, {sku, skv});
            }
            if (a32(STEP_MED)) {
                float *a = q.in(0, dumed), *b = q.in(1, dvmed), *c = q.in(2, dumeds), *d = q.in(3, dvmeds),
                      *uu = q.in(4, u), *vv = q.in(5, v);
                median5<float>(a, b, w, h, s, c, d);
                add<float>(uu, c, (long)h * s, uu);
                add<float>(vv, d, (long)h * s, vv);
                q.out(2, dumeds);
                q.out(3, dvmeds);
                q.out(4, u);
                q.out(5, v);
            } else
            median5(dumed, dvmed, w, h, s, dumeds, dvmeds);
            std::swap(dumed, dumeds);
            std::swap(dvmed, dvmeds);
            if (!a32(STEP_MED)) {
                add(u, dumed, (long)h * s, u);
                add(v, dvmed, (long)h * s, v);
            }
            if (mut != MUT_RECON_NO_UV_MEDIAN) {
                median5(u, v, w, h, s, dumeds, dvmeds);
                std::swap(u, dumeds);
                std::swap(v, dvmeds);
            }
            round_group<T>(P, RND_UV, n, {u, v});
        }
        if (level > 0) {
            const int W = pW[level - 1], H = pH[level - 1], S = pS[level - 1];
            T sc = P.scale;
            if (sizeof(T) == sizeof(double) && (P.f32_steps & STEP_UPS)) {
                // FP32 arithmetic for every upscale of this level change
                auto up32 = [&](T* src, T* dst) {
                    const long m = (long)(H + 1) * S, mc = (long)(h + 1) * s;
                    std::vector<float> a(src, src + std::max(m, mc)), b(dst, dst + m);
                    for (long i = 0; i < (long)a.size(); ++i) a[i] = (float)src[i];
                    for (long i = 0; i < m; ++i) b[i] = (float)dst[i];
                    upscale<float>(a.data(), w, h, s, W, H, S, (float)sc, b.data());
                    for (long i = 0; i < m; ++i) dst[i] = (T)b[i];
                };
                up32(u, us);
                up32(v, vs);
            } else {
            upscale(u, w, h, s, W, H, S, sc, us);
            upscale(v, w, h, s, W, H, S, sc, vs);
            }
            std::swap(u, us);
            std::swap(v, vs);
            if (mut != MUT_RECON_NO_UC_UPSCALE) {
                upscale(uc, w, h, s, W, H, S, sc, us);
                upscale(vc, w, h, s, W, H, S, sc, vs);
                std::swap(uc, us);
                std::swap(vc, vs);
            }
            if (mut != MUT_RECON_NO_PROJ_UPSCALE) {
                upscale(uproj, w, h, s, W, H, S, sc, us);
                upscale(vproj, w, h, s, W, H, S, sc, vs);
                std::swap(uproj, us);
                std::swap(vproj, vs);
            }
            upscale(sku, w, h, s, W, H, S, sc, us);
            upscale(skv, w, h, s, W, H, S, sc, vs);
            std::swap(sku, us);
            std::swap(skv, vs);
            round_group<T>(P, RND_UV, n, {u, v});
            round_group<T>(P, RND_UC, n, {uc, vc});
            round_group<T>(P, RND_PROJ, n, {uproj, vproj});
            round_group<T>(P, RND_SK, n, {sku, skv});
        }
    }
    for (long i = 0; i < n; ++i) {
        u_out[i] = (float)u[i];
        v_out[i] = (float)v[i];
        X_out[i] = (float)X[i];
        Y_out[i] = (float)Y[i];
        Z_out[i] = (float)Z[i];
    }
    return 0;
}

}  // namespace orc

// ---------------------------------------------------------------------------------
// C ABI for the Python test harness (ctypes). Suffix _f64 / _f32 selects T.
extern "C" {

void orc_gray_rgb_f64(const uint8_t* rgb, double* out, int w, int h, int s) { orc::gray_rgb(rgb, out, w, h, s); }
void orc_gray_rgb_f32(const uint8_t* rgb, float* out, int w, int h, int s) { orc::gray_rgb(rgb, out, w, h, s); }
void orc_gray_u8_f64(const uint8_t* in, double* out, int w, int h, int s) { orc::gray_u8(in, out, w, h, s); }
void orc_gray_u8_f32(const uint8_t* in, float* out, int w, int h, int s) { orc::gray_u8(in, out, w, h, s); }
void orc_gray_u16_f64(const uint16_t* in, double* out, int w, int h, int s) { orc::gray_u16(in, out, w, h, s); }
void orc_gray_u16_f32(const uint16_t* in, float* out, int w, int h, int s) { orc::gray_u16(in, out, w, h, s); }

#define DEF2(T, SUF)                                                                                        \
    void orc_downscale_##SUF(const T* src, int W, int H, int S, int w, int h, int s, T* out) {               \
        orc::downscale<T>(src, W, H, S, w, h, s, T(1), false, out);                                          \
    }                                                                                                        \
    void orc_downscale_scaled_##SUF(const T* src, int W, int H, int S, int w, int h, int s, float scale,     \
                                    T* out) {                                                                \
        orc::downscale<T>(src, W, H, S, w, h, s, (T)scale, true, out);                                       \
    }                                                                                                        \
    void orc_upscale_##SUF(const T* src, int w, int h, int s, int W, int H, int S, float scale, T* out) {    \
        orc::upscale<T>(src, w, h, s, W, H, S, (T)scale, out);                                               \
    }                                                                                                        \
    void orc_warp_##SUF(const T* src, int w, int h, int s, const T* u, const T* v, T* out) {                 \
        orc::warp<T>(src, w, h, s, u, v, out);                                                               \
    }                                                                                                        \
    void orc_derivatives_##SUF(const T* I0, const T* I1, int w, int h, int s, T* Ix, T* Iy, T* Iz) {         \
        orc::derivatives<T>(I0, I1, w, h, s, Ix, Iy, Iz);                                                    \
    }                                                                                                        \
    void orc_data_l1_##SUF(const T* du, const T* dv, const T* pu1, const T* pu2, const T* pv1, const T* pv2, \
                           const T* Ix, const T* Iy, const T* Iz, int w, int h, int s, float lambda,         \
                           float theta, T* du1, T* dv1) {                                                    \
        orc::data_l1<T>(du, dv, pu1, pu2, pv1, pv2, Ix, Iy, Iz, w, h, s, (T)lambda, (T)theta, du1, dv1);     \
    }                                                                                                        \
    void orc_dual_tv_##SUF(const T* du, const T* dv, const T* pu1, const T* pu2, const T* pv1, const T* pv2, \
                           int w, int h, int s, float tau, float theta, T* a, T* b, T* c, T* d) {            \
        orc::dual_tv<T>(du, dv, pu1, pu2, pv1, pv2, w, h, s, (T)tau, (T)theta, a, b, c, d);                  \
    }                                                                                                        \
    void orc_median5_##SUF(const T* u, const T* v, int w, int h, int s, T* ou, T* ov) {                      \
        orc::median5<T>(u, v, w, h, s, ou, ov);                                                              \
    }                                                                                                        \
    void orc_median3d_##SUF(const T* X, const T* Y, const T* Z, int w, int h, int s, T* a, T* b, T* c) {     \
        orc::median3d<T>(X, Y, Z, w, h, s, a, b, c);                                                         \
    }                                                                                                        \
    void orc_add_##SUF(const T* a, const T* b, long count, T* out) { orc::add<T>(a, b, count, out); }        \
    void orc_recon_data_l1_##SUF(const T* u, const T* v, const T* du, const T* dv, const T* up, const T* vp, \
                                 const T* sku, const T* skv, const T* pu1, const T* pu2, const T* pv1,       \
                                 const T* pv2, const T* Ix, const T* Iy, const T* Iz, int w, int h, int s,   \
                                 float lambda, float aTv, float aP, T* du1, T* dv1) {                        \
        orc::recon_data_l1<T>(u, v, du, dv, up, vp, sku, skv, pu1, pu2, pv1, pv2, Ix, Iy, Iz, w, h, s,      \
                              (T)lambda, (T)aTv, (T)aP, du1, dv1);                                           \
    }                                                                                                        \
    void orc_recon_data_l1_fn_##SUF(const T* u, const T* v, const T* du, const T* dv, const T* up,           \
                                    const T* vp, const T* ufn, const T* vfn, const T* sku, const T* skv,     \
                                    const T* pu1, const T* pu2, const T* pv1, const T* pv2, const T* Ix,     \
                                    const T* Iy, const T* Iz, int w, int h, int s, float lambda, float aTv,  \
                                    float aP, float aFn, const T* mask, T* du1, T* dv1) {                    \
        orc::recon_data_l1_fn<T>(u, v, du, dv, up, vp, ufn, vfn, sku, skv, pu1, pu2, pv1, pv2, Ix, Iy, Iz,  \
                                 w, h, s, (T)lambda, (T)aTv, (T)aP, (T)aFn, mask, du1, dv1);                 \
    }                                                                                                        \
    void orc_uvproj_##SUF(const T* u, const T* v, const T* du, const T* dv, const T* uc, const T* vc,        \
                          const T* sku, const T* skv, int w, int h, int s, float lf, float aP, T* up, T* vp) {\
        orc::uvproj<T>(u, v, du, dv, uc, vc, sku, skv, w, h, s, (T)lf, (T)aP, up, vp);                      \
    }                                                                                                        \
    void orc_solve3d_##SUF(const T* up, const T* vp, const T* X0, const T* Y0, const T* Z0, const double* P,  \
                           const double* Q, float lf, float lms, int w, int h, int s, T* X, T* Y, T* Z) {    \
        orc::solve3d<T>(up, vp, X0, Y0, Z0, P, Q, (T)lf, (T)lms, w, h, s, X, Y, Z);                          \
    }                                                                                                        \
    void orc_solve3d_fma_##SUF(const T* up, const T* vp, const T* X0, const T* Y0, const T* Z0,             \
                               const double* P, const double* Q, float lf, float lms, int w, int h, int s,   \
                               T* X, T* Y, T* Z) {                                                           \
        orc::solve3d_fma<T>(up, vp, X0, Y0, Z0, P, Q, (T)lf, (T)lms, w, h, s, X, Y, Z);                      \
    }                                                                                                        \
    void orc_cleanup3d_##SUF(const T* X0, const T* Y0, const T* Z0, const T* Xm, const T* Ym, const T* Zm,   \
                             int w, int h, int s, T* X1, T* Y1, T* Z1) {                                     \
        orc::cleanup3d<T>(X0, Y0, Z0, Xm, Ym, Zm, w, h, s, X1, Y1, Z1);                                      \
    }                                                                                                        \
    void orc_solve_uc_##SUF(const T* X, const T* Y, const T* Z, const double* R, const double* t,            \
                            const double* K1, int w, int h, int s, T* uc, T* vc) {                           \
        orc::solve_uc<T>(X, Y, Z, R, t, K1, w, h, s, uc, vc);                                                \
    }                                                                                                        \
    void orc_update_sk_##SUF(const T* sku, const T* skv, const T* u, const T* v, const T* du, const T* dv,   \
                             const T* up, const T* vp, int w, int h, int s, T* sku1, T* skv1) {              \
        orc::update_sk<T>(sku, skv, u, v, du, dv, up, vp, w, h, s, sku1, skv1);                              \
    }                                                                                                        \
    int orc_solve_flow_##SUF(const orc::Params* P, const void* i0, const void* i1, const float* wsu,        \
                             const float* wsv, float* u, float* v) {                                         \
        return orc::solve_flow<T>(*P, i0, i1, u, v, nullptr, wsu, wsv);                                      \
    }                                                                                                        \
    int orc_solve_recon_##SUF(const orc::Params* P, const double* K0, const double* K1, const double* R0,    \
                              const double* t0, const double* R1, const double* t1, const void* i0,          \
                              const void* i1, const float* ufn, const float* vfn, const float* mask,         \
                              const float* wsu, const float* wsv, float* u, float* v, float* X, float* Y,    \
                              float* Z) {                                                                    \
        return orc::solve_recon<T>(*P, K0, K1, R0, t0, R1, t1, i0, i1, ufn, vfn, mask, u, v, X, Y, Z, wsu,  \
                                   wsv);                                                                     \
    }

DEF2(double, f64)
DEF2(float, f32)

int orc_warm_start_usable(const float* wsu, const float* wsv, int w0, int h0, int s0) {
    return orc::warm_start_usable(wsu, wsv, w0, h0, s0) ? 1 : 0;
}

#define DEF3(T, SUF)                                                                                          \
    void orc_scene_flow_to_hsv_##SUF(const T* sf, int w, int h, int s, T* rgb, float fs) {                     \
        orc::scene_flow_to_hsv<T>(sf, w, h, s, rgb, (T)fs);                                                  \
    }                                                                                                        \
    void orc_flow_to_hsv_##SUF(const T* u, const T* v, int w, int h, int s, T* rgb, float fs) {              \
        orc::flow_to_hsv<T>(u, v, w, h, s, rgb, (T)fs);                                                      \
    }                                                                                                        \
    void orc_remove_edge_##SUF(const T* X0, const T* Y0, const T* Z0, int range, float maxarea, int w, int h,  \
                               int s, T* X1, T* Y1, T* Z1) {                                                 \
        orc::remove_edge<T>(X0, Y0, Z0, range, (T)maxarea, w, h, s, X1, Y1, Z1);                              \
    }                                                                                                        \
    void orc_deriv_single_##SUF(const T* I, int w, int h, int s, T* Ix, T* Iy) {                             \
        orc::deriv_single<T>(I, w, h, s, Ix, Iy);                                                            \
    }                                                                                                        \
    void orc_derivatives_patch_##SUF(const T* I0, const T* I1, int w, int h, int s, T* Ix, T* Iy, T* Iz) {   \
        orc::derivatives_patch<T>(I0, I1, w, h, s, Ix, Iy, Iz);                                              \
    }                                                                                                        \
    void orc_data_charb_##SUF(const T* du0, const T* dv0, const T* duh, const T* dvh, const T* pu1,          \
                              const T* pu2, const T* pv1, const T* pv2, const T* Ix, const T* Iy,            \
                              const T* It, const T* Ixx, const T* Ixy, const T* Ixt, const T* Iyx,           \
                              const T* Iyy, const T* Iyt, int w, int h, int s, float lambda, float lgrad,    \
                              float theta, T* du1, T* dv1, T* duh1, T* dvh1) {                               \
        orc::data_charb<T>(du0, dv0, duh, dvh, pu1, pu2, pv1, pv2, Ix, Iy, It, Ixx, Ixy, Ixt, Iyx, Iyy, Iyt, \
                           w, h, s, (T)lambda, (T)lgrad, (T)theta, du1, dv1, duh1, dvh1);                    \
    }                                                                                                        \
    void orc_data_l1_inpaint_##SUF(const T* du, const T* dv, const T* m0, const T* pu1, const T* pu2,        \
                                   const T* pv1, const T* pv2, const T* Ix, const T* Iy, const T* Iz, int w, \
                                   int h, int s, float lambda, float theta, T* du1, T* dv1) {                \
        orc::data_l1_inpaint<T>(du, dv, m0, pu1, pu2, pv1, pv2, Ix, Iy, Iz, w, h, s, (T)lambda, (T)theta,    \
                                du1, dv1);                                                                   \
    }                                                                                                        \
    void orc_scene_flow_##SUF(const T* u, const T* v, const T* d0, const T* d1, float fx, float fy,          \
                              float cx, float cy, int w, int h, int s, T* pc0, T* pc1, T* sf) {              \
        orc::scene_flow<T>(u, v, d0, d1, (T)fx, (T)fy, (T)cx, (T)cy, w, h, s, pc0, pc1, sf);                 \
    }                                                                                                        \
    int orc_solve_flow_masked_##SUF(const orc::Params* P, const void* i0, const void* i1, const float* mask0, \
                                    float* u, float* v) {                                                    \
        return orc::solve_flow<T>(*P, i0, i1, u, v, mask0);                                                  \
    }

DEF3(double, f64)
DEF3(float, f32)

void orc_pyramid(const orc::Params* P, int* W, int* H, int* S) {
    std::vector<int> pW, pH, pS;
    orc::pyramid(*P, pW, pH, pS);
    for (int l = 0; l < P->n_levels; ++l) {
        W[l] = pW[l];
        H[l] = pH[l];
        S[l] = pS[l];
    }
}

int orc_params_size(void) { return (int)sizeof(orc::Params); }

// The binary16 rounding of the FP16 storage modes, elementwise (for its pin).
void orc_round_half(const float* in, float* out, long n) {
    for (long i = 0; i < n; ++i) out[i] = orc::round_half(in[i]);
}

// The per-level intrinsics and camera matrices the joint solve uses (setCameraMatrices,
// Reconflow.cu:78-116, and K*[R|t], Reconflow.cu:225-228): Kl (n_levels x 9) and Pl (n_levels x 12).
void orc_camera_pyramid(const double* K, const double* R, const double* t, int n_levels, float scale,
                        int fix_coarse_k, double* Kl, double* Pl) {
    std::vector<std::vector<double>> pK;
    orc::k_pyramid(K, n_levels, scale, fix_coarse_k, pK);
    for (int l = 0; l < n_levels; ++l) {
        for (int i = 0; i < 9; ++i) Kl[l * 9 + i] = pK[l][i];
        orc::kmul(pK[l].data(), R, t, Pl + l * 12);
    }
}

// Host threads of the OpenMP loops (for timing the restatement at a stated core count).
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
