// Per-pixel device arithmetic of the TV-L1 / joint flow-depth solver for gfx950.
//
// Each function evaluates one reference kernel's formula for one pixel, in the reference's
// operation order; the library is compiled with -ffp-contract=off and IEEE division and
// square root, so the plain and the fused kernels (which share these functions) produce
// identical FP32 results. Citations are to the reference kernels each one follows.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rf {

// Mirror addressing with edge duplication (texture Mirror mode, e.g. warpImage.cu:40-43):
// period 2n, -1 -> 0, n -> n-1.
__device__ __forceinline__ int refl(int i, int n) {
    if ((unsigned)i < (unsigned)n) return i;
    int p = 2 * n;
    int m = i % p;
    if (m < 0) m += p;
    return m < n ? m : p - 1 - m;
}

__device__ __forceinline__ int floor_div(int a, int b) {  // b > 0
    int q = a / b;
    if ((a % b != 0) && (a < 0)) --q;
    return q;
}

// Bilinear texel blend with the texture-fetch weight formula
// (1-a)(1-b)T00 + a(1-b)T10 + (1-a)bT01 + abT11, summed left to right.
__device__ __forceinline__ float blend(float t00, float t10, float t01, float t11, float fx, float fy) {
    float w00 = (1.0f - fx) * (1.0f - fy);
    float w10 = fx * (1.0f - fy);
    float w01 = (1.0f - fx) * fy;
    float w11 = fx * fy;
    float r = w00 * t00;
    r = r + w10 * t10;
    r = r + w01 * t01;
    r = r + w11 * t11;
    return r;
}

template <typename TI>
__device__ __forceinline__ float ld(const TI* p, long i) {
    return (float)p[i];
}

template <typename TI>
__device__ __forceinline__ float bilin(const TI* img, int w, int h, int s, int x0, float fx, int y0, float fy) {
    int xa = refl(x0, w), xb = refl(x0 + 1, w), ya = refl(y0, h), yb = refl(y0 + 1, h);
    return blend(ld(img, xa + (long)ya * s), ld(img, xb + (long)ya * s), ld(img, xa + (long)yb * s),
                 ld(img, xb + (long)yb * s), fx, fy);
}

// Sample at the rational texel coordinate (nx/dx, ny/dy): exact integer split, weight
// fractions rounded once.
template <typename TI>
__device__ __forceinline__ float sample_ratio(const TI* img, int w, int h, int s, int nx, int dx, int ny, int dy) {
    int x0 = floor_div(nx, dx), y0 = floor_div(ny, dy);
    float fx = (float)(nx - x0 * dx) / (float)dx;
    float fy = (float)(ny - y0 * dy) / (float)dy;
    return bilin(img, w, h, s, x0, fx, y0, fy);
}

// Warp sample S(I, ix + u, iy + v) (warpImage.cu:25-28): the fraction u - floor(u) is exact.
template <typename TI>
__device__ __forceinline__ float warp_sample(const TI* img, int w, int h, int s, int ix, int iy, float u, float v) {
    if (!(fabsf(u) <= 16777216.0f) || !(fabsf(v) <= 16777216.0f)) return __builtin_nanf("");
    float fu = floorf(u), fv = floorf(v);
    return bilin(img, w, h, s, ix + (int)fu, u - fu, iy + (int)fv, v - fv);
}

// 5-point derivative (computeDerivatives.cu:26-30): ((a - 8b) + 8c - d) / 12.
__device__ __forceinline__ float d5(float m2, float m1, float p1, float p2) {
    float t = m2;
    t -= m1 * 8.0f;
    t += p1 * 8.0f;
    t -= p2;
    t /= 12.0f;
    return t;
}

// Division by a launch-uniform divisor d (finite, d > 0) that returns the IEEE FP32 quotient n / d
// for every FP32 n, without a per-division reciprocal. With rd = 1/d rounded to nearest in binary64
// (formed once by the host) and C = 1.5 * 2^-124:
//   s = RN64(n rd + C) (one FMA),  q = RN64(s - C),  result RN32(q) with the sign bit of n or-ed in.
// n rd = (n/d)(1 + e), |e| <= 2^-53. Below 2^-125 in magnitude, s lies in [2^-124, 2^-123), so q is
// n rd rounded to the 2^-176 grid (the subtraction is exact); above, q is within 2^-51 |n/d|.
// For FP32 n = N 2^a, d = D 2^b (N, D odd, below 2^24), n/d either lies exactly on an FP32
// rounding boundary, which needs |n/d| < 2^-126 and puts n/d on the 2^-176 grid, at most
// 2^-179 from n rd, so q = n/d and RN32 applies its tie rule to the exact value; or lies at
// least min(2^-174, 2^-24 |n/d|) (below 2^-125) or 2^-49 |n/d| (above, overflow threshold
// included) from every boundary, farther than q. Either way RN32(q) = RN32(n/d). Infinities and
// NaNs pass through; a zero or underflowing quotient comes out +0, and or-ing the sign of n (the
// quotient's sign, as d > 0) restores -0.
struct UDiv {
    double rd;
};

__device__ __forceinline__ float udiv(float n, const UDiv& u) {
    const double C = 0x1.8p-124;
    const double q = fma((double)n, u.rd, C) - C;
    return __uint_as_float(__float_as_uint((float)q) | (__float_as_uint(n) & 0x80000000u));
}

// The UDiv of d (finite, d > 0), formed on the host once per launch.
__host__ __device__ inline UDiv make_udiv(float d) { return UDiv{1.0 / (double)d}; }

__device__ __forceinline__ float divide(float n, float d) { return n / d; }
__device__ __forceinline__ float divide(float n, const UDiv& d) { return udiv(n, d); }

// Divergence with the reference's left/top border rule (solveDataL1.cu:64-88). pl = p1 at
// x-1, pd = p2 at y-1 (ignored on the first column / row).
__device__ __forceinline__ float div_p(float p1, float p1l, float p2, float p2d, bool has_left, bool has_up) {
    if (!has_left) {
        if (!has_up) return p1 + p2;
        return (p1 + p2) - p2d;
    }
    if (!has_up) return (p1 - p1l) + p2;
    return ((p1 - p1l) + p2) - p2d;
}

// L1 data threshold (solveDataL1.cu:27-51 with lt = lambda*theta; solveReconDataL1.cu:38-63
// with the division by A applied to upper/lower/steps). A NaN rho yields NaN.
__device__ __forceinline__ void l1_threshold(float dix, float diy, float dit, float bu, float bv, float lambda,
                                             float theta, float& du, float& dv) {
    float rho = (dix * bu + diy * bv) + dit;
    float lt = lambda * theta;
    float g2 = dix * dix + diy * diy;
    float upper = lt * g2;
    float lower = ((-lambda) * theta) * g2;
    du = __builtin_nanf("");
    dv = du;
    if ((rho <= upper) && (rho >= lower)) {
        float magi = dix * dix + diy * diy;
        if (magi != 0.0f) {
            du = bu - (rho * dix) / magi;
            dv = bv - (rho * diy) / magi;
        } else {
            du = bu;
            dv = bv;
        }
    } else if (rho < lower) {
        du = bu + lt * dix;
        dv = bv + lt * diy;
    } else if (rho > upper) {
        du = bu - lt * dix;
        dv = bv - lt * diy;
    }
}

// lower = ((-lambda) * g2) / A is -upper bit for bit (negation commutes with correctly rounded
// multiplication and division), and bu - t = bu + (-t) exactly, so the two outer cases share
// one pair of divisions. DV: float, or UDiv for a launch-uniform A (same quotients).
template <typename DV>
__device__ __forceinline__ void l1_threshold_div(float dix, float diy, float dit, float bu, float bv, float lambda,
                                                 const DV& A, float& du, float& dv) {
    float rho = (dix * bu + diy * bv) + dit;
    float g2 = dix * dix + diy * diy;
    float upper = divide(lambda * g2, A);
    float lower = -upper;
    if ((rho <= upper) && (rho >= lower)) {
        if (g2 != 0.0f) {
            du = bu - (rho * dix) / g2;
            dv = bv - (rho * diy) / g2;
        } else {
            du = bu;
            dv = bv;
        }
    } else {
        const float tx = divide(lambda * dix, A), ty = divide(lambda * diy, A);
        const bool below = rho < lower;
        du = below ? bu + tx : bu - tx;
        dv = below ? bv + ty : bv - ty;
        if (!(below || rho > upper)) du = dv = __builtin_nanf("");  // NaN rho
    }
}

// sqrtf(x*x + 0.001) of the Charbonnier kernels: 0.001 is a double literal there, so the sum is
// formed in double and rounded to float before the float square root.
__device__ __forceinline__ float charb_root(float x) { return sqrtf((float)((double)(x * x) + 0.001)); }

// SolveDataCharbForTVKernel (solveDataCharbForTV.cu:3-85; GRAD = false: one IRLS pass) and
// SolveDataCharbForTVGradKernel (solveDataCharbForTVGrad.cu:3-103; GRAD: two passes with the
// gradient-constancy terms): the 2x2 reweighted solve for (du, dv) from (du0, dv0).
template <bool GRAD>
__device__ __forceinline__ void charb_px(float du, float dv, float dix, float diy, float dit, float dixx, float dixy,
                                         float dixt, float diyx, float diyy, float diyt, float duhat, float dvhat,
                                         float lambda, float lambdagrad, float theta, float& odu, float& odv) {
#pragma unroll
    for (int k = 0; k < (GRAD ? 2 : 1); ++k) {
        const float it = 1.0f / theta;
        const float innerSum = (dix * du + diy * dv) + dit;
        const float beta = lambda / charb_root(innerSum);
        float a11, a12, a22, b1, b2;
        if (GRAD) {
            const float innerSumx = (dixx * du + dixy * dv) + dixt;
            const float innerSumy = (diyx * du + diyy * dv) + diyt;
            const float betax = lambdagrad / charb_root(innerSumx);
            const float betay = lambdagrad / charb_root(innerSumy);
            a11 = (((dix * dix) * beta + (dixx * dixx) * betax) + (diyx * diyx) * betay) + it;
            a12 = ((dix * diy) * beta + (dixx * dixy) * betax) + (diyx * diyy) * betay;
            a22 = (((diy * diy) * beta + (dixy * dixy) * betax) + (diyy * diyy) * betay) + it;
            b1 = ((it * duhat - (dix * dit) * beta) - (dixx * dixt) * betax) - (diyx * diyt) * betay;
            b2 = ((it * dvhat - (diy * dit) * beta) - (dixy * dixt) * betax) - (diyy * diyt) * betay;
        } else {
            a11 = (dix * dix) * beta + it;
            a12 = (dix * diy) * beta;
            a22 = (diy * diy) * beta + it;
            b1 = it * duhat - (dix * dit) * beta;
            b2 = it * dvhat - (diy * dit) * beta;
        }
        const float a21 = a12;
        const float detA = a11 * a22 - a12 * a21;
        du = (a22 * b1 - a12 * b2) / detA;
        dv = ((-a21) * b1 + a11 * b2) / detA;
    }
    odu = du;
    odv = dv;
}

// Joint data blend (solveReconDataL1.cu:35-36): (aTv*uhat + aP*((uproj - u) - sk)) / A.
template <typename DV>
__device__ __forceinline__ float recon_blend(float aTv, float aP, const DV& A, float uhat, float uproj, float u,
                                             float sk) {
    return divide(aTv * uhat + aP * ((uproj - u) - sk), A);
}

// With a prior flow (solveReconDataL1.cu:172-173).
__device__ __forceinline__ float recon_blend_fn(float aTv, float aP, float aFn, float A, float uhat, float uproj,
                                                float u, float sk, float ufn) {
    return ((aFn * (ufn - u) + aTv * uhat) + aP * ((uproj - u) - sk)) / A;
}

// Dual TV step for one component pair (solveSmoothDualTVGlobal.cu:24-56).
__device__ __forceinline__ void dual_px(float ux, float uy, float p1, float p2, float fac, float& o1, float& o2) {
    float mag = sqrtf(ux * ux + uy * uy);
    o1 = (p1 + fac * ux) / (1.0f + fac * mag);
    o2 = (p2 + fac * uy) / (1.0f + fac * mag);
}

// UVProj blend (solveUVProj.cu:19-20).
__device__ __forceinline__ float uvproj_px(float lf, float aP, float uc, float u, float du, float sk) {
    return (lf * uc + aP * ((u + du) + sk)) / (lf + aP);
}

// The same with the launch-uniform divisor lf + aP as a UDiv.
__device__ __forceinline__ float uvproj_px(float lf, float aP, const UDiv& P, float uc, float u, float du, float sk) {
    return udiv(lf * uc + aP * ((u + du) + sk), P);
}

// Binary64 forms of the two blends for the FP64 mode (this implementation's arithmetic; the
// reference evaluates them in FP32). Both add and subtract values of the size of the flow (u, uproj
// and sk reach 20 px and more at 640x480) to form differences of ~1e-3 px, so an FP32 evaluation
// loses ~1e-6 px each time, and the Sk sum carries the loss through every iteration of a frame.
// Here the stored FP32 values are widened, the launch constants' sums are formed in binary64, the
// launch-uniform divisions are products with the binary64 reciprocals rA = 1 / (aTv + aP) and
// rP = 1 / (lf + aP), and the result is rounded once. The host and the oracle use the same
// expressions in the same order (no contraction).
__device__ __forceinline__ float recon_blend_wide(double aT, double aP, double rA, float uhat, float uproj, float u,
                                                  float sk) {
    return (float)((aT * (double)uhat + aP * (((double)uproj - (double)u) - (double)sk)) * rA);
}

// With a prior flow: AA = (aTv + aP) + aFn in binary64 (aFn = alphaFn times the mask in FP32).
__device__ __forceinline__ float recon_blend_fn_wide(double aT, double aP, double aFn, double AA, float uhat,
                                                     float uproj, float u, float sk, float ufn) {
    return (float)(((aFn * ((double)ufn - (double)u) + aT * (double)uhat) + aP * (((double)uproj - (double)u) - (double)sk)) /
                   AA);
}

__device__ __forceinline__ float uvproj_px_wide(double lf, double aP, double rP, float uc, float u, float du,
                                                float sk) {
    return (float)((lf * (double)uc + aP * (((double)u + (double)du) + (double)sk)) * rP);
}

// Sk update (solveUcSk.cu:93-94).
__device__ __forceinline__ float sk_px(float sk, float u, float du, float uproj) {
    return ((sk + u) + du) - uproj;
}

// Pinhole reprojection into camera 1 minus the pixel (solveUcSk.cu:21-33).
struct UcParams {
    float R11, R12, R13, R21, R22, R23, R31, R32, R33, t1, t2, t3, f, cx, cy;
};
__device__ __forceinline__ void uc_px(const UcParams& c, float X, float Y, float Z, int ix, int iy, float& uc,
                                      float& vc) {
    float X2 = ((c.R11 * X + c.R12 * Y) + c.R13 * Z) + c.t1;
    float Y2 = ((c.R21 * X + c.R22 * Y) + c.R23 * Z) + c.t2;
    float Z2 = ((c.R31 * X + c.R32 * Y) + c.R33 * Z) + c.t3;
    float xproj = (c.f * X2) / Z2 + c.cx;
    float yproj = (c.f * Y2) / Z2 + c.cy;
    uc = xproj - (float)ix;
    vc = yproj - (float)iy;
}

// Camera matrices of Solve3dMinimalArea (solve3dMinimalArea.cu:204-228): cast to float for
// the reference arithmetic, kept in double for the FP64 triangulation.
template <typename R>
struct PQT {
    R P[12], Q[12];
};
using PQ = PQT<float>;
using PQd = PQT<double>;

// Element reads through the constant address space: scalar loads wherever the address is uniform,
// also after barriers and global stores, where a plain read of global memory becomes a vector load
// into registers of every lane. Only for memory no kernel writes (the per-level camera matrices,
// written by the host between launches).
template <typename T>
struct ConstPtr {
    const T* p;
    __device__ __forceinline__ T operator[](int i) const {
        return *(const __attribute__((address_space(4))) T*)(p + i);
    }
};

// The camera matrices of a PQT in device memory, read through ConstPtr.
template <typename R>
struct PQConst {
    ConstPtr<R> P, Q;
};
template <typename R>
__device__ __forceinline__ PQConst<R> const_view(const PQT<R>& m) {
    return PQConst<R>{{m.P}, {m.Q}};
}

// Solve3dMinimalAreaKernel for one pixel (solve3dMinimalArea.cu:21-185), evaluated in R (float
// as the reference, or double). The neighbour values are passed in: r = X0 at ix+1, d = X0 at
// iy+1, l = X0 at ix-1 (used when has_left), u = X0 at iy-1 (used when has_up). Column w and
// row h are legitimately read by the caller (the reference's border guards never fire).
template <typename R, typename M = PQT<R>>
__device__ __forceinline__ void solve3d_px(const M& m, R lambdaf, R lambdams, int ix, int iy, float uproj_f,
                                           float vproj_f, float X0f, float Y0f, float Z0f, float Xrf, float Yrf,
                                           float Zrf, float Xdf, float Ydf, float Zdf, float Xlf, float Ylf, float Zlf,
                                           float Xuf, float Yuf, float Zuf, bool has_left, bool has_up, float& oX,
                                           float& oY, float& oZ) {
    const auto P = m.P;  // a PQT's arrays, or a PQConst's readers
    const auto Q = m.Q;
    const R X0 = X0f, Y0 = Y0f, Z0 = Z0f, Xr = Xrf, Yr = Yrf, Zr = Zrf, Xd = Xdf, Yd = Ydf, Zd = Zdf;
    const R Xl = Xlf, Yl = Ylf, Zl = Zlf, Xu = Xuf, Yu = Yuf, Zu = Zuf;
    R x1 = (R)ix, y1 = (R)iy;
    R x2 = (R)ix + (R)uproj_f;
    R y2 = (R)iy + (R)vproj_f;
    R p1 = x1 * P[8] - P[0], p2 = x1 * P[9] - P[1], p3 = x1 * P[10] - P[2], p4 = x1 * P[11] - P[3];
    R q1 = y1 * P[8] - P[4], q2 = y1 * P[9] - P[5], q3 = y1 * P[10] - P[6], q4 = y1 * P[11] - P[7];
    R r1 = x2 * Q[8] - Q[0], r2 = x2 * Q[9] - Q[1], r3 = x2 * Q[10] - Q[2], r4 = x2 * Q[11] - Q[3];
    R s1 = y2 * Q[8] - Q[4], s2 = y2 * Q[9] - Q[5], s3 = y2 * Q[10] - Q[6], s4 = y2 * Q[11] - Q[7];
#define RF_DOT4(a1, b1, a2, b2, a3, b3, a4, b4) ((((a1) * (b1) + (a2) * (b2)) + (a3) * (b3)) + (a4) * (b4))
    R ax = lambdaf * RF_DOT4(p1, p1, q1, q1, r1, r1, s1, s1);
    R bx = lambdaf * RF_DOT4(p1, p2, q1, q2, r1, r2, s1, s2);
    R cx = lambdaf * RF_DOT4(p1, p3, q1, q3, r1, r3, s1, s3);
    R dx = lambdaf * RF_DOT4(p1, p4, q1, q4, r1, r4, s1, s4);
    R ay = lambdaf * RF_DOT4(p2, p1, q2, q1, r2, r1, s2, s1);
    R by = lambdaf * RF_DOT4(p2, p2, q2, q2, r2, r2, s2, s2);
    R cy = lambdaf * RF_DOT4(p2, p3, q2, q3, r2, r3, s2, s3);
    R dy = lambdaf * RF_DOT4(p2, p4, q2, q4, r2, r4, s2, s4);
    R az = lambdaf * RF_DOT4(p3, p1, q3, q1, r3, r1, s3, s1);
    R bz = lambdaf * RF_DOT4(p3, p2, q3, q2, r3, r2, s3, s2);
    R cz = lambdaf * RF_DOT4(p3, p3, q3, q3, r3, r3, s3, s3);
    R dz = lambdaf * RF_DOT4(p3, p4, q3, q4, r3, r4, s3, s4);
#undef RF_DOT4
    R Xx = Xr - X0, Yx = Yr - Y0, Zx = Zr - Z0;
    R Xy = Xd - X0, Yy = Yd - Y0, Zy = Zd - Z0;
    R c1 = Yx * Zy - Zx * Yy;
    R c2 = Zx * Xy - Xx * Zy;
    R c3 = Xx * Yy - Yx * Xy;
    R lambetams = lambdams / sqrt(((c1 * c1 + c2 * c2) + c3 * c3) + (R)0.001f);
    R Xibar, Yibar, Zibar, Xjbar, Yjbar, Zjbar;
    const R half = (R)0.5;
    if (!has_left) {
        Xibar = Xr;
        Yibar = Yr;
        Zibar = Zr;
    } else {
        Xibar = half * Xl + half * Xr;
        Yibar = half * Yl + half * Yr;
        Zibar = half * Zl + half * Zr;
    }
    if (!has_up) {
        Xjbar = Xd;
        Yjbar = Yd;
        Zjbar = Zd;
    } else {
        Xjbar = half * Xu + half * Xd;
        Yjbar = half * Yu + half * Yd;
        Zjbar = half * Zu + half * Zd;
    }
    R a11 = ax + lambetams * (((Yx * Yx + Zx * Zx) + Yy * Yy) + Zy * Zy);
    R a12 = bx - lambetams * (Xx * Yx + Xy * Yy);
    R a13 = cx - lambetams * (Xx * Zx + Xy * Zy);
    R a21 = ay - lambetams * (Xx * Yx + Xy * Yy);
    R a22 = by + lambetams * (((Xx * Xx + Zx * Zx) + Xy * Xy) + Zy * Zy);
    R a23 = cy - lambetams * (Yx * Zx + Yy * Zy);
    R a31 = az - lambetams * (Xx * Zx + Xy * Zy);
    R a32 = bz - lambetams * (Yx * Zx + Yy * Zy);
    R a33 = cz + lambetams * (((Xx * Xx + Yx * Yx) + Xy * Xy) + Yy * Yy);
    R detA = (a11 * (a22 * a33 - a23 * a32) - a12 * (a21 * a33 - a23 * a31)) + a13 * (a21 * a32 - a22 * a31);
    R E11 = a22 * a33 - a23 * a32, E12 = a13 * a32 - a12 * a33, E13 = a12 * a23 - a13 * a22;
    R E21 = a23 * a31 - a21 * a33, E22 = a11 * a33 - a13 * a31, E23 = a13 * a21 - a11 * a23;
    R E31 = a21 * a32 - a22 * a31, E32 = a12 * a31 - a11 * a32, E33 = a11 * a22 - a12 * a21;
    R b1 = lambetams * ((((((Zy * Zy + Yy * Yy) * Xibar + (Zx * Zx + Yx * Yx) * Xjbar) - (Xy * Yy) * Yibar) -
                          (Xx * Yx) * Yjbar) - (Xy * Zy) * Zibar) - (Xx * Zx) * Zjbar) - dx;
    R b2 = lambetams * (((((-(Xy * Yy) * Xibar - (Xx * Yx) * Xjbar) + (Xy * Xy + Zy * Zy) * Yibar) +
                          (Xx * Xx + Zx * Zx) * Yjbar) - (Yy * Zy) * Zibar) - (Yx * Zx) * Zjbar) - dy;
    R b3 = lambetams * (((((-(Xy * Zy) * Xibar - (Xx * Zx) * Xjbar) - (Yy * Zy) * Yibar) - (Yx * Zx) * Yjbar) +
                          (Xy * Xy + Yy * Yy) * Zibar) + (Xx * Xx + Yx * Yx) * Zjbar) - dz;
    R Xs = ((E11 * b1 + E12 * b2) + E13 * b3) / detA;
    R Ys = ((E21 * b1 + E22 * b2) + E23 * b3) / detA;
    R Zs = ((E31 * b1 + E32 * b2) + E33 * b3) / detA;
    if (isfinite(Xs) && isfinite(Ys) && isfinite(Zs)) {
        oX = (float)Xs;
        oY = (float)Ys;
        oZ = (float)Zs;
    } else {
        oX = X0f;
        oY = Y0f;
        oZ = Z0f;
    }
}

// The FP64 triangulation mode for one pixel: the normal equations of Solve3dMinimalAreaKernel
// (solve3dMinimalArea.cu:21-185) built and solved in FP64 with fused multiply-adds. Same terms
// as solve3d_px; the 4x3 products and every sum of products are FMA chains, the system is scaled
// by sq = sqrt(|Xx x Xy|^2 + 0.001) (data terms weighted lambdaf * sq, the regulariser lambdams,
// so no division forms lambdams / sq), the symmetric normal matrix and its adjugate are formed
// once, the determinant is the cofactor expansion along the first row and the solution is
// adjugate times b times one reciprocal of the determinant. Neighbours and the fallback as in
// solve3d_px.
template <typename M>
__device__ __forceinline__ void solve3d_px_f64(const M& m, double lambdaf, double lambdams, int ix, int iy,
                                               float uproj_f, float vproj_f, float X0f, float Y0f, float Z0f,
                                               float Xrf, float Yrf, float Zrf, float Xdf, float Ydf, float Zdf,
                                               float Xlf, float Ylf, float Zlf, float Xuf, float Yuf, float Zuf,
                                               bool hl, bool hu, float& oX, float& oY, float& oZ) {
    const auto P = m.P;  // a PQd's arrays, or a PQConst<double>'s readers
    const auto Q = m.Q;
    const double x1 = (double)ix, y1 = (double)iy, x2 = (double)ix + (double)uproj_f,
                 y2 = (double)iy + (double)vproj_f;
    double pp[4], qq[4], rr[4], ss[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        pp[k] = fma(x1, P[8 + k], -P[k]);
        qq[k] = fma(y1, P[8 + k], -P[4 + k]);
        rr[k] = fma(x2, Q[8 + k], -Q[k]);
        ss[k] = fma(y2, Q[8 + k], -Q[4 + k]);
    }
    const double X0 = X0f, Y0 = Y0f, Z0 = Z0f, Xr = Xrf, Yr = Yrf, Zr = Zrf, Xd = Xdf, Yd = Ydf, Zd = Zdf;
    const double Xx = Xr - X0, Yx = Yr - Y0, Zx = Zr - Z0;
    const double Xy = Xd - X0, Yy = Yd - Y0, Zy = Zd - Z0;
    const double c1 = fma(Yx, Zy, -(Zx * Yy)), c2 = fma(Zx, Xy, -(Xx * Zy)), c3 = fma(Xx, Yy, -(Yx * Xy));
    // the system scaled by sq = sqrt(|Xx x Xy|^2 + 0.001): data terms weighted lambdaf * sq, the
    // regulariser lambdams (the same solution, no division)
    const double sq = sqrt(fma(c3, c3, fma(c2, c2, c1 * c1)) + (double)0.001f);
    const double ls = lambdaf * sq, lm = lambdams;
    auto dot = [&](int i, int j) {
        return ls * fma(ss[i], ss[j], fma(rr[i], rr[j], fma(qq[i], qq[j], pp[i] * pp[j])));
    };
    const double ax = dot(0, 0), bx = dot(0, 1), cx = dot(0, 2), dx = dot(0, 3);
    const double by = dot(1, 1), cy = dot(1, 2), dy = dot(1, 3);
    const double cz = dot(2, 2), dz = dot(2, 3);
    const double Xib = hl ? fma(0.5, (double)Xlf, 0.5 * Xr) : Xr, Yib = hl ? fma(0.5, (double)Ylf, 0.5 * Yr) : Yr,
                 Zib = hl ? fma(0.5, (double)Zlf, 0.5 * Zr) : Zr;
    const double Xjb = hu ? fma(0.5, (double)Xuf, 0.5 * Xd) : Xd, Yjb = hu ? fma(0.5, (double)Yuf, 0.5 * Yd) : Yd,
                 Zjb = hu ? fma(0.5, (double)Zuf, 0.5 * Zd) : Zd;
    const double xxy = Xx * Yx, xyy = Xy * Yy, xxz = Xx * Zx, xyz = Xy * Zy, yxz = Yx * Zx, yyz = Yy * Zy;
    const double syx = fma(Zx, Zx, Yx * Yx), syy = fma(Zy, Zy, Yy * Yy);
    const double sxx = fma(Zx, Zx, Xx * Xx), sxy = fma(Zy, Zy, Xy * Xy);
    const double txx = fma(Yx, Yx, Xx * Xx), txy = fma(Yy, Yy, Xy * Xy);
    const double a11 = fma(lm, syx + syy, ax), a22 = fma(lm, sxx + sxy, by), a33 = fma(lm, txx + txy, cz);
    const double a12 = fma(-lm, fma(Xy, Yy, xxy), bx), a13 = fma(-lm, fma(Xy, Zy, xxz), cx),
                 a23 = fma(-lm, fma(Yy, Zy, yxz), cy);
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
    if (isfinite(Xs) && isfinite(Ys) && isfinite(Zs)) {
        oX = (float)Xs;
        oY = (float)Ys;
        oZ = (float)Zs;
    } else {
        oX = X0f;
        oY = Y0f;
        oZ = Z0f;
    }
}

// CleanUp3DKernel for one pixel (solve3dMinimalArea.cu:258-295).
__device__ __forceinline__ void cleanup_px(float X, float Y, float Z, float xm, float ym, float zm, float& oX,
                                           float& oY, float& oZ) {
    if (isfinite(X) && isfinite(Y) && isfinite(Z) && isfinite(xm) && isfinite(ym) && isfinite(zm)) {
        oX = (fabsf(xm - X) > 5.0f) ? xm : X;
        oY = (fabsf(ym - Y) > 5.0f) ? ym : Y;
        oZ = (fabsf(zm - Z) > 5.0f) ? zm : Z;
    } else {
        oX = 0.0f;
        oY = 0.0f;
        oZ = 0.0f;
    }
}

// Partial exchange sort of the reference median kernels (medianFilter.cu:52-71, 218-239):
// for j < NSEL, for i > j, swap when m[j] > m[i]; the answer is m[NSEL-1]. Fully unrolled so
// the array lives in registers.
template <int N, int NSEL>
__device__ __forceinline__ float exchange_select(float (&m)[N]) {
#pragma unroll
    for (int j = 0; j < NSEL; ++j) {
#pragma unroll
        for (int i = j + 1; i < N; ++i) {
            float a = m[j], b = m[i];
            bool sw = a > b;
            m[j] = sw ? b : a;
            m[i] = sw ? a : b;
        }
    }
    return m[NSEL - 1];
}

// The value exchange_select<9, 5> returns, for NaN-free taps, with a min/max network: sort the
// three columns (min3 / med3 / max3), then med3(max of the lows, med3 of the middles, min of
// the highs) is the 5th smallest (a min/max network that is right for every 0/1 input is
// right for every totally ordered input). A nonzero 5th smallest has one bit pattern; a zero
// one may be +0 or -0 depending on tap order, so a zero result is recomputed by the exchange
// sort. Callers guarantee NaN-free taps.
__device__ __forceinline__ float median9_finite(float (&m)[9]) {
    float lo[3], mi[3], hi[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float a = m[c], b = m[c + 3], d = m[c + 6];
        lo[c] = fminf(fminf(a, b), d);
        hi[c] = fmaxf(fmaxf(a, b), d);
        mi[c] = __builtin_amdgcn_fmed3f(a, b, d);
    }
    const float r = __builtin_amdgcn_fmed3f(fmaxf(fmaxf(lo[0], lo[1]), lo[2]),
                                            __builtin_amdgcn_fmed3f(mi[0], mi[1], mi[2]),
                                            fminf(fminf(hi[0], hi[1]), hi[2]));
    if (r != 0.0f) return r;
    return exchange_select<9, 5>(m);
}

// The value exchange_select<9, 5> returns, for any taps: the network when no tap is NaN.
__device__ __forceinline__ float median9(float (&m)[9]) {
    bool nan = __builtin_isnan(m[8]);
#pragma unroll
    for (int k = 0; k < 8; k += 2) nan |= __builtin_isunordered(m[k], m[k + 1]);
    if (nan) return exchange_select<9, 5>(m);
    return median9_finite(m);
}

#define RF_CE(a, b)                                \
    do {                                           \
        const float lo_ = fminf(m[a], m[b]);       \
        m[b] = fmaxf(m[a], m[b]);                  \
        m[a] = lo_;                                \
    } while (0)
#define RF_CE_MIN(a, b) m[a] = fminf(m[a], m[b])
#define RF_CE_MAX(a, b) m[b] = fmaxf(m[a], m[b])
#include "rf_median25.hpp"
#undef RF_CE
#undef RF_CE_MIN
#undef RF_CE_MAX

// The value exchange_select<25, 13> returns (the 5x5 flow median, medianFilter.cu:52-71), for
// any taps: with no NaN tap the 13th smallest is the min/max network's output; a nonzero one has
// a single bit pattern. NaN taps (the exchange sort's answer then depends on tap order) and a
// zero result (+0 or -0 by tap order) take the exchange sort itself.
__device__ __forceinline__ float median25(float (&in)[25]) {
    bool nan = __builtin_isnan(in[24]);
#pragma unroll
    for (int k = 0; k < 24; k += 2) nan |= __builtin_isunordered(in[k], in[k + 1]);
    float m[25];
#pragma unroll
    for (int k = 0; k < 25; ++k) m[k] = in[k];
    const float r = median25_network(m);
    if (!nan && r != 0.0f) return r;
    return exchange_select<25, 13>(in);
}

// Tap index of the 5x5 flow median (medianFilter.cu:31-50); -1 = tap keeps its 0.
__device__ __forceinline__ long median5_tap(int col, int row, int w, int h, int s) {
    if (col >= 0 && col < w && row >= 0 && row < h) return col + (long)s * row;
    if (col < 0 && row >= 0 && row < h) return (long)s * row;
    if (col >= w && row >= 0 && row < h) return w - 1 + (long)s * row;
    if (col >= 0 && col < w && row < 0) return col;
    if (col >= 0 && col < w && row >= h) return col + (long)s * (h - 1);
    return -1;
}

// The 5x5 taps of two fields around (ix, iy) under median5_tap's rule, without branches: a tap
// outside the image in one direction only reads the nearest edge pixel (the clamp), one
// outside in both directions keeps 0.
__device__ __forceinline__ void gather5x5(const float* a, const float* b, int ix, int iy, int w, int h, int s,
                                          float (&ma)[25], float (&mb)[25]) {
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        const int row = iy + j - 2;
        const bool rout = row < 0 || row >= h;
        const unsigned rb = (unsigned)min(max(row, 0), h - 1) * (unsigned)s;
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const int col = ix + i - 2;
            const bool zero = rout && (col < 0 || col >= w);
            const unsigned q = rb + (unsigned)min(max(col, 0), w - 1);
            ma[j * 5 + i] = zero ? 0.0f : a[q];
            mb[j * 5 + i] = zero ? 0.0f : b[q];
        }
    }
}

// Tap index of the 3x3 XYZ median (medianFilter.cu:193-214): the right/bottom replacements
// test col > w / row > h, so col == w and row == h keep 0.
__device__ __forceinline__ long median3_tap(int col, int row, int w, int h, int s) {
    if (col >= 0 && col < w && row >= 0 && row < h) return col + (long)s * row;
    if (col < 0 && row >= 0 && row < h) return (long)s * row;
    if (col > w && row >= 0 && row < h) return w - 1 + (long)s * row;
    if (col >= 0 && col < w && row < 0) return col;
    if (col >= 0 && col < w && row > h) return col + (long)s * (h - 1);
    return -1;
}

}  // namespace rf
