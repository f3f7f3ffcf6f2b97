// Fused solve schedule: one kernel per solver iteration.
//
// Joint iteration n (Reconflow.cu:284-407 as one launch). A workgroup owns a 64 x TY output
// tile and works on the tile plus a one-pixel ring:
//   phase 1 (tile + ring): joint data step -> uhat_n, projection blend -> uproj_n (LDS);
//   phase 2 (tile + ring): triangulation -> X'_n (LDS); (tile) dual step -> p_n;
//   phase 3 (tile): 3x3 XYZ median of X'_n, cleanup -> X_n, reprojection -> uc_n,
//                   Sk update -> sk_n; write the tile's state.
// Every field is double-buffered (state n-1 in, state n out), so ring pixels read by a
// neighbouring tile are never overwritten during the launch. The per-pixel arithmetic is the
// one of the per-step kernels (rf_device.hpp), so results equal the plain path bit for bit.
//
// The triangulation reads X_{n-1} at column w and row h of the flat buffer (the reference's
// border guards never fire). The reference reads those cells from one physical buffer that is
// only ever written at valid cells; both X buffers are therefore made identical at the start
// of every level, after which no launch writes outside the valid region.
//
// TV-L1 iteration n (Cudaflow.cu:510-578 as one launch): data step on the tile plus its right
// column and bottom row, dual step on the tile.
//
// Tiles are mapped to workgroups so that each XCD (workgroups are dispatched round-robin over
// the 8 XCDs) receives a contiguous band of tiles, keeping ring re-reads in one L2.

#include <algorithm>

#include "rf_fused.hpp"
#include "rf_kernels.hpp"

namespace rf {
namespace {

constexpr int TXW = 64;  // tile width: one wave64 row
constexpr int TRW = 4;   // thread rows per workgroup (256 threads)
constexpr int NXCD = 8;

// Bijection from the dispatch index onto tiles [0, n): XCD k = b % 8 gets a contiguous range.
__device__ __forceinline__ int xcd_tile(int b, int n) {
    int q = n / NXCD, r = n % NXCD;
    int k = b % NXCD, j = b / NXCD;
    return k * q + min(k, r) + j;
}

struct JointArgs {
    const float *du0, *dv0, *p1u0, *p2u0, *p1v0, *p2v0, *up0, *vp0, *sku0, *skv0, *uc0, *vc0, *X0, *Y0, *Z0;
    float *du1, *dv1, *p1u1, *p2u1, *p1v1, *p2v1, *up1, *vp1, *sku1, *skv1, *uc1, *vc1, *X1, *Y1, *Z1;
    const float *u, *v, *Ix, *Iy, *Iz;
    const float *ufn, *vfn, *mask;
    int w, h, s, tiles_x, ntiles;
    float lambda, aTv, aP, aFn, tau, thetaTv, lf, lms;
    int first;
    const PQ* pq;
    const PQd* pqd;
    const UcParams* ucp;
};

template <int RPT, typename R, bool PRIOR>
__global__ void __launch_bounds__(256) k_joint_iter(JointArgs a) {
    constexpr int TY = TRW * RPT, EX = TXW + 2, EY = TY + 2, NE = EX * EY, NHALO = 2 * EX + 2 * TY;
    static_assert(NHALO <= 256, "ring larger than the workgroup");
    __shared__ float s_du[NE], s_dv[NE], s_up[NE], s_vp[NE], s_X[NE], s_Y[NE], s_Z[NE];

    const int tid = threadIdx.y * TXW + threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, a.ntiles);
    const int x0 = (tile % a.tiles_x) * TXW, y0 = (tile / a.tiles_x) * TY;
    const int w = a.w, h = a.h, s = a.s;
    const float A = PRIOR ? 0.0f : a.aTv + a.aP;

    // the thread's ring pixel (if any), in extended-tile coordinates
    int hex = -1, hey = -1;
    if (tid < NHALO) {
        if (tid < EX) hex = tid, hey = 0;
        else if (tid < 2 * EX) hex = tid - EX, hey = EY - 1;
        else if (tid < 2 * EX + TY) hex = 0, hey = tid - 2 * EX + 1;
        else hex = EX - 1, hey = tid - 2 * EX - TY + 1;
    }

    // phase 1: joint data step + projection blend on tile and ring
    float rp1u[RPT], rp2u[RPT], rp1v[RPT], rp2v[RPT];
    auto phase1 = [&](int ex, int ey, float* keep) {
        const int x = x0 - 1 + ex, y = y0 - 1 + ey;
        if (x < 0 || y < 0 || x >= w || y >= h) return;
        const long q = x + (long)y * s;
        const bool hl = x > 0, hu = y > 0;
        float duh = 0.0f, dvh = 0.0f, p1u = 0.0f, p2u = 0.0f, p1v = 0.0f, p2v = 0.0f, p1ul = 0.0f, p2ud = 0.0f,
              p1vl = 0.0f, p2vd = 0.0f;
        if (!a.first) {
            duh = a.du0[q];
            dvh = a.dv0[q];
            p1u = a.p1u0[q];
            p2u = a.p2u0[q];
            p1v = a.p1v0[q];
            p2v = a.p2v0[q];
            if (hl) p1ul = a.p1u0[q - 1], p1vl = a.p1v0[q - 1];
            if (hu) p2ud = a.p2u0[q - s], p2vd = a.p2v0[q - s];
        }
        const float uu = a.u[q], vv = a.v[q], sku = a.sku0[q], skv = a.skv0[q];
        float AA, bu, bv;
        if (PRIOR) {
            float aFn = a.mask ? a.aFn * a.mask[q] : a.aFn;
            AA = (a.aTv + a.aP) + aFn;
            bu = recon_blend_fn(a.aTv, a.aP, aFn, AA, duh, a.up0[q], uu, sku, a.ufn[q]);
            bv = recon_blend_fn(a.aTv, a.aP, aFn, AA, dvh, a.vp0[q], vv, skv, a.vfn[q]);
        } else {
            AA = A;
            bu = recon_blend(a.aTv, a.aP, AA, duh, a.up0[q], uu, sku);
            bv = recon_blend(a.aTv, a.aP, AA, dvh, a.vp0[q], vv, skv);
        }
        float du, dv;
        l1_threshold_div(a.Ix[q], a.Iy[q], a.Iz[q], bu, bv, a.lambda, AA, du, dv);
        const float nduh = du + div_p(p1u, p1ul, p2u, p2ud, hl, hu) / AA;
        const float ndvh = dv + div_p(p1v, p1vl, p2v, p2vd, hl, hu) / AA;
        const int e = ex + ey * EX;
        s_du[e] = nduh;
        s_dv[e] = ndvh;
        s_up[e] = uvproj_px(a.lf, a.aP, a.uc0[q], uu, nduh, sku);
        s_vp[e] = uvproj_px(a.lf, a.aP, a.vc0[q], vv, ndvh, skv);
        if (keep) {
            keep[0] = p1u;
            keep[1] = p2u;
            keep[2] = p1v;
            keep[3] = p2v;
        }
    };
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        float k4[4] = {0, 0, 0, 0};
        phase1(threadIdx.x + 1, threadIdx.y + TRW * r + 1, k4);
        rp1u[r] = k4[0];
        rp2u[r] = k4[1];
        rp1v[r] = k4[2];
        rp2v[r] = k4[3];
    }
    if (hex >= 0) phase1(hex, hey, nullptr);
    __syncthreads();

    // phase 2: triangulation on tile and ring; dual step on the tile
    auto tri = [&](int ex, int ey) {
        const int x = x0 - 1 + ex, y = y0 - 1 + ey;
        if (x < 0 || y < 0 || x >= w || y >= h) return;
        const long q = x + (long)y * s;
        const bool hl = x > 0, hu = y > 0;
        const long ql = hl ? q - 1 : q, qu = hu ? q - s : q;
        const int e = ex + ey * EX;
        float oX, oY, oZ;
        if (sizeof(R) == 8)
            solve3d_px<double>(*a.pqd, (double)a.lf, (double)a.lms, x, y, s_up[e], s_vp[e], a.X0[q], a.Y0[q], a.Z0[q],
                               a.X0[q + 1], a.Y0[q + 1], a.Z0[q + 1], a.X0[q + s], a.Y0[q + s], a.Z0[q + s], a.X0[ql],
                               a.Y0[ql], a.Z0[ql], a.X0[qu], a.Y0[qu], a.Z0[qu], hl, hu, oX, oY, oZ);
        else
            solve3d_px<float>(*a.pq, a.lf, a.lms, x, y, s_up[e], s_vp[e], a.X0[q], a.Y0[q], a.Z0[q], a.X0[q + 1],
                              a.Y0[q + 1], a.Z0[q + 1], a.X0[q + s], a.Y0[q + s], a.Z0[q + s], a.X0[ql], a.Y0[ql],
                              a.Z0[ql], a.X0[qu], a.Y0[qu], a.Z0[qu], hl, hu, oX, oY, oZ);
        s_X[e] = oX;
        s_Y[e] = oY;
        s_Z[e] = oZ;
    };
    const float fac = a.tau / a.thetaTv;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int ex = threadIdx.x + 1, ey = threadIdx.y + TRW * r + 1;
        tri(ex, ey);
        const int x = x0 + threadIdx.x, y = y0 + threadIdx.y + TRW * r;
        if (x < w && y < h) {
            const long q = x + (long)y * s;
            const int e = ex + ey * EX;
            float dux = 0.0f, dvx = 0.0f, duy = 0.0f, dvy = 0.0f;
            if (x + 1 < w) {
                dux = s_du[e + 1] - s_du[e];
                dvx = s_dv[e + 1] - s_dv[e];
            }
            if (y + 1 < h) {
                duy = s_du[e + EX] - s_du[e];
                dvy = s_dv[e + EX] - s_dv[e];
            }
            float o1, o2;
            dual_px(dux, duy, rp1u[r], rp2u[r], fac, o1, o2);
            a.p1u1[q] = o1;
            a.p2u1[q] = o2;
            dual_px(dvx, dvy, rp1v[r], rp2v[r], fac, o1, o2);
            a.p1v1[q] = o1;
            a.p2v1[q] = o2;
        }
    }
    if (hex >= 0) tri(hex, hey);
    __syncthreads();

    // phase 3: XYZ median + cleanup + reprojection + Sk on the tile; write the state
    const UcParams ucp = *a.ucp;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int x = x0 + threadIdx.x, y = y0 + threadIdx.y + TRW * r;
        if (x >= w || y >= h) continue;
        const long q = x + (long)y * s;
        const int ex = threadIdx.x + 1, ey = threadIdx.y + TRW * r + 1, e = ex + ey * EX;
        float mX[9], mY[9], mZ[9];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                // medianFilter.cu:193-214 tap rule, resolved to a cell of the extended tile
                const int col = x + i - 1, row = y + j - 1;
                int cc = -1, rr = -1;
                if (col >= 0 && col < w && row >= 0 && row < h) cc = col, rr = row;
                else if (col < 0 && row >= 0 && row < h) cc = 0, rr = row;
                else if (col > w && row >= 0 && row < h) cc = w - 1, rr = row;
                else if (col >= 0 && col < w && row < 0) cc = col, rr = 0;
                else if (col >= 0 && col < w && row > h) cc = col, rr = h - 1;
                const int k = j * 3 + i;
                if (cc >= 0) {
                    const int ee = (cc - x0 + 1) + (rr - y0 + 1) * EX;
                    mX[k] = s_X[ee];
                    mY[k] = s_Y[ee];
                    mZ[k] = s_Z[ee];
                } else {
                    mX[k] = 0.0f;
                    mY[k] = 0.0f;
                    mZ[k] = 0.0f;
                }
            }
        const float xm = exchange_select<9, 5>(mX), ym = exchange_select<9, 5>(mY), zm = exchange_select<9, 5>(mZ);
        float Xn, Yn, Zn;
        cleanup_px(s_X[e], s_Y[e], s_Z[e], xm, ym, zm, Xn, Yn, Zn);
        float ucn, vcn;
        uc_px(ucp, Xn, Yn, Zn, x, y, ucn, vcn);
        const float duh = s_du[e], dvh = s_dv[e], upn = s_up[e], vpn = s_vp[e];
        a.X1[q] = Xn;
        a.Y1[q] = Yn;
        a.Z1[q] = Zn;
        a.uc1[q] = ucn;
        a.vc1[q] = vcn;
        a.du1[q] = duh;
        a.dv1[q] = dvh;
        a.up1[q] = upn;
        a.vp1[q] = vpn;
        a.sku1[q] = sk_px(a.sku0[q], a.u[q], duh, upn);
        a.skv1[q] = sk_px(a.skv0[q], a.v[q], dvh, vpn);
    }
}

struct FlowArgs {
    const float *du0, *dv0, *p1u0, *p2u0, *p1v0, *p2v0;
    float *du1, *dv1, *p1u1, *p2u1, *p1v1, *p2v1;
    const float *Ix, *Iy, *Iz;
    int w, h, s, tiles_x, ntiles;
    float lambda, theta, tau;
    int first;
};

// TV-L1 iteration: data step (tile + right column + bottom row) then dual step (tile).
template <int RPT>
__global__ void __launch_bounds__(256) k_flow_iter(FlowArgs a) {
    constexpr int TY = TRW * RPT, EX = TXW + 1, EY = TY + 1, NE = EX * EY, NHALO = EX + TY;
    static_assert(NHALO <= 256, "ring larger than the workgroup");
    __shared__ float s_du[NE], s_dv[NE];
    const int tid = threadIdx.y * TXW + threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, a.ntiles);
    const int x0 = (tile % a.tiles_x) * TXW, y0 = (tile / a.tiles_x) * TY;
    const int w = a.w, h = a.h, s = a.s;
    int hex = -1, hey = -1;
    if (tid < NHALO) {
        if (tid < EX) hex = tid, hey = TY;
        else hex = TXW, hey = tid - EX;
    }
    float rp[RPT][4];
    auto data = [&](int ex, int ey, float* keep) {
        const int x = x0 + ex, y = y0 + ey;
        if (x >= w || y >= h) return;
        const long q = x + (long)y * s;
        const bool hl = x > 0, hu = y > 0;
        float duh = 0.0f, dvh = 0.0f, p1u = 0.0f, p2u = 0.0f, p1v = 0.0f, p2v = 0.0f, p1ul = 0.0f, p2ud = 0.0f,
              p1vl = 0.0f, p2vd = 0.0f;
        if (!a.first) {
            duh = a.du0[q];
            dvh = a.dv0[q];
            p1u = a.p1u0[q];
            p2u = a.p2u0[q];
            p1v = a.p1v0[q];
            p2v = a.p2v0[q];
            if (hl) p1ul = a.p1u0[q - 1], p1vl = a.p1v0[q - 1];
            if (hu) p2ud = a.p2u0[q - s], p2vd = a.p2v0[q - s];
        }
        float du, dv;
        l1_threshold(a.Ix[q], a.Iy[q], a.Iz[q], duh, dvh, a.lambda, a.theta, du, dv);
        const int e = ex + ey * EX;
        s_du[e] = du + a.theta * div_p(p1u, p1ul, p2u, p2ud, hl, hu);
        s_dv[e] = dv + a.theta * div_p(p1v, p1vl, p2v, p2vd, hl, hu);
        if (keep) {
            keep[0] = p1u;
            keep[1] = p2u;
            keep[2] = p1v;
            keep[3] = p2v;
        }
    };
#pragma unroll
    for (int r = 0; r < RPT; ++r) data(threadIdx.x, threadIdx.y + TRW * r, rp[r]);
    if (hex >= 0) data(hex, hey, nullptr);
    __syncthreads();
    const float fac = a.tau / a.theta;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int ex = threadIdx.x, ey = threadIdx.y + TRW * r;
        const int x = x0 + ex, y = y0 + ey;
        if (x >= w || y >= h) continue;
        const long q = x + (long)y * s;
        const int e = ex + ey * EX;
        float dux = 0.0f, dvx = 0.0f, duy = 0.0f, dvy = 0.0f;
        if (x + 1 < w) {
            dux = s_du[e + 1] - s_du[e];
            dvx = s_dv[e + 1] - s_dv[e];
        }
        if (y + 1 < h) {
            duy = s_du[e + EX] - s_du[e];
            dvy = s_dv[e + EX] - s_dv[e];
        }
        float o1, o2;
        dual_px(dux, duy, rp[r][0], rp[r][1], fac, o1, o2);
        a.p1u1[q] = o1;
        a.p2u1[q] = o2;
        dual_px(dvx, dvy, rp[r][2], rp[r][3], fac, o1, o2);
        a.p1v1[q] = o1;
        a.p2v1[q] = o2;
        a.du1[q] = s_du[e];
        a.dv1[q] = s_dv[e];
    }
}

// u += median5(uhat) on valid cells (medianFilter.cu:3-75 then addKernel.cu:3-11); in place.
__global__ void __launch_bounds__(256) k_med5_add(const float* du, const float* dv, float* u, float* v, int w, int h,
                                                   int s) {
    const int ix = blockIdx.x * 64 + threadIdx.x, iy = blockIdx.y * 4 + threadIdx.y;
    if (ix >= w || iy >= h) return;
    float mu[25], mv[25];
#pragma unroll
    for (int j = 0; j < 5; ++j)
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            long q = median5_tap(ix + i - 2, iy + j - 2, w, h, s);
            mu[j * 5 + i] = q >= 0 ? du[q] : 0.0f;
            mv[j * 5 + i] = q >= 0 ? dv[q] : 0.0f;
        }
    const long pos = ix + (long)iy * s;
    u[pos] = u[pos] + exchange_select<25, 13>(mu);
    v[pos] = v[pos] + exchange_select<25, 13>(mv);
}

struct Multi {
    const float* src[8];
    float* dst[8];
};

// Upscale of up to 8 fields in one launch (upscale.cu:14-27 for each).
__global__ void __launch_bounds__(256) k_upscale_multi(Multi m, int nf, int cw, int ch, int cs, int w, int h, int s,
                                                        float scale) {
    const int ix = blockIdx.x * 64 + threadIdx.x, iy = blockIdx.y * 4 + threadIdx.y;
    if (ix >= w || iy >= h) return;
    const int nx = (2 * ix + 1) * cw - w, dx = 2 * w, ny = (2 * iy + 1) * ch - h, dy = 2 * h;
    const int x0 = floor_div(nx, dx), y0 = floor_div(ny, dy);
    const float fx = (float)(nx - x0 * dx) / (float)dx, fy = (float)(ny - y0 * dy) / (float)dy;
    const int xa = refl(x0, cw), xb = refl(x0 + 1, cw), ya = refl(y0, ch), yb = refl(y0 + 1, ch);
    const long pos = ix + (long)iy * s;
    for (int f = 0; f < nf; ++f) {
        const float* c = m.src[f];
        float val = blend(c[xa + (long)ya * cs], c[xb + (long)ya * cs], c[xa + (long)yb * cs], c[xb + (long)yb * cs],
                          fx, fy);
        m.dst[f][pos] = val * scale;
    }
}

// Copies whole flat buffers (valid region, padding and guard row).
__global__ void __launch_bounds__(256) k_copy3(Multi m, int nf, long n) {
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    long stride = (long)gridDim.x * 256;
    for (int f = 0; f < nf; ++f)
        for (long j = i; j < n; j += stride) m.dst[f][j] = m.src[f][j];
}

inline dim3 grid2(int w, int h) { return dim3((w + 63) / 64, (h + 3) / 4); }

#define RF_TRY(x)                          \
    do {                                   \
        hipError_t e__ = (x);              \
        if (e__ != hipSuccess) return e__; \
    } while (0)

constexpr int RPT_JOINT = 2;  // 64 x 8 output tiles
constexpr int RPT_FLOW = 4;   // 64 x 16 output tiles

hipError_t launch_upscale_multi(const float* const* src, float* const* dst, int nf, int cw, int ch, int cs, int w,
                                int h, int s, float scale, hipStream_t st) {
    Multi m{};
    for (int i = 0; i < nf; ++i) {
        m.src[i] = src[i];
        m.dst[i] = dst[i];
    }
    hipLaunchKernelGGL(k_upscale_multi, grid2(w, h), dim3(64, 4), 0, st, m, nf, cw, ch, cs, w, h, s, scale);
    return hipGetLastError();
}

hipError_t launch_copy(const float* const* src, float* const* dst, int nf, long n, hipStream_t st) {
    Multi m{};
    for (int i = 0; i < nf; ++i) {
        m.src[i] = src[i];
        m.dst[i] = dst[i];
    }
    long blocks = std::min<long>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_copy3, dim3((unsigned)std::max<long>(blocks, 1)), dim3(256), 0, st, m, nf, n);
    return hipGetLastError();
}

hipError_t launch_med5_add(const float* du, const float* dv, float* u, float* v, int w, int h, int s,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_med5_add, grid2(w, h), dim3(64, 4), 0, st, du, dv, u, v, w, h, s);
    return hipGetLastError();
}

}  // namespace

hipError_t fused_flow(const FusedCtx& c, Fields& f) {
    const Geometry& g = c.g;
    const rf_params& p = c.p;
    hipStream_t st = c.st;
    // gray conversion and pyramids
    {
        const int w0 = g.W[0], h0 = g.H[0], s0 = g.S[0];
        switch (p.input_type) {
            case RF_INPUT_8UC3:
                RF_TRY(launch_gray_rgb((const uint8_t*)f.in0, f.pI0[0], w0, h0, s0, st));
                RF_TRY(launch_gray_rgb((const uint8_t*)f.in1, f.pI1[0], w0, h0, s0, st));
                break;
            case RF_INPUT_8U:
                RF_TRY(launch_gray_u8((const uint8_t*)f.in0, f.pI0[0], w0, h0, s0, st));
                RF_TRY(launch_gray_u8((const uint8_t*)f.in1, f.pI1[0], w0, h0, s0, st));
                break;
            case RF_INPUT_16U:
                RF_TRY(launch_gray_u16((const uint16_t*)f.in0, f.pI0[0], w0, h0, s0, st));
                RF_TRY(launch_gray_u16((const uint16_t*)f.in1, f.pI1[0], w0, h0, s0, st));
                break;
            default:
                RF_TRY(launch_gray_f32((const float*)f.in0, f.pI0[0], w0, h0, s0, st));
                RF_TRY(launch_gray_f32((const float*)f.in1, f.pI1[0], w0, h0, s0, st));
        }
        for (int l = 1; l < g.L; ++l) {
            RF_TRY(launch_downscale(f.pI0[l - 1], g.W[l - 1], g.H[l - 1], g.S[l - 1], g.W[l], g.H[l], g.S[l], false,
                                    1.0f, f.pI0[l], st));
            RF_TRY(launch_downscale(f.pI1[l - 1], g.W[l - 1], g.H[l - 1], g.S[l - 1], g.W[l], g.H[l], g.S[l], false,
                                    1.0f, f.pI1[l], st));
        }
    }
    {
        float* z[] = {f.u, f.v};
        RF_TRY(launch_fill_zero(z, 2, g.n, st));
    }
    constexpr int TY = TRW * RPT_FLOW;
    for (int level = g.L - 1; level >= 0; --level) {
        const int w = g.W[level], h = g.H[level], s = g.S[level];
        const int tx = (w + TXW - 1) / TXW, ty = (h + TY - 1) / TY;
        for (int wi = 0; wi < p.n_warp_iters; ++wi) {
            RF_TRY(launch_warp(f.pI1[level], w, h, s, f.u, f.v, f.i1warp, st));
            RF_TRY(launch_derivatives(f.pI0[level], f.i1warp, w, h, s, f.Ix, f.Iy, f.Iz, st));
            if (p.n_solver_iters == 0) {
                float* z[] = {f.dumed, f.dvmed};
                RF_TRY(launch_fill_zero(z, 2, g.n, st));
            }
            c.prof->begin(st, RF_KID_FUSED_ITER, level);
            for (int it = 0; it < p.n_solver_iters; ++it) {
                FlowArgs a{f.dumed, f.dvmed, f.pu1, f.pu2, f.pv1, f.pv2, f.dumeds, f.dvmeds, f.pu1s, f.pu2s, f.pv1s,
                           f.pv2s, f.Ix, f.Iy, f.Iz, w, h, s, tx, tx * ty, p.lambda, p.theta, p.tau, it == 0};
                hipLaunchKernelGGL(k_flow_iter<RPT_FLOW>, dim3(tx * ty), dim3(TXW, TRW), 0, st, a);
                RF_TRY(hipGetLastError());
                std::swap(f.dumed, f.dumeds);
                std::swap(f.dvmed, f.dvmeds);
                std::swap(f.pu1, f.pu1s);
                std::swap(f.pu2, f.pu2s);
                std::swap(f.pv1, f.pv1s);
                std::swap(f.pv2, f.pv2s);
            }
            c.prof->end(st, RF_KID_FUSED_ITER, level, p.n_solver_iters);
            RF_TRY(launch_med5_add(f.dumed, f.dvmed, f.u, f.v, w, h, s, st));
        }
        if (level > 0) {
            const float* src[] = {f.u, f.v};
            float* dst[] = {f.us, f.vs};
            RF_TRY(launch_upscale_multi(src, dst, 2, w, h, s, g.W[level - 1], g.H[level - 1], g.S[level - 1], p.scale,
                                        st));
            std::swap(f.u, f.us);
            std::swap(f.v, f.vs);
        }
    }
    if (p.with_visualization)
        RF_TRY(launch_flow_to_hsv(f.u, f.v, g.W[0], g.H[0], g.S[0], f.uvrgb, c.flow_scale, st));
    return hipSuccess;
}

hipError_t fused_recon(const FusedCtx& c, Fields& f) {
    const Geometry& g = c.g;
    const rf_params& p = c.p;
    hipStream_t st = c.st;
    {
        const int w0 = g.W[0], h0 = g.H[0], s0 = g.S[0];
        switch (p.input_type) {
            case RF_INPUT_8UC3:
                RF_TRY(launch_gray_rgb((const uint8_t*)f.in0, f.pI0[0], w0, h0, s0, st));
                RF_TRY(launch_gray_rgb((const uint8_t*)f.in1, f.pI1[0], w0, h0, s0, st));
                break;
            case RF_INPUT_8U:
                RF_TRY(launch_gray_u8((const uint8_t*)f.in0, f.pI0[0], w0, h0, s0, st));
                RF_TRY(launch_gray_u8((const uint8_t*)f.in1, f.pI1[0], w0, h0, s0, st));
                break;
            case RF_INPUT_16U:
                RF_TRY(launch_gray_u16((const uint16_t*)f.in0, f.pI0[0], w0, h0, s0, st));
                RF_TRY(launch_gray_u16((const uint16_t*)f.in1, f.pI1[0], w0, h0, s0, st));
                break;
            default:
                RF_TRY(launch_gray_f32((const float*)f.in0, f.pI0[0], w0, h0, s0, st));
                RF_TRY(launch_gray_f32((const float*)f.in1, f.pI1[0], w0, h0, s0, st));
        }
        for (int l = 1; l < g.L; ++l) {
            RF_TRY(launch_downscale(f.pI0[l - 1], g.W[l - 1], g.H[l - 1], g.S[l - 1], g.W[l], g.H[l], g.S[l], false,
                                    1.0f, f.pI0[l], st));
            RF_TRY(launch_downscale(f.pI1[l - 1], g.W[l - 1], g.H[l - 1], g.S[l - 1], g.W[l], g.H[l], g.S[l], false,
                                    1.0f, f.pI1[l], st));
        }
    }
    {
        // solve-start state (Reconflow.cu:209-220); the zeroed guard rows are never written
        float* z[] = {f.u, f.v, f.uproj, f.vproj, f.uc, f.vc, f.X, f.Y, f.Z, f.sku, f.skv};
        RF_TRY(launch_fill_zero(z, 11, g.n, st));
    }
    const bool fn = c.prior && (p.method == RF_METHODR_TVL1_MS_FN || p.method == RF_METHODR_TVL1_MS_FNSPARSE);
    const float thetaTv = 1.0f / p.alpha_tv;  // Reconflow.cu:32
    constexpr int TY = TRW * RPT_JOINT;
    for (int level = g.L - 1; level >= 0; --level) {
        const int w = g.W[level], h = g.H[level], s = g.S[level];
        if (fn) {
            float sc = (float)g.W[level] / (float)g.W[0];
            RF_TRY(launch_downscale(f.ufn, g.W[0], g.H[0], g.S[0], w, h, s, true, sc, f.ufn_l, st));
            RF_TRY(launch_downscale(f.vfn, g.W[0], g.H[0], g.S[0], w, h, s, true, sc, f.vfn_l, st));
            RF_TRY(launch_downscale(f.mask, g.W[0], g.H[0], g.S[0], w, h, s, false, 1.0f, f.mask_l, st));
        }
        {
            // make both X buffers hold the same cells outside this level's valid region
            const float* src[] = {f.X, f.Y, f.Z};
            float* dst[] = {f.Xs, f.Ys, f.Zs};
            RF_TRY(launch_copy(src, dst, 3, g.bufn, st));
        }
        const int tx = (w + TXW - 1) / TXW, ty = (h + TY - 1) / TY;
        for (int wi = 0; wi < p.n_warp_iters; ++wi) {
            RF_TRY(launch_warp(f.pI1[level], w, h, s, f.u, f.v, f.i1warp, st));
            RF_TRY(launch_derivatives(f.pI0[level], f.i1warp, w, h, s, f.Ix, f.Iy, f.Iz, st));
            if (p.n_solver_iters == 0) {
                float* z[] = {f.dumed, f.dvmed};
                RF_TRY(launch_fill_zero(z, 2, g.n, st));
            }
            c.prof->begin(st, RF_KID_FUSED_ITER, level);
            for (int it = 0; it < p.n_solver_iters; ++it) {
                JointArgs a{};
                a.du0 = f.dumed, a.dv0 = f.dvmed, a.p1u0 = f.pu1, a.p2u0 = f.pu2, a.p1v0 = f.pv1, a.p2v0 = f.pv2;
                a.up0 = f.uproj, a.vp0 = f.vproj, a.sku0 = f.sku, a.skv0 = f.skv, a.uc0 = f.uc, a.vc0 = f.vc;
                a.X0 = f.X, a.Y0 = f.Y, a.Z0 = f.Z;
                a.du1 = f.dumeds, a.dv1 = f.dvmeds, a.p1u1 = f.pu1s, a.p2u1 = f.pu2s, a.p1v1 = f.pv1s, a.p2v1 = f.pv2s;
                a.up1 = f.uprojs, a.vp1 = f.vprojs, a.sku1 = f.skus, a.skv1 = f.skvs, a.uc1 = f.ucs, a.vc1 = f.vcs;
                a.X1 = f.Xs, a.Y1 = f.Ys, a.Z1 = f.Zs;
                a.u = f.u, a.v = f.v, a.Ix = f.Ix, a.Iy = f.Iy, a.Iz = f.Iz;
                a.ufn = fn ? f.ufn_l : nullptr, a.vfn = fn ? f.vfn_l : nullptr;
                a.mask = (fn && c.sparse) ? f.mask_l : nullptr;
                a.w = w, a.h = h, a.s = s, a.tiles_x = tx, a.ntiles = tx * ty;
                a.lambda = p.lambda, a.aTv = p.alpha_tv, a.aP = p.alpha_proj, a.aFn = p.alpha_fn, a.tau = p.tau;
                a.thetaTv = thetaTv, a.lf = p.lambdaf, a.lms = p.lambdams, a.first = it == 0;
                a.pq = &c.dparams->pq[level], a.pqd = &c.dparams->pqd[level], a.ucp = &c.dparams->uc[level];
                const dim3 grid(tx * ty), blk(TXW, TRW);
                if (fn) {
                    if (p.tri_f64) hipLaunchKernelGGL((k_joint_iter<RPT_JOINT, double, true>), grid, blk, 0, st, a);
                    else hipLaunchKernelGGL((k_joint_iter<RPT_JOINT, float, true>), grid, blk, 0, st, a);
                } else {
                    if (p.tri_f64) hipLaunchKernelGGL((k_joint_iter<RPT_JOINT, double, false>), grid, blk, 0, st, a);
                    else hipLaunchKernelGGL((k_joint_iter<RPT_JOINT, float, false>), grid, blk, 0, st, a);
                }
                RF_TRY(hipGetLastError());
                std::swap(f.dumed, f.dumeds);
                std::swap(f.dvmed, f.dvmeds);
                std::swap(f.pu1, f.pu1s);
                std::swap(f.pu2, f.pu2s);
                std::swap(f.pv1, f.pv1s);
                std::swap(f.pv2, f.pv2s);
                std::swap(f.uproj, f.uprojs);
                std::swap(f.vproj, f.vprojs);
                std::swap(f.sku, f.skus);
                std::swap(f.skv, f.skvs);
                std::swap(f.uc, f.ucs);
                std::swap(f.vc, f.vcs);
                std::swap(f.X, f.Xs);
                std::swap(f.Y, f.Ys);
                std::swap(f.Z, f.Zs);
            }
            c.prof->end(st, RF_KID_FUSED_ITER, level, p.n_solver_iters);
            // median of uhat, add to the flow, median of the flow (Reconflow.cu:409-420)
            RF_TRY(launch_med5_add(f.dumed, f.dvmed, f.u, f.v, w, h, s, st));
            RF_TRY(launch_median5(f.u, f.v, w, h, s, f.us, f.vs, st));
            std::swap(f.u, f.us);
            std::swap(f.v, f.vs);
        }
        if (level > 0) {
            const int W = g.W[level - 1], H = g.H[level - 1], S = g.S[level - 1];
            // the eight upscaled fields (Reconflow.cu:424-458) go to the idle partner buffers
            const float* src[] = {f.u, f.v, f.uc, f.vc, f.uproj, f.vproj, f.sku, f.skv};
            float* dst[] = {f.us, f.vs, f.ucs, f.vcs, f.uprojs, f.vprojs, f.skus, f.skvs};
            RF_TRY(launch_upscale_multi(src, dst, 8, w, h, s, W, H, S, p.scale, st));
            std::swap(f.u, f.us);
            std::swap(f.v, f.vs);
            std::swap(f.uc, f.ucs);
            std::swap(f.vc, f.vcs);
            std::swap(f.uproj, f.uprojs);
            std::swap(f.vproj, f.vprojs);
            std::swap(f.sku, f.skus);
            std::swap(f.skv, f.skvs);
        }
    }
    if (p.with_visualization)
        RF_TRY(launch_flow_to_hsv(f.u, f.v, g.W[0], g.H[0], g.S[0], f.uvrgb, c.flow_scale, st));
    return hipSuccess;
}

}  // namespace rf
