// Batched fused solve (rf_batch.hpp) and the rf_batch_* C ABI.
//
// Schedule (shared by every pair of a batch; TV-L1: Cudaflow.cu:403-621, joint:
// Reconflow.cu:147-466), one launch per step over the tiles of all pairs:
//   gray conversion, pyramid levels, zero state;
//   per level: [prior-flow downscale], [X buffers made identical]; per warp: warp, derivatives,
//   n fused iterations, median-of-uhat + add [+ median of the flow]; upscale to the next level.
// Joint iteration n in one kernel: a workgroup owns a 64 x 8 tile plus a one-pixel ring;
//   phase 1 (tile + ring): joint data step -> uhat_n, projection blend -> uproj_n (LDS);
//   phase 2 (tile + ring): triangulation -> X'_n (LDS); (tile) dual step -> p_n;
//   phase 3 (tile): 3x3 XYZ median of X'_n, cleanup -> X_n, reprojection -> uc_n, Sk update.
// TV-L1 iteration n in one kernel: data step on a 64 x 16 tile plus its right column and
// bottom row (LDS), dual step on the tile.
// All iterated fields are double-buffered (state n-1 in, n out), so ring pixels another
// workgroup reads are never overwritten during a launch; the parities live on the host, so a
// captured graph replays the same pointer pattern. The per-pixel arithmetic is the one of the
// per-step kernels (rf_device.hpp): results equal the per-step path bit for bit.
//
// The triangulation reads X_{n-1} at column w and row h of the flat buffer (the reference's
// border guards never fire). The reference reads those cells from one physical buffer that is
// written only at valid cells; here both X buffers are made identical at the start of every
// level, after which no launch writes outside the valid region, so the same values are read.
//
// Tile lists map the dispatch index onto (pair, tile) so that each XCD (workgroups go
// round-robin over the 8 XCDs) receives a contiguous run of tiles, keeping ring re-reads in one
// L2.

#include <hip/hip_fp16.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <string>

#include "rf_batch.hpp"
#include "rf_kernels.hpp"

namespace rf {
namespace {

constexpr int NXCD = 8;
constexpr int TXW = 64;
constexpr int TRW = 4;
constexpr int RPT_JOINT = 2;  // 64 x 8 joint tiles
constexpr int RPT_FLOW = 4;   // 64 x 16 TV-L1 tiles

constexpr int FLAT = kMaxLevels;

__device__ __forceinline__ int xcd_tile(int b, int n) {
    int q = n / NXCD, r = n % NXCD;
    int k = b % NXCD, j = b / NXCD;
    return k * q + min(k, r) + j;
}

struct Launch {
    const PairFields* P;
    const int2* map;
    int ntiles;
};

// Pointwise kernels: 64 x 4 tiles, one pixel per thread.
#define BT_PIX(L, LEVEL)                                                       \
    const int2 pt = L.map[xcd_tile(blockIdx.x, L.ntiles)];                    \
    const PairFields& F = L.P[pt.x];                                           \
    const int w = F.W[LEVEL], h = F.H[LEVEL], s = F.S[LEVEL];                  \
    const int tiles_x = (w + TXW - 1) / TXW;                                   \
    const int ix = (pt.y % tiles_x) * TXW + threadIdx.x;                       \
    const int iy = (pt.y / tiles_x) * TRW + threadIdx.y;                       \
    if (ix >= w || iy >= h) return;                                            \
    const long pos = ix + (long)iy * s;

// Storage types: TI for the images (pyramid, warped image), TS for the derivatives and the TV-L1
// dual fields; float, or __half with every update in FP32 (values rounded to nearest on store,
// widened exactly on load). The joint solver's dual fields are FP32, or binary16 with
// RF_STORAGE_F16_DERIV_DUALS (JDual).
template <typename TS>
__device__ __forceinline__ TS st(float v) {
    return v;
}
template <>
__device__ __forceinline__ __half st<__half>(float v) {
    return __float2half(v);
}
template <typename TS>
__device__ __forceinline__ TS* as(float* p) {
    return reinterpret_cast<TS*>(p);
}
template <typename TS>
__device__ __forceinline__ const TS* as(const float* p) {
    return reinterpret_cast<const TS*>(p);
}

// Element i of a field through a 32-bit byte offset (fields hold at most 2^30 elements), so that
// the access compiles to a scalar base plus one vector offset.
template <typename T>
__device__ __forceinline__ T& el(T* p, unsigned i) {
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + i * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T& el(const T* p, unsigned i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + i * (unsigned)sizeof(T));
}

// The element of a T field that holds element q of a view whose pair sits qb floats further on (the
// field slots are float-sized, so a 16-bit field's pair sits 2 qb elements further on).
template <typename T>
__device__ __forceinline__ unsigned typed_el(unsigned q, unsigned qb) {
    return q + qb * (unsigned)(sizeof(float) / sizeof(T) - 1);
}

// The joint solver's dual fields: FP32, or binary16 (RF_STORAGE_F16_DERIV_DUALS: rounded to nearest
// on store, widened exactly on load, every update in FP32).
template <bool D16>
using JDual = typename std::conditional<D16, __half, float>::type;

template <typename TS>
__global__ void __launch_bounds__(256) kb_gray(Launch L, int type) {
    BT_PIX(L, 0)
    float g0, g1;
    if (type == RF_INPUT_8UC3) {
        const uint8_t *a = (const uint8_t*)F.in0, *b = (const uint8_t*)F.in1;
        g0 = ((float)a[3 * pos] + (float)a[3 * pos + 1] + (float)a[3 * pos + 2]) / 3.0f;
        g1 = ((float)b[3 * pos] + (float)b[3 * pos + 1] + (float)b[3 * pos + 2]) / 3.0f;
        g0 = g0 / 256.0f;
        g1 = g1 / 256.0f;
    } else if (type == RF_INPUT_8U) {
        g0 = (float)((const uint8_t*)F.in0)[pos] / 256.0f;
        g1 = (float)((const uint8_t*)F.in1)[pos] / 256.0f;
    } else if (type == RF_INPUT_16U) {
        g0 = (float)((const uint16_t*)F.in0)[pos] / 65536.0f;
        g1 = (float)((const uint16_t*)F.in1)[pos] / 65536.0f;
    } else {
        g0 = ((const float*)F.in0)[pos];
        g1 = ((const float*)F.in1)[pos];
    }
    as<TS>(F.I0[0])[pos] = st<TS>(g0);
    as<TS>(F.I1[0])[pos] = st<TS>(g1);
}

template <typename TI>
__device__ __forceinline__ float down_px(const TI* src, int W, int H, int S, int ix, int iy, int w, int h) {
    int dxn = 4 * w, dyn = 4 * h;
    int xc = (4 * ix + 2) * W - 2 * w;
    int yc = (4 * iy + 2) * H - 2 * h;
    float a = sample_ratio(src, W, H, S, xc - W, dxn, yc, dyn);
    float b = sample_ratio(src, W, H, S, xc + W, dxn, yc, dyn);
    float c = sample_ratio(src, W, H, S, xc, dxn, yc - H, dyn);
    float d = sample_ratio(src, W, H, S, xc, dxn, yc + H, dyn);
    return ((a + b) + c) + d;
}

// DownscaleKernel for both images of level l from level l-1 (downscale.cu:6-26).
template <typename TS>
__global__ void __launch_bounds__(256) kb_pyramid(Launch L, int l) {
    BT_PIX(L, l)
    const int W = F.W[l - 1], H = F.H[l - 1], S = F.S[l - 1];
    as<TS>(F.I0[l])[pos] = st<TS>(0.25f * down_px(as<TS>(F.I0[l - 1]), W, H, S, ix, iy, w, h));
    as<TS>(F.I1[l])[pos] = st<TS>(0.25f * down_px(as<TS>(F.I1[l - 1]), W, H, S, ix, iy, w, h));
}

// Prior flow at level l, straight from level 0 (Reconflow.cu:232-238; downscale.cu:28-48).
__global__ void __launch_bounds__(256) kb_prior(Launch L, int l) {
    BT_PIX(L, l)
    const int W = F.W[0], H = F.H[0], S = F.S[0];
    const float sc = (float)w / (float)W;
    F.ufn_l[pos] = (sc * 0.25f) * down_px(F.ufn, W, H, S, ix, iy, w, h);
    F.vfn_l[pos] = (sc * 0.25f) * down_px(F.vfn, W, H, S, ix, iy, w, h);
    F.mask_l[pos] = 0.25f * down_px(F.mask, W, H, S, ix, iy, w, h);
}

// Zero the solve-start state over the whole flat level-0 buffer (Reconflow.cu:209-220).
__global__ void __launch_bounds__(256) kb_zero(Launch L, int recon) {
    const int2 pt = L.map[xcd_tile(blockIdx.x, L.ntiles)];
    const PairFields& F = L.P[pt.x];
    const int w = F.S[0], h = F.H[0];
    const int tiles_x = (w + TXW - 1) / TXW;
    const int ix = (pt.y % tiles_x) * TXW + threadIdx.x, iy = (pt.y / tiles_x) * TRW + threadIdx.y;
    if (ix >= w || iy >= h) return;
    const long pos = ix + (long)iy * w;
    F.u[0][pos] = 0.0f;
    F.v[0][pos] = 0.0f;
    if (recon) {
        F.up[0][pos] = 0.0f;
        F.vp[0][pos] = 0.0f;
        F.uc[0][pos] = 0.0f;
        F.vc[0][pos] = 0.0f;
        F.sku[0][pos] = 0.0f;
        F.skv[0][pos] = 0.0f;
        F.X[0][pos] = 0.0f;
        F.Y[0][pos] = 0.0f;
        F.Z[0][pos] = 0.0f;
    }
}

// Make both X buffers identical over the flat level-0 buffer.
__global__ void __launch_bounds__(256) kb_sync_x(Launch L, int px) {
    const int2 pt = L.map[xcd_tile(blockIdx.x, L.ntiles)];
    const PairFields& F = L.P[pt.x];
    const int w = F.S[0], h = F.H[0];
    const int tiles_x = (w + TXW - 1) / TXW;
    const int ix = (pt.y % tiles_x) * TXW + threadIdx.x, iy = (pt.y / tiles_x) * TRW + threadIdx.y;
    if (ix >= w || iy >= h) return;
    const long pos = ix + (long)iy * w;
    F.X[px ^ 1][pos] = F.X[px][pos];
    F.Y[px ^ 1][pos] = F.Y[px][pos];
    F.Z[px ^ 1][pos] = F.Z[px][pos];
}

// WarpingKernel (warpImage.cu:15-29).
template <typename TS>
__global__ void __launch_bounds__(256) kb_warp(Launch L, int l, int pu) {
    BT_PIX(L, l)
    as<TS>(F.i1warp)[pos] = st<TS>(warp_sample(as<TS>(F.I1[l]), w, h, s, ix, iy, F.u[pu][pos], F.v[pu][pos]));
}

// ComputeDerivativesKernel at pixel (ix, iy) (computeDerivatives.cu:8-57): x and y derivatives as
// the mean of the level image's and the warped image's 5-point derivatives over mirrored taps,
// and the time derivative; element type TI, stride s.
template <typename TI>
__device__ __forceinline__ void deriv_px(const TI* I0, const TI* I1, int ix, int iy, int w, int h, int s, float& dx,
                                         float& dy, float& dt) {
    const long r = (long)iy * s;
    const int xm2 = refl(ix - 2, w), xm1 = refl(ix - 1, w), xp1 = refl(ix + 1, w), xp2 = refl(ix + 2, w);
    const int ym2 = refl(iy - 2, h), ym1 = refl(iy - 1, h), yp1 = refl(iy + 1, h), yp2 = refl(iy + 2, h);
    auto g = [](const TI* p, long i) { return (float)p[i]; };
    float t0 = d5(g(I0, r + xm2), g(I0, r + xm1), g(I0, r + xp1), g(I0, r + xp2));
    float t1 = d5(g(I1, r + xm2), g(I1, r + xm1), g(I1, r + xp1), g(I1, r + xp2));
    dx = (t0 + t1) * 0.5f;
    dt = g(I1, r + ix) - g(I0, r + ix);
    t0 = d5(g(I0, ix + (long)ym2 * s), g(I0, ix + (long)ym1 * s), g(I0, ix + (long)yp1 * s), g(I0, ix + (long)yp2 * s));
    t1 = d5(g(I1, ix + (long)ym2 * s), g(I1, ix + (long)ym1 * s), g(I1, ix + (long)yp1 * s), g(I1, ix + (long)yp2 * s));
    dy = (t0 + t1) * 0.5f;
}

// ComputeDerivativesKernel (computeDerivatives.cu:8-57).
template <typename TI, typename TS>
__global__ void __launch_bounds__(256) kb_deriv(Launch L, int l) {
    BT_PIX(L, l)
    float dx, dy, dt;
    deriv_px(as<TI>(F.I0[l]), as<TI>(F.i1warp), ix, iy, w, h, s, dx, dy, dt);
    as<TS>(F.Ix)[pos] = st<TS>(dx);
    as<TS>(F.Iz)[pos] = st<TS>(dt);
    as<TS>(F.Iy)[pos] = st<TS>(dy);
}

// The 5x5 median kernels: a workgroup stages the (TXW + 4) x (TRW + 4) taps around its TXW x TRW
// pixel tile of two fields in LDS, once, under median5_tap's rule (a tap's value depends only on
// its own column and row: the nearest edge pixel when it is outside in one direction, 0 when
// outside in both, as gather5x5 reads them), and each lane reads its 25 taps back at fixed
// offsets. Every lane takes part in the staging; lanes off the image leave after the barrier.
constexpr int MW = TXW + 4, MH = TRW + 4, MN = MW * MH;

#define BT_MED5(L, LEVEL, A, B, MA, MB)                                                        \
    const int2 pt = L.map[xcd_tile(blockIdx.x, L.ntiles)];                                    \
    const PairFields& F = L.P[pt.x];                                                           \
    const int w = F.W[LEVEL], h = F.H[LEVEL], s = F.S[LEVEL];                                  \
    const int tiles_x = (w + TXW - 1) / TXW;                                                   \
    const int x0 = (pt.y % tiles_x) * TXW, y0 = (pt.y / tiles_x) * TRW;                        \
    const int ix = x0 + threadIdx.x, iy = y0 + threadIdx.y;                                    \
    __shared__ float s_ma[MN], s_mb[MN];                                                       \
    med5_stage((A), (B), x0, y0, w, h, s, s_ma, s_mb);                                         \
    __syncthreads();                                                                           \
    if (ix >= w || iy >= h) return;                                                            \
    const long pos = ix + (long)iy * s;                                                        \
    float MA[25], MB[25];                                                                      \
    med5_taps(s_ma, s_mb, MA, MB);

__device__ __forceinline__ void med5_stage(const float* a, const float* b, int x0, int y0, int w, int h, int s,
                                           float* sa, float* sb) {
    constexpr int NT = TXW * TRW;
#pragma unroll
    for (int r = 0; r < (MN + NT - 1) / NT; ++r) {
        const int k = threadIdx.x + threadIdx.y * TXW + r * NT;
        if (k >= MN) break;
        const int row = y0 - 2 + k / MW, col = x0 - 2 + k % MW;
        const bool zero = (row < 0 || row >= h) && (col < 0 || col >= w);
        const unsigned q = (unsigned)min(max(row, 0), h - 1) * (unsigned)s + (unsigned)min(max(col, 0), w - 1);
        sa[k] = zero ? 0.0f : a[q];
        sb[k] = zero ? 0.0f : b[q];
    }
}

// The lane's taps in gather5x5's order (row-major from (ix - 2, iy - 2)).
__device__ __forceinline__ void med5_taps(const float* sa, const float* sb, float (&ma)[25], float (&mb)[25]) {
    const int e = threadIdx.x + threadIdx.y * MW;
#pragma unroll
    for (int j = 0; j < 5; ++j)
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            ma[j * 5 + i] = sa[e + j * MW + i];
            mb[j * 5 + i] = sb[e + j * MW + i];
        }
}

// Median of uhat added to the flow (medianFilter.cu:3-75, addKernel.cu:3-11), in place.
__global__ void __launch_bounds__(256) kb_med5_add(Launch L, int l, int pd, int pu) {
    BT_MED5(L, l, F.du[pd], F.dv[pd], mu, mv)
    F.u[pu][pos] = F.u[pu][pos] + median25(mu);
    F.v[pu][pos] = F.v[pu][pos] + median25(mv);
}

// Median of the flow (Reconflow.cu:418-420): u[pu] -> u[pu^1].
__global__ void __launch_bounds__(256) kb_median5(Launch L, int l, int pu) {
    BT_MED5(L, l, F.u[pu], F.v[pu], mu, mv)
    F.u[pu ^ 1][pos] = median25(mu);
    F.v[pu ^ 1][pos] = median25(mv);
}

// The last kernel of a warp followed by the next warp's WarpingKernel at the same pixel: the warp
// reads the flow at its own pixel only, so it needs no other pixel's result and one launch per
// warp transition is saved. kb_med5_add + kb_warp<TI> (TV-L1):
template <typename TI>
__global__ void __launch_bounds__(256) kb_med5_add_warp(Launch L, int l, int pd, int pu) {
    BT_MED5(L, l, F.du[pd], F.dv[pd], mu, mv)
    const float un = F.u[pu][pos] + median25(mu), vn = F.v[pu][pos] + median25(mv);
    F.u[pu][pos] = un;
    F.v[pu][pos] = vn;
    as<TI>(F.i1warp)[pos] = st<TI>(warp_sample(as<TI>(F.I1[l]), w, h, s, ix, iy, un, vn));
}

// kb_median5 + kb_warp<TI> (joint): the median of the flow into the partner buffers, warped with.
template <typename TI>
__global__ void __launch_bounds__(256) kb_median5_warp(Launch L, int l, int pu) {
    BT_MED5(L, l, F.u[pu], F.v[pu], mu, mv)
    const float un = median25(mu), vn = median25(mv);
    F.u[pu ^ 1][pos] = un;
    F.v[pu ^ 1][pos] = vn;
    as<TI>(F.i1warp)[pos] = st<TI>(warp_sample(as<TI>(F.I1[l]), w, h, s, ix, iy, un, vn));
}

// Upscale from coarse level l+1 to level l (upscale.cu:14-27): the flow, and in joint mode
// uc, uproj and Sk (Reconflow.cu:424-458), into the partner buffers.
__global__ void __launch_bounds__(256) kb_upscale(Launch L, int l, int pu, int pf, int recon, float scale) {
    BT_PIX(L, l)
    const int cw = F.W[l + 1], ch = F.H[l + 1], cs = F.S[l + 1];
    const int nx = (2 * ix + 1) * cw - w, dx = 2 * w, ny = (2 * iy + 1) * ch - h, dy = 2 * h;
    const int xq = floor_div(nx, dx), yq = floor_div(ny, dy);
    const float fx = (float)(nx - xq * dx) / (float)dx, fy = (float)(ny - yq * dy) / (float)dy;
    const long a = refl(xq, cw) + (long)refl(yq, ch) * cs, b = refl(xq + 1, cw) + (long)refl(yq, ch) * cs;
    const long c = refl(xq, cw) + (long)refl(yq + 1, ch) * cs, d = refl(xq + 1, cw) + (long)refl(yq + 1, ch) * cs;
    auto up = [&](const float* src, float* dst) { dst[pos] = blend(src[a], src[b], src[c], src[d], fx, fy) * scale; };
    up(F.u[pu], F.u[pu ^ 1]);
    up(F.v[pu], F.v[pu ^ 1]);
    if (recon) {
        up(F.uc[pf], F.uc[pf ^ 1]);
        up(F.vc[pf], F.vc[pf ^ 1]);
        up(F.up[pf], F.up[pf ^ 1]);
        up(F.vp[pf], F.vp[pf ^ 1]);
        up(F.sku[pf], F.sku[pf ^ 1]);
        up(F.skv[pf], F.skv[pf ^ 1]);
    }
}

__global__ void __launch_bounds__(256) kb_hsv(Launch L, int pu, float flowscale);

// The fields one joint iteration reads (state n-1) and writes (state n) on one pyramid level:
// built on the host for a single pair and passed by value, or resolved in the kernel from the
// pair descriptor for a batch.
struct JointView {
    const float *du0, *dv0, *p1u0, *p2u0, *p1v0, *p2v0, *U, *V, *sku0, *skv0, *up0, *vp0, *uc0, *vc0, *X0, *Y0, *Z0;
    const float *Ix, *Iy, *Iz, *ufn, *vfn, *mask;
    const float *I0, *I1w;  // the level image and the warped image (derivatives computed in place)
    float *du1, *dv1, *p1u1, *p2u1, *p1v1, *p2v1, *sku1, *skv1, *up1, *vp1, *uc1, *vc1, *X1, *Y1, *Z1;
    int w, h, s, tiles_x;
    float inv_tiles_x;  // RN(1/tiles_x): tile row of a tile index without an integer division
    // camera data stay in the device descriptor, so a captured graph sees pose updates
    const PQ* pq;
    const PQd* pqd;
    const UcParams* ucp;
};

// F: the pair's descriptor as the caller sees it; dF: the same descriptor's device address.
// first: the first iteration of a warp, which starts from zero uhat and duals (the pair's zero
// field stands in for them).
__host__ __device__ inline JointView joint_view(const PairFields& F, const PairFields* dF, int l, int pd, int pf,
                                                int px, int pu, int tw, bool first) {
    JointView v;
    v.du0 = F.du[pd], v.dv0 = F.dv[pd], v.p1u0 = F.p1u[pd], v.p2u0 = F.p2u[pd], v.p1v0 = F.p1v[pd],
    v.p2v0 = F.p2v[pd];
    v.U = F.u[pu], v.V = F.v[pu];
    v.sku0 = F.sku[pf], v.skv0 = F.skv[pf], v.up0 = F.up[pf], v.vp0 = F.vp[pf], v.uc0 = F.uc[pf], v.vc0 = F.vc[pf];
    v.X0 = F.X[px], v.Y0 = F.Y[px], v.Z0 = F.Z[px];
    v.Ix = F.Ix, v.Iy = F.Iy, v.Iz = F.Iz, v.ufn = F.ufn_l, v.vfn = F.vfn_l, v.mask = F.mask_l;
    v.I0 = F.I0[l], v.I1w = F.i1warp;
    v.du1 = F.du[pd ^ 1], v.dv1 = F.dv[pd ^ 1], v.p1u1 = F.p1u[pd ^ 1], v.p2u1 = F.p2u[pd ^ 1],
    v.p1v1 = F.p1v[pd ^ 1], v.p2v1 = F.p2v[pd ^ 1];
    v.sku1 = F.sku[pf ^ 1], v.skv1 = F.skv[pf ^ 1], v.up1 = F.up[pf ^ 1], v.vp1 = F.vp[pf ^ 1], v.uc1 = F.uc[pf ^ 1],
    v.vc1 = F.vc[pf ^ 1];
    v.X1 = F.X[px ^ 1], v.Y1 = F.Y[px ^ 1], v.Z1 = F.Z[px ^ 1];
    if (first) v.du0 = v.dv0 = v.p1u0 = v.p2u0 = v.p1v0 = v.p2v0 = F.zero;
    v.w = F.W[l], v.h = F.H[l], v.s = F.S[l];
    v.tiles_x = (v.w + tw - 1) / tw;
    v.inv_tiles_x = 1.0f / (float)v.tiles_x;
    v.pq = &dF->pq[l];
    v.pqd = &dF->pqd[l];
    v.ucp = &dF->ucp[l];
    return v;
}

struct IterArgs {
    Launch L;
    int level, pd, pf, px, pu, first;
    float lambda, theta, tau, aTv, aP, aFn, lf, lms;
    int sparse;
    JointView v;  // pair 0's fields on this level, by value (one-pixel kernels, SINGLE)
    // SINGLE over several pairs of one size: tiles per pair (0: one pair), its reciprocal, and the
    // elements between consecutive pairs' fields in the slab (the fields of pair i are pair 0's
    // fields moved by i * pair_elems elements; the whole slab stays below 2^30 elements)
    int tpp;
    float inv_tpp;
    unsigned pair_elems;
    // the launch-uniform divisors aTv + aP and lf + aP (every joint kernel without a prior flow
    // divides by the first, every joint kernel by the second)
    UDiv uA, uP;
    // FP64 mode: 1 / (aTv + aP) and 1 / (lf + aP) with the sums formed in binary64 (the blends'
    // reciprocals, recon_blend_wide / uvproj_px_wide)
    double rA64, rP64;
};

// Where a tile's pair lives relative to the view: element q of the pair's fields is element
// qb + q of the view's (qb counted in floats: a field of 16-bit elements is qb * 2 elements
// further on), and its camera data are cb bytes after the view's.
struct PairOff {
    unsigned qb;
    unsigned cb;
};
template <typename T>
__device__ __forceinline__ const T& cam(const T* p, const PairOff& o) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + o.cb);
}

// How a launch finds its tiles' fields (SM): 0, from the tile list and the pair descriptors; 1, one
// pair whose field table is in the kernel arguments; 2, pairs of one size: pair 0's table in the
// kernel arguments, the pair = tile / tiles per pair (none when tpp is 0), located by `o`.
template <int SM>
__device__ __forceinline__ JointView tile_view(const IterArgs& a, int tile, int tw, int& t, PairOff& o) {
    o = PairOff{0u, 0u};
    if (SM > 0) {
        t = tile;
        if (SM == 2 && a.tpp > 0) {
            int pair = (int)((float)tile * a.inv_tpp);
            if (pair * a.tpp > tile) --pair;
            else if ((pair + 1) * a.tpp <= tile) ++pair;
            t = tile - pair * a.tpp;
            o = PairOff{(unsigned)pair * a.pair_elems, (unsigned)pair * (unsigned)sizeof(PairFields)};
        }
        return a.v;
    }
    const int2 pt = a.L.map[tile];
    t = pt.y;
    return joint_view(a.L.P[pt.x], a.L.P + pt.x, a.level, a.pd, a.pf, a.px, a.pu, tw, a.first);
}

// Tile index -> (column, row) of tiles for small indices (< 2^24) with a float reciprocal of the
// row length and one correction each way.
__device__ __forceinline__ int2 tile_xy(int t, int tiles_x, float inv) {
    int r = (int)((float)t * inv);
    if (r * tiles_x > t) --r;
    else if ((r + 1) * tiles_x <= t) ++r;
    return make_int2(t - r * tiles_x, r);
}

// One joint iteration (see the file comment). TY = 8.
template <int RPT, bool F64TRI, bool PRIOR, typename TS, bool D16 = false>
__global__ void __launch_bounds__(256) kb_joint_iter(IterArgs a) {
    using TD = JDual<D16>;
    constexpr int TY = TRW * RPT, EX = TXW + 2, EY = TY + 2, NE = EX * EY, NHALO = 2 * EX + 2 * TY;
    static_assert(NHALO <= 256, "ring larger than the workgroup");
    __shared__ float s_du[NE], s_dv[NE], s_up[NE], s_vp[NE], s_X[NE], s_Y[NE], s_Z[NE];
    const int2 pt = a.L.map[xcd_tile(blockIdx.x, a.L.ntiles)];
    const PairFields& F = a.L.P[pt.x];
    const int l = a.level;
    const int w = F.W[l], h = F.H[l], s = F.S[l];
    const int tiles_x = (w + TXW - 1) / TXW;
    const int x0 = (pt.y % tiles_x) * TXW, y0 = (pt.y / tiles_x) * TY;
    const int tid = threadIdx.y * TXW + threadIdx.x;
    const int pd = a.pd, pf = a.pf, px = a.px;
    const float *du0 = F.du[pd], *dv0 = F.dv[pd];
    const TD *p1u0 = as<TD>(F.p1u[pd]), *p2u0 = as<TD>(F.p2u[pd]), *p1v0 = as<TD>(F.p1v[pd]),
             *p2v0 = as<TD>(F.p2v[pd]);
    const float *up0 = F.up[pf], *vp0 = F.vp[pf], *sku0 = F.sku[pf], *skv0 = F.skv[pf], *uc0 = F.uc[pf],
                *vc0 = F.vc[pf];
    const float *X0 = F.X[px], *Y0 = F.Y[px], *Z0 = F.Z[px];
    const float *U = F.u[a.pu], *V = F.v[a.pu];
    const TS *Ix = as<TS>(F.Ix), *Iy = as<TS>(F.Iy), *Iz = as<TS>(F.Iz);
    const float *ufn = F.ufn_l, *vfn = F.vfn_l, *mask = a.sparse ? F.mask_l : nullptr;

    int hex = -1, hey = -1;
    if (tid < NHALO) {
        if (tid < EX) hex = tid, hey = 0;
        else if (tid < 2 * EX) hex = tid - EX, hey = EY - 1;
        else if (tid < 2 * EX + TY) hex = 0, hey = tid - 2 * EX + 1;
        else hex = EX - 1, hey = tid - 2 * EX - TY + 1;
    }

    // phase 1: data step + projection blend (tile and ring)
    float rp[RPT][4];
    auto phase1 = [&](int ex, int ey, float* keep) {
        const int x = x0 - 1 + ex, y = y0 - 1 + ey;
        if (x < 0 || y < 0 || x >= w || y >= h) return;
        const long q = x + (long)y * s;
        const bool hl = x > 0, hu = y > 0;
        float duh = 0.0f, dvh = 0.0f, p1u = 0.0f, p2u = 0.0f, p1v = 0.0f, p2v = 0.0f, p1ul = 0.0f, p2ud = 0.0f,
              p1vl = 0.0f, p2vd = 0.0f;
        if (!a.first) {
            duh = du0[q];
            dvh = dv0[q];
            p1u = (float)p1u0[q];
            p2u = (float)p2u0[q];
            p1v = (float)p1v0[q];
            p2v = (float)p2v0[q];
            if (hl) p1ul = (float)p1u0[q - 1], p1vl = (float)p1v0[q - 1];
            if (hu) p2ud = (float)p2u0[q - s], p2vd = (float)p2v0[q - s];
        }
        const float uu = U[q], vv = V[q], sku = sku0[q], skv = skv0[q];
        float du, dv, nduh, ndvh;
        if (PRIOR) {
            const float aFn = mask ? a.aFn * mask[q] : a.aFn;
            const float AA = (a.aTv + a.aP) + aFn;
            float bu, bv;
            if (F64TRI) {
                const double AAd = ((double)a.aTv + (double)a.aP) + (double)aFn;
                bu = recon_blend_fn_wide(a.aTv, a.aP, aFn, AAd, duh, up0[q], uu, sku, ufn[q]);
                bv = recon_blend_fn_wide(a.aTv, a.aP, aFn, AAd, dvh, vp0[q], vv, skv, vfn[q]);
            } else {
                bu = recon_blend_fn(a.aTv, a.aP, aFn, AA, duh, up0[q], uu, sku, ufn[q]);
                bv = recon_blend_fn(a.aTv, a.aP, aFn, AA, dvh, vp0[q], vv, skv, vfn[q]);
            }
            l1_threshold_div(Ix[q], Iy[q], Iz[q], bu, bv, a.lambda, AA, du, dv);
            nduh = du + div_p(p1u, p1ul, p2u, p2ud, hl, hu) / AA;
            ndvh = dv + div_p(p1v, p1vl, p2v, p2vd, hl, hu) / AA;
        } else {
            const float bu = F64TRI ? recon_blend_wide(a.aTv, a.aP, a.rA64, duh, up0[q], uu, sku)
                                    : recon_blend(a.aTv, a.aP, a.uA, duh, up0[q], uu, sku);
            const float bv = F64TRI ? recon_blend_wide(a.aTv, a.aP, a.rA64, dvh, vp0[q], vv, skv)
                                    : recon_blend(a.aTv, a.aP, a.uA, dvh, vp0[q], vv, skv);
            l1_threshold_div(Ix[q], Iy[q], Iz[q], bu, bv, a.lambda, a.uA, du, dv);
            nduh = du + udiv(div_p(p1u, p1ul, p2u, p2ud, hl, hu), a.uA);
            ndvh = dv + udiv(div_p(p1v, p1vl, p2v, p2vd, hl, hu), a.uA);
        }
        const int e = ex + ey * EX;
        s_du[e] = nduh;
        s_dv[e] = ndvh;
        s_up[e] = F64TRI ? uvproj_px_wide(a.lf, a.aP, a.rP64, uc0[q], uu, nduh, sku)
                         : uvproj_px(a.lf, a.aP, a.uP, uc0[q], uu, nduh, sku);
        s_vp[e] = F64TRI ? uvproj_px_wide(a.lf, a.aP, a.rP64, vc0[q], vv, ndvh, skv)
                         : uvproj_px(a.lf, a.aP, a.uP, vc0[q], vv, ndvh, skv);
        if (keep) {
            keep[0] = p1u;
            keep[1] = p2u;
            keep[2] = p1v;
            keep[3] = p2v;
        }
    };
#pragma unroll
    for (int r = 0; r < RPT; ++r) phase1(threadIdx.x + 1, threadIdx.y + TRW * r + 1, rp[r]);
    if (hex >= 0) phase1(hex, hey, nullptr);
    __syncthreads();

    // phase 2: triangulation (tile and ring), dual step (tile)
    auto tri = [&](int ex, int ey) {
        const int x = x0 - 1 + ex, y = y0 - 1 + ey;
        if (x < 0 || y < 0 || x >= w || y >= h) return;
        const long q = x + (long)y * s;
        const bool hl = x > 0, hu = y > 0;
        const long ql = hl ? q - 1 : q, qu = hu ? q - s : q;
        const int e = ex + ey * EX;
        float oX, oY, oZ;
        if (F64TRI)
            solve3d_px_f64(const_view(F.pqd[l]), (double)a.lf, (double)a.lms, x, y, s_up[e], s_vp[e], X0[q], Y0[q], Z0[q],
                           X0[q + 1], Y0[q + 1], Z0[q + 1], X0[q + s], Y0[q + s], Z0[q + s], X0[ql], Y0[ql], Z0[ql],
                           X0[qu], Y0[qu], Z0[qu], hl, hu, oX, oY, oZ);
        else
            solve3d_px<float>(const_view(F.pq[l]), a.lf, a.lms, x, y, s_up[e], s_vp[e], X0[q], Y0[q], Z0[q], X0[q + 1],
                              Y0[q + 1], Z0[q + 1], X0[q + s], Y0[q + s], Z0[q + s], X0[ql], Y0[ql], Z0[ql], X0[qu],
                              Y0[qu], Z0[qu], hl, hu, oX, oY, oZ);
        s_X[e] = oX;
        s_Y[e] = oY;
        s_Z[e] = oZ;
    };
    const float fac = a.tau / a.theta;
    TD *p1u1 = as<TD>(F.p1u[pd ^ 1]), *p2u1 = as<TD>(F.p2u[pd ^ 1]), *p1v1 = as<TD>(F.p1v[pd ^ 1]),
       *p2v1 = as<TD>(F.p2v[pd ^ 1]);
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
            dual_px(dux, duy, rp[r][0], rp[r][1], fac, o1, o2);
            p1u1[q] = st<TD>(o1);
            p2u1[q] = st<TD>(o2);
            dual_px(dvx, dvy, rp[r][2], rp[r][3], fac, o1, o2);
            p1v1[q] = st<TD>(o1);
            p2v1[q] = st<TD>(o2);
        }
    }
    if (hex >= 0) tri(hex, hey);
    __syncthreads();

    // phase 3: XYZ median, cleanup, reprojection, Sk update; write the tile's state
    const UcParams& ucp = F.ucp[l];
    float *du1 = F.du[pd ^ 1], *dv1 = F.dv[pd ^ 1];
    float *up1 = F.up[pf ^ 1], *vp1 = F.vp[pf ^ 1], *sku1 = F.sku[pf ^ 1], *skv1 = F.skv[pf ^ 1];
    float *uc1 = F.uc[pf ^ 1], *vc1 = F.vc[pf ^ 1], *X1 = F.X[px ^ 1], *Y1 = F.Y[px ^ 1], *Z1 = F.Z[px ^ 1];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int x = x0 + threadIdx.x, y = y0 + threadIdx.y + TRW * r;
        if (x >= w || y >= h) continue;
        const long q = x + (long)y * s;
        const int e = (threadIdx.x + 1) + (threadIdx.y + TRW * r + 1) * EX;
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
        X1[q] = Xn;
        Y1[q] = Yn;
        Z1[q] = Zn;
        uc1[q] = ucn;
        vc1[q] = vcn;
        du1[q] = duh;
        dv1[q] = dvh;
        up1[q] = upn;
        vp1[q] = vpn;
        sku1[q] = sk_px(sku0[q], U[q], duh, upn);
        skv1[q] = sk_px(skv0[q], V[q], dvh, vpn);
    }
}

// The front half of a joint iteration at pixel (xc, yc) (in the image): every global load, the
// data step -> uhat_n, the projection blend -> uproj_n and the triangulation -> X'_n, with the
// values the back half needs.
struct JointFront {
    float p1u, p2u, p1v, p2v, uu, vv, sku, skv, nduh, ndvh, upn, vpn, oX, oY, oZ;
};

// What the front half reads at a pixel: its own state, the duals to the left and above and X0 at
// the four neighbours (X0 to the left and above is the pixel's own on column 0 and row 0).
struct JointIn {
    float duh, dvh, p1u, p2u, p1v, p2v, p1ul, p1vl, p2ud, p2vd;
    float uu, vv, sku, skv, up0, vp0, uc0, vc0, ix, iy, iz;
    float Xc, Yc, Zc, Xr, Yr, Zr, Xd, Yd, Zd, Xl, Yl, Zl, Xu, Yu, Zu;
    float ufn, vfn, msk;
};

// INNER: the pixel has neighbours on every side (not on column 0 or row 0), known at compile time.
// TI (not void): the derivatives are computed at the pixel from the level image and the warped
// image (element type TI) as kb_deriv does and rounded to TS, instead of loaded.
template <bool PRIOR, typename TS, bool INNER = false, typename TI = void, bool D16 = false>
__device__ __forceinline__ JointIn joint_load(const IterArgs& a, const JointView& V, int xc, int yc, unsigned qb) {
    using TD = JDual<D16>;
    const int s = V.s;
    const unsigned q = qb + (unsigned)xc + (unsigned)yc * (unsigned)s;
    const bool hl = INNER || xc > 0, hu = INNER || yc > 0;
    const unsigned ql = hl ? q - 1u : q, qu = hu ? q - (unsigned)s : q;
    // unconditional loads: on the first iteration of a warp the view's uhat and dual inputs are
    // the zero field
    JointIn in;
    in.duh = el(V.du0, q), in.dvh = el(V.dv0, q);
    {
        const unsigned qd = typed_el<TD>(q, qb), qdl = typed_el<TD>(ql, qb), qdu = typed_el<TD>(qu, qb);
        in.p1u = (float)el(as<TD>(V.p1u0), qd), in.p2u = (float)el(as<TD>(V.p2u0), qd);
        in.p1v = (float)el(as<TD>(V.p1v0), qd), in.p2v = (float)el(as<TD>(V.p2v0), qd);
        in.p1ul = (float)el(as<TD>(V.p1u0), qdl), in.p1vl = (float)el(as<TD>(V.p1v0), qdl);
        in.p2ud = (float)el(as<TD>(V.p2u0), qdu), in.p2vd = (float)el(as<TD>(V.p2v0), qdu);
    }
    in.uu = el(V.U, q), in.vv = el(V.V, q);
    in.sku = el(V.sku0, q), in.skv = el(V.skv0, q), in.up0 = el(V.up0, q), in.vp0 = el(V.vp0, q);
    in.uc0 = el(V.uc0, q), in.vc0 = el(V.vc0, q);
    if constexpr (!std::is_void<TI>::value) {
        const unsigned qi = qb * (unsigned)(sizeof(float) / sizeof(TI));  // the pair's images
        float dx, dy, dt;
        deriv_px(as<TI>(V.I0) + qi, as<TI>(V.I1w) + qi, xc, yc, V.w, V.h, s, dx, dy, dt);
        in.ix = (float)st<TS>(dx), in.iy = (float)st<TS>(dy), in.iz = (float)st<TS>(dt);
    } else {
        const unsigned qt = q + qb * (unsigned)(sizeof(float) / sizeof(TS) - 1);  // the same element in TS fields
        in.ix = (float)el(as<TS>(V.Ix), qt), in.iy = (float)el(as<TS>(V.Iy), qt), in.iz = (float)el(as<TS>(V.Iz), qt);
    }
    const unsigned qr = q + 1u, qd = q + (unsigned)s;
    in.Xc = el(V.X0, q), in.Yc = el(V.Y0, q), in.Zc = el(V.Z0, q);
    in.Xr = el(V.X0, qr), in.Yr = el(V.Y0, qr), in.Zr = el(V.Z0, qr);
    in.Xd = el(V.X0, qd), in.Yd = el(V.Y0, qd), in.Zd = el(V.Z0, qd);
    in.Xl = el(V.X0, ql), in.Yl = el(V.Y0, ql), in.Zl = el(V.Z0, ql);
    in.Xu = el(V.X0, qu), in.Yu = el(V.Y0, qu), in.Zu = el(V.Z0, qu);
    in.ufn = 0.0f, in.vfn = 0.0f, in.msk = 1.0f;
    if (PRIOR) {
        in.ufn = el(V.ufn, q);
        in.vfn = el(V.vfn, q);
        if (a.sparse) in.msk = el(V.mask, q);
    }
    return in;
}

template <bool F64TRI, bool PRIOR, bool INNER = false>
__device__ __forceinline__ JointFront joint_compute(const IterArgs& a, const JointView& V, const PairOff& po, int xc,
                                                    int yc, const JointIn& in) {
    const bool hl = INNER || xc > 0, hu = INNER || yc > 0;
    // ---- data step + projection blend (divisions by the launch-uniform A and lf + aP through
    // UDiv: the IEEE quotients)
    float du, dv, nduh, ndvh;
    if (PRIOR) {
        const float aFn = a.sparse ? a.aFn * in.msk : a.aFn;
        const float AA = (a.aTv + a.aP) + aFn;
        float bu, bv;
        if (F64TRI) {
            const double AAd = ((double)a.aTv + (double)a.aP) + (double)aFn;
            bu = recon_blend_fn_wide(a.aTv, a.aP, aFn, AAd, in.duh, in.up0, in.uu, in.sku, in.ufn);
            bv = recon_blend_fn_wide(a.aTv, a.aP, aFn, AAd, in.dvh, in.vp0, in.vv, in.skv, in.vfn);
        } else {
            bu = recon_blend_fn(a.aTv, a.aP, aFn, AA, in.duh, in.up0, in.uu, in.sku, in.ufn);
            bv = recon_blend_fn(a.aTv, a.aP, aFn, AA, in.dvh, in.vp0, in.vv, in.skv, in.vfn);
        }
        l1_threshold_div(in.ix, in.iy, in.iz, bu, bv, a.lambda, AA, du, dv);
        nduh = du + div_p(in.p1u, in.p1ul, in.p2u, in.p2ud, hl, hu) / AA;
        ndvh = dv + div_p(in.p1v, in.p1vl, in.p2v, in.p2vd, hl, hu) / AA;
    } else {
        const float bu = F64TRI ? recon_blend_wide(a.aTv, a.aP, a.rA64, in.duh, in.up0, in.uu, in.sku)
                                : recon_blend(a.aTv, a.aP, a.uA, in.duh, in.up0, in.uu, in.sku);
        const float bv = F64TRI ? recon_blend_wide(a.aTv, a.aP, a.rA64, in.dvh, in.vp0, in.vv, in.skv)
                                : recon_blend(a.aTv, a.aP, a.uA, in.dvh, in.vp0, in.vv, in.skv);
        l1_threshold_div(in.ix, in.iy, in.iz, bu, bv, a.lambda, a.uA, du, dv);
        nduh = du + udiv(div_p(in.p1u, in.p1ul, in.p2u, in.p2ud, hl, hu), a.uA);
        ndvh = dv + udiv(div_p(in.p1v, in.p1vl, in.p2v, in.p2vd, hl, hu), a.uA);
    }
    const float upn = F64TRI ? uvproj_px_wide(a.lf, a.aP, a.rP64, in.uc0, in.uu, nduh, in.sku)
                             : uvproj_px(a.lf, a.aP, a.uP, in.uc0, in.uu, nduh, in.sku);
    const float vpn = F64TRI ? uvproj_px_wide(a.lf, a.aP, a.rP64, in.vc0, in.vv, ndvh, in.skv)
                             : uvproj_px(a.lf, a.aP, a.uP, in.vc0, in.vv, ndvh, in.skv);

    // ---- triangulation
    float oX, oY, oZ;
    if (F64TRI)
        solve3d_px_f64(const_view(cam(V.pqd, po)), (double)a.lf, (double)a.lms, xc, yc, upn, vpn, in.Xc, in.Yc, in.Zc, in.Xr, in.Yr,
                       in.Zr, in.Xd, in.Yd, in.Zd, in.Xl, in.Yl, in.Zl, in.Xu, in.Yu, in.Zu, hl, hu, oX, oY, oZ);
    else
        solve3d_px<float>(const_view(cam(V.pq, po)), a.lf, a.lms, xc, yc, upn, vpn, in.Xc, in.Yc, in.Zc, in.Xr, in.Yr, in.Zr, in.Xd,
                          in.Yd, in.Zd, in.Xl, in.Yl, in.Zl, in.Xu, in.Yu, in.Zu, hl, hu, oX, oY, oZ);
    return JointFront{in.p1u, in.p2u, in.p1v, in.p2v, in.uu, in.vv, in.sku, in.skv, nduh, ndvh, upn, vpn, oX, oY, oZ};
}

// The back half of a joint iteration at the in-image pixel (x, y) whose extended-tile slot is e
// (EXW slots per row; the slot of pixel (x + i, y + j) is e + i + j * EXW): dual step from
// uhat_n in LDS, 3x3 XYZ median of X'_n in LDS, cleanup, reprojection, Sk update. CLAMP keeps
// every LDS index inside the NE slots (for slots whose results are not used). INNER: the pixel
// and its eight neighbours are inside the image (no border rule applies), known at compile time.
struct JointBack {
    float p1u, p2u, p1v, p2v, X, Y, Z, uc, vc, sku, skv;
};

template <int EXW, int NE, bool CLAMP, bool INNER = false>
__device__ __forceinline__ JointBack joint_back_calc(const IterArgs& a, const UcParams& ucp, const JointFront& J,
                                                     int x, int y, int w, int h, int e, const float* s_du,
                                                     const float* s_dv, const float* s_X, const float* s_Y,
                                                     const float* s_Z) {
    auto at = [&](int i) { return CLAMP ? min(max(i, 0), NE - 1) : i; };
    JointBack B;
    // ---- dual step
    const float fac = a.tau / a.theta;
    float dux = 0.0f, dvx = 0.0f, duy = 0.0f, dvy = 0.0f;
    if (INNER || x + 1 < w) {
        dux = s_du[at(e + 1)] - J.nduh;
        dvx = s_dv[at(e + 1)] - J.ndvh;
    }
    if (INNER || y + 1 < h) {
        duy = s_du[at(e + EXW)] - J.nduh;
        dvy = s_dv[at(e + EXW)] - J.ndvh;
    }
    dual_px(dux, duy, J.p1u, J.p2u, fac, B.p1u, B.p2u);
    dual_px(dvx, dvy, J.p1v, J.p2v, fac, B.p1v, B.p2v);

    // ---- XYZ median, cleanup, reprojection, Sk update. Tap rule of medianFilter.cu:193-214 for
    // an interior pixel (taps lie within one pixel of it): a tap in column w or row h, or left
    // of column 0 and above row 0 at once, keeps 0; a tap left of column 0 (above row 0) reads
    // column 0 (row 0); the rest read themselves. X' is finite (the triangulation falls back to
    // the finite X_{n-1}), so the min/max median applies.
    float mX[9], mY[9], mZ[9];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int col = x + i - 1, row = y + j - 1;
            const bool zero = !INNER && (col == w || row == h || (col < 0 && row < 0));
            const int ee = INNER ? e + (i - 1) + (j - 1) * EXW : at(e + (max(col, 0) - x) + (max(row, 0) - y) * EXW);
            const int k = j * 3 + i;
            mX[k] = zero ? 0.0f : s_X[ee];
            mY[k] = zero ? 0.0f : s_Y[ee];
            mZ[k] = zero ? 0.0f : s_Z[ee];
        }
    const float xm = median9_finite(mX), ym = median9_finite(mY), zm = median9_finite(mZ);
    cleanup_px(J.oX, J.oY, J.oZ, xm, ym, zm, B.X, B.Y, B.Z);
    uc_px(ucp, B.X, B.Y, B.Z, x, y, B.uc, B.vc);
    B.sku = sk_px(J.sku, J.uu, J.nduh, J.upn);
    B.skv = sk_px(J.skv, J.vv, J.ndvh, J.vpn);
    return B;
}

// Stores state n of a pixel (element q of the view, whose pair sits qb floats further on).
template <bool D16 = false>
__device__ __forceinline__ void joint_store(const JointView& V, unsigned q, unsigned qb, const JointFront& J,
                                            const JointBack& B) {
    using TD = JDual<D16>;
    const unsigned qd = typed_el<TD>(q, qb);
    el(as<TD>(V.p1u1), qd) = st<TD>(B.p1u);
    el(as<TD>(V.p2u1), qd) = st<TD>(B.p2u);
    el(as<TD>(V.p1v1), qd) = st<TD>(B.p1v);
    el(as<TD>(V.p2v1), qd) = st<TD>(B.p2v);
    el(V.X1, q) = B.X;
    el(V.Y1, q) = B.Y;
    el(V.Z1, q) = B.Z;
    el(V.uc1, q) = B.uc;
    el(V.vc1, q) = B.vc;
    el(V.du1, q) = J.nduh;
    el(V.dv1, q) = J.ndvh;
    el(V.up1, q) = J.upn;
    el(V.vp1, q) = J.vpn;
    el(V.sku1, q) = B.sku;
    el(V.skv1, q) = B.skv;
}

// Stores the derivatives that a launch computed at a pixel (element q of the pair at qb), in TS.
template <typename TS>
__device__ __forceinline__ void joint_store_deriv(const JointView& V, unsigned q, unsigned qb, const JointIn& in) {
    const unsigned qt = q + qb * (unsigned)(sizeof(float) / sizeof(TS) - 1);
    el(as<TS>(const_cast<float*>(V.Ix)), qt) = st<TS>(in.ix);
    el(as<TS>(const_cast<float*>(V.Iy)), qt) = st<TS>(in.iy);
    el(as<TS>(const_cast<float*>(V.Iz)), qt) = st<TS>(in.iz);
}

// One joint iteration with one pixel per lane: the workgroup is an EXW x EXH extended tile whose
// outer ring of lanes only feeds its (EXW-2) x (EXH-2) interior. Every global load of a lane is issued up front
// (coordinates clamped into the image so that the loads are in bounds; lanes off the image
// compute values nobody reads), then
//   data step -> uhat_n, projection blend -> uproj_n (registers), triangulation -> X'_n;
//   uhat_n and X'_n to LDS, one barrier;
//   interior lanes: dual step -> p_n, 3x3 XYZ median of X'_n, cleanup -> X_n, reprojection ->
//   uc_n, Sk update, stores.
// Same per-pixel arithmetic as kb_joint_iter, so the results are identical. SINGLE: the launch
// covers one pair whose tiles are the dispatch indices (no tile list read) and whose field
// table is in the kernel arguments.
// OCC: minimum waves per SIMD the register allocation must allow (1: no constraint, the
// schedule keeps more loads in flight, for latency-bound small levels; 5: at most 96 VGPRs, five
// 4-wave workgroups per CU, for occupancy on large levels; 4 for the 8-wave 32x16 workgroup, of
// which a CU holds two at four or five waves per SIMD alike, so the tighter cap would only spill).
// The body of kb_joint_px for the tile at (x0, y0). INNER: every lane of the extended tile is a pixel
// at least one pixel inside the image, so no coordinate is clamped and no border rule applies.
// TI (not void): the lanes compute the derivatives (joint_load) and interior lanes store them.
template <int EXW, int EXH, bool INNER, bool F64TRI, bool PRIOR, typename TS, typename TI, bool D16>
__device__ __forceinline__ void joint_px_tile(const IterArgs& a, const JointView& V, const PairOff& po, int x0,
                                              int y0, float* s_du, float* s_dv, float* s_X, float* s_Y, float* s_Z) {
    constexpr int NE = EXW * EXH;
    const int w = V.w, h = V.h, s = V.s;
    const int ex = threadIdx.x, ey = threadIdx.y, e = ex + ey * EXW;
    const int x = x0 - 1 + ex, y = y0 - 1 + ey;
    const int xc = INNER ? x : min(max(x, 0), w - 1), yc = INNER ? y : min(max(y, 0), h - 1);
    const JointIn in = joint_load<PRIOR, TS, INNER, TI, D16>(a, V, xc, yc, po.qb);
    const JointFront J = joint_compute<F64TRI, PRIOR, INNER>(a, V, po, xc, yc, in);
    s_du[e] = J.nduh;
    s_dv[e] = J.ndvh;
    s_X[e] = J.oX;
    s_Y[e] = J.oY;
    s_Z[e] = J.oZ;
    __syncthreads();
    // read before any global store of the kernel, so that it compiles to scalar loads
    const UcParams ucp = cam(V.ucp, po);
    if (ex == 0 || ey == 0 || ex == EXW - 1 || ey == EXH - 1 || (!INNER && (x >= w || y >= h))) return;
    // unsigned 32-bit element offsets: every field access becomes a scalar base plus one vector
    // offset
    const unsigned q = po.qb + (unsigned)x + (unsigned)y * (unsigned)s;
    joint_store<D16>(V, q, po.qb, J,
                     joint_back_calc<EXW, NE, false, INNER>(a, ucp, J, x, y, w, h, e, s_du, s_dv, s_X, s_Y, s_Z));
    if constexpr (!std::is_void<TI>::value) joint_store_deriv<TS>(V, q, po.qb, in);
}

// TI (not void): the first launch of a warp, which computes the derivatives at its lanes' pixels
// and stores them for the warp's later launches (kb_deriv is not launched).
// D16: the dual fields in binary16 (RF_STORAGE_F16_DERIV_DUALS).
template <int EXW, int EXH, int OCC, int SM, bool F64TRI, bool PRIOR, typename TS, typename TI = void, bool D16 = false>
__global__ void __launch_bounds__(EXW* EXH) __attribute__((amdgpu_waves_per_eu(OCC))) kb_joint_px(IterArgs a) {
    constexpr int TW = EXW - 2, TH = EXH - 2, NE = EXW * EXH;
    __shared__ float s_du[NE], s_dv[NE], s_X[NE], s_Y[NE], s_Z[NE];
    const int tile = xcd_tile(blockIdx.x, a.L.ntiles);
    int tp;
    PairOff po;
    const JointView V = tile_view<SM>(a, tile, TW, tp, po);
    const int2 txy = tile_xy(tp, V.tiles_x, V.inv_tiles_x);
    const int x0 = txy.x * TW, y0 = txy.y * TH;
    // lanes cover pixels x0 - 1 .. x0 + TW and y0 - 1 .. y0 + TH; the batched kernels keep one path
    // (with more scalars live, a second path makes the registers spill)
    if (SM == 1 && x0 >= 2 && y0 >= 2 && x0 + TW < V.w - 1 && y0 + TH < V.h - 1)
        joint_px_tile<EXW, EXH, true, F64TRI, PRIOR, TS, TI, D16>(a, V, po, x0, y0, s_du, s_dv, s_X, s_Y, s_Z);
    else
        joint_px_tile<EXW, EXH, false, F64TRI, PRIOR, TS, TI, D16>(a, V, po, x0, y0, s_du, s_dv, s_X, s_Y, s_Z);
}

// K joint iterations per launch (temporal blocking) on an EXW x EXH extended tile indexed by flat
// element: lane (lx, ly) holds element E = (x0 - R + lx) + (y0 - R + ly) * s with R = 2K - 1, so
// the lane to the right holds E + 1 and the lane below E + s, the elements the reference reads as
// X0[pos + 1] and X0[pos + s] (solve3dMinimalArea.cu:65-80). At column w - 1 those reach the
// padding or, where the stride has none, the first pixel of the next row, which is then a lane of
// this tile like any other. Lanes whose element is a pixel iterate; the others (padding, the guard
// row) keep their X0, which no kernel writes during a level and which is the same in both X
// buffers, and are read only as those X0 neighbours. The results of iteration k are right on
// lanes at least 2k - 1 lanes inside the tile (the front half reads one lane around, the back
// half one more), so the launch stores its (EXW - 2R) x (EXH - 2R) interior, pixels x0.. and
// y0..: state n + K into the other parity, one flip per launch. Same per-pixel arithmetic as
// kb_joint_px. TI (not void): as for kb_joint_px, the warp's first launch computes and stores the
// derivatives.
// D16: the dual fields in binary16; the duals carried from one in-kernel iteration to the next are
// rounded to binary16 as a store and reload would round them.
template <int EXW, int EXH, int K, int SM, bool F64TRI, bool PRIOR, typename TS, typename TI = void, bool D16 = false>
__global__ void __launch_bounds__(EXW* EXH) kb_joint_multi(IterArgs a) {
    using TD = JDual<D16>;
    constexpr int R = 2 * K - 1, TW = EXW - 2 * R, TH = EXH - 2 * R, NE = EXW * EXH;
    static_assert(TW > 0 && TH > 0, "tile too small for K iterations");
    __shared__ float s_du[NE], s_dv[NE], s_X[NE], s_Y[NE], s_Z[NE];                              // uhat_n, X'_n
    __shared__ float s_p1u[NE], s_p2u[NE], s_p1v[NE], s_p2v[NE], s_Xn[NE], s_Yn[NE], s_Zn[NE];  // p_n, X_n
    const int tile = xcd_tile(blockIdx.x, a.L.ntiles);
    int tp;
    PairOff po;
    const JointView V = tile_view<SM>(a, tile, TW, tp, po);
    const unsigned qb = po.qb;
    const int w = V.w, h = V.h, s = V.s;
    const int2 txy = tile_xy(tp, V.tiles_x, V.inv_tiles_x);
    const int x0 = txy.x * TW, y0 = txy.y * TH;
    const int lx = threadIdx.x, ly = threadIdx.y, e = lx + ly * EXW;
    const int cx = x0 - R + lx, ry = y0 - R + ly;  // flat coordinates: element cx + ry * s
    // the element's pixel (one wrap at most: |cx| < s)
    const int pxw = cx >= s ? cx - s : (cx < 0 ? cx + s : cx);
    const int pyw = cx >= s ? ry + 1 : (cx < 0 ? ry - 1 : ry);
    const bool pix = pxw < w && pyw >= 0 && pyw < h;
    const int xc = min(pxw, w - 1), yc = min(max(pyw, 0), h - 1);  // loads of other lanes stay in bounds
    // X0 of an element that is not a pixel: padding or guard row; elements past the guard row or
    // before the field are never read as neighbours
    const unsigned Ec = qb + (unsigned)min(max(cx + ry * s, 0), (h + 1) * s - 1);
    const float Xk = el(V.X0, Ec), Yk = el(V.Y0, Ec), Zk = el(V.Z0, Ec);
    const UcParams ucp = cam(V.ucp, po);  // before any store: scalar loads
    JointIn in = joint_load<PRIOR, TS, false, TI, D16>(a, V, xc, yc, qb);
    const bool hl = xc > 0, hu = yc > 0;
    const int e_l = max(e - 1, 0), e_u = max(e - EXW, 0), e_r = min(e + 1, NE - 1), e_d = min(e + EXW, NE - 1);
#pragma unroll 1
    for (int k = 0;; ++k) {
        const JointFront J = joint_compute<F64TRI, PRIOR>(a, V, po, xc, yc, in);
        s_du[e] = J.nduh;
        s_dv[e] = J.ndvh;
        s_X[e] = J.oX;
        s_Y[e] = J.oY;
        s_Z[e] = J.oZ;
        __syncthreads();
        const JointBack B = joint_back_calc<EXW, NE, true>(a, ucp, J, xc, yc, w, h, e, s_du, s_dv, s_X, s_Y, s_Z);
        if (k == K - 1) {
            if (lx >= R && lx < EXW - R && ly >= R && ly < EXH - R && cx < w && ry < h) {
                const unsigned q = qb + (unsigned)cx + (unsigned)ry * (unsigned)s;
                joint_store<D16>(V, q, qb, J, B);
                if constexpr (!std::is_void<TI>::value) joint_store_deriv<TS>(V, q, qb, in);
            }
            return;
        }
        const float Xn = pix ? B.X : Xk, Yn = pix ? B.Y : Yk, Zn = pix ? B.Z : Zk;
        // the duals as the next iteration reads them (rounded to the field type)
        const float p1u = (float)st<TD>(B.p1u), p2u = (float)st<TD>(B.p2u), p1v = (float)st<TD>(B.p1v),
                    p2v = (float)st<TD>(B.p2v);
        s_p1u[e] = p1u;
        s_p2u[e] = p2u;
        s_p1v[e] = p1v;
        s_p2v[e] = p2v;
        s_Xn[e] = Xn;
        s_Yn[e] = Yn;
        s_Zn[e] = Zn;
        __syncthreads();
        // state n as the next iteration's inputs (what joint_load reads from memory)
        in.duh = J.nduh, in.dvh = J.ndvh;
        in.p1u = p1u, in.p2u = p2u, in.p1v = p1v, in.p2v = p2v;
        in.p1ul = hl ? s_p1u[e_l] : p1u, in.p1vl = hl ? s_p1v[e_l] : p1v;
        in.p2ud = hu ? s_p2u[e_u] : p2u, in.p2vd = hu ? s_p2v[e_u] : p2v;
        in.sku = B.sku, in.skv = B.skv, in.up0 = J.upn, in.vp0 = J.vpn, in.uc0 = B.uc, in.vc0 = B.vc;
        in.Xc = Xn, in.Yc = Yn, in.Zc = Zn;
        in.Xr = s_Xn[e_r], in.Yr = s_Yn[e_r], in.Zr = s_Zn[e_r];
        in.Xd = s_Xn[e_d], in.Yd = s_Yn[e_d], in.Zd = s_Zn[e_d];
        in.Xl = hl ? s_Xn[e_l] : Xn, in.Yl = hl ? s_Yn[e_l] : Yn, in.Zl = hl ? s_Zn[e_l] : Zn;
        in.Xu = hu ? s_Xn[e_u] : Xn, in.Yu = hu ? s_Yn[e_u] : Yn, in.Zu = hu ? s_Zn[e_u] : Zn;
    }
}

// One TV-L1 iteration: data step (tile + right column + bottom row), dual step (tile).
template <int RPT, typename TS>
__global__ void __launch_bounds__(256) kb_flow_iter(IterArgs a) {
    constexpr int TY = TRW * RPT, EX = TXW + 1, EY = TY + 1, NE = EX * EY, NHALO = EX + TY;
    static_assert(NHALO <= 256, "ring larger than the workgroup");
    __shared__ float s_du[NE], s_dv[NE];
    const int2 pt = a.L.map[xcd_tile(blockIdx.x, a.L.ntiles)];
    const PairFields& F = a.L.P[pt.x];
    const int l = a.level;
    const int w = F.W[l], h = F.H[l], s = F.S[l];
    const int tiles_x = (w + TXW - 1) / TXW;
    const int x0 = (pt.y % tiles_x) * TXW, y0 = (pt.y / tiles_x) * TY;
    const int tid = threadIdx.y * TXW + threadIdx.x;
    const int pd = a.pd;
    const float *du0 = F.du[pd], *dv0 = F.dv[pd];
    const TS *p1u0 = as<TS>(F.p1u[pd]), *p2u0 = as<TS>(F.p2u[pd]), *p1v0 = as<TS>(F.p1v[pd]),
             *p2v0 = as<TS>(F.p2v[pd]);
    const TS *Ix = as<TS>(F.Ix), *Iy = as<TS>(F.Iy), *Iz = as<TS>(F.Iz);
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
            duh = du0[q];
            dvh = dv0[q];
            p1u = p1u0[q];
            p2u = p2u0[q];
            p1v = p1v0[q];
            p2v = p2v0[q];
            if (hl) p1ul = p1u0[q - 1], p1vl = p1v0[q - 1];
            if (hu) p2ud = p2u0[q - s], p2vd = p2v0[q - s];
        }
        float du, dv;
        l1_threshold(Ix[q], Iy[q], Iz[q], duh, dvh, a.lambda, a.theta, du, dv);
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
    TS *p1u1 = as<TS>(F.p1u[pd ^ 1]), *p2u1 = as<TS>(F.p2u[pd ^ 1]), *p1v1 = as<TS>(F.p1v[pd ^ 1]),
       *p2v1 = as<TS>(F.p2v[pd ^ 1]);
    float *du1 = F.du[pd ^ 1], *dv1 = F.dv[pd ^ 1];
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
        p1u1[q] = st<TS>(o1);
        p2u1[q] = st<TS>(o2);
        dual_px(dvx, dvy, rp[r][2], rp[r][3], fac, o1, o2);
        p1v1[q] = st<TS>(o1);
        p2v1[q] = st<TS>(o2);
        du1[q] = s_du[e];
        dv1[q] = s_dv[e];
    }
}

// One TV-L1 iteration with one pixel per lane: the workgroup is an EXW x EXH extended tile whose
// right column and bottom row of lanes only feed the (EXW-1) x (EXH-1) interior (the dual step
// reads uhat_n to the right and below). Every lane loads its inputs up front (coordinates clamped
// into the image), runs the data step -> uhat_n into LDS, passes one barrier; interior lanes run
// the dual step and store. Same per-pixel arithmetic as kb_flow_iter. SINGLE: one pair whose
// tiles are the dispatch indices and whose field table is in the kernel arguments.
template <int EXW, int EXH, bool SINGLE, typename TS>
__global__ void __launch_bounds__(EXW* EXH) kb_flow_px(IterArgs a) {
    constexpr int SM = SINGLE ? 2 : 0;
    constexpr int TW = EXW - 1, TH = EXH - 1, NE = EXW * EXH;
    __shared__ float s_du[NE], s_dv[NE];
    const int tile = xcd_tile(blockIdx.x, a.L.ntiles);
    int tp;
    PairOff po;
    const JointView V = tile_view<SM>(a, tile, TW, tp, po);
    const unsigned qb = po.qb;
    const int w = V.w, h = V.h, s = V.s;
    const int2 txy = tile_xy(tp, V.tiles_x, V.inv_tiles_x);
    const int ex = threadIdx.x, ey = threadIdx.y, e = ex + ey * EXW;
    const int x = txy.x * TW + ex, y = txy.y * TH + ey;
    const int xc = min(x, w - 1), yc = min(y, h - 1);
    const unsigned q = qb + (unsigned)xc + (unsigned)yc * (unsigned)s;
    const bool hl = xc > 0, hu = yc > 0;
    const unsigned ql = hl ? q - 1u : q, qu = hu ? q - (unsigned)s : q;
    const TS *p1u0 = as<TS>(V.p1u0), *p2u0 = as<TS>(V.p2u0), *p1v0 = as<TS>(V.p1v0), *p2v0 = as<TS>(V.p2v0);
    // the same elements in the TS fields (the pair offset counts floats)
    const unsigned dt = qb * (unsigned)(sizeof(float) / sizeof(TS) - 1), qt = q + dt, qlt = ql + dt, qut = qu + dt;
    // unconditional loads: on the first iteration of a warp the view's uhat and dual inputs are
    // the zero field
    const float duh = el(V.du0, q), dvh = el(V.dv0, q);
    const float p1u = (float)el(p1u0, qt), p2u = (float)el(p2u0, qt), p1v = (float)el(p1v0, qt),
                p2v = (float)el(p2v0, qt);
    const float p1ul = (float)el(p1u0, qlt), p1vl = (float)el(p1v0, qlt), p2ud = (float)el(p2u0, qut),
                p2vd = (float)el(p2v0, qut);
    const float ix_ = (float)el(as<TS>(V.Ix), qt), iy_ = (float)el(as<TS>(V.Iy), qt), iz_ = (float)el(as<TS>(V.Iz), qt);
    float du, dv;
    l1_threshold(ix_, iy_, iz_, duh, dvh, a.lambda, a.theta, du, dv);
    const float nduh = du + a.theta * div_p(p1u, p1ul, p2u, p2ud, hl, hu);
    const float ndvh = dv + a.theta * div_p(p1v, p1vl, p2v, p2vd, hl, hu);
    s_du[e] = nduh;
    s_dv[e] = ndvh;
    __syncthreads();
    if (ex == EXW - 1 || ey == EXH - 1 || x >= w || y >= h) return;
    const float fac = a.tau / a.theta;
    float dux = 0.0f, dvx = 0.0f, duy = 0.0f, dvy = 0.0f;
    if (x + 1 < w) {
        dux = s_du[e + 1] - nduh;
        dvx = s_dv[e + 1] - ndvh;
    }
    if (y + 1 < h) {
        duy = s_du[e + EXW] - nduh;
        dvy = s_dv[e + EXW] - ndvh;
    }
    float o1, o2;
    dual_px(dux, duy, p1u, p2u, fac, o1, o2);
    el(as<TS>(V.p1u1), qt) = st<TS>(o1);
    el(as<TS>(V.p2u1), qt) = st<TS>(o2);
    dual_px(dvx, dvy, p1v, p2v, fac, o1, o2);
    el(as<TS>(V.p1v1), qt) = st<TS>(o1);
    el(as<TS>(V.p2v1), qt) = st<TS>(o2);
    el(V.du1, q) = nduh;
    el(V.dv1, q) = ndvh;
}

// K TV-L1 iterations per launch (temporal blocking): the EXW x EXH extended tile starts K - 1
// pixels left of and above its (EXW - 2K + 1) x (EXH - 2K + 1) interior. Iteration k is right on
// lanes k - 1 or more inside the tile on the left and top (the data step reads the duals to the
// left and above) and k or more on the right and bottom (the dual step reads uhat to the right and
// below). State between the iterations stays in registers and LDS; the duals are rounded to the
// storage type between iterations as a store and reload would round them. The launch stores state
// n + K into the other parity, one flip per launch. Same per-pixel arithmetic as kb_flow_px.
// TI (not void): the first launch of a warp, which computes the derivatives at its lanes' pixels
// from the level image and the warped image (element type TI) as kb_deriv does, uses them rounded
// to TS, and stores them for the warp's later launches.
template <int EXW, int EXH, int K, bool SINGLE, typename TS, typename TI = void>
__global__ void __launch_bounds__(EXW* EXH) kb_flow_multi(IterArgs a) {
    constexpr bool DERIV = !std::is_void<TI>::value;
    constexpr int SM = SINGLE ? 2 : 0;
    constexpr int TW = EXW - 2 * K + 1, TH = EXH - 2 * K + 1, NE = EXW * EXH;
    static_assert(TW > 0 && TH > 0, "tile too small for K iterations");
    __shared__ float s_du[NE], s_dv[NE], s_p1u[NE], s_p2u[NE], s_p1v[NE], s_p2v[NE];
    const int tile = xcd_tile(blockIdx.x, a.L.ntiles);
    int tp;
    PairOff po;
    const JointView V = tile_view<SM>(a, tile, TW, tp, po);
    const unsigned qb = po.qb;
    const int w = V.w, h = V.h, s = V.s;
    const int2 txy = tile_xy(tp, V.tiles_x, V.inv_tiles_x);
    const int lx = threadIdx.x, ly = threadIdx.y, e = lx + ly * EXW;
    const int x = txy.x * TW - (K - 1) + lx, y = txy.y * TH - (K - 1) + ly;
    const int xc = min(max(x, 0), w - 1), yc = min(max(y, 0), h - 1);
    const unsigned q = qb + (unsigned)xc + (unsigned)yc * (unsigned)s;
    const bool hl = xc > 0, hu = yc > 0;
    const unsigned ql = hl ? q - 1u : q, qu = hu ? q - (unsigned)s : q;
    const TS *p1u0 = as<TS>(V.p1u0), *p2u0 = as<TS>(V.p2u0), *p1v0 = as<TS>(V.p1v0), *p2v0 = as<TS>(V.p2v0);
    // the same elements in the TS fields (the pair offset counts floats)
    const unsigned dt = qb * (unsigned)(sizeof(float) / sizeof(TS) - 1), qt = q + dt, qlt = ql + dt, qut = qu + dt;
    float duh = el(V.du0, q), dvh = el(V.dv0, q);
    float p1u = (float)el(p1u0, qt), p2u = (float)el(p2u0, qt), p1v = (float)el(p1v0, qt), p2v = (float)el(p2v0, qt);
    float p1ul = (float)el(p1u0, qlt), p1vl = (float)el(p1v0, qlt), p2ud = (float)el(p2u0, qut),
          p2vd = (float)el(p2v0, qut);
    float ix_, iy_, iz_;
    if constexpr (DERIV) {
        const unsigned qi = qb * (unsigned)(sizeof(float) / sizeof(TI));  // the pair's images
        float dx, dy, dtt;
        deriv_px(as<TI>(V.I0) + qi, as<TI>(V.I1w) + qi, xc, yc, w, h, s, dx, dy, dtt);
        ix_ = (float)st<TS>(dx), iy_ = (float)st<TS>(dy), iz_ = (float)st<TS>(dtt);
    } else {
        ix_ = (float)el(as<TS>(V.Ix), qt), iy_ = (float)el(as<TS>(V.Iy), qt), iz_ = (float)el(as<TS>(V.Iz), qt);
    }
    const float fac = a.tau / a.theta;
    const int e_l = max(e - 1, 0), e_u = max(e - EXW, 0), e_r = min(e + 1, NE - 1), e_d = min(e + EXW, NE - 1);
#pragma unroll 1
    for (int k = 0;; ++k) {
        float du, dv;
        l1_threshold(ix_, iy_, iz_, duh, dvh, a.lambda, a.theta, du, dv);
        const float nduh = du + a.theta * div_p(p1u, p1ul, p2u, p2ud, hl, hu);
        const float ndvh = dv + a.theta * div_p(p1v, p1vl, p2v, p2vd, hl, hu);
        s_du[e] = nduh;
        s_dv[e] = ndvh;
        __syncthreads();
        float dux = 0.0f, dvx = 0.0f, duy = 0.0f, dvy = 0.0f;
        if (xc + 1 < w) {
            dux = s_du[e_r] - nduh;
            dvx = s_dv[e_r] - ndvh;
        }
        if (yc + 1 < h) {
            duy = s_du[e_d] - nduh;
            dvy = s_dv[e_d] - ndvh;
        }
        float o1u, o2u, o1v, o2v;
        dual_px(dux, duy, p1u, p2u, fac, o1u, o2u);
        dual_px(dvx, dvy, p1v, p2v, fac, o1v, o2v);
        if (k == K - 1) {
            if (lx >= K - 1 && lx < K - 1 + TW && ly >= K - 1 && ly < K - 1 + TH && x < w && y < h) {
                el(as<TS>(V.p1u1), qt) = st<TS>(o1u);
                el(as<TS>(V.p2u1), qt) = st<TS>(o2u);
                el(as<TS>(V.p1v1), qt) = st<TS>(o1v);
                el(as<TS>(V.p2v1), qt) = st<TS>(o2v);
                el(V.du1, q) = nduh;
                el(V.dv1, q) = ndvh;
                if constexpr (DERIV) {
                    el(as<TS>(const_cast<float*>(V.Ix)), qt) = st<TS>(ix_);
                    el(as<TS>(const_cast<float*>(V.Iy)), qt) = st<TS>(iy_);
                    el(as<TS>(const_cast<float*>(V.Iz)), qt) = st<TS>(iz_);
                }
            }
            return;
        }
        p1u = (float)st<TS>(o1u), p2u = (float)st<TS>(o2u), p1v = (float)st<TS>(o1v), p2v = (float)st<TS>(o2v);
        s_p1u[e] = p1u;
        s_p2u[e] = p2u;
        s_p1v[e] = p1v;
        s_p2v[e] = p2v;
        __syncthreads();
        duh = nduh, dvh = ndvh;
        p1ul = hl ? s_p1u[e_l] : p1u, p1vl = hl ? s_p1v[e_l] : p1v;
        p2ud = hu ? s_p2u[e_u] : p2u, p2vd = hu ? s_p2v[e_u] : p2v;
    }
}

// Zero uhat (a warp with no solver iterations leaves the zeroed uhat of the reference).
__global__ void __launch_bounds__(256) kb_zero_uhat(Launch L, int l, int pd) {
    BT_PIX(L, l)
    F.du[pd][pos] = 0.0f;
    F.dv[pd][pos] = 0.0f;
}

}  // namespace

// ===================================================================================
// Batch

Batch::~Batch() { release(); }

// Teardown: a failing destroy or free leaves nothing to recover, so its status is not checked.
void Batch::drop_graphs() {
    for (auto& g : gexec_)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

void Batch::release() {
    for (auto& e : prof_.ev) (void)hipEventDestroy(e);
    prof_.ev.clear();
    drop_graphs();
    if (ws_bad_) (void)hipFree(ws_bad_);
    ws_bad_ = nullptr;
    if (maps_) (void)hipFree(maps_);
    if (dp_) (void)hipFree(dp_);
    if (slab_) (void)hipFree(slab_);
    if (st_) (void)hipStreamDestroy(st_);
    maps_ = nullptr;
    dp_ = nullptr;
    slab_ = nullptr;
    st_ = nullptr;
}

int Batch::check(hipError_t e, const char* what) {
    if (e == hipSuccess) return RF_OK;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return RF_ERR_HIP;
}

static int align_up_b(int n, int m) {
    int mod = n % m;
    return mod ? n + m - mod : n;
}

int Batch::create(const rf_params& p, int n, const int* widths, const int* heights) {
    if (n <= 0 || !widths || !heights || p.n_levels <= 0 || p.n_levels > kMaxLevels || !(p.scale > 1.0f)) {
        set_error("rf_batch_create: invalid count, sizes or pyramid");
        return RF_ERR_ARG;
    }
    if (p.input_type < RF_INPUT_32F || p.input_type > RF_INPUT_16U) {
        set_error("rf_batch_create: unknown input type");
        return RF_ERR_ARG;
    }
    if (p.joint_tile < RF_JOINT_TILE_AUTO || p.joint_tile > RF_JOINT_TILE_16x16_K3 ||
        p.flow_tile < RF_FLOW_TILE_AUTO || p.flow_tile > RF_FLOW_TILE_32x16_K6) {
        set_error("rf_batch_create: unknown joint or TV-L1 tiling");
        return RF_ERR_ARG;
    }
    if (p.storage < RF_STORAGE_F32 || p.storage > RF_STORAGE_F16_DERIV_DUALS) {
        set_error("rf_batch_create: unknown storage");
        return RF_ERR_ARG;
    }
    const int align = p.stride_alignment > 0 ? p.stride_alignment : 32;
    std::vector<Geo> geo(n);
    for (int i = 0; i < n; ++i) {
        int nh = heights[i], nw = widths[i];
        if (nw <= 0 || nh <= 0) {
            set_error("rf_batch_create: pair " + std::to_string(i) + " has no pixels");
            return RF_ERR_ARG;
        }
        if (nw >= kMaxSide || nh >= kMaxSide) {  // 32-bit texel-coordinate products in the resampling
            set_error("rf_batch_create: pair " + std::to_string(i) + ": width and height must be below " +
                      std::to_string(kMaxSide));
            return RF_ERR_ARG;
        }
        for (int l = 0; l < p.n_levels; ++l) {
            if (nw < 1 || nh < 1) {
                set_error("rf_batch_create: pair " + std::to_string(i) + " runs out of pixels at level " +
                          std::to_string(l));
                return RF_ERR_ARG;
            }
            if (l > 0 && !(resample_fits(geo[i].W[l - 1], nw) && resample_fits(geo[i].H[l - 1], nh))) {
                set_error("rf_batch_create: pair " + std::to_string(i) + ": levels " + std::to_string(l - 1) +
                          " and " + std::to_string(l) + " are too large for 32-bit texel coordinates at this scale");
                return RF_ERR_ARG;
            }
            geo[i].W.push_back(nw);
            geo[i].H.push_back(nh);
            geo[i].S.push_back(align_up_b(nw, align));
            nh = (int)((float)nh / p.scale);
            nw = (int)((float)nw / p.scale);
        }
        geo[i].n0 = (long)geo[i].S[0] * geo[i].H[0];
        geo[i].bufn = ((geo[i].n0 + geo[i].S[0]) + 63) / 64 * 64;
        if (geo[i].bufn > (1L << 30)) {  // the fused kernels address fields with 32-bit byte offsets
            set_error("rf_batch_create: pair " + std::to_string(i) + " exceeds 2^30 elements per field");
            return RF_ERR_ARG;
        }
    }
    release();
    p_ = p;
    geo_ = geo;
    if (check(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "hipStreamCreate")) return RF_ERR_HIP;
    const int L = p.n_levels;
    const long per_fields = 2L * L + 51;  // pyramids, inputs, derivatives, state double buffers, prior, uvrgb(3), scratch, zero
    size_t total = 0;
    for (auto& g : geo_) total += (size_t)per_fields * g.bufn * sizeof(float);
    // pairs of one size: the same field sits per_fields * bufn floats further on in each pair; the
    // kernels reach it through 32-bit byte offsets, so the slab must stay below 2^30 floats
    pair_elems_ = 0;
    bool same = n > 1 && (long long)per_fields * geo_[0].bufn * n < (1LL << 30);
    for (int i = 1; i < n && same; ++i) same = geo_[i].W == geo_[0].W && geo_[i].H == geo_[0].H && geo_[i].S == geo_[0].S;
    if (same) pair_elems_ = (unsigned)(per_fields * geo_[0].bufn);
    slab_bytes_ = total;
    if (hipMalloc(&slab_, total) != hipSuccess) {
        slab_ = nullptr;
        set_error("rf_batch_create: device allocation of " + std::to_string(total) + " bytes failed");
        release();
        return RF_ERR_ALLOC;
    }
    if (check(hipMemset(slab_, 0, total), "hipMemset")) return RF_ERR_HIP;
    wsu_.assign(n, nullptr);
    wsv_.assign(n, nullptr);
    if (hipMalloc(&ws_bad_, sizeof(int) * n) != hipSuccess) {
        ws_bad_ = nullptr;
        set_error("rf_batch_create: device allocation of the warm-start flags failed");
        release();
        return RF_ERR_ALLOC;
    }
    {
        const std::vector<int> unused(n, 1);  // until a solve applies the rule
        if (check(hipMemcpy(ws_bad_, unused.data(), sizeof(int) * n, hipMemcpyHostToDevice), "rf_batch_create"))
            return RF_ERR_HIP;
    }
    hp_.assign(n, PairFields{});
    float* cur = (float*)slab_;
    for (int i = 0; i < n; ++i) {
        PairFields& F = hp_[i];
        const Geo& g = geo_[i];
        auto take = [&](long k) {
            float* r = cur;
            cur += k * g.bufn;
            return r;
        };
        for (int l = 0; l < L; ++l) {
            F.W[l] = g.W[l];
            F.H[l] = g.H[l];
            F.S[l] = g.S[l];
            F.I0[l] = take(1);
            F.I1[l] = take(1);
        }
        F.in0 = take(1);
        F.in1 = take(1);
        F.i1warp = take(1);
        F.Ix = take(1);
        F.Iy = take(1);
        F.Iz = take(1);
        for (int k = 0; k < 2; ++k) {
            F.du[k] = take(1);
            F.dv[k] = take(1);
            F.p1u[k] = take(1);
            F.p2u[k] = take(1);
            F.p1v[k] = take(1);
            F.p2v[k] = take(1);
            F.up[k] = take(1);
            F.vp[k] = take(1);
            F.sku[k] = take(1);
            F.skv[k] = take(1);
            F.uc[k] = take(1);
            F.vc[k] = take(1);
            F.X[k] = take(1);
            F.Y[k] = take(1);
            F.Z[k] = take(1);
            F.u[k] = take(1);
            F.v[k] = take(1);
        }
        F.ufn = take(1);
        F.vfn = take(1);
        F.ufn_l = take(1);
        F.vfn_l = take(1);
        F.mask = take(1);
        F.mask_l = take(1);
        F.uvrgb = take(3);
        F.scratch = take(1);
        F.zero = take(1);  // zero from the slab's clearing, never written
    }
    if (hipMalloc((void**)&dp_, sizeof(PairFields) * n) != hipSuccess) {
        dp_ = nullptr;
        set_error("rf_batch_create: descriptor allocation failed");
        release();
        return RF_ERR_ALLOC;
    }
    // tile lists: levels 0..L-1 and the flat level-0 buffer (pointwise tiles only)
    sets_.assign((kMaxLevels + 1) * TC_COUNT, TileSet{});
    std::vector<int2> all;
    std::vector<long> offs(sets_.size(), -1);
    for (int lv = 0; lv <= kMaxLevels; ++lv) {
        if (lv >= L && lv != FLAT) continue;
        for (int c = 0; c < TC_COUNT; ++c) {
            if (lv == FLAT && c != TC_PIX) continue;
            const int k = lv * TC_COUNT + c;
            offs[k] = (long)all.size();
            for (int i = 0; i < n; ++i) {
                const int w = lv == FLAT ? geo_[i].S[0] : geo_[i].W[lv];
                const int h = lv == FLAT ? geo_[i].H[0] : geo_[i].H[lv];
                const int nt = ((w + kTileW[c] - 1) / kTileW[c]) * ((h + kTileH[c] - 1) / kTileH[c]);
                for (int t = 0; t < nt; ++t) all.push_back(make_int2(i, t));
            }
            sets_[k].ntiles = (int)(all.size() - offs[k]);
        }
    }
    if (hipMalloc((void**)&maps_, sizeof(int2) * std::max<size_t>(all.size(), 1)) != hipSuccess) {
        maps_ = nullptr;
        set_error("rf_batch_create: tile list allocation failed");
        release();
        return RF_ERR_ALLOC;
    }
    if (check(hipMemcpy(maps_, all.data(), sizeof(int2) * all.size(), hipMemcpyHostToDevice), "upload tiles"))
        return RF_ERR_HIP;
    for (size_t k = 0; k < sets_.size(); ++k)
        if (offs[k] >= 0) sets_[k].map = maps_ + offs[k];
    prior_ = false;
    dirty_ = true;
    for (int k = 0; k < 2; ++k) {
        tuned_[k] = false;
        cls_[k].clear();
    }
    return upload_pairs();
}

TileSet& Batch::tiles(int level, int cls) { return sets_[level * TC_COUNT + cls]; }

// Iteration tiling of a level for a warp (TV-L1 or joint): a fixed choice; during the tuning
// solves the candidates in turn over the warps of as many solves as it takes to time each one;
// afterwards the tuned choice (the first candidate until then).
constexpr int kNumCand = 6;
static const int kTuneCand[2][kNumCand] = {{TC_F16, TC_FK4, TC_F32, TC_FK6, TC_FK8, TC_FLOW},
                                           {TC_J16, TC_J32, TC_JK2, TC_J16O, TC_JK3, TC_JK3}};
static const int kCandCount[2] = {6, 5};

// The candidate a tuning solve times on a warp: round r covers candidates r * nw .. r * nw + nw - 1
// (mod the candidate count), so every candidate is timed in at least two rounds, and the warps take
// them in an order rotated by r, so a candidate's samples come from different warps of the level.
int Batch::tune_slot(int k, int round, int warp) const {
    const int nw = p_.n_warp_iters;
    return (round * nw + (warp + round) % nw) % kCandCount[k];
}

// The class of the last `rem` iterations of a warp when the level's class runs more per launch:
// the most iterations per launch that do not pass rem.
static int remainder_class(bool recon, int rem) {
    if (recon) return rem >= 3 ? TC_JK3 : (rem >= 2 ? TC_JK2 : TC_J16);
    return rem >= 4 ? TC_FK4 : (rem >= 2 ? TC_FK2 : TC_F16);
}

int Batch::tile_class(bool recon, int level, int warp) const {
    if (recon) {
        switch (p_.joint_tile) {
            case RF_JOINT_TILE_64x8_RPT2: return TC_J64;
            case RF_JOINT_TILE_32x16: return TC_J32;
            case RF_JOINT_TILE_16x16: return TC_J16;
            case RF_JOINT_TILE_16x16_OCC5: return TC_J16O;
            case RF_JOINT_TILE_16x16_K2: return TC_JK2;
            case RF_JOINT_TILE_16x16_K3: return TC_JK3;
            default: break;
        }
    } else {
        switch (p_.flow_tile) {
            case RF_FLOW_TILE_64x16_RPT4: return TC_FLOW;
            case RF_FLOW_TILE_16x16: return TC_F16;
            case RF_FLOW_TILE_32x16: return TC_F32;
            case RF_FLOW_TILE_32x16_K4: return TC_FK4;
            case RF_FLOW_TILE_32x32_K8: return TC_FK8;
            case RF_FLOW_TILE_32x16_K2: return TC_FK2;
            case RF_FLOW_TILE_32x16_K6: return TC_FK6;
            default: break;
        }
    }
    const int k = recon ? 1 : 0;
    if (tuning_) return kTuneCand[k][tune_slot(k, tune_round_, warp)];
    return tuned_[k] ? cls_[k][level] : kTuneCand[k][0];
}

int Batch::joint_tiling(int level) const {
    if (p_.joint_tile != RF_JOINT_TILE_AUTO) return p_.joint_tile;
    if (!tuned_[1] || level < 0 || level >= (int)cls_[1].size()) return RF_JOINT_TILE_AUTO;
    switch (cls_[1][level]) {
        case TC_J32: return RF_JOINT_TILE_32x16;
        case TC_J16O: return RF_JOINT_TILE_16x16_OCC5;
        case TC_JK2: return RF_JOINT_TILE_16x16_K2;
        case TC_JK3: return RF_JOINT_TILE_16x16_K3;
        default: return RF_JOINT_TILE_16x16;
    }
}

int Batch::flow_tiling(int level) const {
    if (p_.flow_tile != RF_FLOW_TILE_AUTO) return p_.flow_tile;
    if (!tuned_[0] || level < 0 || level >= (int)cls_[0].size()) return RF_FLOW_TILE_AUTO;
    switch (cls_[0][level]) {
        case TC_FLOW: return RF_FLOW_TILE_64x16_RPT4;
        case TC_F32: return RF_FLOW_TILE_32x16;
        case TC_FK4: return RF_FLOW_TILE_32x16_K4;
        case TC_FK8: return RF_FLOW_TILE_32x32_K8;
        case TC_FK2: return RF_FLOW_TILE_32x16_K2;
        case TC_FK6: return RF_FLOW_TILE_32x16_K6;
        default: return RF_FLOW_TILE_16x16;
    }
}

// The tuning solves: eager solves whose warps alternate between the candidate tilings, timed
// per (level, warp) with events; every candidate is timed on at least two warps of each level and
// the level keeps the candidate with the shortest warp (candidates within a few percent of each
// other are common, and one sample per candidate picked the slower one on some runs). Their results
// are those of any other solve.
int Batch::tune_tiling(bool recon) {
    const int L = p_.n_levels, nw = p_.n_warp_iters, k = recon ? 1 : 0;
    const int nc = kCandCount[k], rounds = 2 * ((nc + nw - 1) / nw);  // every candidate timed twice or more
    tev_.assign(2 * L * nw, nullptr);
    int rc = RF_OK;
    for (auto& ev : tev_)
        if ((rc = check(hipEventCreate(&ev), "hipEventCreate")) != RF_OK) break;
    std::vector<double> t((size_t)L * nc, 0.0);
    std::vector<int> cnt((size_t)L * nc, 0);
    for (int r = 0; r < rounds && !rc; ++r) {
        tune_round_ = r;
        tuning_ = true;
        hipError_t e = enqueue(recon);
        tuning_ = false;
        rc = check(e, "tuning solve");
        if (!rc) rc = check(hipStreamSynchronize(st_), "tuning solve");
        for (int l = 0; l < L && !rc; ++l)
            for (int wi = 0; wi < nw; ++wi) {
                float ms = 0;
                if (check(hipEventElapsedTime(&ms, tev_[2 * (l * nw + wi)], tev_[2 * (l * nw + wi) + 1]),
                          "hipEventElapsedTime")) {
                    rc = RF_ERR_HIP;
                    break;
                }
                const size_t i = (size_t)l * nc + tune_slot(k, r, wi);
                t[i] = cnt[i] ? std::min(t[i], (double)ms) : (double)ms;
                ++cnt[i];
            }
    }
    cls_[k].assign(L, kTuneCand[k][0]);
    for (int l = 0; l < L && !rc; ++l) {
        double best = 0;
        bool any = false;
        for (int c = 0; c < nc; ++c) {
            const int i = l * nc + c;
            if (cnt[i] && (!any || t[i] < best)) {
                best = t[i];
                cls_[k][l] = kTuneCand[k][c];
                any = true;
            }
        }
    }
    for (auto& ev : tev_)
        if (ev) (void)hipEventDestroy(ev);
    tev_.clear();
    tuned_[k] = rc == RF_OK;
    return rc;
}

int Batch::upload_pairs() {
    if (check(hipMemcpyAsync(dp_, hp_.data(), sizeof(PairFields) * hp_.size(), hipMemcpyHostToDevice, st_),
              "upload pair descriptors"))
        return RF_ERR_HIP;
    if (check(hipStreamSynchronize(st_), "upload pair descriptors")) return RF_ERR_HIP;
    dirty_ = false;
    return RF_OK;
}

static size_t in_bytes(int t) {
    switch (t) {
        case RF_INPUT_8UC3: return 3;
        case RF_INPUT_8U: return 1;
        case RF_INPUT_16U: return 2;
        default: return 4;
    }
}

int Batch::set_images_host(int i, const void* i0, const void* i1) {
    if (i < 0 || i >= n() || !i0 || !i1) {
        set_error("rf_batch: bad pair index or null image");
        return RF_ERR_ARG;
    }
    size_t bytes = (size_t)geo_[i].n0 * in_bytes(p_.input_type);
    if (check(hipMemcpyAsync((void*)hp_[i].in0, i0, bytes, hipMemcpyHostToDevice, st_), "copy image 0") ||
        check(hipMemcpyAsync((void*)hp_[i].in1, i1, bytes, hipMemcpyHostToDevice, st_), "copy image 1"))
        return RF_ERR_HIP;
    return check(hipStreamSynchronize(st_), "copy images");
}

int Batch::set_images_device(int i, const float* i0, const float* i1) {
    if (i < 0 || i >= n() || !i0 || !i1) {
        set_error("rf_batch_set_images_device: bad pair index or null image");
        return RF_ERR_ARG;
    }
    if (p_.input_type != RF_INPUT_32F) {
        set_error("rf_batch_set_images_device: the batch expects a non-float input type");
        return RF_ERR_ARG;
    }
    size_t bytes = (size_t)geo_[i].n0 * 4;
    if (check(hipMemcpyAsync((void*)hp_[i].in0, i0, bytes, hipMemcpyDeviceToDevice, st_), "copy image 0")) return RF_ERR_HIP;
    return check(hipMemcpyAsync((void*)hp_[i].in1, i1, bytes, hipMemcpyDeviceToDevice, st_), "copy image 1");
}

// setCameraMatrices + the per-level P, Q (Reconflow.cu:78-116, 225-230).
int Batch::set_camera(int i, const double* K0, const double* K1, const double* R0, const double* t0,
                      const double* R1, const double* t1) {
    if (i < 0 || i >= n() || !K0 || !K1 || !R1 || !t1) {
        set_error("rf_batch_set_camera: bad pair index or null matrix");
        return RF_ERR_ARG;
    }
    static const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, Z3[3] = {0, 0, 0};
    if (!R0) R0 = I3;
    if (!t0) t0 = Z3;
    double k0[9], k1[9];
    memcpy(k0, K0, sizeof k0);
    memcpy(k1, K1, sizeof k1);
    for (int l = 0; l < p_.n_levels; ++l) {
        if (l > 0) {
            for (double* K : {k0, k1}) {
                double nk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                nk[0] = K[0] / (double)p_.scale;
                nk[2] = K[2] / (double)p_.scale;
                nk[4] = K[4] / (double)p_.scale;
                nk[5] = K[5] / (double)p_.scale;
                if (p_.fix_coarse_k) nk[8] = 1.0;
                memcpy(K, nk, sizeof nk);
            }
        }
        double P[12], Q[12];
        for (int pass = 0; pass < 2; ++pass) {
            const double* K = pass ? k1 : k0;
            const double* R = pass ? R1 : R0;
            const double* t = pass ? t1 : t0;
            double* out = pass ? Q : P;
            for (int a = 0; a < 3; ++a) {
                for (int b = 0; b < 3; ++b) {
                    double acc = 0;
                    for (int k = 0; k < 3; ++k) acc += K[a * 3 + k] * R[k * 3 + b];
                    out[a * 4 + b] = acc;
                }
                double acc = 0;
                for (int k = 0; k < 3; ++k) acc += K[a * 3 + k] * t[k];
                out[a * 4 + 3] = acc;
            }
        }
        hp_[i].pq[l] = make_pq(P, Q);
        hp_[i].pqd[l] = make_pqd(P, Q);
        hp_[i].ucp[l] = make_uc(R1, t1, k1);
    }
    dirty_ = true;
    return RF_OK;
}

int Batch::set_prior(int i, const float* u, const float* v, const float* mask, bool sparse) {
    if (i < 0 || i >= n() || !u || !v) {
        set_error("rf_batch: bad pair index or null prior");
        return RF_ERR_ARG;
    }
    size_t bytes = (size_t)geo_[i].n0 * 4;
    if (check(hipMemcpyAsync(hp_[i].ufn, u, bytes, hipMemcpyHostToDevice, st_), "copy prior u") ||
        check(hipMemcpyAsync(hp_[i].vfn, v, bytes, hipMemcpyHostToDevice, st_), "copy prior v"))
        return RF_ERR_HIP;
    std::vector<float> ones;
    if (!mask) {
        ones.assign(geo_[i].n0, 1.0f);
        mask = ones.data();
    }
    if (check(hipMemcpyAsync(hp_[i].mask, mask, bytes, hipMemcpyHostToDevice, st_), "copy mask")) return RF_ERR_HIP;
    sparse_ = sparse;
    prior_ = true;
    return check(hipStreamSynchronize(st_), "copy prior");
}

// Warm start (no counterpart in the reference): the schedule gains three launches per pair with
// one, so a captured schedule is captured again when a pair gains or loses its warm start.
int Batch::set_warm_start(int i, const float* u, const float* v) {
    if (i < 0 || i >= n()) {
        set_error("rf_batch_set_warm_start_device: bad pair index");
        return RF_ERR_ARG;
    }
    if (!u || !v) u = v = nullptr;
    // a solve clears the batch's state fields before it reads the warm start, so the warm start
    // may not live in them (a result handed straight back, for one): copy it out first
    const char *b0 = (const char*)slab_, *b1 = b0 + slab_bytes_;
    for (const float* q : {u, v})
        if (q && (const char*)q >= b0 && (const char*)q < b1) {
            set_error("rf_batch_set_warm_start_device: the warm start lies in the batch's own fields, which a solve "
                      "clears first; copy it to a buffer of its own");
            return RF_ERR_ARG;
        }
    if ((u != nullptr) != (wsu_[i] != nullptr) || u != wsu_[i] || v != wsv_[i]) drop_graphs();
    wsu_[i] = u;
    wsv_[i] = v;
    return RF_OK;
}

int Batch::warm_start_in_use(int i) {
    if (i < 0 || i >= n()) {
        set_error("rf_batch_warm_start_in_use: bad pair index");
        return RF_ERR_ARG;
    }
    if (!wsu_[i]) return 0;
    int bad = 1;
    if (check(hipStreamSynchronize(st_), "warm start flag") ||
        check(hipMemcpy(&bad, ws_bad_ + i, sizeof bad, hipMemcpyDeviceToHost), "warm start flag"))
        return RF_ERR_HIP;
    return bad ? 0 : 1;
}

// TK: the image type when the launch also computes the warp's derivatives, else void
template <typename TK, typename TS, int SM, bool F64, bool PRIOR, bool D16>
static void launch_joint_t(int cls, hipStream_t st, const IterArgs& a) {
    const dim3 grid(a.L.ntiles);
    switch (cls) {
        case TC_J64:  // never asked to compute the derivatives (fuses_deriv)
            hipLaunchKernelGGL((kb_joint_iter<RPT_JOINT, F64, PRIOR, TS, D16>), grid, dim3(TXW, TRW), 0, st, a);
            break;
        case TC_J32:
            hipLaunchKernelGGL((kb_joint_px<32, 16, 4, SM, F64, PRIOR, TS, TK, D16>), grid, dim3(32, 16), 0, st, a);
            break;
        case TC_J16O:
            hipLaunchKernelGGL((kb_joint_px<16, 16, 5, SM, F64, PRIOR, TS, TK, D16>), grid, dim3(16, 16), 0, st, a);
            break;
        case TC_JK2:
            hipLaunchKernelGGL((kb_joint_multi<16, 16, 2, SM, F64, PRIOR, TS, TK, D16>), grid, dim3(16, 16), 0, st, a);
            break;
        case TC_JK3:
            hipLaunchKernelGGL((kb_joint_multi<16, 16, 3, SM, F64, PRIOR, TS, TK, D16>), grid, dim3(16, 16), 0, st, a);
            break;
        default:
            hipLaunchKernelGGL((kb_joint_px<16, 16, 1, SM, F64, PRIOR, TS, TK, D16>), grid, dim3(16, 16), 0, st, a);
            break;
    }
}

// deriv: the launch also computes the warp's derivatives (the multi-iteration classes; the caller
// then skips kb_deriv)
template <typename TI, typename TS>
static void launch_flow(int cls, bool single, bool deriv, hipStream_t st, const IterArgs& a) {
    const dim3 grid(a.L.ntiles);
#define RF_FM(EXW, EXH, K)                                                                                           \
    do {                                                                                                             \
        const dim3 blk(EXW, EXH);                                                                                    \
        if (deriv) {                                                                                                 \
            if (single) hipLaunchKernelGGL((kb_flow_multi<EXW, EXH, K, true, TS, TI>), grid, blk, 0, st, a);         \
            else hipLaunchKernelGGL((kb_flow_multi<EXW, EXH, K, false, TS, TI>), grid, blk, 0, st, a);               \
        } else {                                                                                                     \
            if (single) hipLaunchKernelGGL((kb_flow_multi<EXW, EXH, K, true, TS>), grid, blk, 0, st, a);             \
            else hipLaunchKernelGGL((kb_flow_multi<EXW, EXH, K, false, TS>), grid, blk, 0, st, a);                   \
        }                                                                                                            \
    } while (0)
    switch (cls) {
        case TC_F16:
            if (single) hipLaunchKernelGGL((kb_flow_px<16, 16, true, TS>), grid, dim3(16, 16), 0, st, a);
            else hipLaunchKernelGGL((kb_flow_px<16, 16, false, TS>), grid, dim3(16, 16), 0, st, a);
            break;
        case TC_F32:
            if (single) hipLaunchKernelGGL((kb_flow_px<32, 16, true, TS>), grid, dim3(32, 16), 0, st, a);
            else hipLaunchKernelGGL((kb_flow_px<32, 16, false, TS>), grid, dim3(32, 16), 0, st, a);
            break;
        case TC_FK4: RF_FM(32, 16, 4); break;
        case TC_FK8: RF_FM(32, 32, 8); break;
        case TC_FK2: RF_FM(32, 16, 2); break;
        case TC_FK6: RF_FM(32, 16, 6); break;
        default: hipLaunchKernelGGL((kb_flow_iter<RPT_FLOW, TS>), grid, dim3(TXW, TRW), 0, st, a); break;
    }
#undef RF_FM
}

// The iteration classes whose kernel can compute the warp's derivatives in its first launch.
static bool fuses_deriv(int cls) {
    return cls == TC_FK2 || cls == TC_FK4 || cls == TC_FK6 || cls == TC_FK8 || cls == TC_J16 || cls == TC_J32 ||
           cls == TC_J16O || cls == TC_JK2 || cls == TC_JK3;
}

// sm: how the kernels find the fields (see tile_view); deriv: the launch also computes the warp's
// derivatives (the caller then skips kb_deriv); d16: the dual fields in binary16 (with 16-bit
// derivative fields only)
template <typename TI, typename TS>
static void launch_joint(int cls, int sm, bool f64, bool prior, bool deriv, bool d16, hipStream_t st,
                         const IterArgs& a) {
    const int k = (f64 ? 2 : 0) | (prior ? 1 : 0);
#define RF_J(TK, SM, D)                                                            \
    switch (k) {                                                                   \
        case 0: launch_joint_t<TK, TS, SM, false, false, D>(cls, st, a); break;    \
        case 1: launch_joint_t<TK, TS, SM, false, true, D>(cls, st, a); break;     \
        case 2: launch_joint_t<TK, TS, SM, true, false, D>(cls, st, a); break;     \
        default: launch_joint_t<TK, TS, SM, true, true, D>(cls, st, a); break;     \
    }
#define RF_JS(TK, D)       \
    if (sm == 1) {         \
        RF_J(TK, 1, D)     \
    } else if (sm == 2) {  \
        RF_J(TK, 2, D)     \
    } else {               \
        RF_J(TK, 0, D)     \
    }
#define RF_JD(D)          \
    if (deriv) {          \
        RF_JS(TI, D)      \
    } else {              \
        RF_JS(void, D)    \
    }
    if constexpr (std::is_same<TS, __half>::value) {
        if (d16) {
            RF_JD(true)
            return;
        }
    }
    RF_JD(false)
#undef RF_JD
#undef RF_JS
#undef RF_J
}

#define RF_TRY(x)                          \
    do {                                   \
        hipError_t e__ = (x);              \
        if (e__ != hipSuccess) return e__; \
    } while (0)

#define RF_LAUNCH(K, TS, ...)                                                          \
    do {                                                                               \
        const TileSet& ts__ = (TS);                                                    \
        if (ts__.ntiles > 0) {                                                         \
            Launch L__{dp_, ts__.map, ts__.ntiles};                                    \
            hipLaunchKernelGGL(K, dim3(ts__.ntiles), dim3(TXW, TRW), 0, st_, L__, ##__VA_ARGS__); \
            RF_TRY(hipGetLastError());                                                 \
        }                                                                              \
    } while (0)

hipError_t Batch::enqueue(bool recon) {
    switch (p_.storage) {
        case RF_STORAGE_F16: return enqueue_t<__half, __half>(recon);
        case RF_STORAGE_F16_DERIV:
        case RF_STORAGE_F16_DERIV_DUALS: return enqueue_t<float, __half>(recon);
        default: return enqueue_t<float, float>(recon);
    }
}

template <typename TI, typename TS>
hipError_t Batch::enqueue_t(bool recon) {
    const int L = p_.n_levels;
    Parity pr{};
    RF_LAUNCH(kb_gray<TI>, tiles(0, TC_PIX), p_.input_type);
    for (int l = 1; l < L; ++l) RF_LAUNCH(kb_pyramid<TI>, tiles(l, TC_PIX), l);
    RF_LAUNCH(kb_zero, tiles(FLAT, TC_PIX), recon ? 1 : 0);
    for (int i = 0; i < n(); ++i)
        if (wsu_[i]) {
            const Geo& g = geo_[i];
            RF_TRY(launch_warm_start(wsu_[i], wsv_[i], g.W[0], g.H[0], g.S[0], g.W[L - 1], g.H[L - 1], g.S[L - 1],
                                     ws_bad_ + i, hp_[i].u[0], hp_[i].v[0], st_));
        }
    const bool fn = recon && prior_ && (p_.method == RF_METHODR_TVL1_MS_FN || p_.method == RF_METHODR_TVL1_MS_FNSPARSE);
    const float theta = recon ? 1.0f / p_.alpha_tv : p_.theta;  // Reconflow.cu:32
    for (int level = L - 1; level >= 0; --level) {
        if (recon) {
            if (fn) RF_LAUNCH(kb_prior, tiles(level, TC_PIX), level);
            RF_LAUNCH(kb_sync_x, tiles(FLAT, TC_PIX), pr.x);
        }
        bool warped = false;  // this warp's image was warped by the previous warp's last kernel
        for (int wi = 0; wi < p_.n_warp_iters; ++wi) {
            if (!warped) RF_LAUNCH(kb_warp<TI>, tiles(level, TC_PIX), level, pr.u);
            // the warp's first iteration launch computes the derivatives itself where its kernel can
            int cls0 = tile_class(recon, level, wi);
            if (kIters[cls0] > p_.n_solver_iters) cls0 = remainder_class(recon, p_.n_solver_iters);
            const bool fused_deriv = p_.n_solver_iters > 0 && fuses_deriv(cls0);
            if (!fused_deriv) RF_LAUNCH((kb_deriv<TI, TS>), tiles(level, TC_PIX), level);
            if (p_.n_solver_iters == 0) RF_LAUNCH(kb_zero_uhat, tiles(level, TC_PIX), level, pr.d);
            prof_.begin(st_, RF_KID_FUSED_ITER, level);
            if (tuning_) RF_TRY(hipEventRecord(tev_[2 * (level * p_.n_warp_iters + wi)], st_));
            for (int it = 0; it < p_.n_solver_iters;) {
                int cls = tile_class(recon, level, wi);
                if (kIters[cls] > p_.n_solver_iters - it) cls = remainder_class(recon, p_.n_solver_iters - it);
                const TileSet& ts = tiles(level, cls);
                IterArgs a{Launch{dp_, ts.map, ts.ntiles}, level, pr.d, pr.f, pr.x, pr.u, it == 0, p_.lambda, theta,
                           p_.tau, p_.alpha_tv, p_.alpha_proj, p_.alpha_fn, p_.lambdaf, p_.lambdams,
                           (fn && sparse_) ? 1 : 0, JointView{}, 0, 0.0f, 0u,
                           make_udiv(p_.alpha_tv + p_.alpha_proj), make_udiv(p_.lambdaf + p_.alpha_proj),
                           1.0 / ((double)p_.alpha_tv + (double)p_.alpha_proj),
                           1.0 / ((double)p_.lambdaf + (double)p_.alpha_proj)};
                // one pair, or pairs of one size: the fields travel in the kernel arguments (no
                // dependent descriptor load)
                const bool single = (n() == 1 || pair_elems_ > 0) && cls != TC_J64 && cls != TC_FLOW;
                if (single) {
                    a.v = joint_view(hp_[0], dp_, level, pr.d, pr.f, pr.x, pr.u, kTileW[cls], it == 0);
                    if (n() > 1) {
                        a.tpp = ts.ntiles / n();
                        a.inv_tpp = 1.0f / (float)a.tpp;
                        a.pair_elems = pair_elems_;
                    }
                }
                if (ts.ntiles > 0) {
                    if (!recon)
                        launch_flow<TI, TS>(cls, single, fused_deriv && it == 0, st_, a);
                    else
                        launch_joint<TI, TS>(cls, single ? (n() == 1 ? 1 : 2) : 0, p_.tri_f64 != 0, fn,
                                             fused_deriv && it == 0, p_.storage == RF_STORAGE_F16_DERIV_DUALS, st_,
                                             a);
                    RF_TRY(hipGetLastError());
                }
                pr.d ^= 1;  // one flip per launch, whatever its iteration count
                if (recon) {
                    pr.f ^= 1;
                    pr.x ^= 1;
                }
                it += kIters[cls];
            }
            if (tuning_) RF_TRY(hipEventRecord(tev_[2 * (level * p_.n_warp_iters + wi) + 1], st_));
            prof_.end(st_, RF_KID_FUSED_ITER, level, p_.n_solver_iters);
            warped = wi + 1 < p_.n_warp_iters;
            if (recon) {
                RF_LAUNCH(kb_med5_add, tiles(level, TC_PIX), level, pr.d, pr.u);
                if (warped) RF_LAUNCH(kb_median5_warp<TI>, tiles(level, TC_PIX), level, pr.u);
                else RF_LAUNCH(kb_median5, tiles(level, TC_PIX), level, pr.u);
                pr.u ^= 1;
            } else if (warped) {
                RF_LAUNCH(kb_med5_add_warp<TI>, tiles(level, TC_PIX), level, pr.d, pr.u);
            } else {
                RF_LAUNCH(kb_med5_add, tiles(level, TC_PIX), level, pr.d, pr.u);
            }
        }
        if (level > 0) {
            RF_LAUNCH(kb_upscale, tiles(level - 1, TC_PIX), level - 1, pr.u, pr.f, recon ? 1 : 0, p_.scale);
            pr.u ^= 1;
            if (recon) pr.f ^= 1;
        }
    }
    if (p_.with_visualization) RF_LAUNCH(kb_hsv, tiles(0, TC_PIX), pr.u, flow_scale_);
    final_ = pr;
    return hipSuccess;
}

int Batch::solve(bool recon, float flow_scale) {
    if (!st_) {
        set_error("rf_batch_solve: batch not created");
        return RF_ERR_STATE;
    }
    if (recon) {
        // the joint kernels divide by these two launch-uniform sums through UDiv, which needs them
        // finite and positive (the plain path takes any value)
        const float A = p_.alpha_tv + p_.alpha_proj, P = p_.lambdaf + p_.alpha_proj;
        if (!(A > 0.0f) || !std::isfinite(A) || !(P > 0.0f) || !std::isfinite(P)) {
            set_error("fused joint solve: alpha_tv + alpha_proj and lambdaf + alpha_proj must be finite and "
                      "positive; use the plain path");
            return RF_ERR_UNSUPPORTED;
        }
    }
    if (dirty_ && upload_pairs()) return RF_ERR_HIP;
    const int kind = recon ? 1 : 0;
    if (p_.with_visualization && flow_scale != flow_scale_ && gexec_[kind]) {
        (void)hipGraphExecDestroy(gexec_[kind]);  // recaptured below with the new scale
        gexec_[kind] = nullptr;
    }
    flow_scale_ = flow_scale;
    const bool auto_tile = recon ? p_.joint_tile == RF_JOINT_TILE_AUTO : p_.flow_tile == RF_FLOW_TILE_AUTO;
    if (auto_tile && !tuned_[kind] && p_.n_warp_iters > 0 && p_.n_solver_iters > 0 && !prof_.armed)
        return tune_tiling(recon);
    if (!p_.use_graph || prof_.armed) {
        if (check(enqueue(recon), "batch solve")) return RF_ERR_HIP;
        return RF_OK;
    }
    if (!gexec_[kind]) {
        hipGraph_t graph = nullptr;
        if (check(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
            return RF_ERR_HIP;
        hipError_t e = enqueue(recon);
        hipError_t e2 = hipStreamEndCapture(st_, &graph);
        if (check(e, "batch solve (capture)") || check(e2, "hipStreamEndCapture")) {
            if (graph) (void)hipGraphDestroy(graph);
            return RF_ERR_HIP;
        }
        hipError_t e3 = hipGraphInstantiate(&gexec_[kind], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (check(e3, "hipGraphInstantiate")) {
            gexec_[kind] = nullptr;
            return RF_ERR_HIP;
        }
        gfinal_[kind] = final_;
    }
    final_ = gfinal_[kind];
    return check(hipGraphLaunch(gexec_[kind], st_), "hipGraphLaunch");
}

int Batch::sync() { return check(hipStreamSynchronize(st_), "hipStreamSynchronize"); }

void Batch::result(int i, const float** u, const float** v, const float** X, const float** Y,
                   const float** Z) const {
    const PairFields& F = hp_[i];
    if (u) *u = F.u[final_.u];
    if (v) *v = F.v[final_.u];
    if (X) *X = F.X[final_.x];
    if (Y) *Y = F.Y[final_.x];
    if (Z) *Z = F.Z[final_.x];
}

// removeEdgeArtifacts3D on the result (Reconflow.cu:468-477): filtered into the partner X
// buffers, which then become the result.
int Batch::remove_edge(int i, int range, float maxarea) {
    const PairFields& F = hp_[i];
    const int x = final_.x;
    hipError_t e = launch_remove_edge(F.X[x], F.Y[x], F.Z[x], range, maxarea, geo_[i].W[0], geo_[i].H[0], geo_[i].S[0],
                                      F.X[x ^ 1], F.Y[x ^ 1], F.Z[x ^ 1], st_);
    if (check(e, "remove edge artifacts")) return RF_ERR_HIP;
    final_.x ^= 1;
    return RF_OK;
}

namespace {
// Middlebury colour wheel (flowToHSV.cu:194-302), batched form of the per-step kernel.
__constant__ float cb_wheel[3][55] = {
    {255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 213, 170, 128,
     85,  43,  0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   19,
     39,  58,  78,  98,  117, 137, 156, 176, 196, 215, 235, 255, 255, 255, 255, 255, 255},
    {0,   17,  34,  51,  68,  85,  102, 119, 136, 153, 170, 187, 204, 221, 238, 255, 255, 255, 255,
     255, 255, 255, 255, 255, 255, 255, 232, 209, 186, 163, 140, 116, 93,  70,  47,  24,  0,   0,
     0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0},
    {0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,
     0,   0,   0,   63,  127, 191, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255,
     255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 213, 170, 128, 85,  43}};
__global__ void __launch_bounds__(256) kb_hsv(Launch L, int pu, float flowscale) {
    BT_PIX(L, 0)
    float du = F.u[pu][pos] / flowscale, dv = F.v[pu][pos] / flowscale;
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
    for (int c = 0; c < 3; ++c) {
        float cc = (1.0f - f) * (cb_wheel[c][k0] / 255.0f) + f * (cb_wheel[c][k1] / 255.0f);
        col[c] = rad <= 1.0f ? 1.0f - rad * (1.0f - cc) : cc * 0.75f;
    }
    F.uvrgb[3 * pos + 0] = col[2];
    F.uvrgb[3 * pos + 1] = col[1];
    F.uvrgb[3 * pos + 2] = col[0];
}
}  // namespace

}  // namespace rf

// ===================================================================================
// C ABI
struct rf_batch {
    rf::Batch b;
};

extern "C" {

int rf_batch_create(rf_batch** out, const rf_params* params, int n, const int* widths, const int* heights) {
    if (!out || !params) {
        rf::set_error("rf_batch_create: null argument");
        return RF_ERR_ARG;
    }
    *out = nullptr;
    auto* b = new rf_batch;
    int rc = b->b.create(*params, n, widths, heights);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return RF_OK;
}

int rf_batch_destroy(rf_batch* b) {
    delete b;
    return RF_OK;
}

int rf_batch_stride(const rf_batch* b, int i) {
    if (!b || i < 0 || i >= b->b.n()) return RF_ERR_ARG;
    return b->b.stride(i);
}

int rf_batch_set_images_device(rf_batch* b, int i, const float* d_i0, const float* d_i1) {
    if (!b) return RF_ERR_ARG;
    return b->b.set_images_device(i, d_i0, d_i1);
}

int rf_batch_set_camera(rf_batch* b, int i, const double K0[9], const double K1[9], const double R1[9],
                        const double t1[3]) {
    if (!b) return RF_ERR_ARG;
    return b->b.set_camera(i, K0, K1, nullptr, nullptr, R1, t1);
}

int rf_batch_solve(rf_batch* b) {
    if (!b) return RF_ERR_ARG;
    const int m = b->b.params().method;
    return b->b.solve(m >= RF_METHODR_TVL1_MS, 5.0f);
}

int rf_batch_synchronize(rf_batch* b) {
    if (!b) return RF_ERR_ARG;
    return b->b.sync();
}

int rf_batch_get_result_device(const rf_batch* b, int i, const float** u, const float** v, const float** X,
                               const float** Y, const float** Z) {
    if (!b || i < 0 || i >= b->b.n()) return RF_ERR_ARG;
    b->b.result(i, u, v, X, Y, Z);
    return RF_OK;
}

void* rf_batch_get_stream(const rf_batch* b) { return b ? (void*)b->b.stream() : nullptr; }

int rf_batch_set_warm_start_device(rf_batch* b, int i, const float* d_u, const float* d_v) {
    if (!b) return RF_ERR_ARG;
    return b->b.set_warm_start(i, d_u, d_v);
}

int rf_batch_warm_start_in_use(rf_batch* b, int i) {
    if (!b) return RF_ERR_ARG;
    return b->b.warm_start_in_use(i);
}

int rf_batch_tiling(const rf_batch* b, int* tiling, int n) {
    if (!b || !tiling || n < 0) return RF_ERR_ARG;
    const bool recon = b->b.params().method >= RF_METHODR_TVL1_MS;
    for (int l = 0; l < n; ++l) tiling[l] = recon ? b->b.joint_tiling(l) : b->b.flow_tiling(l);
    return RF_OK;
}

}  // extern "C"
