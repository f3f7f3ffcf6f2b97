// Learned warm start: a small convolutional encoder-decoder that predicts the flow at the
// level-3 resolution of the joint solver's pyramid (1/8 of the image) from the two coarsest
// image levels of both frames, trained on the solver's own results. Its prediction, scaled up to
// level 0, is the solver's warm start (rf_copy_warm_start / rf_batch_set_warm_start_device), so a
// solve can start at a finer level and skip the latency-bound coarse levels.
//
// Network (H x W = the level-3 size, both even; activations NHWC, channels padded to multiples
// of 8; BF16 weights and activations, FP32 accumulation and biases):
//   x0 [H, W, 8]   : I0_3, I1_3, I0_4 and I1_4 (level 4 by pixel replication), 4 zero channels
//   a1 = relu(conv3x3(x0) + b1)            [H, W, 32]
//   a2 = relu(conv3x3 stride 2(a1) + b2)   [H/2, W/2, 32]
//   a3 = relu(conv3x3(a2) + b3)            [H/2, W/2, 32]
//   a4 = relu(conv3x3(a3) + b4)            [H/2, W/2, 32]
//   u  = up2(a4) + a1                      [H, W, 32]   (pixel replication, skip connection)
//   a5 = relu(conv3x3(u) + b5)             [H, W, 32]
//   f  = conv3x3(a5) + b6                  [H, W, 2]    FP32: the flow in level-3 pixels
// Convolutions pad with zeros. Every convolution is an implicit GEMM on the matrix cores: a wave
// computes 32 output pixels x 32 output channels with v_mfma_f32_32x32x8_bf16, the K dimension
// running over the 9 taps x input channels (forward), 9 taps x output channels (gradient of the
// input) or output pixels (gradient of the weights). Training: Huber loss (delta 1 px) against the
// target flow, Adam on FP32 master weights, which are cast to the BF16 weights after each step.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rfhip.h"
#include "rf_kernels.hpp"
#include "rf_solver.hpp"

namespace rf {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int kLayers = 6;
constexpr int kCoPad = 32;  // output channels of every weight tensor, padded
// real input / output channels and strides of the six convolutions
constexpr int kCin[kLayers] = {4, 32, 32, 32, 32, 32};
constexpr int kCinP[kLayers] = {8, 32, 32, 32, 32, 32};
constexpr int kCout[kLayers] = {32, 32, 32, 32, 32, 2};
constexpr int kStride[kLayers] = {1, 2, 1, 1, 1, 1};

// BF16 bits of x, rounded to nearest even (x finite).
__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t b = __float_as_uint(x);
    b += 0x7fffu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

struct ConvGeo {
    int N, Hi, Wi, Ci;  // input (Ci padded to a multiple of 8)
    int Ho, Wo;         // output
    int stride;
};

// Forward: y[p][j] = act(sum_t sum_c x[in(p, t)][c] w[t][j][c] + b[j]); one wave per 32 output
// pixels. A (32 pixels x 8 channels of one tap): lane l holds pixel l % 32, channels
// 4 (l / 32) .. +3; B (8 channels x 32 outputs): lane l holds output l % 32, the same channels;
// D: lane l holds output l % 32 of pixels 8 (r / 4) + 4 (l / 32) + r % 4, r = 0..15.
// y (bf16, Cy channels) and / or yf (float, Cf channels) may be null.
__global__ void __launch_bounds__(256) k_conv_fwd(const uint16_t* __restrict__ x, const uint16_t* __restrict__ w,
                                                  const float* __restrict__ bias, ConvGeo g, int relu,
                                                  uint16_t* __restrict__ y, int Cy, float* __restrict__ yf, int Cf) {
    const int l = threadIdx.x & 63;
    const long P = (long)g.N * g.Ho * g.Wo;
    const long p0 = (long)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 32;
    if (p0 >= P) return;  // the whole wave
    const long p = p0 + (l & 31);
    const bool pv = p < P;
    int n = 0, oy = 0, ox = 0;
    if (pv) {
        n = (int)(p / ((long)g.Ho * g.Wo));
        const int r = (int)(p - (long)n * g.Ho * g.Wo);
        oy = r / g.Wo;
        ox = r - oy * g.Wo;
    }
    const int kq = (l >> 5) * 4;
    f16v acc = {};
    for (int t = 0; t < 9; ++t) {
        const int iy = oy * g.stride + t / 3 - 1, ix = ox * g.stride + t % 3 - 1;
        const bool in = pv && iy >= 0 && iy < g.Hi && ix >= 0 && ix < g.Wi;
        const long xo = in ? (((long)n * g.Hi + iy) * g.Wi + ix) * g.Ci : 0;
        const uint16_t* wp = w + ((long)t * kCoPad + (l & 31)) * g.Ci;
        for (int c0 = 0; c0 < g.Ci; c0 += 8) {
            const s4 a = in ? *reinterpret_cast<const s4*>(x + xo + c0 + kq) : s4{0, 0, 0, 0};
            const s4 b = *reinterpret_cast<const s4*>(wp + c0 + kq);
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
    }
    const int j = l & 31;
    const float bj = bias[j];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const long q = p0 + (r >> 2) * 8 + (l >> 5) * 4 + (r & 3);
        if (q >= P) continue;
        float v = acc[r] + bj;
        if (relu) v = fmaxf(v, 0.0f);
        if (y && j < Cy) y[q * Cy + j] = f2bf(v);
        if (yf && j < Cf) yf[q * Cf + j] = v;
    }
}

// Gradient of the input: dx[p][j] = sum_t sum_c dy[o(p, t)][c] wt[t][j][c], o(p, t) the output
// pixel whose tap t reads input pixel p. dy bf16 [P_out][32], wt[t][ci][co] (ci padded to 32),
// dx float [P_in][Cx] (Cx <= 32 real channels). One wave per 32 input pixels.
__global__ void __launch_bounds__(256) k_conv_bwd_x(const uint16_t* __restrict__ dy, const uint16_t* __restrict__ wt,
                                                    ConvGeo g, float* __restrict__ dx, int Cx) {
    const int l = threadIdx.x & 63;
    const long P = (long)g.N * g.Hi * g.Wi;
    const long p0 = (long)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 32;
    if (p0 >= P) return;
    const long p = p0 + (l & 31);
    const bool pv = p < P;
    int n = 0, iy = 0, ix = 0;
    if (pv) {
        n = (int)(p / ((long)g.Hi * g.Wi));
        const int r = (int)(p - (long)n * g.Hi * g.Wi);
        iy = r / g.Wi;
        ix = r - iy * g.Wi;
    }
    const int kq = (l >> 5) * 4;
    f16v acc = {};
    for (int t = 0; t < 9; ++t) {
        // forward: iy = oy * stride + t / 3 - 1
        const int ny = iy - t / 3 + 1, nx = ix - t % 3 + 1;
        const int oy = ny / g.stride, ox = nx / g.stride;
        const bool in = pv && ny >= 0 && nx >= 0 && oy * g.stride == ny && ox * g.stride == nx && oy < g.Ho &&
                        ox < g.Wo;
        const long yo = in ? (((long)n * g.Ho + oy) * g.Wo + ox) * kCoPad : 0;
        const uint16_t* wp = wt + ((long)t * kCoPad + (l & 31)) * kCoPad;
        for (int c0 = 0; c0 < kCoPad; c0 += 8) {
            const s4 a = in ? *reinterpret_cast<const s4*>(dy + yo + c0 + kq) : s4{0, 0, 0, 0};
            const s4 b = *reinterpret_cast<const s4*>(wp + c0 + kq);
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
    }
    const int j = l & 31;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const long q = p0 + (r >> 2) * 8 + (l >> 5) * 4 + (r & 3);
        if (q < P && j < Cx) dx[q * Cx + j] = acc[r];
    }
}

// Gradient of the weights for tap t = blockIdx.y: dw[t][co][ci] += sum_o dy[o][co] x[in(o, t)][ci]
// over the chunk of CH output pixels of this wave (A: 32 output channels x 8 pixels, B: 8 pixels x
// 32 input channels), added into the FP32 gradient with atomics. dy bf16 [P_out][32], x bf16
// [P_in][Ci].
constexpr int kWChunk = 512;
__global__ void __launch_bounds__(256) k_conv_bwd_w(const uint16_t* __restrict__ dy, const uint16_t* __restrict__ x,
                                                    ConvGeo g, int Co, int Cr, float* __restrict__ dw) {
    const int l = threadIdx.x & 63;
    const int t = blockIdx.y;
    const long P = (long)g.N * g.Ho * g.Wo;
    const long o0 = (long)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * kWChunk;
    if (o0 >= P) return;
    const int kq = (l >> 5) * 4, i = l & 31;
    f16v acc = {};
    for (int k0 = 0; k0 < kWChunk; k0 += 8) {
        s4 a, b;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const long o = o0 + k0 + kq + r;
            uint16_t av = 0, bv = 0;
            if (o < P) {
                av = dy[o * kCoPad + i];
                const int n = (int)(o / ((long)g.Ho * g.Wo));
                const int rr = (int)(o - (long)n * g.Ho * g.Wo);
                const int oy = rr / g.Wo, ox = rr - oy * g.Wo;
                const int iy = oy * g.stride + t / 3 - 1, ix = ox * g.stride + t % 3 - 1;
                if (i < g.Ci && iy >= 0 && iy < g.Hi && ix >= 0 && ix < g.Wi)
                    bv = x[(((long)n * g.Hi + iy) * g.Wi + ix) * g.Ci + i];
            }
            a[r] = (short)av;
            b[r] = (short)bv;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int co = (r >> 2) * 8 + (l >> 5) * 4 + (r & 3), ci = l & 31;
        if (co < Co && ci < Cr && acc[r] != 0.0f) atomicAdd(&dw[((long)t * Co + co) * Cr + ci], acc[r]);
    }
}

// ReLU backward and bias gradient: g = dout * (y > 0) (no mask when y is null), written as the BF16
// gradient of the pre-activation (32 channels, padding zero); db[c] += sum of g (one atomic per
// channel and workgroup).
__global__ void __launch_bounds__(256) k_act_bwd(const float* __restrict__ dout, int Cd, const uint16_t* __restrict__ y,
                                                 int Cy, long P, uint16_t* __restrict__ dyb, float* __restrict__ db) {
    __shared__ float s_db[kCoPad];
    if (threadIdx.x < kCoPad) s_db[threadIdx.x] = 0.0f;
    __syncthreads();
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;  // one (pixel, channel) per thread
    const long p = e / kCoPad;
    const int c = (int)(e % kCoPad);
    if (p < P) {
        float gv = 0.0f;
        if (c < Cd) {
            gv = dout[p * Cd + c];
            if (y && !(bf2f(y[p * Cy + c]) > 0.0f)) gv = 0.0f;
        }
        dyb[p * kCoPad + c] = f2bf(gv);
        if (gv != 0.0f) atomicAdd(&s_db[c], gv);
    }
    __syncthreads();
    if (threadIdx.x < kCoPad && s_db[threadIdx.x] != 0.0f) atomicAdd(&db[threadIdx.x], s_db[threadIdx.x]);
}

// u = up2(a4) + a1 (pixel replication), BF16, 32 channels.
__global__ void __launch_bounds__(256) k_up2_add(const uint16_t* __restrict__ a4, const uint16_t* __restrict__ a1, int N,
                                                 int H, int W, uint16_t* __restrict__ u) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)N * H * W * kCoPad;
    if (e >= total) return;
    const int c = (int)(e % kCoPad);
    const long p = e / kCoPad;
    const int n = (int)(p / ((long)H * W));
    const int r = (int)(p - (long)n * H * W);
    const int yy = r / W, xx = r - yy * W;
    const long q = (((long)n * (H / 2) + yy / 2) * (W / 2) + xx / 2) * kCoPad + c;
    u[e] = f2bf(bf2f(a4[q]) + bf2f(a1[e]));
}

// Backward of u = up2(a4) + a1: da1 += du; da4 = sum of du over each 2x2 block (float, 32 channels).
__global__ void __launch_bounds__(256) k_up2_add_bwd(const float* __restrict__ du, int N, int H, int W,
                                                     float* __restrict__ da1, float* __restrict__ da4) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;  // over the coarse grid x channels
    const int Hc = H / 2, Wc = W / 2;
    const long total = (long)N * Hc * Wc * kCoPad;
    if (e >= total) return;
    const int c = (int)(e % kCoPad);
    const long p = e / kCoPad;
    const int n = (int)(p / ((long)Hc * Wc));
    const int r = (int)(p - (long)n * Hc * Wc);
    const int yy = r / Wc, xx = r - yy * Wc;
    float s = 0.0f;
    for (int dy2 = 0; dy2 < 2; ++dy2)
        for (int dx2 = 0; dx2 < 2; ++dx2) {
            const long f = (((long)n * H + 2 * yy + dy2) * W + 2 * xx + dx2) * kCoPad + c;
            s += du[f];
            da1[f] += du[f];
        }
    da4[e] = s;
}

// Huber loss (delta 1) of the predicted flow against the target, mean over N H W x 2 components:
// loss (atomic sum) and its gradient. Components whose target is not finite do not count.
__global__ void __launch_bounds__(256) k_huber(const float* __restrict__ pred, const float* __restrict__ target, long n,
                                               float inv_count, float* __restrict__ grad, float* __restrict__ loss) {
    __shared__ float s_l[256];
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float lv = 0.0f;
    if (e < n) {
        const float tg = target[e];
        float gv = 0.0f;
        if (isfinite(tg)) {
            const float d = pred[e] - tg, a = fabsf(d);
            lv = a <= 1.0f ? 0.5f * d * d : a - 0.5f;
            gv = fminf(fmaxf(d, -1.0f), 1.0f) * inv_count;
        }
        grad[e] = gv;
    }
    s_l[threadIdx.x] = lv;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s_l[threadIdx.x] += s_l[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0 && s_l[0] != 0.0f) atomicAdd(loss, s_l[0] * inv_count);
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the FP32 master parameters; c1 = 1 / (1 - beta1^t),
// c2 = 1 / (1 - beta2^t).
__global__ void __launch_bounds__(256) k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, long n, float lr, float c1, float c2) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gv = g[e];
    const float mv = 0.9f * m[e] + 0.1f * gv;
    const float vv = 0.999f * v[e] + 0.001f * gv * gv;
    m[e] = mv;
    v[e] = vv;
    w[e] -= lr * (mv * c1) / (sqrtf(vv * c2) + 1e-8f);
}

// The FP32 master weights of one layer [9][Co][Ci] to the BF16 forward layout w[t][32][CiP] and the
// transposed layout wt[t][32 (ci)][32 (co)] of the input gradient, zero padded; bias to 32 floats.
__global__ void __launch_bounds__(256) k_cast_layer(const float* __restrict__ wm, const float* __restrict__ bm, int Co,
                                                    int Ci, int CiP, uint16_t* __restrict__ w, uint16_t* __restrict__ wt,
                                                    float* __restrict__ b) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;  // over 9 x 32 x 32
    if (e >= 9 * kCoPad * kCoPad) return;
    const int t = e / (kCoPad * kCoPad), co = (e / kCoPad) % kCoPad, ci = e % kCoPad;
    const float v = (co < Co && ci < Ci) ? wm[((long)t * Co + co) * Ci + ci] : 0.0f;
    if (ci < CiP) w[((long)t * kCoPad + co) * CiP + ci] = f2bf(v);
    wt[((long)t * kCoPad + ci) * kCoPad + co] = f2bf(v);
    if (t == 0 && ci == 0) b[co] = co < Co ? bm[co] : 0.0f;
}

// The network input from the two coarsest levels: x0[n][y][x] = (I0_3, I1_3, I0_4, I1_4, 0, 0, 0, 0)
// with level 4 read at (y / 2, x / 2), each image (h x w, stride s) minus 0.5.
__global__ void __launch_bounds__(256) k_assemble(const float* __restrict__ i03, const float* __restrict__ i13,
                                                  int s3, const float* __restrict__ i04, const float* __restrict__ i14,
                                                  int s4, int H, int W, long nstride3, long nstride4, int N,
                                                  uint16_t* __restrict__ x0) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)N * H * W) return;
    const int n = (int)(e / ((long)H * W));
    const int r = (int)(e - (long)n * H * W);
    const int yy = r / W, xx = r - yy * W;
    const long q3 = n * nstride3 + (long)yy * s3 + xx, q4 = n * nstride4 + (long)(yy / 2) * s4 + xx / 2;
    uint16_t* o = x0 + e * 8;
    o[0] = f2bf(i03[q3] - 0.5f);
    o[1] = f2bf(i13[q3] - 0.5f);
    o[2] = f2bf(i04[q4] - 0.5f);
    o[3] = f2bf(i14[q4] - 0.5f);
    o[4] = o[5] = o[6] = o[7] = 0;
}

// Deinterleave the predicted flow (H x W x 2) into two planes of stride s.
__global__ void __launch_bounds__(256) k_split_flow(const float* __restrict__ f, int H, int W, int s, float* __restrict__ u,
                                                    float* __restrict__ v) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)H * W) return;
    const int yy = (int)(e / W), xx = (int)(e - (long)yy * W);
    u[(long)yy * s + xx] = f[2 * e];
    v[(long)yy * s + xx] = f[2 * e + 1];
}

unsigned blocks(long n, int per = 256) { return (unsigned)((n + per - 1) / per); }

}  // namespace

// ---------------------------------------------------------------------------------------------
struct WarmNet {
    int max_n = 0, H = 0, W = 0;
    long nparams = 0;
    long woff[kLayers] = {}, boff[kLayers] = {};
    void* mem = nullptr;  // one device allocation for everything below
    float *master = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    uint16_t *w[kLayers] = {}, *wt[kLayers] = {};
    float* b[kLayers] = {};
    uint16_t *x0 = nullptr, *a[6] = {}, *u = nullptr, *dyb = nullptr;  // a[1..5]
    float *pred = nullptr, *dpred = nullptr, *g0 = nullptr, *g1 = nullptr, *gskip = nullptr, *loss = nullptr;
    float* pyr = nullptr;  // level images of the warm-start path: 2 frames x levels 1..4
    int steps = 0;

    // layers 3 and 4 run at half the size; layer 2 halves it (stride 2)
    ConvGeo geo(int layer, int n) const {
        const int Hi = (layer == 3 || layer == 4) ? H / 2 : H;
        const int Wi = (layer == 3 || layer == 4) ? W / 2 : W;
        ConvGeo g{};
        g.N = n;
        g.Ci = kCinP[layer - 1];
        g.stride = kStride[layer - 1];
        g.Hi = Hi;
        g.Wi = Wi;
        g.Ho = layer == 2 ? H / 2 : Hi;
        g.Wo = layer == 2 ? W / 2 : Wi;
        return g;
    }
};

namespace {

hipError_t cast_weights(WarmNet& m, hipStream_t st) {
    for (int L = 0; L < kLayers; ++L) {
        hipLaunchKernelGGL(k_cast_layer, dim3(blocks(9 * kCoPad * kCoPad)), dim3(256), 0, st, m.master + m.woff[L],
                           m.master + m.boff[L], kCout[L], kCin[L], kCinP[L], m.w[L], m.wt[L], m.b[L]);
    }
    return hipGetLastError();
}

hipError_t forward(WarmNet& m, int n, hipStream_t st) {
    const long P1 = (long)n * m.H * m.W;
    auto conv = [&](int layer, const uint16_t* in, uint16_t* out, float* outf, int relu) {
        const ConvGeo g = m.geo(layer, n);
        const long P = (long)g.N * g.Ho * g.Wo;
        hipLaunchKernelGGL(k_conv_fwd, dim3(blocks((P + 31) / 32 * 64)), dim3(256), 0, st, in, m.w[layer - 1],
                           m.b[layer - 1], g, relu, out, out ? kCoPad : 0, outf, outf ? kCout[layer - 1] : 0);
    };
    conv(1, m.x0, m.a[1], nullptr, 1);
    conv(2, m.a[1], m.a[2], nullptr, 1);
    conv(3, m.a[2], m.a[3], nullptr, 1);
    conv(4, m.a[3], m.a[4], nullptr, 1);
    hipLaunchKernelGGL(k_up2_add, dim3(blocks(P1 * kCoPad)), dim3(256), 0, st, m.a[4], m.a[1], n, m.H, m.W, m.u);
    conv(5, m.u, m.a[5], nullptr, 1);
    conv(6, m.a[5], nullptr, m.pred, 0);
    return hipGetLastError();
}

// Gradients of every parameter into m.grad (zeroed first) from dpred (the loss gradient).
hipError_t backward(WarmNet& m, int n, hipStream_t st) {
    hipError_t e0 = hipMemsetAsync(m.grad, 0, m.nparams * sizeof(float), st);
    if (e0 != hipSuccess) return e0;
    const long P1 = (long)n * m.H * m.W, P2 = (long)n * (m.H / 2) * (m.W / 2);
    // one layer: activation backward from dout (Cd channels, masked by y unless null), weight and
    // bias gradients from input x, and the input gradient into dx (Cx channels) unless null
    auto layer_bwd = [&](int layer, const float* dout, const uint16_t* y, const uint16_t* x, float* dx) {
        const ConvGeo g = m.geo(layer, n);
        const long Po = (long)g.N * g.Ho * g.Wo, Pi = (long)g.N * g.Hi * g.Wi;
        const int Co = kCout[layer - 1];
        hipLaunchKernelGGL(k_act_bwd, dim3(blocks(Po * kCoPad)), dim3(256), 0, st, dout, Co, y, kCoPad, Po, m.dyb,
                           m.grad + m.boff[layer - 1]);
        hipLaunchKernelGGL(k_conv_bwd_w, dim3(blocks((Po + kWChunk - 1) / kWChunk * 64), 9), dim3(256), 0, st, m.dyb, x,
                           g, Co, kCin[layer - 1], m.grad + m.woff[layer - 1]);
        if (dx)
            hipLaunchKernelGGL(k_conv_bwd_x, dim3(blocks((Pi + 31) / 32 * 64)), dim3(256), 0, st, m.dyb,
                               m.wt[layer - 1], g, dx, kCoPad);
    };
    // g0, g1: ping-pong float gradients of 32-channel activations (P1 x 32 each)
    layer_bwd(6, m.dpred, nullptr, m.a[5], m.g0);               // d a5 -> g0
    layer_bwd(5, m.g0, m.a[5], m.u, m.g1);                      // d u -> g1
    hipLaunchKernelGGL(k_up2_add_bwd, dim3(blocks(P2 * kCoPad)), dim3(256), 0, st, m.g1, n, m.H, m.W, m.gskip, m.g0);
    // gskip holds d a1 from the skip (zeroed below before the accumulation); g0 = d a4
    layer_bwd(4, m.g0, m.a[4], m.a[3], m.g1);                   // d a3 -> g1
    layer_bwd(3, m.g1, m.a[3], m.a[2], m.g0);                   // d a2 -> g0
    layer_bwd(2, m.g0, m.a[2], m.a[1], m.g1);                   // d a1 (through the encoder) -> g1
    e0 = launch_add(m.g1, m.gskip, P1 * kCoPad, m.g1, st);
    if (e0 != hipSuccess) return e0;
    layer_bwd(1, m.g1, m.a[1], m.x0, nullptr);
    return hipGetLastError();
}

}  // namespace
}  // namespace rf

using rf::WarmNet;

struct rf_warmnet {
    WarmNet m;
};

extern "C" {

int rf_warmnet_create(rf_warmnet** out, int max_batch, int height, int width) {
    if (!out || max_batch < 1 || height < 2 || width < 2 || (height & 1) || (width & 1) || height > 4096 ||
        width > 4096 || (long)max_batch * height * width > (1L << 26)) {
        rf::set_error("rf_warmnet_create: invalid batch or level-3 size (positive, even, at most 4096)");
        return RF_ERR_ARG;
    }
    auto* h = new rf_warmnet();
    WarmNet& m = h->m;
    m.max_n = max_batch;
    m.H = height;
    m.W = width;
    long np = 0;
    for (int L = 0; L < rf::kLayers; ++L) {
        m.woff[L] = np;
        np += 9L * rf::kCout[L] * rf::kCin[L];
        m.boff[L] = np;
        np += rf::kCout[L];
    }
    m.nparams = np;
    const long P1 = (long)max_batch * height * width, P2 = P1 / 4;
    // byte layout, each piece 256-byte aligned
    std::vector<std::pair<void**, size_t>> pieces;
    auto add = [&](void** p, size_t bytes) { pieces.push_back({p, (bytes + 255) / 256 * 256}); };
    add((void**)&m.master, np * 4);
    add((void**)&m.grad, np * 4);
    add((void**)&m.adam_m, np * 4);
    add((void**)&m.adam_v, np * 4);
    for (int L = 0; L < rf::kLayers; ++L) {
        add((void**)&m.w[L], 9L * rf::kCoPad * rf::kCinP[L] * 2);
        add((void**)&m.wt[L], 9L * rf::kCoPad * rf::kCoPad * 2);
        add((void**)&m.b[L], rf::kCoPad * 4);
    }
    add((void**)&m.x0, P1 * 8 * 2);
    add((void**)&m.a[1], P1 * rf::kCoPad * 2);
    add((void**)&m.a[2], P2 * rf::kCoPad * 2);
    add((void**)&m.a[3], P2 * rf::kCoPad * 2);
    add((void**)&m.a[4], P2 * rf::kCoPad * 2);
    add((void**)&m.u, P1 * rf::kCoPad * 2);
    add((void**)&m.a[5], P1 * rf::kCoPad * 2);
    add((void**)&m.dyb, P1 * rf::kCoPad * 2);
    add((void**)&m.pred, P1 * 2 * 4);
    add((void**)&m.dpred, P1 * 2 * 4);
    add((void**)&m.g0, P1 * rf::kCoPad * 4);
    add((void**)&m.g1, P1 * rf::kCoPad * 4);
    add((void**)&m.gskip, P1 * rf::kCoPad * 4);
    add((void**)&m.loss, 256);
    size_t total = 0;
    for (auto& pc : pieces) total += pc.second;
    if (hipMalloc(&m.mem, total) != hipSuccess) {
        delete h;
        rf::set_error("rf_warmnet_create: device allocation of " + std::to_string(total) + " bytes failed");
        return RF_ERR_ALLOC;
    }
    if (hipMemset(m.mem, 0, total) != hipSuccess) {
        (void)hipFree(m.mem);
        delete h;
        rf::set_error("rf_warmnet_create: hipMemset failed");
        return RF_ERR_HIP;
    }
    char* cur = (char*)m.mem;
    for (auto& pc : pieces) {
        *pc.first = cur;
        cur += pc.second;
    }
    *out = h;
    return RF_OK;
}

int rf_warmnet_destroy(rf_warmnet* h) {
    if (!h) return RF_ERR_ARG;
    if (h->m.mem) (void)hipFree(h->m.mem);
    if (h->m.pyr) (void)hipFree(h->m.pyr);
    delete h;
    return RF_OK;
}

int rf_warmnet_param_count(const rf_warmnet* h) { return h ? (int)h->m.nparams : RF_ERR_ARG; }

int rf_warmnet_set_params(rf_warmnet* h, const float* params, int n) {
    if (!h || !params || n != (int)h->m.nparams) {
        rf::set_error("rf_warmnet_set_params: null or wrong parameter count");
        return RF_ERR_ARG;
    }
    WarmNet& m = h->m;
    if (hipMemcpy(m.master, params, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m.adam_m, 0, (size_t)n * 4) != hipSuccess || hipMemset(m.adam_v, 0, (size_t)n * 4) != hipSuccess ||
        rf::cast_weights(m, nullptr) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        rf::set_error("rf_warmnet_set_params: HIP error");
        return RF_ERR_HIP;
    }
    m.steps = 0;
    return RF_OK;
}

int rf_warmnet_get_params(rf_warmnet* h, float* params, int n) {
    if (!h || !params || n != (int)h->m.nparams) {
        rf::set_error("rf_warmnet_get_params: null or wrong parameter count");
        return RF_ERR_ARG;
    }
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(params, h->m.master, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        rf::set_error("rf_warmnet_get_params: HIP error");
        return RF_ERR_HIP;
    }
    return RF_OK;
}

int rf_warmnet_get_grads(rf_warmnet* h, float* grads, int n) {
    if (!h || !grads || n != (int)h->m.nparams) {
        rf::set_error("rf_warmnet_get_grads: null or wrong parameter count");
        return RF_ERR_ARG;
    }
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(grads, h->m.grad, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        rf::set_error("rf_warmnet_get_grads: HIP error");
        return RF_ERR_HIP;
    }
    return RF_OK;
}

int rf_warmnet_input(rf_warmnet* h, const float* d_i03, const float* d_i13, int s3, long nstride3, const float* d_i04,
                     const float* d_i14, int s4, long nstride4, int n, void* stream) {
    if (!h || !d_i03 || !d_i13 || !d_i04 || !d_i14 || n < 1 || n > h->m.max_n || s3 < h->m.W || s4 < h->m.W / 2) {
        rf::set_error("rf_warmnet_input: null image, bad batch or stride");
        return RF_ERR_ARG;
    }
    WarmNet& m = h->m;
    hipLaunchKernelGGL(rf::k_assemble, dim3(rf::blocks((long)n * m.H * m.W)), dim3(256), 0, (hipStream_t)stream, d_i03,
                       d_i13, s3, d_i04, d_i14, s4, m.H, m.W, nstride3, nstride4, n, m.x0);
    return hipGetLastError() == hipSuccess ? RF_OK : RF_ERR_HIP;
}

int rf_warmnet_forward(rf_warmnet* h, int n, float* d_flow, void* stream) {
    if (!h || n < 1 || n > h->m.max_n) {
        rf::set_error("rf_warmnet_forward: bad batch");
        return RF_ERR_ARG;
    }
    WarmNet& m = h->m;
    hipStream_t st = (hipStream_t)stream;
    if (rf::forward(m, n, st) != hipSuccess) return RF_ERR_HIP;
    if (d_flow &&
        hipMemcpyAsync(d_flow, m.pred, (size_t)n * m.H * m.W * 2 * 4, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return RF_ERR_HIP;
    return RF_OK;
}

int rf_warmnet_train_step(rf_warmnet* h, int n, const float* d_target, float lr, float* loss, void* stream) {
    if (!h || !d_target || n < 1 || n > h->m.max_n || !(lr >= 0.0f)) {
        rf::set_error("rf_warmnet_train_step: bad batch, target or learning rate");
        return RF_ERR_ARG;
    }
    WarmNet& m = h->m;
    hipStream_t st = (hipStream_t)stream;
    const long ncomp = (long)n * m.H * m.W * 2;
    if (rf::forward(m, n, st) != hipSuccess || hipMemsetAsync(m.loss, 0, 4, st) != hipSuccess) return RF_ERR_HIP;
    hipLaunchKernelGGL(rf::k_huber, dim3(rf::blocks(ncomp)), dim3(256), 0, st, m.pred, d_target, ncomp,
                       1.0f / (float)ncomp, m.dpred, m.loss);
    if (hipMemsetAsync(m.gskip, 0, (size_t)n * m.H * m.W * rf::kCoPad * 4, st) != hipSuccess) return RF_ERR_HIP;
    if (rf::backward(m, n, st) != hipSuccess) return RF_ERR_HIP;
    if (lr > 0.0f) {
        ++m.steps;
        const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)m.steps));
        const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)m.steps));
        hipLaunchKernelGGL(rf::k_adam, dim3(rf::blocks(m.nparams)), dim3(256), 0, st, m.master, m.grad, m.adam_m,
                           m.adam_v, m.nparams, lr, c1, c2);
        if (rf::cast_weights(m, st) != hipSuccess) return RF_ERR_HIP;
    }
    if (loss) {
        if (hipMemcpyAsync(loss, m.loss, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return RF_ERR_HIP;
    }
    return hipGetLastError() == hipSuccess ? RF_OK : RF_ERR_HIP;
}

// The warm start of one pair: the image pyramid of both level-0 frames (w0 x h0, stride s0) to
// level 4 (scale 2, the solver's downscale), the network's level-3 flow, scaled up by the solver's
// upscale (bilinear, x2 per level) to a level-0 flow (stride s0) in d_u, d_v.
int rf_warmnet_warm_start(rf_warmnet* h, const float* d_i0, const float* d_i1, int w0, int h0, int s0, float* d_u,
                          float* d_v, void* stream) {
    if (!h || !d_i0 || !d_i1 || !d_u || !d_v || w0 < 1 || h0 < 1 || s0 < w0) {
        rf::set_error("rf_warmnet_warm_start: null pointer or bad geometry");
        return RF_ERR_ARG;
    }
    WarmNet& m = h->m;
    int W[5], H[5], S[5];
    W[0] = w0, H[0] = h0, S[0] = s0;
    for (int l = 1; l < 5; ++l) {
        W[l] = (int)((float)W[l - 1] / 2.0f);
        H[l] = (int)((float)H[l - 1] / 2.0f);
        S[l] = (W[l] + 31) / 32 * 32;
    }
    if (W[3] != m.W || H[3] != m.H) {
        rf::set_error("rf_warmnet_warm_start: the image's level 3 is " + std::to_string(W[3]) + "x" +
                      std::to_string(H[3]) + ", the network's " + std::to_string(m.W) + "x" + std::to_string(m.H));
        return RF_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    // level images 1..4 of both frames, then the level-3 flow planes and their upscaled copies
    const long field = ((long)S[1] * (H[1] + 1) + 63) / 64 * 64;
    if (!m.pyr && hipMalloc(&m.pyr, (size_t)field * 12 * 4) != hipSuccess) {
        m.pyr = nullptr;
        rf::set_error("rf_warmnet_warm_start: allocation of the pyramid buffers failed");
        return RF_ERR_ALLOC;
    }
    float* lv[2][5];
    for (int f = 0; f < 2; ++f) {
        lv[f][0] = const_cast<float*>(f ? d_i1 : d_i0);
        for (int l = 1; l < 5; ++l) lv[f][l] = m.pyr + field * (f * 4 + l - 1);
        for (int l = 1; l < 5; ++l)
            if (rf::launch_downscale(lv[f][l - 1], W[l - 1], H[l - 1], S[l - 1], W[l], H[l], S[l], false, 1.0f,
                                     lv[f][l], st) != hipSuccess)
                return RF_ERR_HIP;
    }
    int rc = rf_warmnet_input(h, lv[0][3], lv[1][3], S[3], 0, lv[0][4], lv[1][4], S[4], 0, 1, stream);
    if (rc) return rc;
    if (rf::forward(m, 1, st) != hipSuccess) return RF_ERR_HIP;
    float *u3 = m.pyr + field * 8, *v3 = m.pyr + field * 9, *ua = m.pyr + field * 10, *va = m.pyr + field * 11;
    hipLaunchKernelGGL(rf::k_split_flow, dim3(rf::blocks((long)m.H * m.W)), dim3(256), 0, st, m.pred, m.H, m.W, S[3],
                       u3, v3);
    // level 3 -> 2 -> 1 -> 0, each x2 in size and value
    float *cu = u3, *cv = v3;
    for (int l = 3; l >= 1; --l) {
        float* nu = l == 1 ? d_u : (cu == u3 ? ua : u3);
        float* nv = l == 1 ? d_v : (cv == v3 ? va : v3);
        if (rf::launch_upscale(cu, W[l], H[l], S[l], W[l - 1], H[l - 1], S[l - 1], 2.0f, nu, st) != hipSuccess ||
            rf::launch_upscale(cv, W[l], H[l], S[l], W[l - 1], H[l - 1], S[l - 1], 2.0f, nv, st) != hipSuccess)
            return RF_ERR_HIP;
        cu = nu;
        cv = nv;
    }
    return hipGetLastError() == hipSuccess ? RF_OK : RF_ERR_HIP;
}

}  // extern "C"
