// Seeded synthetic image pairs for the TV-L1 and joint flow/depth workloads.
//
// Deterministic on any x86-64 host (own xoshiro256** stream, double arithmetic), so the
// development container and the GPU host produce identical inputs from one seed.
//
//   * flow pairs: frame0 is band-limited noise (white N(0,1), Gaussian blur sigma 1.5 with
//     radius 5 and mirrored borders, rescaled to [0,1]); frame1(y) = frame0(A^-1 (y - b)),
//     i.e. frame0 moved by the affine flow u(x) = (A - I) x + b with max |u| <= 4 px,
//     plus N(0, 0.01^2) noise. The reference convention I0(x) ~ I1(x + u(x)) holds
//     (warpImage.cu:25-28).
//   * recon pairs: pinhole intrinsics f = 500 * W / 640, c = (W/2, H/2); camera-0 at the
//     origin, camera-1 pose X1 = R X0 + t with |rotation| <= 3 deg and |t| = 0.1; the scene
//     is 8 planes over a Voronoi partition of the camera-0 image, depths in [2, 10].
//     frame1 is ray-cast from camera 1 into the planes and samples frame0.
//   * sequences: the same kind of scene with 16 planes at depths [2, 20], seen by a camera moving
//     along a smooth oscillating trajectory (per-frame rotation <= 0.5 deg, translation <= 0.02);
//     every frame is ray-cast into the planes and samples the reference view's texture.
// Images are float, row stride s (elements), padding columns zero.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto& v : s) v = splitmix(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

long refl(long i, long n) {
    long p = 2 * n;
    long m = i % p;
    if (m < 0) m += p;
    return m < n ? m : p - 1 - m;
}

double sample(const std::vector<double>& img, int w, int h, double x, double y) {
    double fx0 = std::floor(x), fy0 = std::floor(y);
    long x0 = (long)fx0, y0 = (long)fy0;
    double ax = x - fx0, ay = y - fy0;
    long xa = refl(x0, w), xb = refl(x0 + 1, w), ya = refl(y0, h), yb = refl(y0 + 1, h);
    return (1 - ax) * (1 - ay) * img[xa + ya * w] + ax * (1 - ay) * img[xb + ya * w] +
           (1 - ax) * ay * img[xa + yb * w] + ax * ay * img[xb + yb * w];
}

// Band-limited noise in [0,1], w*h, dense row-major.
std::vector<double> texture(Rng& rng, int w, int h) {
    std::vector<double> a((size_t)w * h), b((size_t)w * h);
    for (auto& v : a) v = rng.normal();
    const int R = 5;
    const double sigma = 1.5;
    double k[2 * R + 1], ks = 0;
    for (int i = -R; i <= R; ++i) ks += (k[i + R] = std::exp(-0.5 * i * i / (sigma * sigma)));
    for (auto& v : k) v /= ks;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            double acc = 0;
            for (int i = -R; i <= R; ++i) acc += k[i + R] * a[refl(x + i, w) + (long)y * w];
            b[x + (size_t)y * w] = acc;
        }
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            double acc = 0;
            for (int i = -R; i <= R; ++i) acc += k[i + R] * b[x + refl(y + i, h) * w];
            a[x + (size_t)y * w] = acc;
        }
    double lo = *std::min_element(a.begin(), a.end()), hi = *std::max_element(a.begin(), a.end());
    for (auto& v : a) v = (v - lo) / (hi - lo);
    return a;
}

void store(const std::vector<double>& img, int w, int h, int s, float* out) {
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) out[x + (long)y * s] = (float)img[x + (size_t)y * w];
        for (int x = w; x < s; ++x) out[x + (long)y * s] = 0.0f;
    }
}

void rodrigues(const double axis[3], double ang, double R[9]) {
    double c = std::cos(ang), s = std::sin(ang), C = 1 - c;
    double x = axis[0], y = axis[1], z = axis[2];
    double m[9] = {c + x * x * C,     x * y * C - z * s, x * z * C + y * s,
                   y * x * C + z * s, c + y * y * C,     y * z * C - x * s,
                   z * x * C - y * s, z * y * C + x * s, c + z * z * C};
    std::memcpy(R, m, sizeof m);
}

void unit(Rng& rng, double v[3]) {
    double n;
    do {
        for (int i = 0; i < 3; ++i) v[i] = rng.normal();
        n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    } while (n < 1e-9);
    for (int i = 0; i < 3; ++i) v[i] /= n;
}

}  // namespace

extern "C" {

// Piecewise-planar scene in the reference camera's frame: plane k holds the points X with
// n_k . X = 1 (1/Z = n_k . (xn, yn, 1)) and is seen where the reference image's Voronoi cell k is.
struct Scene {
    int np = 0;
    double f = 0, cx = 0, cy = 0;
    std::vector<double> sx, sy;
    std::vector<std::array<double, 3>> n;
    int cell(double x, double y) const {
        int best = 0;
        double bd = 1e300;
        for (int k = 0; k < np; ++k) {
            double d = (x - sx[k]) * (x - sx[k]) + (y - sy[k]) * (y - sy[k]);
            if (d < bd) bd = d, best = k;
        }
        return best;
    }
};

// np planes with inverse depth at the cell seed in [dmin, dmax] and small slopes.
Scene make_scene(Rng& rng, int np, int w, int h, double f, double cx, double cy, double dmin, double dmax) {
    Scene sc;
    sc.np = np, sc.f = f, sc.cx = cx, sc.cy = cy;
    sc.sx.resize(np), sc.sy.resize(np), sc.n.resize(np);
    for (int k = 0; k < np; ++k) {
        sc.sx[k] = rng.uniform(0, w);
        sc.sy[k] = rng.uniform(0, h);
        double d = rng.uniform(dmin, dmax), gx = rng.uniform(-0.02, 0.02), gy = rng.uniform(-0.02, 0.02);
        double xn = (sc.sx[k] - cx) / f, yn = (sc.sy[k] - cy) / f;
        sc.n[k] = {gx, gy, d - gx * xn - gy * yn};
    }
    return sc;
}

// The view of a camera with pose X_cam = R X_ref + t: each pixel's ray hits the nearest plane whose
// cell contains the hit's projection into the reference camera, whose texture f0 it samples,
// plus N(0, 0.01^2) noise drawn from rng in row-major order.
std::vector<double> render(const Scene& sc, const std::vector<double>& f0, int w, int h, const double* R,
                           const double* t, Rng& rng) {
    const double f = sc.f, cx = sc.cx, cy = sc.cy;
    // R^T
    double Rt[9] = {R[0], R[3], R[6], R[1], R[4], R[7], R[2], R[5], R[8]};
    double Rtt[3] = {Rt[0] * t[0] + Rt[1] * t[1] + Rt[2] * t[2], Rt[3] * t[0] + Rt[4] * t[1] + Rt[5] * t[2],
                     Rt[6] * t[0] + Rt[7] * t[1] + Rt[8] * t[2]};
    std::vector<double> f1((size_t)w * h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            double r1[3] = {(x - cx) / f, (y - cy) / f, 1.0};
            double d0[3] = {Rt[0] * r1[0] + Rt[1] * r1[1] + Rt[2] * r1[2],
                            Rt[3] * r1[0] + Rt[4] * r1[1] + Rt[5] * r1[2],
                            Rt[6] * r1[0] + Rt[7] * r1[1] + Rt[8] * r1[2]};
            double bestl = 1e300, bx = x, by = y;
            for (int k = 0; k < sc.np; ++k) {
                const auto& n = sc.n[k];
                double den = n[0] * d0[0] + n[1] * d0[1] + n[2] * d0[2];
                if (den <= 1e-12) continue;
                double lam = (1.0 + n[0] * Rtt[0] + n[1] * Rtt[1] + n[2] * Rtt[2]) / den;
                if (lam <= 0 || lam >= bestl) continue;
                double X0[3] = {lam * d0[0] - Rtt[0], lam * d0[1] - Rtt[1], lam * d0[2] - Rtt[2]};
                if (X0[2] <= 1e-9) continue;
                double px = f * X0[0] / X0[2] + cx, py = f * X0[1] / X0[2] + cy;
                if (sc.cell(px, py) != k) continue;
                bestl = lam;
                bx = px;
                by = py;
            }
            // texel-centre coordinates of frame0
            f1[x + (size_t)y * w] = sample(f0, w, h, bx, by) + 0.01 * rng.normal();
        }
    return f1;
}

// Plain TV-L1 pair. gt_u/gt_v (may be null) receive the generating flow on frame0's grid.
int rfs_make_flow_pair(int w, int h, int s, uint64_t seed, float* i0, float* i1, float* gt_u, float* gt_v) {
    if (w <= 0 || h <= 0 || s < w || !i0 || !i1) return -1;
    Rng rng(seed * 2654435761ull + 17);
    std::vector<double> f0 = texture(rng, w, h);
    double cx = 0.5 * (w - 1), cy = 0.5 * (h - 1);
    double M[4], b[2];
    for (double& v : M) v = rng.uniform(-1, 1);
    for (double& v : b) v = rng.uniform(-1, 1);
    // scale so that max |u| over the corners (the extremes of an affine field) is umax
    double umax = rng.uniform(2.0, 4.0), peak = 0;
    for (int cxn = 0; cxn < 2; ++cxn)
        for (int cyn = 0; cyn < 2; ++cyn) {
            double dx = (cxn ? w - 1 : 0) - cx, dy = (cyn ? h - 1 : 0) - cy;
            double ux = M[0] * dx / w + M[1] * dy / h + b[0], uy = M[2] * dx / w + M[3] * dy / h + b[1];
            peak = std::max(peak, std::sqrt(ux * ux + uy * uy));
        }
    double g = umax / peak;
    // u(x) = G (x - c) + b' with G = g*M/size; y = x + u(x) = A x + const
    double G[4] = {g * M[0] / w, g * M[1] / h, g * M[2] / w, g * M[3] / h};
    double bb[2] = {g * b[0], g * b[1]};
    double A[4] = {1 + G[0], G[1], G[2], 1 + G[3]};
    double det = A[0] * A[3] - A[1] * A[2];
    double Ai[4] = {A[3] / det, -A[1] / det, -A[2] / det, A[0] / det};
    std::vector<double> f1((size_t)w * h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            // y_img = A (x_src - c) + c + b  =>  x_src = Ai (y_img - c - b) + c
            double ry = x - cx - bb[0], ryy = y - cy - bb[1];
            double sx = Ai[0] * ry + Ai[1] * ryy + cx, sy = Ai[2] * ry + Ai[3] * ryy + cy;
            f1[x + (size_t)y * w] = sample(f0, w, h, sx, sy) + 0.01 * rng.normal();
        }
    store(f0, w, h, s, i0);
    store(f1, w, h, s, i1);
    if (gt_u && gt_v)
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < s; ++x) {
                double dx = x - cx, dy = y - cy;
                gt_u[x + (long)y * s] = x < w ? (float)(G[0] * dx + G[1] * dy + bb[0]) : 0.0f;
                gt_v[x + (long)y * s] = x < w ? (float)(G[2] * dx + G[3] * dy + bb[1]) : 0.0f;
            }
    return 0;
}

// Joint flow/depth pair. K (3x3), R (3x3) and t (3) are row-major doubles; the pose maps
// camera-0 points into camera 1 (X1 = R X0 + t). gt_* may be null.
int rfs_make_recon_pair(int w, int h, int s, uint64_t seed, double* K, double* R, double* t, float* i0,
                        float* i1, float* gt_u, float* gt_v, float* gt_Z) {
    if (w <= 0 || h <= 0 || s < w || !i0 || !i1 || !K || !R || !t) return -1;
    Rng rng(seed * 2654435761ull + 29);
    std::vector<double> f0 = texture(rng, w, h);
    const double f = 500.0 * w / 640.0, cx = 0.5 * w, cy = 0.5 * h;
    double Km[9] = {f, 0, cx, 0, f, cy, 0, 0, 1};
    std::memcpy(K, Km, sizeof Km);
    double axis[3];
    unit(rng, axis);
    rodrigues(axis, rng.uniform(0.0, 3.0) * 3.141592653589793 / 180.0, R);
    double tdir[3];
    unit(rng, tdir);
    for (int i = 0; i < 3; ++i) t[i] = 0.1 * tdir[i];

    Scene sc = make_scene(rng, 8, w, h, f, cx, cy, 0.15, 0.45);
    std::vector<double> f1 = render(sc, f0, w, h, R, t, rng);
    store(f0, w, h, s, i0);
    store(f1, w, h, s, i1);
    if (gt_u || gt_v || gt_Z)
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < s; ++x) {
                long p = x + (long)y * s;
                if (x >= w) {
                    if (gt_u) gt_u[p] = 0;
                    if (gt_v) gt_v[p] = 0;
                    if (gt_Z) gt_Z[p] = 0;
                    continue;
                }
                int k = sc.cell(x, y);
                double xn = (x - cx) / f, yn = (y - cy) / f;
                double Z = 1.0 / (sc.n[k][0] * xn + sc.n[k][1] * yn + sc.n[k][2]);
                double X[3] = {Z * xn, Z * yn, Z};
                double X1[3] = {R[0] * X[0] + R[1] * X[1] + R[2] * X[2] + t[0],
                                R[3] * X[0] + R[4] * X[1] + R[5] * X[2] + t[1],
                                R[6] * X[0] + R[7] * X[1] + R[8] * X[2] + t[2]};
                if (gt_u) gt_u[p] = (float)(f * X1[0] / X1[2] + cx - x);
                if (gt_v) gt_v[p] = (float)(f * X1[1] / X1[2] + cy - y);
                if (gt_Z) gt_Z[p] = (float)Z;
            }
    return 0;
}

// Sequence of n frames of one scene (16 planes, depths 2-20) along a smooth trajectory: camera pose
// X_k = R_k X_ref + t_k with R_k = rot(theta_k) about a fixed random axis, theta_k = 4.5 deg *
// sin(2 pi k / 60), t_k = 0.153 sin(2 pi k / 60) e1 + 0.0765 sin(2 pi k / 45) e2 + 0.046
// sin(2 pi k / 75) e3 for a random orthonormal basis e1, e2, e3; per-frame rotation <= 0.5 deg
// and camera displacement <= 0.02. K is row-major 3x3; Rrel/trel receive the n - 1 relative poses
// (X_{k+1} = Rrel_k X_k + trel_k), frames the n images (n x h x s floats). Frames are rendered in
// parallel; each draws its noise from its own stream.
int rfs_make_sequence(int w, int h, int s, uint64_t seed, int n, double* K, double* Rrel, double* trel,
                      float* frames) {
    if (w <= 0 || h <= 0 || s < w || n < 2 || !K || !Rrel || !trel || !frames) return -1;
    Rng rng(seed * 2654435761ull + 71);
    std::vector<double> f0 = texture(rng, w, h);
    const double f = 500.0 * w / 640.0, cx = 0.5 * w, cy = 0.5 * h;
    double Km[9] = {f, 0, cx, 0, f, cy, 0, 0, 1};
    std::memcpy(K, Km, sizeof Km);
    Scene sc = make_scene(rng, 16, w, h, f, cx, cy, 0.05, 0.5);
    double axis[3], e1[3], e2[3], e3[3];
    unit(rng, axis);
    unit(rng, e1);
    unit(rng, e2);
    const double d12 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];  // Gram-Schmidt
    for (int i = 0; i < 3; ++i) e2[i] -= d12 * e1[i];
    const double n2 = std::sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
    for (int i = 0; i < 3; ++i) e2[i] /= n2;
    e3[0] = e1[1] * e2[2] - e1[2] * e2[1];
    e3[1] = e1[2] * e2[0] - e1[0] * e2[2];
    e3[2] = e1[0] * e2[1] - e1[1] * e2[0];
    std::vector<double> R((size_t)n * 9), t((size_t)n * 3);
    const double pi = 3.141592653589793;
    for (int k = 0; k < n; ++k) {
        rodrigues(axis, 4.5 * pi / 180.0 * std::sin(2 * pi * k / 60.0), &R[(size_t)k * 9]);
        const double a = 0.153 * std::sin(2 * pi * k / 60.0), b = 0.0765 * std::sin(2 * pi * k / 45.0),
                     c = 0.046 * std::sin(2 * pi * k / 75.0);
        for (int i = 0; i < 3; ++i) t[(size_t)k * 3 + i] = a * e1[i] + b * e2[i] + c * e3[i];
    }
    for (int k = 0; k + 1 < n; ++k) {  // X_{k+1} = R_{k+1} R_k^T X_k + t_{k+1} - R_{k+1} R_k^T t_k
        const double *Ra = &R[(size_t)k * 9], *Rb = &R[(size_t)(k + 1) * 9];
        double* Rr = &Rrel[(size_t)k * 9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                Rr[i * 3 + j] = Rb[i * 3 + 0] * Ra[j * 3 + 0] + Rb[i * 3 + 1] * Ra[j * 3 + 1] + Rb[i * 3 + 2] * Ra[j * 3 + 2];
        for (int i = 0; i < 3; ++i)
            trel[(size_t)k * 3 + i] = t[(size_t)(k + 1) * 3 + i] -
                                      (Rr[i * 3 + 0] * t[(size_t)k * 3 + 0] + Rr[i * 3 + 1] * t[(size_t)k * 3 + 1] +
                                       Rr[i * 3 + 2] * t[(size_t)k * 3 + 2]);
    }
    const int nt = std::max(1, std::min<int>(n, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int th = 0; th < nt; ++th)
        pool.emplace_back([&, th] {
            for (int k = th; k < n; k += nt) {
                Rng fr(seed * 0x9E3779B97F4A7C15ull + 1000003ull * (uint64_t)k + 7);
                store(render(sc, f0, w, h, &R[(size_t)k * 9], &t[(size_t)k * 3], fr), w, h, s,
                      frames + (size_t)k * h * s);
            }
        });
    for (auto& p : pool) p.join();
    return 0;
}

}  // extern "C"
