// Solver context behind the C ABI: device buffers of one image pair, the static solve
// schedule (plain per-step kernels or fused kernels), and an optional hipGraph of it.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/rfhip.h"
#include "rf_device.hpp"

namespace rf {

constexpr int kMaxLevels = 16;

void set_error(const std::string& msg);

// Per-level camera data that the solve kernels read from device memory, so a captured
// schedule stays valid when the pose changes between solves.
struct DevParams {
    PQ pq[kMaxLevels];
    PQd pqd[kMaxLevels];
    UcParams uc[kMaxLevels];
};

// The pointer table the schedule rotates (the reference's d_* members, Cudaflow.h:195-352,
// Reconflow.h:15-31). Each field buffer holds s0*h0 floats plus one zero guard row.
struct Fields {
    float* pI0[kMaxLevels];
    float* pI1[kMaxLevels];
    float *i1warp, *dumed, *dvmed, *dumeds, *dvmeds;
    float *pu1, *pu2, *pv1, *pv2, *pu1s, *pu2s, *pv1s, *pv2s;
    float *Ix, *Iy, *Iz, *u, *v, *us, *vs;
    float *X, *Y, *Z, *Xs, *Ys, *Zs, *Xm, *Ym, *Zm;
    float *uproj, *vproj, *uc, *vc, *sku, *skv, *skus, *skvs;
    float *ufn, *vfn, *ufn_l, *vfn_l, *mask, *mask_l;
    float* uvrgb;
    void *in0, *in1;
    // alternate data terms (allocated for the methods that use them): gradient-image pyramids and
    // their warped / differentiated fields (Charbonnier-gradient), the IRLS iterate du, dv
    // (Charbonnier), mask pyramids (masked TV-L1), depth and scene-flow fields (Kinect scene flow)
    float* pIx0[kMaxLevels];
    float* pIx1[kMaxLevels];
    float* pIy0[kMaxLevels];
    float* pIy1[kMaxLevels];
    float* pM0[kMaxLevels];
    float* pM1[kMaxLevels];
    float *ix1warp, *iy1warp, *Ixx, *Ixy, *Ixz, *Iyx, *Iyy, *Iyz;
    float *du, *dv, *dus, *dvs;
    float *mask0, *mask1;
    void *depth0_16, *depth1_16;
    float *depth032f, *depth132f, *sceneflow, *sceneflowrgb, *pcloud0, *pcloud1;
};

// HIP-event timing of one kernel kind at one level (eager solves only).
struct Profiler {
    bool armed = false;
    int kid = -1, level = -1;
    std::vector<hipEvent_t> ev;  // start/stop pairs
    int used = 0;                // events recorded
    int launches = 0;
    bool open = false;
    void begin(hipStream_t st, int k, int l) {
        if (!armed || k != kid || l != level || open) return;
        if (used + 2 > (int)ev.size()) return;
        (void)hipEventRecord(ev[used], st);  // a failed record shows when the times are read
        open = true;
    }
    void end(hipStream_t st, int k, int l, int n) {
        if (!armed || k != kid || l != level || !open) return;
        (void)hipEventRecord(ev[used + 1], st);
        used += 2;
        launches += n;
        open = false;
    }
};

struct Geometry {
    int L = 0;
    int W[kMaxLevels] = {}, H[kMaxLevels] = {}, S[kMaxLevels] = {};
    long n = 0;     // s0 * h0
    long bufn = 0;  // floats per field buffer (n + guard row, rounded to 64)
};

// Widths and heights at or above this are rejected: the resampling kernels form texel coordinates
// as 32-bit products such as (4 x + 2) * W.
constexpr int kMaxSide = 32768;

// The same products between a level of side `fine` and the next coarser one of side `coarse`:
// downscale forms (4 x + 2) * fine + fine with x < coarse, upscale (2 x + 1) * coarse with
// x < fine. Below kMaxSide they fit for a scale of 2 or more; a smaller scale needs this check.
inline bool resample_fits(int fine, int coarse) {
    const long long lim = 2147483647LL;
    return 4LL * coarse * fine + fine <= lim && 2LL * fine * coarse <= lim;
}

class Batch;

// One image pair. The per-step path runs here; the fused path runs as a batch of one pair.
class Solver {
  public:
    Solver(int bw, int bh, int align);
    ~Solver();
    int initialize(const rf_params& p);
    int init_large_displacement();
    int set_camera(const double* K0, const double* K1);
    int copy_images(const void* i0, const void* i1);
    int copy_images_device(const float* i0, const float* i1);
    int copy_prior(const float* u, const float* v, const float* mask, bool sparse);
    int copy_warm_start(const float* u, const float* v);  // host, level-0 layout; null clears
    int warm_start_in_use();
    int copy_masks(const float* m0, const float* m1);
    int copy_depth(const uint16_t* d0, const uint16_t* d1);
    int solve_flow(float flow_scale);
    int solve_inpainting(float flow_scale);
    int solve_scene_flow(float fx, float fy, float cx, float cy, float flow_scale, float sf_scale, bool point_cloud);
    int copy_scene_flow(float* sf, float* sfrgb, float* pc0, float* pc1);
    int solve_recon(const double* R0, const double* t0, const double* R1, const double* t1, float flow_scale);
    int copy_flow(float* u, float* v, float* uvrgb);
    int copy_3d(float* X, float* Y, float* Z);
    int remove_edge(int range, float maxarea);
    int sync();
    int profile_arm(int kid, int level);
    int profile_read(double* total_ms, int* launches);
    int tiling(bool recon, int* tiling, int n);  // joint (recon) or TV-L1 tilings per level
    // rf_solve_flow_noalloc: the plain TV-L1 solve of params on `st`, its fields carved from the
    // caller's device scratch (nfloats floats), which it zeroes first; allocates nothing. For a
    // Solver that was never initialized.
    int solve_flow_noalloc(const rf_params& p, const float* i0, const float* i1, float* u, float* v,
                           hipStream_t st, float* scratch, size_t nfloats);
    Profiler& prof() { return prof_; }

    const Geometry& geom() const { return g_; }
    const rf_params& params() const { return p_; }
    hipStream_t stream() const;
    void result(const float** u, const float** v, const float** X, const float** Y, const float** Z) const;

  private:
    int release();
    int check_params(const rf_params& p, const char* who);
    int set_geometry(const rf_params& p, int align, const char* who);
    int run_schedule(int kind);  // 0: flow, 1: joint, 2: masked flow
    hipError_t enqueue_flow_plain(Fields& f);
    hipError_t enqueue_recon_plain(Fields& f);
    hipError_t enqueue_gray_pyramid(Fields& f);
    int check(hipError_t e, const char* what);

    // the reference constructor's launch block shape (Cudaflow.cu:10-14), kept for its interface:
    // the kernels here choose their own tiles
    [[maybe_unused]] int bw_, bh_;
    int align_;
    rf_params p_{};
    bool inited_ = false, prior_ = false, sparse_ = false, camera_ = false;
    bool masks_ = false, depth_ = false, scene_ = false, masked_run_ = false;
    // the colour map (uvrgb) holds the last solve's flow: that solve ran with visualization
    bool uvrgb_valid_ = false;
    Geometry g_;
    hipStream_t st_ = nullptr;
    void* slab_ = nullptr;
    size_t slab_bytes_ = 0;
    Fields init_{}, cur_{};
    DevParams hparams_{};
    DevParams* dparams_ = nullptr;
    std::vector<std::vector<double>> K0_, K1_;
    double R0_[9], t0_[3], R1_[9], t1_[3];
    float flow_scale_ = 5.0f;
    // graph cache: one per schedule kind (flow / joint / masked flow)
    hipGraphExec_t gexec_[3] = {nullptr, nullptr, nullptr};
    Fields gfinal_[3];
    Profiler prof_;
    std::unique_ptr<Batch> batch_;  // the fused path
    double K0raw_[9], K1raw_[9];
    // warm start: level-0 flow (device, allocated on first use) and the rule's flag of the last
    // solve that applied it
    float* ws_ = nullptr;
    int* ws_bad_ = nullptr;
    bool ws_set_ = false;
    void drop_graphs();
};

}  // namespace rf
