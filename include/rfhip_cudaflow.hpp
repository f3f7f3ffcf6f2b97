// rfhip_cudaflow.hpp — header-only C++ drop-in for the reference's CudaFlow / ReconFlow classes
// (reconflow/Cudaflow.h:17-361, reconflow/Reconflow.h:9-159), implemented over the C ABI of
// rfhip.h. Method names, argument order and defaults follow the reference; cv::Mat arguments
// become rfhip::Mat, a non-owning view (data pointer, rows, cols, row step in bytes) that a
// cv::Mat converts to with rfhip::Mat(m.data, m.rows, m.cols, m.step). Methods return the C ABI
// status (0 = success) instead of exiting on error; lastError() describes a failure.
#ifndef RFHIP_CUDAFLOW_HPP
#define RFHIP_CUDAFLOW_HPP

#include <cstddef>
#include <cstring>
#include <string>
#include <vector>

#include "rfhip.h"

namespace rfhip {

// Non-owning 2-D host buffer: rows x cols elements with `step` bytes between rows. Images and
// fields handed to the solver must be padded to the solver stride, as the reference requires
// (test_reconflow/main.cpp:57-59).
struct Mat {
    void* data = nullptr;
    int rows = 0, cols = 0;
    size_t step = 0;
    Mat() = default;
    Mat(void* d, int r, int c, size_t s) : data(d), rows(r), cols(c), step(s) {}
    template <typename T>
    T* ptr() const {
        return static_cast<T*>(data);
    }
};

inline std::string lastError() { return rf_last_error(); }

class CudaFlow {
  public:
    static const int METHOD_TVL1 = 0;
    static const int METHOD_TVCHARB = 1;
    static const int METHOD_TVCHARBGRAD = 2;
    static const int METHOD_TVL1PATCH = 3;
    static const int SCENEFLOW_KINECT_TVL1 = 4;
    static const int METHOD_TVL1_FN = 5;
    static const int METHOD_TVCHARB_FN = 6;
    static const int METHOD_TVCHARBGRAD_FN = 7;
    static const int METHOD_TVL1_INPAINTING = 8;

    CudaFlow() : CudaFlow(32, 12, 32) {}
    CudaFlow(int BlockWidth, int BlockHeight, int StrideAlignment) {
        rf_create(&ctx_, BlockWidth, BlockHeight, StrideAlignment);
        StrideAlignment_ = StrideAlignment;
    }
    virtual ~CudaFlow() { rf_destroy(ctx_); }
    CudaFlow(const CudaFlow&) = delete;
    CudaFlow& operator=(const CudaFlow&) = delete;

    // CudaFlow::initialize (Cudaflow.cu:99-127). cvType: RF_INPUT_32F / 8UC3 / 8U / 16U.
    int initialize(int width, int height, int channels, int cvType, int nLevels, float scale, int method,
                   float lambda = 100.0f, float lambdagrad = 400.0f, float theta = 0.33f, float tau = 0.25f,
                   int nWarpIters = 1, int nSolverIters = 100, bool withVisualization = true) {
        rf_params p;
        rf_params_default(&p, width, height);
        p.channels = channels;
        p.input_type = cvType;
        p.n_levels = nLevels;
        p.scale = scale;
        p.method = method;
        p.lambda = lambda;
        p.lambdagrad = lambdagrad;
        p.theta = theta;
        p.tau = tau;
        p.n_warp_iters = nWarpIters;
        p.n_solver_iters = nSolverIters;
        p.with_visualization = withVisualization ? 1 : 0;
        p.stride_alignment = StrideAlignment_;
        applyOptions(p);
        return commit(p);
    }
    // The overload without cvType assumes 8-bit RGB (Cudaflow.cu:16-25).
    int initialize(int width, int height, int channels, int nLevels, float scale, int method, float lambda = 100.0f,
                   float lambdagrad = 400.0f, float theta = 0.33f, float tau = 0.25f, int nWarpIters = 1,
                   int nSolverIters = 100) {
        return initialize(width, height, channels, RF_INPUT_8UC3, nLevels, scale, method, lambda, lambdagrad, theta,
                          tau, nWarpIters, nSolverIters, true);
    }
    int close() { return 0; }  // Cudaflow.cu:980-983

    // Execution options of this implementation (set before initialize).
    void setPath(int path) { path_ = path; }
    void setUseGraph(bool g) { graph_ = g; }
    void setStorage(int storage) { storage_ = storage; }

    int copyImagesToDevice(const Mat& i0, const Mat& i1) { return rf_copy_images(ctx_, i0.data, i1.data); }
    int copyOpticalFlowToHost(Mat& u, Mat& v) { return rf_copy_flow(ctx_, u.ptr<float>(), v.ptr<float>(), nullptr); }
    int copyOpticalFlowToHost(Mat& u, Mat& v, Mat& uvrgb) {
        return rf_copy_flow(ctx_, u.ptr<float>(), v.ptr<float>(), uvrgb.ptr<float>());
    }
    int copyInitialOpticalFlowToDevice(const Mat& u, const Mat& v) {
        return rf_copy_initial_flow(ctx_, u.ptr<float>(), v.ptr<float>());
    }
    int copySparseOpticalFlowToDevice(const Mat& u, const Mat& v) {
        return rf_copy_sparse_flow(ctx_, u.ptr<float>(), v.ptr<float>(), nullptr);
    }
    int copySparseOpticalFlowToDevice(const Mat& u, const Mat& v, const Mat& mask) {
        return rf_copy_sparse_flow(ctx_, u.ptr<float>(), v.ptr<float>(), mask.ptr<float>());
    }
    // Warm-start input (not in the reference): the flow that replaces the zero initial flow at
    // the coarsest level of the following solves; clearWarmStart goes back to zero.
    int copyWarmStartToDevice(const Mat& u, const Mat& v) {
        return rf_copy_warm_start(ctx_, u.ptr<float>(), v.ptr<float>());
    }
    int clearWarmStart() { return rf_copy_warm_start(ctx_, nullptr, nullptr); }
    bool warmStartInUse() { return rf_warm_start_in_use(ctx_) == 1; }

    int solveOpticalFlow() { return solveOpticalFlow(5.0f); }
    int solveOpticalFlow(float flowScale) {
        int rc = rf_solve_flow(ctx_, flowScale);
        return rc ? rc : rf_synchronize(ctx_);
    }

    // Masked TV-L1 (Inpainting.cu:3-152) and Kinect scene flow (Cudaflow.cu:282-286, 345-362, 816-871).
    int copyMaskToDevice(const Mat& mask0, const Mat& mask1) {
        return rf_copy_masks(ctx_, mask0.ptr<float>(), mask1.ptr<float>());
    }
    int solveInpainting(float flowScale) {
        int rc = rf_solve_inpainting(ctx_, flowScale);
        return rc ? rc : rf_synchronize(ctx_);
    }
    int copyDepthToDevice(const Mat& depth0, const Mat& depth1) {
        return rf_copy_depth(ctx_, depth0.ptr<uint16_t>(), depth1.ptr<uint16_t>());
    }
    int solveSceneFlow() { return solveSceneFlow(fx_, fy_, cx_, cy_, 5.0f, 1.0f); }
    int solveSceneFlow(float focalx, float focaly, float cx, float cy, float flowScale) {
        return solveSceneFlow(focalx, focaly, cx, cy, flowScale, 1.0f);
    }
    int solveSceneFlow(float focalx, float focaly, float cx, float cy, float opticalFlowScale, float sceneFlowScale) {
        fx_ = focalx, fy_ = focaly, cx_ = cx, cy_ = cy;
        int rc = rf_solve_scene_flow(ctx_, focalx, focaly, cx, cy, opticalFlowScale, sceneFlowScale, 0);
        return rc ? rc : rf_synchronize(ctx_);
    }
    int solveSceneFlowAndPointCloud(float focalx, float focaly, float cx, float cy, float opticalFlowScale,
                                    float sceneFlowScale) {
        fx_ = focalx, fy_ = focaly, cx_ = cx, cy_ = cy;
        int rc = rf_solve_scene_flow(ctx_, focalx, focaly, cx, cy, opticalFlowScale, sceneFlowScale, 1);
        return rc ? rc : rf_synchronize(ctx_);
    }
    int copySceneFlowToHost(Mat& sceneflow) {
        return rf_copy_scene_flow(ctx_, sceneflow.ptr<float>(), nullptr, nullptr, nullptr);
    }
    int copySceneFlowToHost(Mat& sceneflow, Mat& sceneflowrgb) {
        return rf_copy_scene_flow(ctx_, sceneflow.ptr<float>(), sceneflowrgb.ptr<float>(), nullptr, nullptr);
    }
    int copySceneFlowAndPointCloudToHost(Mat& sceneflow, Mat& pcloud0, Mat& pcloud1, Mat& sceneflowrgb) {
        return rf_copy_scene_flow(ctx_, sceneflow.ptr<float>(), sceneflowrgb.ptr<float>(), pcloud0.ptr<float>(),
                                  pcloud1.ptr<float>());
    }

    int iAlignUp(int n) const {
        int m = StrideAlignment_, mod = n % m;
        return mod ? n + m - mod : n;
    }
    int computePyramidLevels(int width, int height, int minWidth, float scale) const {
        return rf_compute_pyramid_levels(width, height, minWidth, scale);
    }

    // public per-kernel wrappers on device pointers (Cudaflow.h:101-193), null stream
    void MedianFilter(float* inputu, float* inputv, int w, int h, int s, float* outputu, float* outputv,
                      int kernelsize) {
        rf_k_median_filter(inputu, inputv, w, h, s, outputu, outputv, kernelsize, nullptr);
    }
    void MedianFilter3D(float* X, float* Y, float* Z, int w, int h, int s, float* X1, float* Y1, float* Z1,
                        int kernelsize) {
        rf_k_median_filter_3d(X, Y, Z, w, h, s, X1, Y1, Z1, kernelsize, nullptr);
    }
    void Add(const float* op1, const float* op2, int count, float* sum) { rf_k_add(op1, op2, count, sum, nullptr); }
    void ComputeDerivatives(float* I0, float* I1, int w, int h, int s, float* Ix, float* Iy, float* Iz) {
        rf_k_compute_derivatives(I0, I1, w, h, s, Ix, Iy, Iz, nullptr);
    }
    void Downscale(const float* src, int width, int height, int stride, int newWidth, int newHeight, int newStride,
                   float* out) {
        rf_k_downscale(src, width, height, stride, newWidth, newHeight, newStride, out, nullptr);
    }
    void Downscale(const float* src, int width, int height, int stride, int newWidth, int newHeight, int newStride,
                   float scale, float* out) {
        rf_k_downscale_scaled(src, width, height, stride, newWidth, newHeight, newStride, scale, out, nullptr);
    }
    void Upscale(const float* src, int width, int height, int stride, int newWidth, int newHeight, int newStride,
                 float scale, float* out) {
        rf_k_upscale(src, width, height, stride, newWidth, newHeight, newStride, scale, out, nullptr);
    }
    void WarpImage(const float* src, int w, int h, int s, const float* u, const float* v, float* out) {
        rf_k_warp_image(src, w, h, s, u, v, out, nullptr);
    }
    void SolveDataL1(const float* duhat0, const float* dvhat0, const float* pu1, const float* pu2, const float* pv1,
                     const float* pv2, const float* Ix, const float* Iy, const float* Iz, int w, int h, int s,
                     float lambda, float theta, float* duhat1, float* dvhat1) {
        rf_k_solve_data_l1(duhat0, dvhat0, pu1, pu2, pv1, pv2, Ix, Iy, Iz, w, h, s, lambda, theta, duhat1, dvhat1,
                           nullptr);
    }
    void SolveSmoothDualTVGlobal(float* duhat, float* dvhat, float* pu1, float* pu2, float* pv1, float* pv2, int w,
                                 int h, int s, float tau, float theta, float* pu1s, float* pu2s, float* pv1s,
                                 float* pv2s) {
        rf_k_solve_smooth_dual_tv_global(duhat, dvhat, pu1, pu2, pv1, pv2, w, h, s, tau, theta, pu1s, pu2s, pv1s,
                                         pv2s, nullptr);
    }
    void ComputeDeriv(float* I0, int w, int h, int s, float* Ix0, float* Iy0) {
        rf_k_compute_deriv(I0, w, h, s, Ix0, Iy0, nullptr);
    }
    void ComputeDerivativesPatch(float* I0, float* I1, int w, int h, int s, float* Ix, float* Iy, float* Iz) {
        rf_k_compute_derivatives_patch(I0, I1, w, h, s, Ix, Iy, Iz, nullptr);
    }
    void SolveDataCharbForTV(float* du0, float* dv0, float* duhat0, float* dvhat0, float* pu1, float* pu2,
                             float* pv1, float* pv2, float* Ix, float* Iy, float* Iz, int w, int h, int s,
                             float lambda, float theta, float* du1, float* dv1, float* duhat1, float* dvhat1) {
        rf_k_solve_data_charb_for_tv(du0, dv0, duhat0, dvhat0, pu1, pu2, pv1, pv2, Ix, Iy, Iz, nullptr, nullptr,
                                     nullptr, nullptr, nullptr, nullptr, w, h, s, lambda, 0.0f, theta, du1, dv1,
                                     duhat1, dvhat1, nullptr);
    }
    void SolveDataCharbForTVGrad(float* du0, float* dv0, float* duhat0, float* dvhat0, float* pu1, float* pu2,
                                 float* pv1, float* pv2, float* Ix, float* Iy, float* Iz, float* Ixx, float* Ixy,
                                 float* Ixz, float* Iyx, float* Iyy, float* Iyz, int w, int h, int s, float lambda,
                                 float lambdagrad, float theta, float* du1, float* dv1, float* duhat1,
                                 float* dvhat1) {
        rf_k_solve_data_charb_for_tv(du0, dv0, duhat0, dvhat0, pu1, pu2, pv1, pv2, Ix, Iy, Iz, Ixx, Ixy, Ixz, Iyx,
                                     Iyy, Iyz, w, h, s, lambda, lambdagrad, theta, du1, dv1, duhat1, dvhat1, nullptr);
    }
    void SolveDataL1Inpaint(const float* duhat0, const float* dvhat0, const float* mask0, const float* mask1,
                            const float* pu1, const float* pu2, const float* pv1, const float* pv2, const float* Ix,
                            const float* Iy, const float* Iz, int w, int h, int s, float lambda, float theta,
                            float* duhat1, float* dvhat1) {
        rf_k_solve_data_l1_inpaint(duhat0, dvhat0, mask0, mask1, pu1, pu2, pv1, pv2, Ix, Iy, Iz, w, h, s, lambda,
                                   theta, duhat1, dvhat1, nullptr);
    }
    void Cv16uTo32f(const uint16_t* d_iCv16u, float* d_iGray, int w, int h, int s) {
        rf_k_cv16u_to_32f(d_iCv16u, d_iGray, w, h, s, nullptr);
    }
    void SolveSceneFlow_(float* u, float* v, float* depth0, float* depth1, float fx, float fy, float cx, float cy,
                         int w, int h, int s, float* sceneflow) {
        rf_k_solve_scene_flow(u, v, depth0, depth1, fx, fy, cx, cy, w, h, s, nullptr, nullptr, sceneflow, nullptr);
    }
    void SolveSceneFlow_(float* u, float* v, float* depth0, float* depth1, float fx, float fy, float cx, float cy,
                         int w, int h, int s, float* pcloud0, float* pcloud1, float* sceneflow) {
        rf_k_solve_scene_flow(u, v, depth0, depth1, fx, fy, cx, cy, w, h, s, pcloud0, pcloud1, sceneflow, nullptr);
    }
    void SceneFlowToHSV(float* sceneflow, int w, int h, int s, float* uRGB, float flowscale) {
        rf_k_scene_flow_to_hsv(sceneflow, w, h, s, uRGB, flowscale, nullptr);
    }
    void FlowToHSV(float* u, float* v, int w, int h, int s, float* uRGB, float flowscale) {
        rf_k_flow_to_hsv(u, v, w, h, s, uRGB, flowscale, nullptr);
    }
    void rgbToGray(const uint8_t* d_iRgb, float* d_iGray, int w, int h, int s) {
        rf_k_rgb_to_gray(d_iRgb, d_iGray, w, h, s, nullptr);
    }
    void Cv8uToGray(const uint8_t* d_iCv8u, float* d_iGray, int w, int h, int s) {
        rf_k_cv8u_to_gray(d_iCv8u, d_iGray, w, h, s, nullptr);
    }
    void Cv16uToGray(const uint16_t* d_iCv16u, float* d_iGray, int w, int h, int s) {
        rf_k_cv16u_to_gray(d_iCv16u, d_iGray, w, h, s, nullptr);
    }
    void Cv32fToGray(const float* d_iCv32f, float* d_iGray, int w, int h, int s) {
        rf_k_cv32f_to_gray(d_iCv32f, d_iGray, w, h, s, nullptr);
    }

    int width = 0, height = 0, stride = 0, nLevels = 0;
    rf_context* context() const { return ctx_; }

  protected:
    void applyOptions(rf_params& p) const {
        // the fused path runs METHOD_TVL1 and the joint methods; the others run per step
        const bool fusable = p.method == RF_METHOD_TVL1 || p.method == RF_METHODR_TVL1_MS ||
                             p.method == RF_METHODR_TVL1_MS_FN || p.method == RF_METHODR_TVL1_MS_FNSPARSE;
        p.path = fusable ? path_ : RF_PATH_PLAIN;
        p.use_graph = graph_ ? 1 : 0;
        p.storage = storage_;
    }
    int commit(const rf_params& p) {
        int rc = rf_initialize(ctx_, &p);
        if (rc) return rc;
        width = p.width;
        height = p.height;
        nLevels = p.n_levels;
        stride = rf_get_stride(ctx_);
        return 0;
    }
    rf_context* ctx_ = nullptr;
    int StrideAlignment_ = 32;
    int path_ = RF_PATH_FUSED;
    bool graph_ = false;
    int storage_ = RF_STORAGE_F32;
    float fx_ = 0.0f, fy_ = 0.0f, cx_ = 0.0f, cy_ = 0.0f;  // depth camera of solveSceneFlow()
};

class ReconFlow : public CudaFlow {
  public:
    static const int METHODR_TVL1_MS = 14;
    static const int METHODR_TVL1_MSPLANAR = 15;
    static const int METHODR_TVL1_3D = 16;
    static const int METHODR_TVL1_MS_FN = 17;
    static const int METHODR_TVCHARBGRAD_MS = 18;
    static const int METHODR_TVCHARBGRAD_MS_FN = 19;
    static const int METHODR_TVL1_MS_FNSPARSE = 20;

    ReconFlow() : CudaFlow(32, 12, 32) {}
    ReconFlow(int BlockWidth, int BlockHeight, int StrideAlignment) : CudaFlow(BlockWidth, BlockHeight, StrideAlignment) {}

    // ReconFlow::_initializeR (Reconflow.cu:27-63). The reference takes 8-bit RGB only
    // (Reconflow.cu:162-165); setInputType() selects another CudaFlow input type.
    int _initializeR(int width, int height, int channels, int nLevels, float scale, int method, float lambda = 100.0f,
                     float lambdagrad = 400.0f, float lambdaf = 100.0f, float lambdams = 100.0f, float alphaTv = 3.0f,
                     float alphaProj = 1.0f, float tau = 0.25f, int nWarpIters = 1, int nSolverIters = 100) {
        return init(width, height, channels, nLevels, scale, method, lambda, lambdagrad, lambdaf, lambdams, alphaTv,
                    alphaProj, 0.5f, tau, nWarpIters, nSolverIters, false);
    }
    // ReconFlow::initializeR (Reconflow.cu:15-25): with the prior-flow (large displacement) buffers.
    int initializeR(int width, int height, int channels, int nLevels, float scale, int method, float lambda = 100.0f,
                    float lambdagrad = 400.0f, float lambdaf = 100.0f, float lambdams = 100.0f, float alphaTv = 3.0f,
                    float alphaProj = 1.0f, float alphaFn = 0.5f, float tau = 0.25f, int nWarpIters = 1,
                    int nSolverIters = 100) {
        return init(width, height, channels, nLevels, scale, method, lambda, lambdagrad, lambdaf, lambdams, alphaTv,
                    alphaProj, alphaFn, tau, nWarpIters, nSolverIters, true);
    }
    void setInputType(int cvType) { inputType_ = cvType; }
    // the FP64 mode (rf_params::tri_f64): binary64 triangulation, projection blend and data blend
    void setTriangulationFp64(bool on) { triF64_ = on; }
    void setFixCoarseIntrinsics(bool on) { fixK_ = on; }
    // solveReconFlow ends with FlowToHSV (Reconflow.cu:460), so copyOpticalFlowToHost(u, v, uvrgb)
    // returns the colour map; false skips that kernel (uvrgb is then unavailable), e.g. for timing.
    void setVisualization(bool on) { visualization_ = on; }

    int setCameraMatrices(const Mat&) { return 0; }  // Reconflow.cu:72-76 is a no-op
    // K0, K1: 3x3 row-major CV_64F
    int setCameraMatrices(const Mat& K0, const Mat& K1) {
        return rf_set_camera(ctx_, K0.ptr<const double>(), K1.ptr<const double>());
    }
    // solveR(R1, t1, flowScale) / solveR(R0, t0, R1, t1, flowScale) (Reconflow.cu:130-136).
    int solveR(const Mat& R1, const Mat& t1, float flowScale) {
        int rc = rf_solve_recon(ctx_, nullptr, nullptr, R1.ptr<const double>(), t1.ptr<const double>(), flowScale);
        return rc ? rc : rf_synchronize(ctx_);
    }
    int solveR(const Mat& R0, const Mat& t0, const Mat& R1, const Mat& t1, float flowScale) {
        int rc = rf_solve_recon(ctx_, R0.ptr<const double>(), t0.ptr<const double>(), R1.ptr<const double>(),
                                t1.ptr<const double>(), flowScale);
        return rc ? rc : rf_synchronize(ctx_);
    }
    int solveReconFlow(const Mat& R1, const Mat& t1, float flowScale) { return solveR(R1, t1, flowScale); }
    int solveReconFlow(const Mat& R0, const Mat& t0, const Mat& R1, const Mat& t1, float flowScale) {
        return solveR(R0, t0, R1, t1, flowScale);
    }
    int copy3dToHost(Mat& X, Mat& Y, Mat& Z) { return rf_copy_3d(ctx_, X.ptr<float>(), Y.ptr<float>(), Z.ptr<float>()); }
    int removeEdgeArtifacts3D(int surfaceWidth, float maxSurfaceArea) {
        int rc = rf_remove_edge_artifacts_3d(ctx_, surfaceWidth, maxSurfaceArea);
        return rc ? rc : rf_synchronize(ctx_);
    }

    // public per-kernel wrappers (Reconflow.h:86-158), null stream; P, Q are 3x4 row-major
    void SolveReconDataL1(const float* u0, const float* v0, const float* duhat0, const float* dvhat0,
                          const float* uproj0, const float* vproj0, const float* sku0, const float* skv0,
                          const float* pu1, const float* pu2, const float* pv1, const float* pv2, const float* Ix,
                          const float* Iy, const float* Iz, int w, int h, int s, float lambda, float alphaTv,
                          float alphaProj, float* duhat1, float* dvhat1) {
        rf_k_solve_recon_data_l1(u0, v0, duhat0, dvhat0, uproj0, vproj0, sku0, skv0, pu1, pu2, pv1, pv2, Ix, Iy, Iz,
                                 w, h, s, lambda, alphaTv, alphaProj, duhat1, dvhat1, nullptr);
    }
    void SolveUVProj(const float* u0, const float* v0, const float* du0, const float* dv0, const float* uc0,
                     const float* vc0, const float* sku0, const float* skv0, int w, int h, int s, float lambdaf,
                     float alphaProj, float* uproj1, float* vproj1) {
        rf_k_solve_uvproj(u0, v0, du0, dv0, uc0, vc0, sku0, skv0, w, h, s, lambdaf, alphaProj, uproj1, vproj1,
                          nullptr);
    }
    void Solve3dMinimalArea(const float* uproj0, const float* vproj0, const float* X0, const float* Y0,
                            const float* Z0, double* P, double* Q, float lambdaf, float lambdams, int w, int h, int s,
                            float* X, float* Y, float* Z) {
        rf_k_solve_3d_minimal_area(uproj0, vproj0, X0, Y0, Z0, P, Q, lambdaf, lambdams, w, h, s, X, Y, Z, nullptr);
    }
    void SolveUc(const float* X, const float* Y, const float* Z, double* R, double* t, double* K1, int w, int h,
                 int s, float* uc, float* vc) {
        rf_k_solve_uc(X, Y, Z, R, t, K1, w, h, s, uc, vc, nullptr);
    }
    void UpdateSk(const float* sku0, const float* skv0, const float* u0, const float* v0, const float* du0,
                  const float* dv0, const float* uproj0, const float* vproj0, int width, int height, int stride,
                  float* sku1, float* skv1) {
        rf_k_update_sk(sku0, skv0, u0, v0, du0, dv0, uproj0, vproj0, width, height, stride, sku1, skv1, nullptr);
    }
    void CleanUp3D(const float* X0, const float* Y0, const float* Z0, const float* Xmed, const float* Ymed,
                   const float* Zmed, int w, int h, int s, float* X1, float* Y1, float* Z1) {
        rf_k_cleanup_3d(X0, Y0, Z0, Xmed, Ymed, Zmed, w, h, s, X1, Y1, Z1, nullptr);
    }
    void RemoveEdgeArtifacts3D(const float* X0, const float* Y0, const float* Z0, int surfaceWidth,
                               float maxSurfaceArea, int w, int h, int s, float* X1, float* Y1, float* Z1) {
        rf_k_remove_edge_artifacts_3d(X0, Y0, Z0, surfaceWidth, maxSurfaceArea, w, h, s, X1, Y1, Z1, nullptr);
    }

    float thetaTv = 0, alphaTv = 0, alphaProj = 0, lambdaf = 0, lambdams = 0, alphaFn = 0;

  private:
    int init(int width, int height, int channels, int nLevels, float scale, int method, float lambda, float lambdagrad,
             float lambdaf_, float lambdams_, float alphaTv_, float alphaProj_, float alphaFn_, float tau,
             int nWarpIters, int nSolverIters, bool prior) {
        rf_params p;
        rf_params_default(&p, width, height);
        p.channels = channels;
        p.input_type = inputType_;
        p.n_levels = nLevels;
        p.scale = scale;
        p.method = method;
        p.lambda = lambda;
        p.lambdagrad = lambdagrad;
        p.theta = 1.0f / alphaTv_;  // Reconflow.cu:32
        p.tau = tau;
        p.n_warp_iters = nWarpIters;
        p.n_solver_iters = nSolverIters;
        p.lambdaf = lambdaf_;
        p.lambdams = lambdams_;
        p.alpha_tv = alphaTv_;
        p.alpha_proj = alphaProj_;
        p.alpha_fn = alphaFn_;
        p.stride_alignment = StrideAlignment_;
        p.with_visualization = visualization_ ? 1 : 0;
        p.tri_f64 = triF64_ ? 1 : 0;
        p.fix_coarse_k = fixK_ ? 1 : 0;
        applyOptions(p);
        int rc = commit(p);
        if (rc) return rc;
        thetaTv = p.theta;
        alphaTv = alphaTv_;
        alphaProj = alphaProj_;
        lambdaf = lambdaf_;
        lambdams = lambdams_;
        alphaFn = alphaFn_;
        return prior ? rf_initialize_large_displacement(ctx_) : 0;
    }
    int inputType_ = RF_INPUT_8UC3;
    bool triF64_ = false, fixK_ = false, visualization_ = true;
};

}  // namespace rfhip

#endif  // RFHIP_CUDAFLOW_HPP
