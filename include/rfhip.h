/*
 * rfhip.h — C ABI of the MI355X-native TV-L1 optical-flow and joint flow/depth solver.
 *
 * This is the drop-in boundary for the reference's CudaFlow / ReconFlow classes
 * (reconflow/Cudaflow.h:17-361, reconflow/Reconflow.h:9-159). Every entry point is
 * extern "C", takes plain pointers and sizes, and returns an int status (0 = success,
 * negative = error; rf_last_error() describes the most recent failure on the calling
 * thread). Nothing here exits the process.
 *
 * Layout conventions kept from the reference:
 *   - images and fields are row-major float arrays of stride x height elements, with
 *     stride = width rounded up to the stride alignment (Cudaflow.cu:985-994, default 32);
 *   - host images handed to rf_copy_images are already padded to the stride
 *     (test_reconflow/main.cpp:57-59); padding columns of outputs are undefined;
 *   - flow maps frame-0 pixels into frame 1 (I0(x) ~ I1(x + u(x)), warpImage.cu:25-28);
 *   - the pose of camera 1 is X1 = R X0 + t (solveUcSk.cu:25-27), 3x3 and 3x1 row-major
 *     doubles; intrinsics are 3x3 row-major doubles.
 *
 * Device pointers (rf_k_* and the *_device calls) are HIP device addresses; `stream` is a
 * hipStream_t passed as void* (NULL = the null stream).
 */
#ifndef RFHIP_H
#define RFHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- method identifiers (Cudaflow.h:20-28, Reconflow.h:46-52) ----
 * rf_solve_flow follows _solveOpticalFlow's data-term choice (Cudaflow.cu:512-557): the
 * Charbonnier-gradient step for TVCHARBGRAD and TVCHARBGRAD_FN, the Charbonnier step for
 * TVCHARB, patch derivatives for TVL1PATCH, the L1 step for every other method (TVL1,
 * SCENEFLOW_KINECT_TVL1, TVL1_FN, TVCHARB_FN, TVL1_INPAINTING). The masked step of
 * TVL1_INPAINTING runs through rf_solve_inpainting. Methods other than TVL1 run on the plain
 * path only. */
#define RF_METHOD_TVL1 0
#define RF_METHOD_TVCHARB 1
#define RF_METHOD_TVCHARBGRAD 2
#define RF_METHOD_TVL1PATCH 3
#define RF_SCENEFLOW_KINECT_TVL1 4
#define RF_METHOD_TVL1_FN 5
#define RF_METHOD_TVCHARB_FN 6
#define RF_METHOD_TVCHARBGRAD_FN 7
#define RF_METHOD_TVL1_INPAINTING 8
#define RF_METHODR_TVL1_MS 14
#define RF_METHODR_TVL1_MS_FN 17
#define RF_METHODR_TVL1_MS_FNSPARSE 20

/* ---- input pixel types (the cvType argument of CudaFlow::initialize, Cudaflow.cu:99) ---- */
#define RF_INPUT_32F 0   /* CV_32F gray, copied (rgbToGray.cu:66-77) */
#define RF_INPUT_8UC3 1  /* CV_8UC3, ((r+g+b)/3)/256 (rgbToGray.cu:3-19) */
#define RF_INPUT_8U 2    /* CV_8U, /256 (rgbToGray.cu:21-34) */
#define RF_INPUT_16U 3   /* CV_16U, /65536 (rgbToGray.cu:36-49) */

/* ---- execution paths ---- */
#define RF_PATH_PLAIN 0  /* one kernel per reference step, reference buffer rotation */
#define RF_PATH_FUSED 1  /* one fused kernel per solver iteration, same arithmetic */

/* ---- storage of images, derivatives and dual fields ---- */
#define RF_STORAGE_F32 0
#define RF_STORAGE_F16 1       /* FP16 images, derivatives (and TV-L1 duals); updates in FP32 */
#define RF_STORAGE_F16_DERIV 2 /* FP32 images; FP16 derivatives (and TV-L1 duals); updates in FP32 */
#define RF_STORAGE_F16_DERIV_DUALS 3 /* FP32 images; FP16 derivatives and duals (TV-L1 and joint); updates in FP32 */

/* Status codes. */
#define RF_OK 0
#define RF_ERR_ARG -1
#define RF_ERR_STATE -2
#define RF_ERR_HIP -3
#define RF_ERR_ALLOC -4
#define RF_ERR_UNSUPPORTED -5

/* Solver parameters: the arguments of CudaFlow::initialize (Cudaflow.cu:99-127) and
 * ReconFlow::_initializeR / initializeR (Reconflow.cu:15-41). */
typedef struct rf_params {
    int width, height;
    int channels;          /* kept for interface parity; the pixel type decides the layout */
    int input_type;        /* RF_INPUT_* */
    int n_levels;
    float scale;           /* pyramid divisor: 2.0 halves each level (Cudaflow.cu:167-168) */
    int method;            /* RF_METHOD_* / RF_METHODR_* */
    float lambda, lambdagrad, theta, tau;
    int n_warp_iters, n_solver_iters;
    int with_visualization; /* 1: each solve ends with the FlowToHSV colour map of the flow (uvrgb),
                               as CudaFlow does with withVisualization (Cudaflow.cu:614-618) and
                               ReconFlow::solveReconFlow always does (Reconflow.cu:460) */
    /* ReconFlow (joint) parameters; theta is replaced by 1/alpha_tv (Reconflow.cu:32) */
    float lambdaf, lambdams, alpha_tv, alpha_proj, alpha_fn;
    /* launch geometry of the reference constructor (Cudaflow.cu:10-14) */
    int stride_alignment;
    /* this implementation's execution options */
    int path;              /* RF_PATH_* */
    int storage;           /* RF_STORAGE_* */
    int use_graph;         /* capture the static solve schedule in a hipGraph and replay it */
    int fix_coarse_k;      /* 1: K[2][2] = 1 on every pyramid level (the reference leaves it 0 on
                              levels >= 1, Reconflow.cu:99-110, which breaks the coarse triangulation) */
    int tri_f64;           /* 1: the FP64 mode: binary64 arithmetic where FP32 loses the joint solve's
                              accuracy: the triangulation (Solve3dMinimalArea), the projection blend
                              (SolveUVProj) and the data step's blend (SolveReconDataL1); fields stay
                              FP32. 0: every step in FP32 in the reference's operation order */
    int joint_tile;        /* fused joint iteration tiling: RF_JOINT_TILE_* (results are identical) */
    int flow_tile;         /* fused TV-L1 iteration tiling: RF_FLOW_TILE_* (results are identical) */
} rf_params;

/* Tilings of the fused joint iteration (all give the same results). AUTO: the first joint solves
 * time 16x16, 32x16, 16x16_K2, 16x16_OCC5 and 16x16_K3 on alternating warps of every pyramid
 * level (as many solves as it takes to time each on at least two warps) and keep, per level, the
 * one with the shortest warp. Those solves synchronise. A K2/K3 launch runs two/three solver
 * iterations on an extended tile that recomputes a ring of 3/5 pixels (for latency-bound small
 * levels); an iteration count that K does not divide ends with launches of fewer iterations. */
#define RF_JOINT_TILE_AUTO 0
#define RF_JOINT_TILE_64x8_RPT2 1  /* 64x4 lanes, two rows per lane plus a ring pass */
#define RF_JOINT_TILE_32x16 2      /* one pixel per lane, 30x14 pixels per workgroup */
#define RF_JOINT_TILE_16x16 3      /* one pixel per lane, 14x14 pixels per workgroup */
#define RF_JOINT_TILE_16x16_OCC5 4 /* as 16x16, registers capped for five waves per SIMD */
#define RF_JOINT_TILE_16x16_K2 5   /* two iterations per launch, 10x10 pixels per 16x16 workgroup */
#define RF_JOINT_TILE_16x16_K3 6   /* three iterations per launch, 6x6 pixels per 16x16 workgroup */

/* Tilings of the fused TV-L1 iteration (all give the same results). AUTO: the first solves time
 * 16x16, 32x16_K4, 32x16, 32x16_K6, 32x32_K8 and 64x16_RPT4 on alternating warps of every level
 * (as many solves as it takes to time each on at least two warps) and keep, per level, the one with
 * the shortest warp. Those solves synchronise. A K2/K4/K6/K8 launch runs that many solver
 * iterations on an extended tile that recomputes a ring of pixels; an iteration count that K does
 * not divide ends with launches of fewer iterations. */
#define RF_FLOW_TILE_AUTO 0
#define RF_FLOW_TILE_64x16_RPT4 1 /* 64x4 lanes, four rows per lane plus a ring pass */
#define RF_FLOW_TILE_16x16 2      /* one pixel per lane, 15x15 pixels per workgroup */
#define RF_FLOW_TILE_32x16 3      /* one pixel per lane, 31x15 pixels per workgroup */
#define RF_FLOW_TILE_32x16_K4 4   /* four iterations per launch, 25x9 pixels per 32x16 workgroup */
#define RF_FLOW_TILE_32x32_K8 5   /* eight iterations per launch, 17x17 pixels per 32x32 workgroup */
#define RF_FLOW_TILE_32x16_K2 6   /* two iterations per launch, 29x13 pixels per 32x16 workgroup */
#define RF_FLOW_TILE_32x16_K6 7   /* six iterations per launch, 21x5 pixels per 32x16 workgroup */

/* Fills p with the reference defaults (Cudaflow.cu:99-101, Reconflow.cu:27-30) for
 * width x height: 1 warp, 100 iterations, lambda 100, theta 0.33, tau 0.25, alignment 32,
 * visualization on, plain FP32 path without a graph. */
void rf_params_default(rf_params* p, int width, int height);

typedef struct rf_context rf_context;

/* Last error message of the calling thread ("" if none). */
const char* rf_last_error(void);

/* Number of levels the reference demo derives for a height (main.cpp:50-56 and
 * CudaFlow::computePyramidLevels, Cudaflow.cu:1023-1031). */
int rf_compute_pyramid_levels(int width, int height, int min_width, float scale);

/* ---- lifecycle: CudaFlow(), initialize(...), ~CudaFlow() ---- */
/* CudaFlow::CudaFlow(BlockWidth, BlockHeight, StrideAlignment) (Cudaflow.cu:10-14). The block
 * shape is recorded for interface parity; kernels use gfx950 wave64 tiles. */
int rf_create(rf_context** out, int block_width, int block_height, int stride_alignment);
/* ~CudaFlow / ~ReconFlow (Cudaflow.cu:873-978, Reconflow.cu:479-508). */
int rf_destroy(rf_context* ctx);
/* CudaFlow::initialize / ReconFlow::_initializeR. Re-initialising releases the previous
 * buffers (the reference leaks them). Allocates every device buffer once. */
int rf_initialize(rf_context* ctx, const rf_params* params);
/* ReconFlow::initializeR with a prior flow (Reconflow.cu:15-25): initialize + prior buffers. */
int rf_initialize_large_displacement(rf_context* ctx);
/* Pyramid geometry chosen at initialize (Cudaflow.cu:146-171). Arrays hold n_levels ints. */
int rf_get_pyramid(const rf_context* ctx, int* widths, int* heights, int* strides);
int rf_get_stride(const rf_context* ctx);
/* The stream every kernel of this context runs on (hipStream_t as void*). */
void* rf_get_stream(const rf_context* ctx);

/* ---- inputs ---- */
/* CudaFlow::copyImagesToDevice (Cudaflow.cu:288-311): stride x height pixels of the
 * initialize() input type from host memory. Synchronous. */
int rf_copy_images(rf_context* ctx, const void* i0, const void* i1);
/* Same from device memory (stride x height float gray each), enqueued on the stream. */
int rf_copy_images_device(rf_context* ctx, const float* d_i0, const float* d_i1);
/* ReconFlow::setCameraMatrices(K0, K1) (Reconflow.cu:78-116). */
int rf_set_camera(rf_context* ctx, const double K0[9], const double K1[9]);
/* CudaFlow::copyInitialOpticalFlowToDevice (Cudaflow.cu:326-330) and
 * copySparseOpticalFlowToDevice (Cudaflow.cu:332-343); mask may be NULL (all ones). */
int rf_copy_initial_flow(rf_context* ctx, const float* u, const float* v);
int rf_copy_sparse_flow(rf_context* ctx, const float* u, const float* v, const float* mask);
/* Warm-start input, no counterpart in the reference (whose solves start from zero flow,
 * Cudaflow.cu:467-468, Reconflow.cu:209-220): stride x height host flow (u, v) that replaces the
 * zero initial flow at the coarsest level of the following solves, downscaled there as the prior
 * flow is (DownscaleScalingKernel from level 0, downscale.cu:28-48). A warm start with a
 * non-finite component, or one larger than max(width, height), over the width x height pixels
 * is not used: that solve starts from zero. NULL u or v clears it. Synchronous. */
int rf_copy_warm_start(rf_context* ctx, const float* u, const float* v);
/* 1 when a warm start is set and the last solve used it, 0 otherwise; synchronises the stream. */
int rf_warm_start_in_use(rf_context* ctx);
/* CudaFlow::copyMaskToDevice (Inpainting.cu:3-7): stride x height float masks of
 * METHOD_TVL1_INPAINTING (0 = no data term at the pixel); mask1 is kept and, as in the
 * reference, not used by the data step. Synchronous. */
int rf_copy_masks(rf_context* ctx, const float* mask0, const float* mask1);
/* CudaFlow::copyDepthToDevice (Cudaflow.cu:282-286): stride x height 16-bit depth maps of
 * SCENEFLOW_KINECT_TVL1. Synchronous. */
int rf_copy_depth(rf_context* ctx, const uint16_t* depth0, const uint16_t* depth1);

/* ---- solve ---- */
/* CudaFlow::solveOpticalFlow(flowScale) (Cudaflow.cu:395-621). Asynchronous on the stream. */
int rf_solve_flow(rf_context* ctx, float flow_scale);
/* ReconFlow::solveR(R0, t0, R1, t1, flowScale) (Reconflow.cu:134-466). Pass R0 = t0 = NULL
 * for camera 0 at the origin (solveR(R1, t1), Reconflow.cu:130-144). Asynchronous. */
int rf_solve_recon(rf_context* ctx, const double* R0, const double* t0, const double* R1, const double* t1,
                   float flow_scale);
/* CudaFlow::solveInpainting(flowScale) (Inpainting.cu:9-152): the TV-L1 solve with the data
 * step skipped where mask0 == 0 (METHOD_TVL1_INPAINTING, masks copied). Asynchronous. */
int rf_solve_inpainting(rf_context* ctx, float flow_scale);
/* CudaFlow::solveSceneFlow(fx, fy, cx, cy, flowScale, sceneFlowScale) and
 * solveSceneFlowAndPointCloud (Cudaflow.cu:816-871): the TV-L1 solve, the depth0 map converted
 * to float, back-projection of the depth at the pixel and at the flow target (both from depth0,
 * as the reference passes it twice) and the scene-flow colour map. With point_cloud = 0 the
 * point clouds are not written. SCENEFLOW_KINECT_TVL1 with depth copied. Asynchronous. */
int rf_solve_scene_flow(rf_context* ctx, float fx, float fy, float cx, float cy, float flow_scale,
                        float scene_flow_scale, int point_cloud);
/* Blocks until the context's stream is idle. */
int rf_synchronize(rf_context* ctx);

/* Allocation-free TV-L1 solve. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It gives the bits of rf_initialize + rf_copy_images_device + rf_solve_flow on a fresh
 * context (METHOD_TVL1, float gray input; params->path, storage, use_graph and with_visualization
 * are ignored: plain FP32 schedule, no colour map) and calls no allocating routine. d_i0, d_i1:
 * stride x height float device images; d_u, d_v: stride x height device outputs; enqueued on
 * `stream` (hipStream_t as void*, NULL = the null stream). The working fields (2 * n_levels + 20
 * buffers of stride x (height + 1) floats, rounded up to 64) live in a static device array of
 * RF_NOALLOC_SCRATCH_BYTES per GPU, held from the library's load on that GPU until the process ends;
 * a solve that does not fit (e.g. above 960x540 with 5 levels) is RF_ERR_ARG. Every call on a GPU
 * shares that array and zeroes the part it uses: calls that overlap on one GPU race. */
#define RF_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int rf_solve_flow_noalloc(const rf_params* params, const float* d_i0, const float* d_i1, float* d_u, float* d_v,
                          void* stream);

/* ---- outputs ---- */
/* CudaFlow::copyOpticalFlowToHost(u, v[, uvrgb]) (Cudaflow.cu:313-324): stride x height
 * floats each; uvrgb (3 floats per pixel, BGR) may be NULL. A non-NULL uvrgb after a solve that
 * ran without visualization is RF_ERR_STATE (no colour map was computed). Synchronous. */
int rf_copy_flow(rf_context* ctx, float* u, float* v, float* uvrgb);
/* ReconFlow::copy3dToHost(X, Y, Z) (Reconflow.cu:65-70): camera-0 frame points. */
int rf_copy_3d(rf_context* ctx, float* X, float* Y, float* Z);
/* Device addresses of the current result fields (valid until the next solve). */
int rf_get_result_device(const rf_context* ctx, const float** u, const float** v, const float** X,
                         const float** Y, const float** Z);
/* CudaFlow::copySceneFlowToHost / copySceneFlowAndPointCloudToHost (Cudaflow.cu:345-362): 3 floats
 * (x, y, z) per pixel, stride x height; any pointer may be NULL. Synchronous. */
int rf_copy_scene_flow(rf_context* ctx, float* sceneflow, float* sceneflow_rgb, float* pcloud0, float* pcloud1);
/* ReconFlow::removeEdgeArtifacts3D(surfaceWidth, maxSurfaceArea) (Reconflow.cu:468-477). */
int rf_remove_edge_artifacts_3d(rf_context* ctx, int surface_width, float max_surface_area);

/* ---- host I/O formats ----
 * Middlebury .flo reader / writer: the prior-flow file of the reference demo
 * (cv::optflow::readOpticalFlow, test_reconflow/main.cpp:82-90). With u or v NULL, the reader
 * only returns the size; otherwise it fills stride-padded planes (zero padding, main.cpp:96). */
int rf_read_flo(const char* path, int* width, int* height, float* u, float* v, int stride);
int rf_write_flo(const char* path, const float* u, const float* v, int width, int height, int stride);
/* ASCII PLY of the reconstruction with the demo's mask (main.cpp:108-139): keeps points with
 * zmin < Z <= zmax, X <= xymax, -xymax < Y <= xymax, written as (X, -Y, -Z); rgb (BGR, 3 bytes
 * per pixel, stride-padded) may be NULL. */
int rf_write_ply(const char* path, const float* X, const float* Y, const float* Z, const uint8_t* rgb, int width,
                 int height, int stride, float zmin, float zmax, float xymax);

/* ---- kernel timing (no counterpart in the reference, which has no timers) ----
 * Arms HIP-event timing of one kernel kind at one pyramid level for the next solves that run
 * without a graph: events are recorded on the context's stream around every launch of that
 * kind (around each uninterrupted run of launches for the fused iteration kernel).
 * rf_profile_read returns the summed kernel time and the number of launches timed since the
 * last arm, and stops timing. */
#define RF_KID_FUSED_ITER 0
#define RF_KID_DATA 1
#define RF_KID_DUAL 2
#define RF_KID_UVPROJ 3
#define RF_KID_SOLVE3D 4
#define RF_KID_MEDIAN3D 5
#define RF_KID_CLEANUP 6
#define RF_KID_SOLVE_UC 7
#define RF_KID_UPDATE_SK 8
int rf_profile_arm(rf_context* ctx, int kernel_id, int level);
int rf_profile_read(rf_context* ctx, double* total_ms, int* launches);

/* The fused joint tiling (RF_JOINT_TILE_*) each pyramid level uses, for levels 0..n-1: the
 * fixed choice, or what the tuning solves picked (RF_JOINT_TILE_AUTO before it ran). Fused path
 * only (no counterpart in the reference). */
int rf_joint_tiling(rf_context* ctx, int* tiling, int n);
/* The same for the fused TV-L1 iteration (RF_FLOW_TILE_*). */
int rf_flow_tiling(rf_context* ctx, int* tiling, int n);

/* ---- batch API: many pairs of uneven sizes per launch ---- */
typedef struct rf_batch rf_batch;
/* Creates a batch of n pairs that share params (except width/height/input_type, taken
 * from the arrays). Pairs are solved together level by level: every kernel launch covers
 * the tiles of all pairs at that step. */
int rf_batch_create(rf_batch** out, const rf_params* params, int n, const int* widths, const int* heights);
int rf_batch_destroy(rf_batch* b);
int rf_batch_stride(const rf_batch* b, int i);
/* Device float gray images (stride_i x height_i) of pair i; enqueued on the batch stream. */
int rf_batch_set_images_device(rf_batch* b, int i, const float* d_i0, const float* d_i1);
int rf_batch_set_camera(rf_batch* b, int i, const double K0[9], const double K1[9], const double R1[9],
                        const double t1[3]);
int rf_batch_solve(rf_batch* b);
int rf_batch_synchronize(rf_batch* b);
int rf_batch_get_result_device(const rf_batch* b, int i, const float** u, const float** v, const float** X,
                               const float** Y, const float** Z);
void* rf_batch_get_stream(const rf_batch* b);
/* Warm-start input of pair i (no counterpart in the reference; the rule of rf_copy_warm_start):
 * device flow (stride_i x height_i floats each), read by every solve until replaced, for example
 * a copy of the previous frame's result in a stream; NULL clears. It may not point into the
 * batch's own fields (a solve clears them before reading it): RF_ERR_ARG. rf_batch_warm_start_in_use: 1 when the
 * last solve used it, 0 otherwise (synchronises the batch stream). */
int rf_batch_set_warm_start_device(rf_batch* b, int i, const float* d_u, const float* d_v);
int rf_batch_warm_start_in_use(rf_batch* b, int i);
/* The iteration tiling per level for levels 0..n-1: RF_JOINT_TILE_* for a joint batch,
 * RF_FLOW_TILE_* for a TV-L1 batch (AUTO until the tuning solves ran). */
int rf_batch_tiling(const rf_batch* b, int* tiling, int n);

/* ---- learned warm start (no counterpart in the reference) ----
 * A six-layer 3x3 convolutional encoder-decoder (32 channels, BF16 weights and activations on the
 * matrix cores, FP32 accumulation) that predicts the flow at pyramid level 3 (1/8 of the image) from
 * levels 3 and 4 of both frames; its prediction, scaled up to level 0, is a warm start for
 * rf_copy_warm_start / rf_batch_set_warm_start_device, which lets a solve start at a finer level.
 * The reference's own prior is a flow computed outside (the FlowNet input of test_reconflow/main.cpp:
 * 82-101, consumed by SolveReconDataL1Fn); this network is trained on the joint solver's results.
 * Parameters are FP32 masters, flat: per layer the weights [9][cout][cin] then the bias [cout]. */
typedef struct rf_warmnet rf_warmnet;
/* height x width: the level-3 size (even, at most 4096); max_batch: samples per forward / step. */
int rf_warmnet_create(rf_warmnet** out, int max_batch, int height, int width);
int rf_warmnet_destroy(rf_warmnet* net);
int rf_warmnet_param_count(const rf_warmnet* net);
/* Host FP32 parameters in / out (setting them resets the optimizer state). Synchronous. */
int rf_warmnet_set_params(rf_warmnet* net, const float* params, int count);
int rf_warmnet_get_params(rf_warmnet* net, float* params, int count);
/* The gradients of the last rf_warmnet_train_step (FP32, the parameter layout). Synchronous. */
int rf_warmnet_get_grads(rf_warmnet* net, float* grads, int count);
/* The network input of n samples from device level images: level 3 (height x width, stride s3,
 * nstride3 floats between samples) and level 4 (stride s4, nstride4) of frames 0 and 1. */
int rf_warmnet_input(rf_warmnet* net, const float* d_i0_l3, const float* d_i1_l3, int s3, long nstride3,
                     const float* d_i0_l4, const float* d_i1_l4, int s4, long nstride4, int n, void* stream);
/* Forward pass of the last input: the predicted level-3 flow (n x height x width x 2 floats, u v
 * interleaved, level-3 pixels) into d_flow (may be NULL). */
int rf_warmnet_forward(rf_warmnet* net, int n, float* d_flow, void* stream);
/* One training step on the last input against d_target (n x height x width x 2; non-finite entries
 * are ignored): Huber loss (delta 1 px, mean), gradients, and with lr > 0 an Adam step (beta 0.9,
 * 0.999). *loss (may be NULL) receives the loss; then the call synchronises the stream. */
int rf_warmnet_train_step(rf_warmnet* net, int n, const float* d_target, float lr, float* loss, void* stream);
/* The warm start of one pair: level-0 float images (w0 x h0, stride s0) -> pyramid (the solver's
 * downscale) -> the network's level-3 flow -> the solver's upscale to level 0 (d_u, d_v, stride s0).
 * The network's size must be the images' level 3. */
int rf_warmnet_warm_start(rf_warmnet* net, const float* d_i0, const float* d_i1, int w0, int h0, int s0, float* d_u,
                          float* d_v, void* stream);

/* ---- per-kernel entry points on device pointers (the public kernel wrappers of
 * Cudaflow.h:101-193 and Reconflow.h:86-158), enqueued on `stream` ---- */
int rf_k_rgb_to_gray(const uint8_t* rgb, float* gray, int w, int h, int s, void* stream);       /* rgbToGray.cu:80-86 */
int rf_k_cv8u_to_gray(const uint8_t* in, float* gray, int w, int h, int s, void* stream);      /* rgbToGray.cu:88-94 */
int rf_k_cv16u_to_gray(const uint16_t* in, float* gray, int w, int h, int s, void* stream);    /* rgbToGray.cu:96-102 */
int rf_k_cv32f_to_gray(const float* in, float* gray, int w, int h, int s, void* stream);       /* rgbToGray.cu:112-118 */
int rf_k_downscale(const float* src, int W, int H, int S, int w, int h, int s, float* out,
                   void* stream);                                                            /* downscale.cu:50-67 */
int rf_k_downscale_scaled(const float* src, int W, int H, int S, int w, int h, int s, float scale,
                          float* out, void* stream);                                         /* downscale.cu:69-85 */
int rf_k_upscale(const float* src, int w, int h, int s, int W, int H, int S, float scale, float* out,
                 void* stream);                                                              /* upscale.cu:42-59 */
int rf_k_warp_image(const float* src, int w, int h, int s, const float* u, const float* v, float* out,
                    void* stream);                                                           /* warpImage.cu:33-50 */
int rf_k_compute_derivatives(const float* I0, const float* I1, int w, int h, int s, float* Ix, float* Iy,
                             float* Iz, void* stream);                                       /* computeDerivatives.cu:60-84 */
int rf_k_solve_data_l1(const float* duhat0, const float* dvhat0, const float* pu1, const float* pu2,
                       const float* pv1, const float* pv2, const float* Ix, const float* Iy, const float* Iz,
                       int w, int h, int s, float lambda, float theta, float* duhat1, float* dvhat1,
                       void* stream);                                                        /* solveDataL1.cu:94-114 */
int rf_k_solve_smooth_dual_tv_global(const float* duhat, const float* dvhat, const float* pu1, const float* pu2,
                                     const float* pv1, const float* pv2, int w, int h, int s, float tau,
                                     float theta, float* pu1s, float* pu2s, float* pv1s, float* pv2s,
                                     void* stream);                                          /* solveSmoothDualTVGlobal.cu:80-96 */
int rf_k_median_filter(const float* u, const float* v, int w, int h, int s, float* ou, float* ov,
                       int kernel_size, void* stream);                                       /* medianFilter.cu:145-161 */
int rf_k_median_filter_3d(const float* X, const float* Y, const float* Z, int w, int h, int s, float* X1,
                          float* Y1, float* Z1, int kernel_size, void* stream);              /* medianFilter.cu:246-258 */
int rf_k_add(const float* a, const float* b, int count, float* sum, void* stream);           /* addKernel.cu:13-18 */
/* No counterpart in the reference: q[i] = n[i] / d by the sequence the fused joint kernels use for
 * their launch-uniform divisors (alpha_tv + alpha_proj, lambdaf + alpha_proj), for tests that
 * compare it with IEEE FP32 division. d must be finite and positive. */
int rf_k_uniform_divide(const float* n, int count, float d, float* q, void* stream);
int rf_k_solve_recon_data_l1(const float* u, const float* v, const float* duhat0, const float* dvhat0,
                             const float* uproj, const float* vproj, const float* sku, const float* skv,
                             const float* pu1, const float* pu2, const float* pv1, const float* pv2,
                             const float* Ix, const float* Iy, const float* Iz, int w, int h, int s,
                             float lambda, float alpha_tv, float alpha_proj, float* duhat1, float* dvhat1,
                             void* stream);                                                  /* solveReconDataL1.cu:105-131 */
int rf_k_solve_recon_data_l1_fn(const float* u, const float* v, const float* duhat0, const float* dvhat0,
                                const float* uproj, const float* vproj, const float* ufn, const float* vfn,
                                const float* sku, const float* skv, const float* pu1, const float* pu2,
                                const float* pv1, const float* pv2, const float* Ix, const float* Iy,
                                const float* Iz, int w, int h, int s, float lambda, float alpha_tv,
                                float alpha_proj, float alpha_fn, const float* fnmask, float* duhat1,
                                float* dvhat1, void* stream);                                /* solveReconDataL1.cu:243-274, 381-409; fnmask NULL = dense */
int rf_k_solve_uvproj(const float* u, const float* v, const float* du, const float* dv, const float* uc,
                      const float* vc, const float* sku, const float* skv, int w, int h, int s, float lambdaf,
                      float alpha_proj, float* uproj, float* vproj, void* stream);          /* solveUVProj.cu:25-45 */
int rf_k_solve_3d_minimal_area(const float* uproj, const float* vproj, const float* X0, const float* Y0,
                               const float* Z0, const double P[12], const double Q[12], float lambdaf,
                               float lambdams, int w, int h, int s, float* X, float* Y, float* Z,
                               void* stream);                                                /* solve3dMinimalArea.cu:197-241 */
int rf_k_cleanup_3d(const float* X0, const float* Y0, const float* Z0, const float* Xm, const float* Ym,
                    const float* Zm, int w, int h, int s, float* X1, float* Y1, float* Z1,
                    void* stream);                                                           /* solve3dMinimalArea.cu:299-313 */
int rf_k_solve_uc(const float* X, const float* Y, const float* Z, const double R[9], const double t[3],
                  const double K1[9], int w, int h, int s, float* uc, float* vc, void* stream); /* solveUcSk.cu:37-74 */
int rf_k_update_sk(const float* sku0, const float* skv0, const float* u, const float* v, const float* du,
                   const float* dv, const float* uproj, const float* vproj, int w, int h, int s, float* sku1,
                   float* skv1, void* stream);                                               /* solveUcSk.cu:99-115 */
int rf_k_flow_to_hsv(const float* u, const float* v, int w, int h, int s, float* uvrgb, float flow_scale,
                     void* stream);                                                          /* flowToHSV.cu:304-311 */
int rf_k_remove_edge_artifacts_3d(const float* X0, const float* Y0, const float* Z0, int surface_width,
                                  float max_surface_area, int w, int h, int s, float* X1, float* Y1, float* Z1,
                                  void* stream);                                             /* solve3dMinimalArea.cu:380-393 */
/* ComputeDeriv (computeDeriv.cu:6-55): 5-point gradient of one image. */
int rf_k_compute_deriv(const float* I, int w, int h, int s, float* Ix, float* Iy, void* stream);
/* ComputeDerivativesPatch (computeDerivativesPatch.cu:8-85): ComputeDerivatives summed over the
 * 5x5 patch. */
int rf_k_compute_derivatives_patch(const float* I0, const float* I1, int w, int h, int s, float* Ix, float* Iy,
                                   float* Iz, void* stream);
/* SolveDataCharbForTV (solveDataCharbForTV.cu:87-112) and, with Ixx..Iyt non-NULL,
 * SolveDataCharbForTVGrad (solveDataCharbForTVGrad.cu:105-135). */
int rf_k_solve_data_charb_for_tv(const float* du0, const float* dv0, const float* duhat0, const float* dvhat0,
                                 const float* pu1, const float* pu2, const float* pv1, const float* pv2,
                                 const float* Ix, const float* Iy, const float* It, const float* Ixx,
                                 const float* Ixy, const float* Ixt, const float* Iyx, const float* Iyy,
                                 const float* Iyt, int w, int h, int s, float lambda, float lambdagrad, float theta,
                                 float* du1, float* dv1, float* duhat1, float* dvhat1, void* stream);
/* SolveDataL1Inpaint (solveDataL1Inpaint.cu:100-125); mask1 is not read. */
int rf_k_solve_data_l1_inpaint(const float* duhat0, const float* dvhat0, const float* mask0, const float* mask1,
                               const float* pu1, const float* pu2, const float* pv1, const float* pv2,
                               const float* Ix, const float* Iy, const float* It, int w, int h, int s, float lambda,
                               float theta, float* duhat1, float* dvhat1, void* stream);
/* Cv16uTo32f (rgbToGray.cu:51-64, 104-110). */
int rf_k_cv16u_to_32f(const uint16_t* in, float* out, int w, int h, int s, void* stream);
/* SolveSceneFlow_ (solveSceneFlow.cu:95-160): 3 floats per pixel; pcloud0 / pcloud1 may be NULL. */
int rf_k_solve_scene_flow(const float* u, const float* v, const float* depth0, const float* depth1, float fx,
                          float fy, float cx, float cy, int w, int h, int s, float* pcloud0, float* pcloud1,
                          float* sceneflow, void* stream);
/* SceneFlowToHSV (flowToHSV.cu:315-321). */
int rf_k_scene_flow_to_hsv(const float* sceneflow, int w, int h, int s, float* rgb, float flow_scale, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RFHIP_H */
