// Solver context and the C ABI (rfhip.h).
//
// The plain schedule mirrors the reference host loops step by step, including the pointer
// swaps (Cudaflow.cu:403-621, Reconflow.cu:147-466), so every buffer that a kernel reads out
// of its valid region holds what the reference's buffer would hold. All device memory is one
// slab allocated at initialize and zeroed once; a solve allocates nothing. Every launch of a
// solve goes to the context's stream; with use_graph the whole static schedule is captured
// once into a hipGraph and replayed (camera data lives in device memory, so a new pose
// needs no re-capture).

#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>

#include "rf_batch.hpp"
#include "rf_kernels.hpp"
#include "rf_solver.hpp"

namespace rf {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

static int align_up(int n, int m) {
    int mod = n % m;
    return mod ? n + m - mod : n;
}

Solver::Solver(int bw, int bh, int align) : bw_(bw), bh_(bh), align_(align > 0 ? align : 32) {}

Solver::~Solver() { release(); }

int Solver::check(hipError_t e, const char* what) {
    if (e == hipSuccess) return RF_OK;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return RF_ERR_HIP;
}

// Teardown: a failing destroy or free leaves nothing to recover, so its status is not checked.
void Solver::drop_graphs() {
    for (auto& g : gexec_)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

int Solver::release() {
    batch_.reset();
    for (auto& e : prof_.ev) (void)hipEventDestroy(e);
    prof_.ev.clear();
    prof_.armed = false;
    drop_graphs();
    if (ws_) {
        (void)hipFree(ws_);
        ws_ = nullptr;
        ws_bad_ = nullptr;
    }
    ws_set_ = false;
    uvrgb_valid_ = false;
    if (slab_) {
        (void)hipFree(slab_);
        slab_ = nullptr;
    }
    if (dparams_) {
        (void)hipFree(dparams_);
        dparams_ = nullptr;
    }
    if (st_) {
        (void)hipStreamDestroy(st_);
        st_ = nullptr;
    }
    inited_ = false;
    return RF_OK;
}

// Argument checks shared by rf_initialize and rf_solve_flow_noalloc.
int Solver::check_params(const rf_params& p, const char* who) {
    if (p.width <= 0 || p.height <= 0 || p.n_levels <= 0 || p.n_levels > kMaxLevels || !(p.scale > 1.0f) ||
        p.n_warp_iters < 0 || p.n_solver_iters < 0) {
        set_error(std::string(who) + ": invalid size, level count or scale");
        return RF_ERR_ARG;
    }
    if (p.width >= kMaxSide || p.height >= kMaxSide) {
        set_error(std::string(who) + ": width and height must be below " + std::to_string(kMaxSide));
        return RF_ERR_ARG;
    }
    return RF_OK;
}

// Pyramid geometry (Cudaflow.cu:146-171): w_{l+1} = (int)(w_l / fScale) in float.
int Solver::set_geometry(const rf_params& p, int align, const char* who) {
    Geometry g;
    g.L = p.n_levels;
    int nh = p.height, nw = p.width, ns = align_up(p.width, align);
    for (int l = 0; l < p.n_levels; ++l) {
        if (nw < 1 || nh < 1) {
            set_error(std::string(who) + ": pyramid level " + std::to_string(l) + " has no pixels");
            return RF_ERR_ARG;
        }
        if (l > 0 && !(resample_fits(g.W[l - 1], nw) && resample_fits(g.H[l - 1], nh))) {
            set_error(std::string(who) + ": levels " + std::to_string(l - 1) + " and " + std::to_string(l) +
                      " are too large for 32-bit texel coordinates at this scale");
            return RF_ERR_ARG;
        }
        g.W[l] = nw;
        g.H[l] = nh;
        g.S[l] = ns;
        nh = (int)((float)nh / p.scale);
        nw = (int)((float)nw / p.scale);
        ns = align_up(nw, align);
    }
    g.n = (long)g.S[0] * g.H[0];
    g.bufn = ((g.n + g.S[0]) + 63) / 64 * 64;
    g_ = g;
    return RF_OK;
}

// CudaFlow::initialize (Cudaflow.cu:99-269) and ReconFlow::_initializeR (Reconflow.cu:27-63).
int Solver::initialize(const rf_params& p) {
    if (int rc = check_params(p, "rf_initialize")) return rc;
    if (p.input_type < RF_INPUT_32F || p.input_type > RF_INPUT_16U) {
        set_error("rf_initialize: unknown input type");
        return RF_ERR_ARG;
    }
    if (p.path != RF_PATH_PLAIN && p.path != RF_PATH_FUSED) {
        set_error("rf_initialize: unknown path");
        return RF_ERR_ARG;
    }
    if (p.joint_tile < RF_JOINT_TILE_AUTO || p.joint_tile > RF_JOINT_TILE_16x16_K3 ||
        p.flow_tile < RF_FLOW_TILE_AUTO || p.flow_tile > RF_FLOW_TILE_32x16_K6) {
        set_error("rf_initialize: unknown joint or TV-L1 tiling");
        return RF_ERR_ARG;
    }
    if (p.storage < RF_STORAGE_F32 || p.storage > RF_STORAGE_F16_DERIV_DUALS) {
        set_error("rf_initialize: unknown storage");
        return RF_ERR_ARG;
    }
    if (p.path == RF_PATH_FUSED && p.method != RF_METHOD_TVL1 && p.method != RF_METHODR_TVL1_MS &&
        p.method != RF_METHODR_TVL1_MS_FN && p.method != RF_METHODR_TVL1_MS_FNSPARSE) {
        set_error("rf_initialize: the fused path runs METHOD_TVL1 and the joint methods; use the plain path");
        return RF_ERR_UNSUPPORTED;
    }
    if (p.storage != RF_STORAGE_F32 && p.path != RF_PATH_FUSED) {
        set_error("rf_initialize: FP16 storage is an option of the fused path");
        return RF_ERR_UNSUPPORTED;
    }
    release();
    g_ = Geometry();
    const int align = p.stride_alignment > 0 ? p.stride_alignment : align_;
    if (int rc = set_geometry(p, align, "rf_initialize")) {
        g_ = Geometry();
        return rc;
    }
    p_ = p;
    align_ = align;
    if (p.path == RF_PATH_FUSED) {
        batch_.reset(new Batch());
        int rc = batch_->create(p, 1, &p.width, &p.height);
        if (rc) {
            batch_.reset();
            g_ = Geometry();
            return rc;
        }
        inited_ = true;
        prior_ = false;
        camera_ = false;
        return RF_OK;
    }
    if (check(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "hipStreamCreate")) return RF_ERR_HIP;

    // field buffers: 2 per level (image pyramids), the fixed set, the uvrgb triple, the raw inputs,
    // and the buffers of the alternate data terms the method uses (Cudaflow.cu:181-266)
    const int L = g_.L, m = p.method;
    const bool cg = (m == RF_METHOD_TVCHARBGRAD || m == RF_METHOD_TVCHARBGRAD_FN);
    const bool charb = cg || m == RF_METHOD_TVCHARB;
    const bool inpaint = (m == RF_METHOD_TVL1_INPAINTING), scene = (m == RF_SCENEFLOW_KINECT_TVL1);
    const int nfixed = 43;
    long nf = (long)(2 * L + nfixed + 3);
    if (cg) nf += 4L * L + 8;
    if (charb) nf += 4;
    if (inpaint) nf += 2L * L + 2;
    if (scene) nf += 2 + 4 * 3;
    const long inbytes = g_.n * 4;  // raw input (up to 4 B per pixel)
    const long infl = (inbytes / 4 + 63) / 64 * 64;
    const long nin = 2 + (scene ? 2 : 0);
    slab_bytes_ = (size_t)(nf * ((g_.bufn + 63) / 64 * 64) + nin * infl) * sizeof(float) + 256 * 64;
    if (hipMalloc(&slab_, slab_bytes_) != hipSuccess) {
        slab_ = nullptr;
        set_error("rf_initialize: device allocation of " + std::to_string(slab_bytes_) + " bytes failed");
        release();
        return RF_ERR_ALLOC;
    }
    if (check(hipMemset(slab_, 0, slab_bytes_), "hipMemset")) return RF_ERR_HIP;
    if (hipMalloc((void**)&dparams_, sizeof(DevParams)) != hipSuccess) {
        dparams_ = nullptr;
        set_error("rf_initialize: parameter block allocation failed");
        release();
        return RF_ERR_ALLOC;
    }
    float* cursor = (float*)slab_;
    auto take = [&](long nfl) {
        float* r = cursor;
        cursor += (nfl + 63) / 64 * 64;
        return r;
    };
    Fields f{};
    for (int l = 0; l < L; ++l) {
        f.pI0[l] = take(g_.bufn);
        f.pI1[l] = take(g_.bufn);
    }
    float** fixed[] = {&f.i1warp, &f.dumed, &f.dvmed, &f.dumeds, &f.dvmeds, &f.pu1,  &f.pu2,  &f.pv1,  &f.pv2,
                       &f.pu1s,   &f.pu2s,  &f.pv1s,  &f.pv2s,   &f.Ix,     &f.Iy,   &f.Iz,   &f.u,    &f.v,
                       &f.us,     &f.vs,    &f.X,     &f.Y,      &f.Z,      &f.Xs,   &f.Ys,   &f.Zs,   &f.Xm,
                       &f.Ym,     &f.Zm,    &f.uproj, &f.vproj,  &f.uc,     &f.vc,   &f.sku,  &f.skv,  &f.skus,
                       &f.skvs,   &f.ufn,   &f.vfn,   &f.ufn_l,  &f.vfn_l,  &f.mask, &f.mask_l};
    static_assert(sizeof fixed / sizeof fixed[0] == 43, "fixed field count");
    for (float** pp : fixed) *pp = take(g_.bufn);
    f.uvrgb = take(3 * g_.bufn);
    f.in0 = take(infl);
    f.in1 = take(infl);
    if (cg) {
        for (int l = 0; l < L; ++l) {
            f.pIx0[l] = take(g_.bufn);
            f.pIx1[l] = take(g_.bufn);
            f.pIy0[l] = take(g_.bufn);
            f.pIy1[l] = take(g_.bufn);
        }
        for (float** pp : {&f.ix1warp, &f.iy1warp, &f.Ixx, &f.Ixy, &f.Ixz, &f.Iyx, &f.Iyy, &f.Iyz})
            *pp = take(g_.bufn);
    }
    if (charb)
        for (float** pp : {&f.du, &f.dv, &f.dus, &f.dvs}) *pp = take(g_.bufn);
    if (inpaint) {
        for (int l = 0; l < L; ++l) {
            f.pM0[l] = take(g_.bufn);
            f.pM1[l] = take(g_.bufn);
        }
        f.mask0 = take(g_.bufn);
        f.mask1 = take(g_.bufn);
    }
    if (scene) {
        f.depth032f = take(g_.bufn);
        f.depth132f = take(g_.bufn);
        f.sceneflow = take(3 * g_.bufn);
        f.sceneflowrgb = take(3 * g_.bufn);
        f.pcloud0 = take(3 * g_.bufn);
        f.pcloud1 = take(3 * g_.bufn);
        f.depth0_16 = take(infl);
        f.depth1_16 = take(infl);
    }
    if ((char*)cursor > (char*)slab_ + slab_bytes_) {
        set_error("rf_initialize: internal slab accounting error");
        release();
        return RF_ERR_STATE;
    }
    init_ = cur_ = f;
    gfinal_[0] = gfinal_[1] = gfinal_[2] = f;
    inited_ = true;
    prior_ = false;
    camera_ = false;
    masks_ = depth_ = scene_ = false;
    return RF_OK;
}

int Solver::init_large_displacement() {
    if (!inited_) {
        set_error("rf_initialize_large_displacement: not initialized");
        return RF_ERR_STATE;
    }
    prior_ = true;
    if (batch_) batch_->enable_prior();
    return RF_OK;
}

// ReconFlow::setCameraMatrices (Reconflow.cu:78-116): levels >= 1 keep fx, fy, cx, cy divided
// by fScale and leave K[2][2] = 0.
int Solver::set_camera(const double* K0, const double* K1) {
    if (!inited_ || !K0 || !K1) {
        set_error("rf_set_camera: not initialized or null matrix");
        return inited_ ? RF_ERR_ARG : RF_ERR_STATE;
    }
    memcpy(K0raw_, K0, sizeof K0raw_);
    memcpy(K1raw_, K1, sizeof K1raw_);
    K0_.assign(g_.L, std::vector<double>(9, 0.0));
    K1_.assign(g_.L, std::vector<double>(9, 0.0));
    for (int i = 0; i < 9; ++i) {
        K0_[0][i] = K0[i];
        K1_[0][i] = K1[i];
    }
    for (int l = 1; l < g_.L; ++l)
        for (auto* K : {&K0_, &K1_}) {
            (*K)[l][0] = (*K)[l - 1][0] / (double)p_.scale;
            (*K)[l][2] = (*K)[l - 1][2] / (double)p_.scale;
            (*K)[l][4] = (*K)[l - 1][4] / (double)p_.scale;
            (*K)[l][5] = (*K)[l - 1][5] / (double)p_.scale;
            if (p_.fix_coarse_k) (*K)[l][8] = 1.0;
        }
    camera_ = true;
    return RF_OK;
}

static size_t input_bytes_per_px(int t) {
    switch (t) {
        case RF_INPUT_8UC3: return 3;
        case RF_INPUT_8U: return 1;
        case RF_INPUT_16U: return 2;
        default: return 4;
    }
}

// CudaFlow::copyImagesToDevice (Cudaflow.cu:288-311): stride x height pixels each.
int Solver::copy_images(const void* i0, const void* i1) {
    if (!inited_ || !i0 || !i1) {
        set_error("rf_copy_images: not initialized or null image");
        return inited_ ? RF_ERR_ARG : RF_ERR_STATE;
    }
    if (batch_) return batch_->set_images_host(0, i0, i1);
    size_t bytes = (size_t)g_.n * input_bytes_per_px(p_.input_type);
    if (check(hipMemcpyAsync(init_.in0, i0, bytes, hipMemcpyHostToDevice, st_), "copy image 0")) return RF_ERR_HIP;
    if (check(hipMemcpyAsync(init_.in1, i1, bytes, hipMemcpyHostToDevice, st_), "copy image 1")) return RF_ERR_HIP;
    return check(hipStreamSynchronize(st_), "copy images");
}

int Solver::copy_images_device(const float* i0, const float* i1) {
    if (!inited_ || !i0 || !i1) {
        set_error("rf_copy_images_device: not initialized or null image");
        return inited_ ? RF_ERR_ARG : RF_ERR_STATE;
    }
    if (p_.input_type != RF_INPUT_32F) {
        set_error("rf_copy_images_device: the context expects a non-float input type");
        return RF_ERR_ARG;
    }
    if (batch_) return batch_->set_images_device(0, i0, i1);
    size_t bytes = (size_t)g_.n * 4;
    if (check(hipMemcpyAsync(init_.in0, i0, bytes, hipMemcpyDeviceToDevice, st_), "copy image 0")) return RF_ERR_HIP;
    return check(hipMemcpyAsync(init_.in1, i1, bytes, hipMemcpyDeviceToDevice, st_), "copy image 1");
}

// CudaFlow::copyInitialOpticalFlowToDevice / copySparseOpticalFlowToDevice (Cudaflow.cu:326-343).
int Solver::copy_prior(const float* u, const float* v, const float* mask, bool sparse) {
    if (!inited_ || !prior_ || !u || !v) {
        set_error("rf_copy_*_flow: call rf_initialize_large_displacement first and pass u and v");
        return RF_ERR_STATE;
    }
    if (batch_) {
        sparse_ = sparse;
        return batch_->set_prior(0, u, v, mask, sparse);
    }
    size_t bytes = (size_t)g_.n * 4;
    if (check(hipMemcpyAsync(init_.ufn, u, bytes, hipMemcpyHostToDevice, st_), "copy prior u")) return RF_ERR_HIP;
    if (check(hipMemcpyAsync(init_.vfn, v, bytes, hipMemcpyHostToDevice, st_), "copy prior v")) return RF_ERR_HIP;
    if (mask) {
        if (check(hipMemcpyAsync(init_.mask, mask, bytes, hipMemcpyHostToDevice, st_), "copy mask")) return RF_ERR_HIP;
    } else {
        std::vector<float> ones(g_.n, 1.0f);
        if (check(hipMemcpyAsync(init_.mask, ones.data(), bytes, hipMemcpyHostToDevice, st_), "copy mask"))
            return RF_ERR_HIP;
        if (check(hipStreamSynchronize(st_), "copy mask")) return RF_ERR_HIP;
    }
    sparse_ = sparse;
    return check(hipStreamSynchronize(st_), "copy prior");
}

// Warm start (no counterpart in the reference): the flow that replaces the zero initial u, v at
// the coarsest level of the next solves (see launch_warm_start). The schedule gains three launches
// when one is set, so a captured schedule is captured again.
int Solver::copy_warm_start(const float* u, const float* v) {
    if (!inited_) {
        set_error("rf_copy_warm_start: not initialized");
        return RF_ERR_STATE;
    }
    const bool set = u && v;
    if (set != ws_set_) drop_graphs();
    if (!set) {
        ws_set_ = false;
        return batch_ ? batch_->set_warm_start(0, nullptr, nullptr) : RF_OK;
    }
    const size_t bytes = (size_t)g_.n * 4;
    if (!ws_) {
        if (check(hipMalloc(&ws_, 2 * bytes + 256), "rf_copy_warm_start: allocation")) return RF_ERR_HIP;
        ws_bad_ = reinterpret_cast<int*>(reinterpret_cast<char*>(ws_) + 2 * bytes);
        const int unused = 1;  // until a solve applies the rule
        if (check(hipMemcpy(ws_bad_, &unused, sizeof unused, hipMemcpyHostToDevice), "rf_copy_warm_start"))
            return RF_ERR_HIP;
    }
    if (check(hipMemcpyAsync(ws_, u, bytes, hipMemcpyHostToDevice, st_), "rf_copy_warm_start u")) return RF_ERR_HIP;
    if (check(hipMemcpyAsync(ws_ + g_.n, v, bytes, hipMemcpyHostToDevice, st_), "rf_copy_warm_start v"))
        return RF_ERR_HIP;
    if (check(hipStreamSynchronize(st_), "rf_copy_warm_start")) return RF_ERR_HIP;
    ws_set_ = true;
    return batch_ ? batch_->set_warm_start(0, ws_, ws_ + g_.n) : RF_OK;
}

// 1 when a warm start is set and the last solve found it usable, else 0.
int Solver::warm_start_in_use() {
    if (!inited_ || !ws_set_) return 0;
    if (batch_) return batch_->warm_start_in_use(0);
    int bad = 1;
    if (check(hipStreamSynchronize(st_), "rf_warm_start_in_use")) return RF_ERR_HIP;
    if (check(hipMemcpy(&bad, ws_bad_, sizeof bad, hipMemcpyDeviceToHost), "rf_warm_start_in_use")) return RF_ERR_HIP;
    return bad ? 0 : 1;
}

#define RF_TRY(x)                         \
    do {                                  \
        hipError_t e__ = (x);             \
        if (e__ != hipSuccess) return e__; \
    } while (0)

// Launch with optional event timing of kernel kind KID at the current level.
#define RF_PROF(KID, x)                      \
    do {                                     \
        prof_.begin(st_, KID, level);        \
        RF_TRY(x);                           \
        prof_.end(st_, KID, level, 1);       \
    } while (0)

// Gray conversion of both inputs and the image pyramids (Cudaflow.cu:405-443).
hipError_t Solver::enqueue_gray_pyramid(Fields& f) {
    const int w0 = g_.W[0], h0 = g_.H[0], s0 = g_.S[0];
    switch (p_.input_type) {
        case RF_INPUT_8UC3:
            RF_TRY(launch_gray_rgb((const uint8_t*)f.in0, f.pI0[0], w0, h0, s0, st_));
            RF_TRY(launch_gray_rgb((const uint8_t*)f.in1, f.pI1[0], w0, h0, s0, st_));
            break;
        case RF_INPUT_8U:
            RF_TRY(launch_gray_u8((const uint8_t*)f.in0, f.pI0[0], w0, h0, s0, st_));
            RF_TRY(launch_gray_u8((const uint8_t*)f.in1, f.pI1[0], w0, h0, s0, st_));
            break;
        case RF_INPUT_16U:
            RF_TRY(launch_gray_u16((const uint16_t*)f.in0, f.pI0[0], w0, h0, s0, st_));
            RF_TRY(launch_gray_u16((const uint16_t*)f.in1, f.pI1[0], w0, h0, s0, st_));
            break;
        default:
            RF_TRY(launch_gray_f32((const float*)f.in0, f.pI0[0], w0, h0, s0, st_));
            RF_TRY(launch_gray_f32((const float*)f.in1, f.pI1[0], w0, h0, s0, st_));
    }
    const bool cg = f.pIx0[0] != nullptr;
    if (cg) {  // gradient images of both frames (Cudaflow.cu:426-429)
        RF_TRY(launch_deriv_single(f.pI0[0], w0, h0, s0, f.pIx0[0], f.pIy0[0], st_));
        RF_TRY(launch_deriv_single(f.pI1[0], w0, h0, s0, f.pIx1[0], f.pIy1[0], st_));
    }
    if (masked_run_) {  // masks (Inpainting.cu:36-37)
        RF_TRY(launch_gray_f32(f.mask0, f.pM0[0], w0, h0, s0, st_));
        RF_TRY(launch_gray_f32(f.mask1, f.pM1[0], w0, h0, s0, st_));
    }
    for (int l = 1; l < g_.L; ++l) {
        auto down = [&](float* const* py) {
            return launch_downscale(py[l - 1], g_.W[l - 1], g_.H[l - 1], g_.S[l - 1], g_.W[l], g_.H[l], g_.S[l], false,
                                    1.0f, py[l], st_);
        };
        RF_TRY(down(f.pI0));
        RF_TRY(down(f.pI1));
        if (cg) {
            RF_TRY(down(f.pIx0));
            RF_TRY(down(f.pIx1));
            RF_TRY(down(f.pIy0));
            RF_TRY(down(f.pIy1));
        }
        if (masked_run_) {
            RF_TRY(down(f.pM0));
            RF_TRY(down(f.pM1));
        }
    }
    return hipSuccess;
}

// CudaFlow::_solveOpticalFlow (Cudaflow.cu:403-621) with its data-term branches (:489-557), and
// CudaFlow::_solveInpainting (Inpainting.cu:13-152) when masked_run_ is set.
hipError_t Solver::enqueue_flow_plain(Fields& f) {
    RF_TRY(enqueue_gray_pyramid(f));
    {
        float* z[] = {f.u, f.v};
        RF_TRY(launch_fill_zero(z, 2, g_.n, st_));
    }
    if (ws_set_) {
        const int c = g_.L - 1;
        RF_TRY(launch_warm_start(ws_, ws_ + g_.n, g_.W[0], g_.H[0], g_.S[0], g_.W[c], g_.H[c], g_.S[c], ws_bad_, f.u,
                                 f.v, st_));
    }
    const int m = p_.method;
    const bool cg = (m == RF_METHOD_TVCHARBGRAD || m == RF_METHOD_TVCHARBGRAD_FN);
    const bool charb = (m == RF_METHOD_TVCHARB);
    for (int level = g_.L - 1; level >= 0; --level) {
        const int w = g_.W[level], h = g_.H[level], s = g_.S[level];
        for (int wi = 0; wi < p_.n_warp_iters; ++wi) {
            float* z[] = {f.dumed, f.dvmed, f.dumeds, f.dvmeds, f.pu1, f.pu2, f.pv1, f.pv2};
            RF_TRY(launch_fill_zero(z, 8, g_.n, st_));
            if (f.du) {
                float* zd[] = {f.du, f.dv, f.dus, f.dvs};
                RF_TRY(launch_fill_zero(zd, 4, g_.n, st_));
            }
            RF_TRY(launch_warp(f.pI1[level], w, h, s, f.u, f.v, f.i1warp, st_));
            if (cg) {
                RF_TRY(launch_warp(f.pIx1[level], w, h, s, f.u, f.v, f.ix1warp, st_));
                RF_TRY(launch_warp(f.pIy1[level], w, h, s, f.u, f.v, f.iy1warp, st_));
            }
            if (m == RF_METHOD_TVL1PATCH && !masked_run_)
                RF_TRY(launch_derivatives_patch(f.pI0[level], f.i1warp, w, h, s, f.Ix, f.Iy, f.Iz, st_));
            else
                RF_TRY(launch_derivatives(f.pI0[level], f.i1warp, w, h, s, f.Ix, f.Iy, f.Iz, st_));
            if (cg) {
                RF_TRY(launch_derivatives(f.pIx0[level], f.ix1warp, w, h, s, f.Ixx, f.Ixy, f.Ixz, st_));
                RF_TRY(launch_derivatives(f.pIy0[level], f.iy1warp, w, h, s, f.Iyx, f.Iyy, f.Iyz, st_));
            }
            for (int it = 0; it < p_.n_solver_iters; ++it) {
                if (masked_run_) {
                    RF_PROF(RF_KID_DATA, launch_data_l1_inpaint(f.dumed, f.dvmed, f.pM0[level], f.pu1, f.pu2, f.pv1,
                                                                f.pv2, f.Ix, f.Iy, f.Iz, w, h, s, p_.lambda,
                                                                p_.theta, f.dumeds, f.dvmeds, st_));
                } else if (cg || charb) {
                    RF_PROF(RF_KID_DATA, launch_data_charb(f.du, f.dv, f.dumed, f.dvmed, f.pu1, f.pu2, f.pv1, f.pv2,
                                                           f.Ix, f.Iy, f.Iz, cg ? f.Ixx : nullptr, f.Ixy, f.Ixz,
                                                           f.Iyx, f.Iyy, f.Iyz, w, h, s, p_.lambda, p_.lambdagrad,
                                                           p_.theta, f.dus, f.dvs, f.dumeds, f.dvmeds, st_));
                    std::swap(f.du, f.dus);
                    std::swap(f.dv, f.dvs);
                } else {
                    RF_PROF(RF_KID_DATA, launch_data_l1(f.dumed, f.dvmed, f.pu1, f.pu2, f.pv1, f.pv2, f.Ix, f.Iy,
                                                        f.Iz, w, h, s, p_.lambda, p_.theta, f.dumeds, f.dvmeds, st_));
                }
                std::swap(f.dumed, f.dumeds);
                std::swap(f.dvmed, f.dvmeds);
                RF_PROF(RF_KID_DUAL, launch_dual_tv(f.dumed, f.dvmed, f.pu1, f.pu2, f.pv1, f.pv2, w, h, s, p_.tau,
                                                    p_.theta, f.pu1s, f.pu2s, f.pv1s, f.pv2s, st_));
                std::swap(f.pu1, f.pu1s);
                std::swap(f.pu2, f.pu2s);
                std::swap(f.pv1, f.pv1s);
                std::swap(f.pv2, f.pv2s);
            }
            RF_TRY(launch_median5(f.dumed, f.dvmed, w, h, s, f.dumeds, f.dvmeds, st_));
            std::swap(f.dumed, f.dumeds);
            std::swap(f.dvmed, f.dvmeds);
            RF_TRY(launch_add(f.u, f.dumed, (long)h * s, f.u, st_));
            RF_TRY(launch_add(f.v, f.dvmed, (long)h * s, f.v, st_));
        }
        if (level > 0) {
            RF_TRY(launch_upscale(f.u, w, h, s, g_.W[level - 1], g_.H[level - 1], g_.S[level - 1], p_.scale, f.us, st_));
            RF_TRY(launch_upscale(f.v, w, h, s, g_.W[level - 1], g_.H[level - 1], g_.S[level - 1], p_.scale, f.vs, st_));
            std::swap(f.u, f.us);
            std::swap(f.v, f.vs);
        }
    }
    if (p_.with_visualization)
        RF_TRY(launch_flow_to_hsv(f.u, f.v, g_.W[0], g_.H[0], g_.S[0], f.uvrgb, flow_scale_, st_));
    return hipSuccess;
}

// ReconFlow::solveReconFlow, METHODR_TVL1_MS and the prior-flow variants (Reconflow.cu:147-466).
hipError_t Solver::enqueue_recon_plain(Fields& f) {
    RF_TRY(enqueue_gray_pyramid(f));
    {
        float* z[] = {f.u, f.v, f.uproj, f.vproj, f.uc, f.vc, f.X, f.Y, f.Z, f.sku, f.skv};
        RF_TRY(launch_fill_zero(z, 11, g_.n, st_));
    }
    if (ws_set_) {
        const int c = g_.L - 1;
        RF_TRY(launch_warm_start(ws_, ws_ + g_.n, g_.W[0], g_.H[0], g_.S[0], g_.W[c], g_.H[c], g_.S[c], ws_bad_, f.u,
                                 f.v, st_));
    }
    const bool fn = prior_ && (p_.method == RF_METHODR_TVL1_MS_FN || p_.method == RF_METHODR_TVL1_MS_FNSPARSE);
    const float thetaTv = 1.0f / p_.alpha_tv;  // Reconflow.cu:32
    for (int level = g_.L - 1; level >= 0; --level) {
        const int w = g_.W[level], h = g_.H[level], s = g_.S[level];
        if (fn) {
            float sc = (float)g_.W[level] / (float)g_.W[0];
            RF_TRY(launch_downscale(f.ufn, g_.W[0], g_.H[0], g_.S[0], w, h, s, true, sc, f.ufn_l, st_));
            RF_TRY(launch_downscale(f.vfn, g_.W[0], g_.H[0], g_.S[0], w, h, s, true, sc, f.vfn_l, st_));
            RF_TRY(launch_downscale(f.mask, g_.W[0], g_.H[0], g_.S[0], w, h, s, false, 1.0f, f.mask_l, st_));
        }
        const PQ* dpq = &dparams_->pq[level];
        const PQd* dpqd = &dparams_->pqd[level];
        const UcParams* duc = &dparams_->uc[level];
        for (int wi = 0; wi < p_.n_warp_iters; ++wi) {
            float* z[] = {f.dumed, f.dvmed, f.dumeds, f.dvmeds, f.pu1, f.pu2, f.pv1, f.pv2};
            RF_TRY(launch_fill_zero(z, 8, g_.n, st_));
            RF_TRY(launch_warp(f.pI1[level], w, h, s, f.u, f.v, f.i1warp, st_));
            RF_TRY(launch_derivatives(f.pI0[level], f.i1warp, w, h, s, f.Ix, f.Iy, f.Iz, st_));
            for (int it = 0; it < p_.n_solver_iters; ++it) {
                RF_PROF(RF_KID_DATA, launch_recon_data_l1(f.u, f.v, f.dumed, f.dvmed, f.uproj, f.vproj, f.sku, f.skv, f.pu1, f.pu2,
                                            f.pv1, f.pv2, f.Ix, f.Iy, f.Iz, w, h, s, p_.lambda, p_.alpha_tv,
                                            p_.alpha_proj, fn ? f.ufn_l : nullptr, fn ? f.vfn_l : nullptr,
                                            p_.alpha_fn, (fn && sparse_) ? f.mask_l : nullptr, f.dumeds, f.dvmeds,
                                            p_.tri_f64 != 0, st_));
                std::swap(f.dumed, f.dumeds);
                std::swap(f.dvmed, f.dvmeds);
                RF_PROF(RF_KID_DUAL, launch_dual_tv(f.dumed, f.dvmed, f.pu1, f.pu2, f.pv1, f.pv2, w, h, s, p_.tau,
                                                    thetaTv, f.pu1s, f.pu2s, f.pv1s, f.pv2s, st_));
                std::swap(f.pu1, f.pu1s);
                std::swap(f.pu2, f.pu2s);
                std::swap(f.pv1, f.pv1s);
                std::swap(f.pv2, f.pv2s);
                RF_PROF(RF_KID_UVPROJ, launch_uvproj(f.u, f.v, f.dumed, f.dvmed, f.uc, f.vc, f.sku, f.skv, w, h, s, p_.lambdaf,
                                     p_.alpha_proj, f.uproj, f.vproj, p_.tri_f64 != 0, st_));
                if (p_.tri_f64)
                    RF_PROF(RF_KID_SOLVE3D, launch_solve3d_f64(f.uproj, f.vproj, f.X, f.Y, f.Z, PQd{}, dpqd, p_.lambdaf, p_.lambdams, w,
                                              h, s, f.Xs, f.Ys, f.Zs, st_));
                else
                    RF_PROF(RF_KID_SOLVE3D, launch_solve3d(f.uproj, f.vproj, f.X, f.Y, f.Z, PQ{}, dpq, p_.lambdaf, p_.lambdams, w, h,
                                          s, f.Xs, f.Ys, f.Zs, st_));
                std::swap(f.X, f.Xs);
                std::swap(f.Y, f.Ys);
                std::swap(f.Z, f.Zs);
                RF_PROF(RF_KID_MEDIAN3D, launch_median3d(f.X, f.Y, f.Z, w, h, s, f.Xm, f.Ym, f.Zm, st_));
                RF_PROF(RF_KID_CLEANUP, launch_cleanup3d(f.X, f.Y, f.Z, f.Xm, f.Ym, f.Zm, w, h, s, f.Xs, f.Ys, f.Zs, st_));
                std::swap(f.X, f.Xs);
                std::swap(f.Y, f.Ys);
                std::swap(f.Z, f.Zs);
                RF_PROF(RF_KID_SOLVE_UC, launch_solve_uc(f.X, f.Y, f.Z, UcParams{}, duc, w, h, s, f.uc, f.vc, st_));
                RF_PROF(RF_KID_UPDATE_SK, launch_update_sk(f.sku, f.skv, f.u, f.v, f.dumed, f.dvmed, f.uproj, f.vproj, w, h, s, f.skus,
                                        f.skvs, st_));
                std::swap(f.sku, f.skus);
                std::swap(f.skv, f.skvs);
            }
            RF_TRY(launch_median5(f.dumed, f.dvmed, w, h, s, f.dumeds, f.dvmeds, st_));
            std::swap(f.dumed, f.dumeds);
            std::swap(f.dvmed, f.dvmeds);
            RF_TRY(launch_add(f.u, f.dumed, (long)h * s, f.u, st_));
            RF_TRY(launch_add(f.v, f.dvmed, (long)h * s, f.v, st_));
            RF_TRY(launch_median5(f.u, f.v, w, h, s, f.dumeds, f.dvmeds, st_));
            std::swap(f.u, f.dumeds);
            std::swap(f.v, f.dvmeds);
        }
        if (level > 0) {
            const int W = g_.W[level - 1], H = g_.H[level - 1], S = g_.S[level - 1];
            RF_TRY(launch_upscale(f.u, w, h, s, W, H, S, p_.scale, f.us, st_));
            RF_TRY(launch_upscale(f.v, w, h, s, W, H, S, p_.scale, f.vs, st_));
            std::swap(f.u, f.us);
            std::swap(f.v, f.vs);
            RF_TRY(launch_upscale(f.uc, w, h, s, W, H, S, p_.scale, f.us, st_));
            RF_TRY(launch_upscale(f.vc, w, h, s, W, H, S, p_.scale, f.vs, st_));
            std::swap(f.uc, f.us);
            std::swap(f.vc, f.vs);
            RF_TRY(launch_upscale(f.uproj, w, h, s, W, H, S, p_.scale, f.us, st_));
            RF_TRY(launch_upscale(f.vproj, w, h, s, W, H, S, p_.scale, f.vs, st_));
            std::swap(f.uproj, f.us);
            std::swap(f.vproj, f.vs);
            RF_TRY(launch_upscale(f.sku, w, h, s, W, H, S, p_.scale, f.us, st_));
            RF_TRY(launch_upscale(f.skv, w, h, s, W, H, S, p_.scale, f.vs, st_));
            std::swap(f.sku, f.us);
            std::swap(f.skv, f.vs);
        }
    }
    if (p_.with_visualization)
        RF_TRY(launch_flow_to_hsv(f.u, f.v, g_.W[0], g_.H[0], g_.S[0], f.uvrgb, flow_scale_, st_));
    return hipSuccess;
}

int Solver::profile_arm(int kid, int level) {
    if (!inited_) {
        set_error("rf_profile_arm: not initialized");
        return RF_ERR_STATE;
    }
    Profiler& pr = batch_ ? batch_->prof() : prof_;
    if (pr.ev.empty()) {
        pr.ev.resize(8192);
        for (auto& e : pr.ev)
            if (check(hipEventCreate(&e), "hipEventCreate")) return RF_ERR_HIP;
    }
    pr.armed = true;
    pr.kid = kid;
    pr.level = level;
    pr.used = 0;
    pr.launches = 0;
    pr.open = false;
    return RF_OK;
}

int Solver::tiling(bool recon, int* tiling, int n) {
    const char* name = recon ? "rf_joint_tiling" : "rf_flow_tiling";
    if (!tiling || n < 0) {
        set_error(std::string(name) + ": null output or negative count");
        return RF_ERR_ARG;
    }
    if (!batch_) {
        set_error(std::string(name) + ": the fused path is not in use");
        return RF_ERR_STATE;
    }
    for (int l = 0; l < n; ++l) tiling[l] = recon ? batch_->joint_tiling(l) : batch_->flow_tiling(l);
    return RF_OK;
}

int Solver::profile_read(double* total_ms, int* launches) {
    Profiler& pr = batch_ ? batch_->prof() : prof_;
    if (check(hipStreamSynchronize(stream()), "rf_profile_read")) return RF_ERR_HIP;
    double tot = 0;
    for (int i = 0; i + 1 < pr.used; i += 2) {
        float ms = 0;
        if (check(hipEventElapsedTime(&ms, pr.ev[i], pr.ev[i + 1]), "hipEventElapsedTime")) return RF_ERR_HIP;
        tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = pr.launches;
    pr.armed = false;
    return RF_OK;
}

hipStream_t Solver::stream() const { return batch_ ? batch_->stream() : st_; }

void Solver::result(const float** u, const float** v, const float** X, const float** Y, const float** Z) const {
    if (batch_) {
        batch_->result(0, u, v, X, Y, Z);
        return;
    }
    if (u) *u = cur_.u;
    if (v) *v = cur_.v;
    if (X) *X = cur_.X;
    if (Y) *Y = cur_.Y;
    if (Z) *Z = cur_.Z;
}

// Runs one solve. The pointer table restarts from its initial assignment every solve, so the
// schedule (and a graph captured from it) is identical from solve to solve.
int Solver::run_schedule(int kind) {
    auto enqueue = [&](Fields& f) -> hipError_t {
        masked_run_ = (kind == 2);
        return kind == 1 ? enqueue_recon_plain(f) : enqueue_flow_plain(f);
    };
    if (!p_.use_graph || prof_.armed) {
        Fields f = init_;
        if (check(enqueue(f), "solve")) return RF_ERR_HIP;
        cur_ = f;
        return RF_OK;
    }
    if (!gexec_[kind]) {
        Fields f = init_;
        hipGraph_t graph = nullptr;
        if (check(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
            return RF_ERR_HIP;
        hipError_t e = enqueue(f);
        hipError_t e2 = hipStreamEndCapture(st_, &graph);
        if (check(e, "solve (capture)") || check(e2, "hipStreamEndCapture")) {
            if (graph) (void)hipGraphDestroy(graph);
            return RF_ERR_HIP;
        }
        hipError_t e3 = hipGraphInstantiate(&gexec_[kind], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (check(e3, "hipGraphInstantiate")) {
            gexec_[kind] = nullptr;
            return RF_ERR_HIP;
        }
        gfinal_[kind] = f;
    }
    if (check(hipGraphLaunch(gexec_[kind], st_), "hipGraphLaunch")) return RF_ERR_HIP;
    cur_ = gfinal_[kind];
    return RF_OK;
}

int Solver::solve_flow(float flow_scale) {
    if (!inited_) {
        set_error("rf_solve_flow: not initialized");
        return RF_ERR_STATE;
    }
    if (p_.method < RF_METHOD_TVL1 || p_.method > RF_METHOD_TVL1_INPAINTING) {
        set_error("rf_solve_flow: not a CudaFlow method");
        return RF_ERR_UNSUPPORTED;
    }
    uvrgb_valid_ = false;
    int rc;
    if (batch_) {
        rc = batch_->solve(false, flow_scale);
    } else {
        if (p_.use_graph && p_.with_visualization && flow_scale != flow_scale_ && gexec_[0]) {
            (void)hipGraphExecDestroy(gexec_[0]);  // recaptured with the new scale
            gexec_[0] = nullptr;
        }
        flow_scale_ = flow_scale;
        rc = run_schedule(0);
    }
    uvrgb_valid_ = rc == RF_OK && p_.with_visualization;
    return rc;
}

int Solver::solve_recon(const double* R0, const double* t0, const double* R1, const double* t1, float flow_scale) {
    if (!inited_) {
        set_error("rf_solve_recon: not initialized");
        return RF_ERR_STATE;
    }
    if (!camera_) {
        set_error("rf_solve_recon: call rf_set_camera first");
        return RF_ERR_STATE;
    }
    if (!R1 || !t1) {
        set_error("rf_solve_recon: R1 and t1 are required");
        return RF_ERR_ARG;
    }
    if (p_.method != RF_METHODR_TVL1_MS && p_.method != RF_METHODR_TVL1_MS_FN &&
        p_.method != RF_METHODR_TVL1_MS_FNSPARSE) {
        set_error("rf_solve_recon: method not implemented");
        return RF_ERR_UNSUPPORTED;
    }
    if ((p_.method == RF_METHODR_TVL1_MS_FN || p_.method == RF_METHODR_TVL1_MS_FNSPARSE) && !prior_) {
        set_error("rf_solve_recon: prior-flow methods need rf_initialize_large_displacement and a prior");
        return RF_ERR_STATE;
    }
    static const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, Z3[3] = {0, 0, 0};
    memcpy(R0_, R0 ? R0 : I3, sizeof R0_);
    memcpy(t0_, t0 ? t0 : Z3, sizeof t0_);
    memcpy(R1_, R1, sizeof R1_);
    memcpy(t1_, t1, sizeof t1_);
    uvrgb_valid_ = false;
    if (batch_) {
        int rc = batch_->set_camera(0, K0raw_, K1raw_, R0_, t0_, R1_, t1_);
        rc = rc ? rc : batch_->solve(true, flow_scale);
        uvrgb_valid_ = rc == RF_OK && p_.with_visualization;
        return rc;
    }
    // P = K0[l]*[R0|t0], Q = K1[l]*[R1|t1] in double (Reconflow.cu:225-226), cast to float
    for (int l = 0; l < g_.L; ++l) {
        double P[12], Q[12];
        for (int pass = 0; pass < 2; ++pass) {
            const double* K = pass ? K1_[l].data() : K0_[l].data();
            const double* R = pass ? R1_ : R0_;
            const double* t = pass ? t1_ : t0_;
            double* out = pass ? Q : P;
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
        hparams_.pq[l] = make_pq(P, Q);
        hparams_.pqd[l] = make_pqd(P, Q);
        hparams_.uc[l] = make_uc(R1_, t1_, K1_[l].data());
    }
    if (check(hipMemcpyAsync(dparams_, &hparams_, sizeof(DevParams), hipMemcpyHostToDevice, st_), "upload pose"))
        return RF_ERR_HIP;
    // the stream orders the upload before the solve; make sure the host copy is consumed
    if (check(hipStreamSynchronize(st_), "upload pose")) return RF_ERR_HIP;
    if (p_.use_graph && p_.with_visualization && flow_scale != flow_scale_ && gexec_[1]) {
        (void)hipGraphExecDestroy(gexec_[1]);  // recaptured with the new scale
        gexec_[1] = nullptr;
    }
    flow_scale_ = flow_scale;
    int rc = run_schedule(1);
    uvrgb_valid_ = rc == RF_OK && p_.with_visualization;
    return rc;
}

// rf_solve_flow_noalloc: the TV-L1 solve of enqueue_flow_plain on fields carved from the caller's
// scratch in the order rf_initialize carves them, every one zeroed first (as a fresh slab is), the
// images read in place and the flow copied out. No allocation: no slab, parameter block or stream.
int Solver::solve_flow_noalloc(const rf_params& p, const float* i0, const float* i1, float* u, float* v,
                               hipStream_t st, float* scratch, size_t nfloats) {
    const char* who = "rf_solve_flow_noalloc";
    if (inited_ || st_) {
        set_error(std::string(who) + ": internal: the solver must be uninitialized");
        return RF_ERR_STATE;
    }
    if (int rc = check_params(p, who)) return rc;
    if (!i0 || !i1 || !u || !v) {
        set_error(std::string(who) + ": null image or flow pointer");
        return RF_ERR_ARG;
    }
    if (p.method != RF_METHOD_TVL1 || p.input_type != RF_INPUT_32F) {
        set_error(std::string(who) + ": runs METHOD_TVL1 on float gray images only");
        return RF_ERR_UNSUPPORTED;
    }
    const int align = p.stride_alignment > 0 ? p.stride_alignment : align_;
    if (int rc = set_geometry(p, align, who)) return rc;
    // the fields enqueue_flow_plain uses for METHOD_TVL1: two image pyramids and 20 state fields
    const long per = (g_.bufn + 63) / 64 * 64;
    const long nfields = 2L * g_.L + 20;
    if ((size_t)(nfields * per) > nfloats) {
        set_error(std::string(who) + ": " + std::to_string(p.width) + "x" + std::to_string(p.height) + " with " +
                  std::to_string(p.n_levels) + " levels needs " + std::to_string(nfields * per * 4) +
                  " bytes of scratch; the scratch holds " + std::to_string(nfloats * 4));
        return RF_ERR_ARG;
    }
    p_ = p;
    p_.with_visualization = 0;
    p_.path = RF_PATH_PLAIN;
    p_.storage = RF_STORAGE_F32;
    p_.use_graph = 0;
    float* cursor = scratch;
    auto take = [&]() {
        float* r = cursor;
        cursor += per;
        return r;
    };
    Fields f{};
    for (int l = 0; l < g_.L; ++l) {
        f.pI0[l] = take();
        f.pI1[l] = take();
    }
    for (float** pp : {&f.i1warp, &f.dumed, &f.dvmed, &f.dumeds, &f.dvmeds, &f.pu1, &f.pu2, &f.pv1, &f.pv2, &f.pu1s,
                       &f.pu2s, &f.pv1s, &f.pv2s, &f.Ix, &f.Iy, &f.Iz, &f.u, &f.v, &f.us, &f.vs})
        *pp = take();
    f.in0 = const_cast<float*>(i0);  // read only (Cv32fToGray)
    f.in1 = const_cast<float*>(i1);
    st_ = st;
    float* all[] = {scratch};
    hipError_t e = launch_fill_zero(all, 1, nfields * per, st_);
    if (e == hipSuccess) e = enqueue_flow_plain(f);
    if (e == hipSuccess) e = hipMemcpyAsync(u, f.u, (size_t)g_.n * 4, hipMemcpyDeviceToDevice, st_);
    if (e == hipSuccess) e = hipMemcpyAsync(v, f.v, (size_t)g_.n * 4, hipMemcpyDeviceToDevice, st_);
    st_ = nullptr;  // the caller's stream: not this solver's to destroy
    return check(e, who);
}

int Solver::sync() { return check(hipStreamSynchronize(stream()), "hipStreamSynchronize"); }

// CudaFlow::copyOpticalFlowToHost (Cudaflow.cu:313-324).
int Solver::copy_flow(float* u, float* v, float* uvrgb) {
    if (!inited_ || !u || !v) {
        set_error("rf_copy_flow: not initialized or null output");
        return inited_ ? RF_ERR_ARG : RF_ERR_STATE;
    }
    size_t bytes = (size_t)g_.n * 4;
    const float *du, *dv;
    result(&du, &dv, nullptr, nullptr, nullptr);
    const float* drgb = batch_ ? batch_->result_uvrgb(0) : cur_.uvrgb;
    hipStream_t st = stream();
    if (check(hipMemcpyAsync(u, du, bytes, hipMemcpyDeviceToHost, st), "copy u")) return RF_ERR_HIP;
    if (check(hipMemcpyAsync(v, dv, bytes, hipMemcpyDeviceToHost, st), "copy v")) return RF_ERR_HIP;
    if (uvrgb && !uvrgb_valid_) {
        set_error("rf_copy_flow: no colour map: the last solve ran without visualization");
        return RF_ERR_STATE;
    }
    if (uvrgb && check(hipMemcpyAsync(uvrgb, drgb, bytes * 3, hipMemcpyDeviceToHost, st), "copy uvrgb"))
        return RF_ERR_HIP;
    return check(hipStreamSynchronize(st), "copy flow");
}

// CudaFlow::copyMaskToDevice (Inpainting.cu:3-7).
int Solver::copy_masks(const float* m0, const float* m1) {
    if (!inited_ || !m0 || !m1) {
        set_error("rf_copy_masks: not initialized or null mask");
        return inited_ ? RF_ERR_ARG : RF_ERR_STATE;
    }
    if (!init_.mask0) {
        set_error("rf_copy_masks: the context was not initialized with METHOD_TVL1_INPAINTING");
        return RF_ERR_STATE;
    }
    size_t bytes = (size_t)g_.n * 4;
    if (check(hipMemcpyAsync(init_.mask0, m0, bytes, hipMemcpyHostToDevice, st_), "copy mask0") ||
        check(hipMemcpyAsync(init_.mask1, m1, bytes, hipMemcpyHostToDevice, st_), "copy mask1"))
        return RF_ERR_HIP;
    masks_ = true;
    return check(hipStreamSynchronize(st_), "copy masks");
}

// CudaFlow::copyDepthToDevice (Cudaflow.cu:282-286).
int Solver::copy_depth(const uint16_t* d0, const uint16_t* d1) {
    if (!inited_ || !d0 || !d1) {
        set_error("rf_copy_depth: not initialized or null depth");
        return inited_ ? RF_ERR_ARG : RF_ERR_STATE;
    }
    if (!init_.depth0_16) {
        set_error("rf_copy_depth: the context was not initialized with SCENEFLOW_KINECT_TVL1");
        return RF_ERR_STATE;
    }
    size_t bytes = (size_t)g_.n * 2;
    if (check(hipMemcpyAsync(init_.depth0_16, d0, bytes, hipMemcpyHostToDevice, st_), "copy depth0") ||
        check(hipMemcpyAsync(init_.depth1_16, d1, bytes, hipMemcpyHostToDevice, st_), "copy depth1"))
        return RF_ERR_HIP;
    depth_ = true;
    return check(hipStreamSynchronize(st_), "copy depth");
}

// CudaFlow::solveInpainting (Inpainting.cu:9-152).
int Solver::solve_inpainting(float flow_scale) {
    if (!inited_ || !init_.mask0) {
        set_error("rf_solve_inpainting: needs a context initialized with METHOD_TVL1_INPAINTING");
        return RF_ERR_STATE;
    }
    if (!masks_) {
        set_error("rf_solve_inpainting: call rf_copy_masks first");
        return RF_ERR_STATE;
    }
    if (p_.use_graph && p_.with_visualization && flow_scale != flow_scale_ && gexec_[2]) {
        (void)hipGraphExecDestroy(gexec_[2]);  // recaptured with the new scale
        gexec_[2] = nullptr;
    }
    flow_scale_ = flow_scale;
    uvrgb_valid_ = false;
    int rc = run_schedule(2);
    uvrgb_valid_ = rc == RF_OK && p_.with_visualization;
    return rc;
}

// CudaFlow::_solveSceneFlow / _solveSceneFlowAndPointCloud (Cudaflow.cu:836-871): the flow solve,
// then depth0 converted to float and passed as both depth maps (Cudaflow.cu:843, 865).
int Solver::solve_scene_flow(float fx, float fy, float cx, float cy, float flow_scale, float sf_scale,
                             bool point_cloud) {
    if (!inited_ || !init_.depth0_16) {
        set_error("rf_solve_scene_flow: needs a context initialized with SCENEFLOW_KINECT_TVL1");
        return RF_ERR_STATE;
    }
    if (!depth_) {
        set_error("rf_solve_scene_flow: call rf_copy_depth first");
        return RF_ERR_STATE;
    }
    int rc = solve_flow(flow_scale);
    if (rc) return rc;
    const int w = g_.W[0], h = g_.H[0], s = g_.S[0];
    const Fields& f = init_;
    if (check(launch_cv16u_to_32f((const uint16_t*)f.depth0_16, f.depth032f, w, h, s, st_), "depth0 to float") ||
        check(launch_cv16u_to_32f((const uint16_t*)f.depth1_16, f.depth132f, w, h, s, st_), "depth1 to float") ||
        check(launch_scene_flow(cur_.u, cur_.v, f.depth032f, f.depth032f, fx, fy, cx, cy, w, h, s,
                                point_cloud ? f.pcloud0 : nullptr, point_cloud ? f.pcloud1 : nullptr, f.sceneflow,
                                st_),
              "scene flow") ||
        check(launch_scene_flow_to_hsv(f.sceneflow, w, h, s, f.sceneflowrgb, sf_scale, st_), "scene flow colour"))
        return RF_ERR_HIP;
    scene_ = true;
    return RF_OK;
}

// CudaFlow::copySceneFlowToHost / copySceneFlowAndPointCloudToHost (Cudaflow.cu:345-362).
int Solver::copy_scene_flow(float* sf, float* sfrgb, float* pc0, float* pc1) {
    if (!inited_ || !scene_) {
        set_error("rf_copy_scene_flow: no scene flow solved");
        return RF_ERR_STATE;
    }
    const size_t bytes = (size_t)g_.n * 12;
    const Fields& f = init_;
    const float* src[4] = {f.sceneflow, f.sceneflowrgb, f.pcloud0, f.pcloud1};
    float* dst[4] = {sf, sfrgb, pc0, pc1};
    for (int i = 0; i < 4; ++i)
        if (dst[i] && check(hipMemcpyAsync(dst[i], src[i], bytes, hipMemcpyDeviceToHost, st_), "copy scene flow"))
            return RF_ERR_HIP;
    return check(hipStreamSynchronize(st_), "copy scene flow");
}

// ReconFlow::copy3dToHost (Reconflow.cu:65-70).
int Solver::copy_3d(float* X, float* Y, float* Z) {
    if (!inited_ || !X || !Y || !Z) {
        set_error("rf_copy_3d: not initialized or null output");
        return inited_ ? RF_ERR_ARG : RF_ERR_STATE;
    }
    size_t bytes = (size_t)g_.n * 4;
    const float *dX, *dY, *dZ;
    result(nullptr, nullptr, &dX, &dY, &dZ);
    hipStream_t st = stream();
    if (check(hipMemcpyAsync(X, dX, bytes, hipMemcpyDeviceToHost, st), "copy X")) return RF_ERR_HIP;
    if (check(hipMemcpyAsync(Y, dY, bytes, hipMemcpyDeviceToHost, st), "copy Y")) return RF_ERR_HIP;
    if (check(hipMemcpyAsync(Z, dZ, bytes, hipMemcpyDeviceToHost, st), "copy Z")) return RF_ERR_HIP;
    return check(hipStreamSynchronize(st), "copy 3d");
}

// ReconFlow::removeEdgeArtifacts3D (Reconflow.cu:468-477): X -> Xs, then swap.
int Solver::remove_edge(int range, float maxarea) {
    if (!inited_) {
        set_error("rf_remove_edge_artifacts_3d: not initialized");
        return RF_ERR_STATE;
    }
    if (range < 1 || range >= g_.W[0] || range >= g_.H[0]) {
        set_error("rf_remove_edge_artifacts_3d: surface width must lie in [1, min(width, height))");
        return RF_ERR_ARG;
    }
    if (batch_) return batch_->remove_edge(0, range, maxarea);
    Fields& f = cur_;
    if (check(launch_remove_edge(f.X, f.Y, f.Z, range, maxarea, g_.W[0], g_.H[0], g_.S[0], f.Xs, f.Ys, f.Zs, st_),
              "remove edge artifacts"))
        return RF_ERR_HIP;
    std::swap(f.X, f.Xs);
    std::swap(f.Y, f.Ys);
    std::swap(f.Z, f.Zs);
    return RF_OK;
}

}  // namespace rf

// ===================================================================================
// C ABI
using rf::Solver;

struct rf_context {
    Solver* s;
};

#define RF_CTX(c)                                  \
    if (!(c) || !(c)->s) {                         \
        rf::set_error("null context");             \
        return RF_ERR_ARG;                         \
    }

extern "C" {

void rf_params_default(rf_params* p, int width, int height) {
    if (!p) return;
    memset(p, 0, sizeof *p);
    p->width = width;
    p->height = height;
    p->channels = 1;
    p->input_type = RF_INPUT_32F;
    p->n_levels = 1;
    p->scale = 2.0f;
    p->method = RF_METHOD_TVL1;
    p->lambda = 100.0f;
    p->lambdagrad = 400.0f;
    p->theta = 0.33f;
    p->tau = 0.25f;
    p->n_warp_iters = 1;
    p->n_solver_iters = 100;
    p->with_visualization = 1;
    p->lambdaf = 100.0f;
    p->lambdams = 100.0f;
    p->alpha_tv = 3.0f;
    p->alpha_proj = 1.0f;
    p->alpha_fn = 0.5f;
    p->stride_alignment = 32;
    p->path = RF_PATH_PLAIN;
    p->storage = RF_STORAGE_F32;
    p->use_graph = 0;
    p->fix_coarse_k = 0;
    p->tri_f64 = 0;
}

const char* rf_last_error(void) { return rf::g_err.c_str(); }

int rf_compute_pyramid_levels(int width, int height, int min_width, float scale) {
    (void)width;
    if (height <= 0 || !(scale > 1.0f)) return RF_ERR_ARG;
    int n = 1;
    int ph = (int)((float)height / scale);
    while (ph > min_width) {
        ++n;
        ph = (int)((float)ph / scale);
    }
    return n;
}

int rf_create(rf_context** out, int block_width, int block_height, int stride_alignment) {
    if (!out) {
        rf::set_error("rf_create: null output");
        return RF_ERR_ARG;
    }
    *out = new rf_context{new Solver(block_width, block_height, stride_alignment)};
    return RF_OK;
}

int rf_destroy(rf_context* ctx) {
    if (!ctx) return RF_OK;
    delete ctx->s;
    delete ctx;
    return RF_OK;
}

int rf_initialize(rf_context* ctx, const rf_params* params) {
    RF_CTX(ctx);
    if (!params) {
        rf::set_error("rf_initialize: null params");
        return RF_ERR_ARG;
    }
    return ctx->s->initialize(*params);
}

int rf_initialize_large_displacement(rf_context* ctx) {
    RF_CTX(ctx);
    return ctx->s->init_large_displacement();
}

int rf_get_pyramid(const rf_context* ctx, int* widths, int* heights, int* strides) {
    RF_CTX(ctx);
    const auto& g = ctx->s->geom();
    if (g.L == 0) {
        rf::set_error("rf_get_pyramid: not initialized");
        return RF_ERR_STATE;
    }
    for (int l = 0; l < g.L; ++l) {
        if (widths) widths[l] = g.W[l];
        if (heights) heights[l] = g.H[l];
        if (strides) strides[l] = g.S[l];
    }
    return RF_OK;
}

int rf_get_stride(const rf_context* ctx) {
    RF_CTX(ctx);
    const auto& g = ctx->s->geom();
    return g.L ? g.S[0] : RF_ERR_STATE;
}

void* rf_get_stream(const rf_context* ctx) { return (ctx && ctx->s) ? (void*)ctx->s->stream() : nullptr; }

int rf_copy_images(rf_context* ctx, const void* i0, const void* i1) {
    RF_CTX(ctx);
    return ctx->s->copy_images(i0, i1);
}

int rf_copy_images_device(rf_context* ctx, const float* d_i0, const float* d_i1) {
    RF_CTX(ctx);
    return ctx->s->copy_images_device(d_i0, d_i1);
}

int rf_set_camera(rf_context* ctx, const double K0[9], const double K1[9]) {
    RF_CTX(ctx);
    return ctx->s->set_camera(K0, K1);
}

int rf_copy_initial_flow(rf_context* ctx, const float* u, const float* v) {
    RF_CTX(ctx);
    return ctx->s->copy_prior(u, v, nullptr, false);
}

int rf_copy_sparse_flow(rf_context* ctx, const float* u, const float* v, const float* mask) {
    RF_CTX(ctx);
    return ctx->s->copy_prior(u, v, mask, true);
}

int rf_copy_warm_start(rf_context* ctx, const float* u, const float* v) {
    RF_CTX(ctx);
    return ctx->s->copy_warm_start(u, v);
}

int rf_warm_start_in_use(rf_context* ctx) {
    RF_CTX(ctx);
    return ctx->s->warm_start_in_use();
}

int rf_solve_flow(rf_context* ctx, float flow_scale) {
    RF_CTX(ctx);
    return ctx->s->solve_flow(flow_scale);
}

int rf_copy_masks(rf_context* ctx, const float* mask0, const float* mask1) {
    RF_CTX(ctx);
    return ctx->s->copy_masks(mask0, mask1);
}

int rf_copy_depth(rf_context* ctx, const uint16_t* depth0, const uint16_t* depth1) {
    RF_CTX(ctx);
    return ctx->s->copy_depth(depth0, depth1);
}

int rf_solve_inpainting(rf_context* ctx, float flow_scale) {
    RF_CTX(ctx);
    return ctx->s->solve_inpainting(flow_scale);
}

int rf_solve_scene_flow(rf_context* ctx, float fx, float fy, float cx, float cy, float flow_scale,
                        float scene_flow_scale, int point_cloud) {
    RF_CTX(ctx);
    return ctx->s->solve_scene_flow(fx, fy, cx, cy, flow_scale, scene_flow_scale, point_cloud != 0);
}

int rf_copy_scene_flow(rf_context* ctx, float* sceneflow, float* sceneflow_rgb, float* pcloud0, float* pcloud1) {
    RF_CTX(ctx);
    return ctx->s->copy_scene_flow(sceneflow, sceneflow_rgb, pcloud0, pcloud1);
}

int rf_solve_recon(rf_context* ctx, const double* R0, const double* t0, const double* R1, const double* t1,
                   float flow_scale) {
    RF_CTX(ctx);
    return ctx->s->solve_recon(R0, t0, R1, t1, flow_scale);
}

int rf_synchronize(rf_context* ctx) {
    RF_CTX(ctx);
    return ctx->s->sync();
}

int rf_copy_flow(rf_context* ctx, float* u, float* v, float* uvrgb) {
    RF_CTX(ctx);
    return ctx->s->copy_flow(u, v, uvrgb);
}

int rf_copy_3d(rf_context* ctx, float* X, float* Y, float* Z) {
    RF_CTX(ctx);
    return ctx->s->copy_3d(X, Y, Z);
}

int rf_get_result_device(const rf_context* ctx, const float** u, const float** v, const float** X, const float** Y,
                         const float** Z) {
    RF_CTX(ctx);
    ctx->s->result(u, v, X, Y, Z);
    return RF_OK;
}

int rf_remove_edge_artifacts_3d(rf_context* ctx, int surface_width, float max_surface_area) {
    RF_CTX(ctx);
    return ctx->s->remove_edge(surface_width, max_surface_area);
}

int rf_profile_arm(rf_context* ctx, int kernel_id, int level) {
    RF_CTX(ctx);
    return ctx->s->profile_arm(kernel_id, level);
}

int rf_profile_read(rf_context* ctx, double* total_ms, int* launches) {
    RF_CTX(ctx);
    return ctx->s->profile_read(total_ms, launches);
}

int rf_joint_tiling(rf_context* ctx, int* tiling, int n) {
    RF_CTX(ctx);
    return ctx->s->tiling(true, tiling, n);
}

int rf_flow_tiling(rf_context* ctx, int* tiling, int n) {
    RF_CTX(ctx);
    return ctx->s->tiling(false, tiling, n);
}

}  // extern "C"
