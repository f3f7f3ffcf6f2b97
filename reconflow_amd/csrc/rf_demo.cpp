// Command-line program over the C++ drop-in classes (include/rfhip_cudaflow.hpp), in the shape of
// the reference demo (test_reconflow/main.cpp:16-141): build the solver, copy a pair in, solve,
// copy the results out. The pair comes from the seeded generator; results are written as raw
// float32 planes (u, v[, X, Y, Z, uvrgb]) of stride x height each (uvrgb: 3 floats per pixel).
//
//   rf_demo flow  <w> <h> <levels> <warps> <iters> <seed> <out.bin> [fused]
//   rf_demo recon <w> <h> <levels> <warps> <iters> <seed> <out.bin> [fused] [ply]

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rfhip_cudaflow.hpp"

extern "C" int rfs_make_flow_pair(int w, int h, int s, uint64_t seed, float* i0, float* i1, float* gt_u, float* gt_v);
extern "C" int rfs_make_recon_pair(int w, int h, int s, uint64_t seed, double* K, double* R, double* t, float* i0,
                                   float* i1, float* gt_u, float* gt_v, float* gt_Z);

static int fail(const char* what) {
    std::fprintf(stderr, "%s failed: %s\n", what, rfhip::lastError().c_str());
    return 1;
}

int main(int argc, char** argv) {
    if (argc < 9) {
        std::fprintf(stderr, "usage: %s flow|recon w h levels warps iters seed out.bin [fused] [ply]\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    const int w = std::atoi(argv[2]), h = std::atoi(argv[3]), L = std::atoi(argv[4]), nw = std::atoi(argv[5]),
              ni = std::atoi(argv[6]);
    const uint64_t seed = std::strtoull(argv[7], nullptr, 10);
    const char* out = argv[8];
    bool fused = false, ply = false;
    for (int i = 9; i < argc; ++i) {
        if (!std::strcmp(argv[i], "fused")) fused = true;
        if (!std::strcmp(argv[i], "ply")) ply = true;
    }
    const int s = (w + 31) / 32 * 32;
    std::vector<float> i0((size_t)s * h), i1((size_t)s * h), u((size_t)s * h), v((size_t)s * h);
    FILE* f = std::fopen(out, "wb");
    if (!f) return 2;
    if (mode == "flow") {
        rfs_make_flow_pair(w, h, s, seed, i0.data(), i1.data(), nullptr, nullptr);
        rfhip::CudaFlow flow(32, 12, 32);
        flow.setPath(fused ? RF_PATH_FUSED : RF_PATH_PLAIN);
        if (flow.initialize(w, h, 1, RF_INPUT_32F, L, 2.0f, rfhip::CudaFlow::METHOD_TVL1, 50.0f, 0.0f, 0.3f, 0.25f,
                            nw, ni, false))
            return fail("initialize");
        rfhip::Mat m0(i0.data(), h, s, s * 4), m1(i1.data(), h, s, s * 4), mu(u.data(), h, s, s * 4),
            mv(v.data(), h, s, s * 4);
        if (flow.copyImagesToDevice(m0, m1)) return fail("copyImagesToDevice");
        if (flow.solveOpticalFlow(5.0f)) return fail("solveOpticalFlow");
        if (flow.copyOpticalFlowToHost(mu, mv)) return fail("copyOpticalFlowToHost");
        std::fwrite(u.data(), 4, u.size(), f);
        std::fwrite(v.data(), 4, v.size(), f);
    } else {
        double K[9], R[9], t[3];
        rfs_make_recon_pair(w, h, s, seed, K, R, t, i0.data(), i1.data(), nullptr, nullptr, nullptr);
        rfhip::ReconFlow rec(32, 12, 32);
        rec.setPath(fused ? RF_PATH_FUSED : RF_PATH_PLAIN);
        rec.setInputType(RF_INPUT_32F);
        // demo parameters of the reference (test_reconflow/main.cpp:30-36)
        if (rec._initializeR(w, h, 1, L, 2.0f, rfhip::ReconFlow::METHODR_TVL1_MS, 50.0f, 0.0f, 0.1f, 100.0f, 33.3f,
                             60.0f, 0.125f, nw, ni))
            return fail("_initializeR");
        rfhip::Mat mK(K, 3, 3, 24), mR(R, 3, 3, 24), mt(t, 3, 1, 8);
        if (rec.setCameraMatrices(mK, mK)) return fail("setCameraMatrices");
        rfhip::Mat m0(i0.data(), h, s, s * 4), m1(i1.data(), h, s, s * 4), mu(u.data(), h, s, s * 4),
            mv(v.data(), h, s, s * 4);
        if (rec.copyImagesToDevice(m0, m1)) return fail("copyImagesToDevice");
        if (rec.solveR(mR, mt, 50.0f)) return fail("solveR");
        std::vector<float> X((size_t)s * h), Y((size_t)s * h), Z((size_t)s * h), rgb((size_t)s * h * 3);
        rfhip::Mat mX(X.data(), h, s, s * 4), mY(Y.data(), h, s, s * 4), mZ(Z.data(), h, s, s * 4),
            mrgb(rgb.data(), h, s, s * 12);
        // the reference demo's copy-out (test_reconflow/main.cpp:105-106)
        if (rec.copyOpticalFlowToHost(mu, mv, mrgb) || rec.copy3dToHost(mX, mY, mZ)) return fail("copy out");
        for (auto* p : {&u, &v, &X, &Y, &Z, &rgb}) std::fwrite(p->data(), 4, p->size(), f);
        if (ply) {
            std::string plyname = std::string(out) + ".ply";
            if (rf_write_ply(plyname.c_str(), X.data(), Y.data(), Z.data(), nullptr, w, h, s, 10.0f, 75.0f, 75.0f))
                return fail("rf_write_ply");
        }
    }
    std::fclose(f);
    std::printf("%s %dx%d done\n", mode.c_str(), w, h);
    return 0;
}
