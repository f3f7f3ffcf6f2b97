// Batched fused solve: many image pairs of uneven sizes advance through one static schedule,
// every launch covering the tiles of all pairs at that step. A single pair is a batch of one.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/rfhip.h"
#include "rf_device.hpp"
#include "rf_solver.hpp"

namespace rf {

// Device-resident descriptor of one pair. Double-buffered fields are indexed by a parity
// the host tracks (the same for every pair of a batch, since the schedule is shared).
struct PairFields {
    int W[kMaxLevels], H[kMaxLevels], S[kMaxLevels];
    float* I0[kMaxLevels];
    float* I1[kMaxLevels];
    const void *in0, *in1;
    float *i1warp, *Ix, *Iy, *Iz;
    float *du[2], *dv[2], *p1u[2], *p2u[2], *p1v[2], *p2v[2];
    float *up[2], *vp[2], *sku[2], *skv[2], *uc[2], *vc[2];
    float *X[2], *Y[2], *Z[2];
    float *u[2], *v[2];
    float *ufn, *vfn, *ufn_l, *vfn_l, *mask, *mask_l;
    float* uvrgb;
    float* scratch;          // spare field (edge-artifact filter output)
    float* zero;             // all zero: the uhat and dual inputs of a warp's first iteration
    PQ pq[kMaxLevels];
    PQd pqd[kMaxLevels];
    UcParams ucp[kMaxLevels];
};

// Tile shapes (pixels per workgroup) of the batched launches: pointwise 64x4; joint iteration
// 64x8 (two rows per lane), 30x14 and 14x14 (one pixel per lane of a 32x16 / 16x16 extended
// tile; TC_J16O: the 14x14 tiling compiled for occupancy rather than latency; TC_JK2, TC_JK3: 10x10
// and 6x6 pixels of a 16x16 extended tile, two and three iterations per launch); TV-L1 iteration
// 64x16 (four rows per lane), 15x15 and 31x15 (one pixel per lane of a 16x16 / 32x16 extended
// tile), 25x9 and 17x17 (four / eight iterations per launch on a 32x16 / 32x32 extended tile),
// 29x13 and 21x5 (two / six iterations per launch on a 32x16 extended tile).
enum TileClass {
    TC_PIX = 0,
    TC_J64 = 1,
    TC_FLOW = 2,
    TC_J32 = 3,
    TC_J16 = 4,
    TC_J16O = 5,
    TC_F16 = 6,
    TC_F32 = 7,
    TC_JK2 = 8,
    TC_JK3 = 9,
    TC_FK4 = 10,
    TC_FK8 = 11,
    TC_FK2 = 12,
    TC_FK6 = 13,
    TC_COUNT = 14
};
constexpr int kTileW[TC_COUNT] = {64, 64, 64, 30, 14, 14, 15, 31, 10, 6, 25, 17, 29, 21};
constexpr int kTileH[TC_COUNT] = {4, 8, 16, 14, 14, 14, 15, 15, 10, 6, 9, 17, 13, 5};
// solver iterations one launch of the class runs
constexpr int kIters[TC_COUNT] = {1, 1, 1, 1, 1, 1, 1, 1, 2, 3, 4, 8, 2, 6};

// One launch's tile list: (pair, tile index within the pair).
struct TileSet {
    int2* map = nullptr;  // device
    int ntiles = 0;
};

// Buffer parities of the shared schedule.
struct Parity {
    int d = 0;  // uhat, p
    int f = 0;  // uproj, sk, uc
    int x = 0;  // X, Y, Z
    int u = 0;  // u, v
};

class Batch {
  public:
    Batch() = default;
    ~Batch();
    int create(const rf_params& p, int n, const int* widths, const int* heights);
    int n() const { return (int)geo_.size(); }
    int stride(int i) const { return geo_[i].S[0]; }
    int set_images_host(int i, const void* i0, const void* i1);
    int set_images_device(int i, const float* i0, const float* i1);
    int set_camera(int i, const double* K0, const double* K1, const double* R0, const double* t0, const double* R1,
                   const double* t1);
    int set_prior(int i, const float* u, const float* v, const float* mask, bool sparse);
    void enable_prior() { prior_ = true; }
    // warm start of pair i: device flow in the pair's level-0 layout, read at every solve until
    // replaced (null clears); warm_start_in_use: 1 when the last solve found it usable
    int set_warm_start(int i, const float* u, const float* v);
    int warm_start_in_use(int i);
    int solve(bool recon, float flow_scale);
    int sync();
    void result(int i, const float** u, const float** v, const float** X, const float** Y, const float** Z) const;
    float* result_uvrgb(int i) const { return hp_[i].uvrgb; }
    int remove_edge(int i, int range, float maxarea);
    hipStream_t stream() const { return st_; }
    Profiler& prof() { return prof_; }
    const rf_params& params() const { return p_; }
    int joint_tiling(int level) const;  // RF_JOINT_TILE_* in use at a level
    int flow_tiling(int level) const;   // RF_FLOW_TILE_* in use at a level

  private:
    struct Geo {
        std::vector<int> W, H, S;
        long n0 = 0, bufn = 0;
    };
    hipError_t enqueue(bool recon);
    template <typename TI, typename TS>
    hipError_t enqueue_t(bool recon);
    int check(hipError_t e, const char* what);
    int upload_pairs();
    TileSet& tiles(int level, int cls);
    int tile_class(bool recon, int level, int warp) const;
    int tune_slot(int k, int round, int warp) const;
    int tune_tiling(bool recon);
    void release();

    rf_params p_{};
    std::vector<Geo> geo_;
    std::vector<PairFields> hp_;
    PairFields* dp_ = nullptr;
    void* slab_ = nullptr;
    size_t slab_bytes_ = 0;
    hipStream_t st_ = nullptr;
    bool prior_ = false, sparse_ = false;
    float flow_scale_ = 5.0f;
    // tile sets: [level][TileClass]; level kMaxLevels is the flat level-0 buffer (stride x height)
    std::vector<TileSet> sets_;
    int2* maps_ = nullptr;
    Parity final_{};
    bool dirty_ = true;
    // iteration tiling per level, [0] TV-L1 and [1] joint: fixed by the parameters, or chosen by
    // timing the candidates on alternating warps of the first solve (every tiling gives the same
    // results)
    std::vector<int> cls_[2];
    bool tuned_[2] = {false, false}, tuning_ = false;
    int tune_round_ = 0;
    unsigned pair_elems_ = 0;  // pairs of one size: floats between their fields (0 otherwise)
    std::vector<hipEvent_t> tev_;
    hipGraphExec_t gexec_[2] = {nullptr, nullptr};
    std::vector<const float*> wsu_, wsv_;  // warm starts (null: none)
    int* ws_bad_ = nullptr;                // per pair: the warm-start rule failed in the last solve
    void drop_graphs();
    Parity gfinal_[2];
    Profiler prof_;
};

}  // namespace rf
