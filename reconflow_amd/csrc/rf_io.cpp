// Host I/O formats of the solve path (rfhip.h rf_read_flo / rf_write_flo / rf_write_ply).
//
//  * Middlebury .flo: float32 magic 202021.25, int32 width, int32 height, then width*height
//    (u, v) float32 pairs row-major — the prior-flow file the reference demo loads
//    (test_reconflow/main.cpp:82-90, data/flo319.flo).
//  * PLY point cloud of the reconstruction with the demo's depth-range mask
//    (test_reconflow/main.cpp:108-139): a point is kept when zmin < Z <= zmax, X <= xymax and
//    -xymax < Y <= xymax, and is written as (X, -Y, -Z). Rejected points are left out of the
//    file (the demo writes them as NaN vertices).

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rfhip.h"

namespace rf {
void set_error(const std::string& msg);
}

extern "C" {

int rf_read_flo(const char* path, int* width, int* height, float* u, float* v, int stride) {
    if (!path || !width || !height) {
        rf::set_error("rf_read_flo: null argument");
        return RF_ERR_ARG;
    }
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        rf::set_error(std::string("rf_read_flo: cannot open ") + path);
        return RF_ERR_ARG;
    }
    float magic = 0;
    int32_t wh[2] = {0, 0};
    if (std::fread(&magic, 4, 1, f) != 1 || std::fread(wh, 4, 2, f) != 2 || magic != 202021.25f || wh[0] <= 0 ||
        wh[1] <= 0 || wh[0] > (1 << 16) || wh[1] > (1 << 16)) {
        std::fclose(f);
        rf::set_error("rf_read_flo: not a Middlebury .flo file");
        return RF_ERR_ARG;
    }
    *width = wh[0];
    *height = wh[1];
    if (!u || !v) {
        std::fclose(f);
        return RF_OK;
    }
    if (stride < wh[0]) {
        std::fclose(f);
        rf::set_error("rf_read_flo: stride smaller than the width");
        return RF_ERR_ARG;
    }
    std::vector<float> row(2 * (size_t)wh[0]);
    for (int y = 0; y < wh[1]; ++y) {
        if (std::fread(row.data(), 4, row.size(), f) != row.size()) {
            std::fclose(f);
            rf::set_error("rf_read_flo: truncated file");
            return RF_ERR_ARG;
        }
        for (int x = 0; x < stride; ++x) {  // zero padding to the stride (main.cpp:96)
            u[(size_t)y * stride + x] = x < wh[0] ? row[2 * x] : 0.0f;
            v[(size_t)y * stride + x] = x < wh[0] ? row[2 * x + 1] : 0.0f;
        }
    }
    std::fclose(f);
    return RF_OK;
}

int rf_write_flo(const char* path, const float* u, const float* v, int width, int height, int stride) {
    if (!path || !u || !v || width <= 0 || height <= 0 || stride < width) {
        rf::set_error("rf_write_flo: invalid argument");
        return RF_ERR_ARG;
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        rf::set_error(std::string("rf_write_flo: cannot open ") + path);
        return RF_ERR_ARG;
    }
    const float magic = 202021.25f;
    const int32_t wh[2] = {width, height};
    std::fwrite(&magic, 4, 1, f);
    std::fwrite(wh, 4, 2, f);
    std::vector<float> row(2 * (size_t)width);
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            row[2 * x] = u[(size_t)y * stride + x];
            row[2 * x + 1] = v[(size_t)y * stride + x];
        }
        std::fwrite(row.data(), 4, row.size(), f);
    }
    return std::fclose(f) == 0 ? RF_OK : RF_ERR_ARG;
}

int rf_write_ply(const char* path, const float* X, const float* Y, const float* Z, const uint8_t* rgb, int width,
                 int height, int stride, float zmin, float zmax, float xymax) {
    if (!path || !X || !Y || !Z || width <= 0 || height <= 0 || stride < width) {
        rf::set_error("rf_write_ply: invalid argument");
        return RF_ERR_ARG;
    }
    std::vector<size_t> keep;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            size_t p = (size_t)y * stride + x;
            float px = X[p], py = Y[p], pz = Z[p];
            if ((pz <= zmax && pz > zmin) && (px <= xymax && py > -xymax) && (py <= xymax && py > -xymax))
                keep.push_back(p);
        }
    FILE* f = std::fopen(path, "w");
    if (!f) {
        rf::set_error(std::string("rf_write_ply: cannot open ") + path);
        return RF_ERR_ARG;
    }
    std::fprintf(f, "ply\nformat ascii 1.0\nelement vertex %zu\nproperty float x\nproperty float y\nproperty float z\n",
                 keep.size());
    if (rgb) std::fprintf(f, "property uchar red\nproperty uchar green\nproperty uchar blue\n");
    std::fprintf(f, "end_header\n");
    for (size_t p : keep) {
        std::fprintf(f, "%.7g %.7g %.7g", X[p], -Y[p], -Z[p]);
        if (rgb) std::fprintf(f, " %u %u %u", rgb[3 * p + 2], rgb[3 * p + 1], rgb[3 * p]);  // BGR input
        std::fprintf(f, "\n");
    }
    return std::fclose(f) == 0 ? RF_OK : RF_ERR_ARG;
}

}  // extern "C"
