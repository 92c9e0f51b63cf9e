// Internal interface between the partition (partition.cpp), the thumbnail
// morphology (morph.cpp) and the reference-compatible layer (compat.cpp).
#pragma once

#include <cstdint>
#include <vector>

#include "TaiChi.h"


namespace tc {

// getSdf's image pipeline: x1,y1,x2,y2 per kept contour
std::vector<int> sdf_from_image(const uint8_t *mat, int H, int W);
// its two halves: the morphology (binary image out) and the contour points
void morph_thin(const uint8_t *mat, int H, int W, uint8_t *thin);
std::vector<int> contour_points(const uint8_t *thin, int H, int W);
// contours alone: [n, (len, hole, x0, y0, ...)...]
std::vector<int> contours_flat(const uint8_t *img, int H, int W);

// One square shape's diagonal partition (partition_dia_shape semantics).
struct DiaShape {
    int xs = 0, ys = 0, side = 0, width = 0;
    std::vector<int> offs;                 // dense diagonal offsets, increasing
    std::vector<int> st, en;               // row range [st, en) of each dense diagonal
    std::vector<std::vector<int>> zeros;   // per dense diagonal: zero rows (local)
    long long nnz = 0;                     // claimed nonzeros
};
void partition_square(const int *rowptr, const int *col, std::vector<uint8_t> &claimed, int x, int y, int side,
                      DiaShape &sh);

// Steps of the partition shared by the host and the device builds.
struct Square {
    int x, y, side;
};
int width_tier(int side);
int candidate_squares(int M, int N, const unsigned char *img, std::vector<Square> &out);
void candidate_squares_from_pts(int M, int N, const std::vector<int> &pts, std::vector<Square> &out);
void decide_dense(int side, int w, const long long *cnt, DiaShape &sh, std::vector<int> &dense_of);
void shape_arrays(const std::vector<DiaShape> &kept, int nnz, TaiChi_new *out);

// tc_partition_device, and when claimed (nnz bytes, host) is not null it
// receives 1 for every entry a dense diagonal claimed: the entries the
// reference marks csrVal[j] = FLT_MAX (TaiChi.cu:866). partition_gpu.hip.
int partition_device_claims(int M, int N, int nnz, int *rowptr, int *col, const int *d_rowptr, const int *d_col,
                            TaiChi_new *out, void *stream, uint8_t *claimed);

}  // namespace tc
