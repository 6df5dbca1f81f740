// curve_keys.h -- internal to librbf_amd: Morton and Hilbert keys of one
// point, shared by the host node layer (host_nodes.cpp, g++) and the device
// node layer (knn_device.hip, hipcc), so both give the same keys.
//
// Midpoint rule of zidx / hidx (rbf_ordering.F90:82-226): at each of ndiv
// levels the box is halved at lo + 0.5 * (hi - lo) and a coordinate goes
// right / up when it is strictly greater than the midpoint.
#ifndef RBF_CURVE_KEYS_H
#define RBF_CURVE_KEYS_H

#include <cstdint>

#if defined(__HIPCC__)
#define RBF_HD __host__ __device__
#else
#define RBF_HD
#endif

namespace rbf_amd_internal {

RBF_HD inline int64_t morton_key2(double px, double py, double x0, double y0, double x1,
                                  double y1, int ndiv) {
    int64_t z = 0;
    for (int i = 0; i < ndiv; ++i) {
        z <<= 2;
        const double cx = x0 + 0.5 * (x1 - x0);
        const double cy = y0 + 0.5 * (y1 - y0);
        const bool right = px > cx, up = py > cy;
        z += (right ? 1 : 0) + (up ? 2 : 0);
        if (right) x0 = cx; else x1 = cx;
        if (up) y0 = cy; else y1 = cy;
    }
    return z;
}

// Hilbert state machine: the 2-bit digit emitted for quadrant q (bit 0:
// right half, bit 1: upper half) in orientation rot, and the orientation of
// that sub-square. These are the tables of hidx (rbf_ordering.F90:164-181),
// digit[rot][q] = {0,3,1,2}, {2,3,1,0}, {2,1,3,0}, {0,1,3,2} and
// next[rot][q] = {3,1,0,0}, {1,0,1,2}, {2,2,3,1}, {0,3,2,3}, packed two
// bits per entry at bit 2 * (4 * rot + q).
constexpr uint32_t kHilbertDigit = 0xb4361e9cu;
constexpr uint32_t kHilbertNext = 0xec7a9107u;

RBF_HD inline int64_t hilbert_key2(double px, double py, double x0, double y0, double x1,
                                   double y1, int ndiv) {
    int64_t h = 0;
    int rot = 0;
    for (int i = 0; i < ndiv; ++i) {
        h <<= 2;
        const double cx = x0 + 0.5 * (x1 - x0);
        const double cy = y0 + 0.5 * (y1 - y0);
        const bool right = px > cx, up = py > cy;
        const int q = (right ? 1 : 0) + (up ? 2 : 0);
        const int sh = 2 * (4 * rot + q);
        h += (kHilbertDigit >> sh) & 3u;
        rot = static_cast<int>((kHilbertNext >> sh) & 3u);
        if (right) x0 = cx; else x1 = cx;
        if (up) y0 = cy; else y1 = cy;
    }
    return h;
}

RBF_HD inline int64_t morton_key3(double px, double py, double pz, const double* bb, int ndiv) {
    double lo[3] = {bb[0], bb[1], bb[2]}, hi[3] = {bb[3], bb[4], bb[5]};
    const double p[3] = {px, py, pz};
    int64_t z = 0;
    for (int i = 0; i < ndiv; ++i) {
        z <<= 3;
        for (int d = 0; d < 3; ++d) {
            const double c = lo[d] + 0.5 * (hi[d] - lo[d]);
            if (p[d] > c) {
                z += int64_t(1) << d;
                lo[d] = c;
            } else {
                hi[d] = c;
            }
        }
    }
    return z;
}

}  // namespace rbf_amd_internal

#endif  // RBF_CURVE_KEYS_H
