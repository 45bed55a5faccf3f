// vs_plane.hip -- single-plane-symmetric cross-sections: the contour of each
// tube section (a 3:1 ellipse joined to a circle, grown to the section area),
// its maximum chord and alignment, and the per-column depth profile read from
// it by vertical-line intersections.
//
// Host part (float32, the reference's operation order):
//   ShapeGenerator::combineRegularShapes   src/ShapeGenerator.cpp:35-197
//   ShapeGenerator::generateEllipseCoordinates / finalContourCoordinates
//                                          src/ShapeGenerator.cpp:220-318
//   Utils::linspace / computeTotalArea     src/Utils.cpp:59-77, 104-122
//   ShapeAlignment::rotateShapeAboutOrigin src/ShapeAlignment.cu:73-113
// Device part:
//   max-chord search (findShapeDiameter)   src/ShapeAlignment.cu:119-152
//   depth per tube cell                    src/TubeDepth.cu:493-640 (in k_depth_raw)
//
// Boost.Geometry's linestring intersection and ring area are replaced by a
// segment-pair intersection and the shoelace formula, both evaluated in
// double and rounded to float.

#include "vs_common.h"
#include "../../include/vocalsim_abi.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <string>
#include <vector>

namespace vs {

namespace {

constexpr float kMillimeter = 1e-3f;

std::vector<float> linspace_f(float start, float end, int n) {
    std::vector<float> r;
    if (n <= 0) return r;
    if (n == 1) { r.push_back(start); return r; }
    const float step = (end - start) / (n - 1);
    r.resize(n);
    for (int i = 0; i < n; ++i) r[i] = start + step * i;
    return r;
}

// Upper half of an ellipse centred on (yc, 0), sampled at n abscissae.
void half_ellipse(float y0, float y1, float a, float b, int n, float yc, std::vector<float>& ys,
                  std::vector<float>& zs) {
    ys = linspace_f(y0, y1, n);
    std::vector<float> d2(n);
    for (int i = 0; i < n; ++i) {
        const float y = ys[i] - yc;
        d2[i] = y * y;
    }
    const float a2 = a * a, b2 = b * b;
    d2[0] = a2;
    for (int i = 0; i < n / 2; ++i) d2[n - i - 1] = d2[i];
    zs.resize(n);
    for (int i = 0; i < n; ++i) zs[i] = std::sqrt(std::max(0.0f, b2 * (1.0f - d2[i] / a2)));
}

// First crossing of two polylines (points ordered by ascending y), found by
// walking the y-overlap of both; the crossing point of the two segments is
// evaluated in double.  Returns false if they do not cross.
bool first_crossing(const std::vector<float>& y1, const std::vector<float>& z1, const std::vector<float>& y2,
                    const std::vector<float>& z2, float& yi, float& zi) {
    const int n1 = (int)y1.size(), n2 = (int)y2.size();
    for (int i = 0; i + 1 < n1; ++i) {
        const double ay = y1[i], az = z1[i], by = y1[i + 1], bz = z1[i + 1];
        const double lo = std::min(ay, by), hi = std::max(ay, by);
        // segments of the second polyline whose y-range meets [lo, hi]
        int j0 = int(std::lower_bound(y2.begin(), y2.end(), (float)lo) - y2.begin()) - 1;
        j0 = std::max(j0, 0);
        for (int j = j0; j + 1 < n2 && y2[j] <= hi; ++j) {
            const double cy = y2[j], cz = z2[j], dy = y2[j + 1], dz = z2[j + 1];
            // solve a + s (b - a) = c + t (d - c)
            const double rx = by - ay, rz = bz - az, sx = dy - cy, sz = dz - cz;
            const double den = rx * sz - rz * sx;
            if (den == 0.0) continue;
            const double qx = cy - ay, qz = cz - az;
            const double s = (qx * sz - qz * sx) / den;
            const double t = (qx * rz - qz * rx) / den;
            if (s < 0.0 || s > 1.0 || t < 0.0 || t > 1.0) continue;
            yi = (float)(ay + s * rx);
            zi = (float)(az + s * rz);
            return true;
        }
    }
    return false;
}

// |area| of the closed ring through the points (shoelace, double).
float ring_area(const std::vector<float>& y, const std::vector<float>& z) {
    const size_t n = y.size();
    double acc = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const size_t k = (i + 1) % n;
        acc += (double)y[i] * (double)z[k] - (double)y[k] * (double)z[i];
    }
    return (float)std::fabs(0.5 * acc);
}

// Joined contour (upper half, 2 * nh - 1 points) for the given semi-axes.
void joined_half(float a1, float b1, float a2, float b2, float yc1, float yc2, int nh, std::vector<float>& hy,
                 std::vector<float>& hz, bool& ok) {
    std::vector<float> s1y, s1z, s2y, s2z;
    const float s1_start = yc1 - a1, s1_end = yc1 + a1;
    const float s2_start = yc2 - a2, s2_end = yc2 + a2;
    half_ellipse(s1_start, s1_end, a1, b1, nh, yc1, s1y, s1z);
    half_ellipse(s2_start, s2_end, a2, b2, nh, yc2, s2y, s2z);
    float iy = 0, iz = 0;
    ok = first_crossing(s1y, s1z, s2y, s2z, iy, iz);
    if (!ok) return;
    const int seg = nh - 1;
    const std::vector<float> ya = linspace_f(s1_start, iy, seg);
    const std::vector<float> yb = linspace_f(iy, s2_end, seg);
    hy.assign(2 * nh - 1, 0.0f);
    hz.assign(2 * nh - 1, 0.0f);
    int p = 0;
    for (int i = 0; i < seg; ++i, ++p) {
        const float d = ya[i] - yc1;
        hy[p] = ya[i];
        hz[p] = std::sqrt(std::max(0.0f, (b1 * b1) * (1.0f - (d * d) / (a1 * a1))));
    }
    hy[p] = iy;
    hz[p] = iz;
    ++p;
    for (int i = 0; i < seg; ++i, ++p) {
        const float d = yb[i] - yc2;
        hy[p] = yb[i];
        hz[p] = std::sqrt(std::max(0.0f, (b2 * b2) * (1.0f - (d * d) / (a2 * a2))));
    }
}

// For every point of every section, the farthest point of the same section
// (first index attaining the strict maximum) and that distance.  One lane per
// point; the section's points stream through LDS in tiles of kChordTile, which
// every lane of the work-group then reads as a broadcast.  The candidates are
// visited in index order, so the first maximum is kept as in a plain loop.
constexpr int kChordTile = 256;
__global__ __launch_bounds__(kChordTile) void k_max_chord(const float* __restrict__ py, const float* __restrict__ pz,
                                                          int npt, float* __restrict__ dist, int* __restrict__ far_idx) {
    __shared__ float ty[kChordTile], tz[kChordTile];
    const int sec = blockIdx.y;
    const int i = blockIdx.x * kChordTile + threadIdx.x;
    const float* y = py + (size_t)sec * npt;
    const float* z = pz + (size_t)sec * npt;
    const bool live = i < npt;
    const float yi = live ? y[i] : 0.0f, zi = live ? z[i] : 0.0f;
    float best = 0.0f;
    int arg = 0;
    for (int k0 = 0; k0 < npt; k0 += kChordTile) {
        const int k = k0 + threadIdx.x;
        __syncthreads();   // the previous tile has been read
        if (k < npt) {
            ty[threadIdx.x] = y[k];
            tz[threadIdx.x] = z[k];
        }
        __syncthreads();
        const int m = min(kChordTile, npt - k0);
        for (int j = 0; j < m; ++j) {
            const float dy = yi - ty[j], dz = zi - tz[j];
            const float d = __builtin_sqrtf(dy * dy + dz * dz);
            if (d > best) { best = d; arg = k0 + j; }
        }
    }
    if (live) {
        dist[(size_t)sec * npt + i] = best;
        far_idx[(size_t)sec * npt + i] = arg;
    }
}

}  // namespace

int plane_sections(const float* area_cm2, int n, std::vector<PlaneSection>& out, std::string& err, hipStream_t st) {
    constexpr int kCoords = 1999;               // TubeGeometry.cu:373
    constexpr int nh = (kCoords + 1) / 2;       // points per half shape
    out.assign(n, PlaneSection{});
    std::vector<float> all_y, all_z;
    all_y.reserve((size_t)n * kPlanePoints);
    all_z.reserve((size_t)n * kPlanePoints);
    for (int s = 0; s < n; ++s) {
        const float A = area_cm2[s] * 1e-2f * 1e-2f;
        const float m = 3.0f;
        const float x = A / (m + 1.0f);
        const float A1 = m * x, A2 = 1.0f * x;
        const float r1 = 3.0f, r2 = 1.0f;
        const float l1 = std::sqrt((float)(A1 / (M_PI * r1 * 1.0f)));
        float a1 = l1 * r1, b1 = l1 * 1.0f;
        const float l2 = std::sqrt((float)(A2 / (M_PI * r2 * 1.0f)));
        float a2 = l2 * r2, b2 = l2 * 1.0f;
        const float yc1 = 0.0f;
        const float yc2 = yc1 + a1 + 0.003f * kMillimeter;
        std::vector<float> hy, hz;
        bool ok = false;
        joined_half(a1, b1, a2, b2, yc1, yc2, nh, hy, hz, ok);
        if (!ok) { err = "section " + std::to_string(s) + ": contour shapes do not intersect"; return VS_E_GEOMETRY; }
        float area = 2 * ring_area(hy, hz);
        const float grow = kMillimeter * 0.01f;
        int iters = 0;
        while (area < A) {
            a1 = a1 + grow; b1 = b1 + grow;
            a2 = a2 + grow; b2 = b2 + grow;
            joined_half(a1, b1, a2, b2, yc1, yc2, nh, hy, hz, ok);
            if (!ok) { err = "section " + std::to_string(s) + ": contour shapes do not intersect"; return VS_E_GEOMETRY; }
            area = 2 * ring_area(hy, hz);
            if (++iters > 1000000) { err = "section " + std::to_string(s) + ": contour growth does not converge"; return VS_E_GEOMETRY; }
        }
        hz.front() = 0.0f;
        hz.back() = 0.0f;
        PlaneSection& P = out[s];
        P.y = hy;
        P.z = hz;
        for (int i = (int)hy.size() - 2; i >= 0; --i) {
            P.y.push_back(hy[i]);
            P.z.push_back(-1 * hz[i]);
        }
        P.grow_steps = iters;
        all_y.insert(all_y.end(), P.y.begin(), P.y.end());
        all_z.insert(all_z.end(), P.z.begin(), P.z.end());
    }
    // maximum chord on the device: O(points^2) per section, on the caller's
    // stream with stream-ordered buffers (no device-wide synchronisation)
    const int npt = kPlanePoints;
    const size_t np = (size_t)n * npt;
    float* buf = nullptr;   // y, z, distance (float) and far index (int), np each
    std::vector<float> hd(np);
    std::vector<int> hi(np);
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&buf), 4 * np * sizeof(float), st);
    if (e == hipSuccess) {
        float *dy = buf, *dz = buf + np, *dd = buf + 2 * np;
        int* di = reinterpret_cast<int*>(buf + 3 * np);
        e = hipMemcpyAsync(dy, all_y.data(), np * sizeof(float), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dz, all_z.data(), np * sizeof(float), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_max_chord, dim3((npt + kChordTile - 1) / kChordTile, n), dim3(kChordTile), 0, st,
                               dy, dz, npt, dd, di);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(hd.data(), dd, np * sizeof(float), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipMemcpyAsync(hi.data(), di, np * sizeof(int), hipMemcpyDeviceToHost, st);
        const hipError_t ef = hipFreeAsync(buf, st);
        if (e == hipSuccess) e = ef;
        const hipError_t es = hipStreamSynchronize(st);
        if (e == hipSuccess) e = es;
    }
    if (e != hipSuccess) { err = std::string("max chord: ") + hipGetErrorString(e); return VS_E_DEVICE; }
    for (int s = 0; s < n; ++s) {
        PlaneSection& P = out[s];
        const float* d = hd.data() + (size_t)s * npt;
        const int* f = hi.data() + (size_t)s * npt;
        float best = 0.0f;
        int arg = 0;
        for (int i = 0; i < npt; ++i)
            if (d[i] > best) { best = d[i]; arg = i; }
        P.diameter = best;
        const float p1y = P.y[arg], p1z = P.z[arg], p2y = P.y[f[arg]], p2z = P.z[f[arg]];
        // rotate the maximum chord onto the y axis, then shift into the first
        // quadrant with a 0.05 margin
        const float theta = std::atan2(p2z - p1z, p2y - p1y);
        const float c = std::cos(-1 * theta), sn = std::sin(-1 * theta);
        for (int i = 0; i < npt; ++i) {
            const float y = P.y[i], z = P.z[i];
            P.y[i] = c * y - sn * z;
            P.z[i] = sn * y + c * z;
        }
        float mnY = P.y[0], mnZ = P.z[0];
        for (int i = 0; i < npt; ++i) { mnY = std::min(mnY, P.y[i]); mnZ = std::min(mnZ, P.z[i]); }
        const float shY = (mnY < 0.0f ? -mnY : 0.0f) + 0.05f;
        const float shZ = (mnZ < 0.0f ? -mnZ : 0.0f) + 0.05f;
        P.minY = std::numeric_limits<float>::infinity();
        P.minZ = std::numeric_limits<float>::infinity();
        P.maxZ = -std::numeric_limits<float>::infinity();
        for (int i = 0; i < npt; ++i) {
            P.y[i] += shY;
            P.z[i] += shZ;
            P.minY = std::min(P.minY, P.y[i]);
            P.minZ = std::min(P.minZ, P.z[i]);
            P.maxZ = std::max(P.maxZ, P.z[i]);
        }
    }
    return 0;
}

}  // namespace vs
