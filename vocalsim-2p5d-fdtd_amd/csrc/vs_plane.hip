// vs_plane.hip -- maximum chord of the single-plane-symmetric contours on
// the device (findShapeDiameter, src/ShapeAlignment.cu:119-152).  The
// contours themselves, their alignment and the per-section bookkeeping are
// host code (plane_sections in vs_host.cpp); the depth per tube cell read
// from them by vertical-line intersections is in k_depth_raw
// (src/TubeDepth.cu:493-640).

#include "vs_common.h"

#include <string>
#include <vector>

namespace vs {

namespace {

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

int max_chord_on_device(hipStream_t st, const std::vector<float>& y, const std::vector<float>& z, int sections,
                        int npt, std::vector<float>& dist, std::vector<int>& far_idx, std::string& err) {
    // O(points^2) per section, on the caller's stream with stream-ordered
    // buffers (no device-wide synchronisation)
    const size_t np = (size_t)sections * npt;
    if (y.size() != np || z.size() != np) { err = "max chord: contour size mismatch"; return VS_E_INVALID; }
    dist.assign(np, 0.0f);
    far_idx.assign(np, 0);
    float* buf = nullptr;   // y, z, distance (float) and far index (int), np each
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&buf), 4 * np * sizeof(float), st);
    if (e == hipSuccess) {
        float *dy = buf, *dz = buf + np, *dd = buf + 2 * np;
        int* di = reinterpret_cast<int*>(buf + 3 * np);
        e = hipMemcpyAsync(dy, y.data(), np * sizeof(float), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dz, z.data(), np * sizeof(float), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_max_chord, dim3((npt + kChordTile - 1) / kChordTile, sections), dim3(kChordTile), 0, st,
                               dy, dz, npt, dd, di);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(dist.data(), dd, np * sizeof(float), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipMemcpyAsync(far_idx.data(), di, np * sizeof(int), hipMemcpyDeviceToHost, st);
        const hipError_t ef = hipFreeAsync(buf, st);
        if (e == hipSuccess) e = ef;
        const hipError_t es = hipStreamSynchronize(st);
        if (e == hipSuccess) e = es;
    }
    if (e != hipSuccess) { err = std::string("max chord: ") + hipGetErrorString(e); return VS_E_DEVICE; }
    return VS_OK;
}

}  // namespace vs
