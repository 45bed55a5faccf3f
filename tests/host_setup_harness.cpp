// host_setup_harness.cpp -- C entry points over the library's host-only
// set-up code (vocalsim-2p5d-fdtd_amd/csrc/vs_host.cpp) for the CPU tests,
// which build the two files together, plainly and with AddressSanitizer and
// UndefinedBehaviorSanitizer.  The maximum chord of the single-plane
// contours, which the library finds on the device, is found here by a CPU
// search in the device kernel's order (for each point the first strict
// maximum over the points in index order, distances as sqrtf(dy^2 + dz^2)
// with no contraction).

#include "../vocalsim-2p5d-fdtd_amd/csrc/vs_host.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

int cpu_max_chord(const std::vector<float>& y, const std::vector<float>& z, int sections, int npt,
                  std::vector<float>& dist, std::vector<int>& far_idx, std::string& err) {
    const size_t np = (size_t)sections * npt;
    if (y.size() != np || z.size() != np) { err = "contour size mismatch"; return VS_E_INVALID; }
    dist.assign(np, 0.0f);
    far_idx.assign(np, 0);
    for (int s = 0; s < sections; ++s) {
        const float* ys = y.data() + (size_t)s * npt;
        const float* zs = z.data() + (size_t)s * npt;
        for (int i = 0; i < npt; ++i) {
            float best = 0.0f;
            int arg = 0;
            for (int j = 0; j < npt; ++j) {
                const float dy = ys[i] - ys[j], dz = zs[i] - zs[j];
                const float d = std::sqrt(dy * dy + dz * dz);
                if (d > best) { best = d; arg = j; }
            }
            dist[(size_t)s * npt + i] = best;
            far_idx[(size_t)s * npt + i] = arg;
        }
    }
    return VS_OK;
}

void put_err(const std::string& e, char* err, int errlen) {
    if (!err || errlen <= 0) return;
    std::strncpy(err, e.c_str(), (size_t)errlen - 1);
    err[errlen - 1] = 0;
}

}  // namespace

extern "C" {

// Prelude of one tube.  dims[15] in vs_dims order; sec[5 * n] per section:
// cumulative length, diameter in cells, minor axis, z_inv, depth scale.
int hs_prelude(const vs_tube_spec* s, int fp64, int* dims, double* sec, int sec_cap, char* err, int errlen) {
    vs::Prelude p{};
    std::string e;
    const int rc = vs::prelude(*s, fp64 != 0, p, cpu_max_chord, e);
    if (rc != VS_OK) { put_err(e, err, errlen); return rc; }
    const int d[15] = {p.H, p.W, p.domainH, p.domainW, p.listenerX, p.listenerY, p.excX, p.excY, p.num_steps,
                       p.ts, p.te, p.L, p.midY, s->open_space ? 0 : s->num_sections, p.mic};
    std::memcpy(dims, d, sizeof d);
    const int n = (int)p.diam.size();
    if (n > sec_cap) { put_err("section buffer too small", err, errlen); return VS_E_RANGE; }
    for (int i = 0; i < n; ++i) {
        sec[5 * i + 0] = p.cum[i];
        sec[5 * i + 1] = p.diam[i];
        sec[5 * i + 2] = p.minor[i];
        sec[5 * i + 3] = p.zsec[i];
        sec[5 * i + 4] = p.dscale[i];
    }
    return VS_OK;
}

// Aligned contour of one single-plane-symmetric section (kPlanePoints
// points) and info[5] = minY, minZ, maxZ, diameter, growth steps.
int hs_plane_section(float area_cm2, float* y, float* z, double* info, char* err, int errlen) {
    std::vector<vs::PlaneSection> out;
    std::string e;
    const int rc = vs::plane_sections(&area_cm2, 1, out, cpu_max_chord, e);
    if (rc != VS_OK) { put_err(e, err, errlen); return rc; }
    const vs::PlaneSection& P = out[0];
    std::memcpy(y, P.y.data(), P.y.size() * sizeof(float));
    std::memcpy(z, P.z.data(), P.z.size() * sizeof(float));
    info[0] = P.minY; info[1] = P.minZ; info[2] = P.maxZ; info[3] = P.diameter; info[4] = P.grow_steps;
    return (int)P.y.size();
}

int hs_plane_points(void) { return vs::kPlanePoints; }
int hs_exact_max(void) { return vs::kExactClasses; }
int hs_fast_max(void) { return vs::kFastClasses; }

int hs_exact_classes(const uint8_t* cls, int H, int W, uint8_t* xcls, int* keys, int* nk, char* err, int errlen) {
    std::string e;
    const int rc = vs::enumerate_exact_classes(cls, H, W, xcls, keys, *nk, e);
    if (rc != VS_OK) put_err(e, err, errlen);
    return rc;
}

// Cells [bounds[g], bounds[g + 1]) read their damping factors from
// luts[8 g .. 8 g + 8), g < groups.
int hs_fast_classes(const uint16_t* code, const long* bounds, const float* luts, int groups, uint8_t* c8, float* tab,
                    int* ncls, char* err, int errlen) {
    std::string e;
    std::vector<size_t> b(bounds, bounds + groups + 1);
    const int rc = vs::enumerate_fast_classes(code, b.data(), luts, groups, c8, tab, *ncls, e);
    if (rc != VS_OK) put_err(e, err, errlen);
    return rc;
}

int hs_fast_pairs_max(void) { return vs::kFastPairs; }

// Pair classes of the wide-access on-chip step over `groups` tube frames
// H[g] x W[g] whose class bytes start at off[g] in c8, and each tube's
// source column.
int hs_fast_pairs(const uint8_t* c8, const float* tab, const int* H, const int* W, const long long* off, int groups,
                  uint8_t* p8, float* ptab, int* npairs, int* src_col, char* err, int errlen) {
    std::string e;
    const int rc = vs::enumerate_fast_pairs(c8, tab, H, W, off, groups, p8, ptab, *npairs, src_col, e);
    if (rc != VS_OK) put_err(e, err, errlen);
    return rc;
}

// Reads one element past the end of a heap array: the sanitizer build must
// report it (the check that the instrumentation is live).
int hs_planted_overflow(int n) {
    std::vector<int> v((size_t)n, 1);
    volatile const int* p = v.data();
    return p[n];
}

}  // extern "C"
