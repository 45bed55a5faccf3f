// vs_host.cpp -- host-only set-up code of the library (no HIP): the float32
// prelude that integerises one tube, the single-plane-symmetric contours, and
// the cell-class enumeration of both step forms.
//
// Prelude: O(sections) float32 integerisation of one tube
// (TubeGeometry.cu:53-179, 241-286; FDTDSolver.cu:479-617).  Unit constants
// are float32 as in include/SimulationUnits.h:11-25.
//
// Single-plane contours (float32, the reference's operation order):
//   ShapeGenerator::combineRegularShapes   src/ShapeGenerator.cpp:35-197
//   ShapeGenerator::generateEllipseCoordinates / finalContourCoordinates
//                                          src/ShapeGenerator.cpp:220-318
//   Utils::linspace / computeTotalArea     src/Utils.cpp:59-77, 104-122
//   ShapeAlignment::rotateShapeAboutOrigin src/ShapeAlignment.cu:73-113
// Boost.Geometry's linestring intersection and ring area are replaced by a
// segment-pair intersection and the shoelace formula, both evaluated in
// double and rounded to float.

#include "vs_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace vs {

namespace {

constexpr float kCm = 1e-2f;
constexpr float kMm = 1e-3f;

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

// Open-space domain (FDTDSolver.cu:479-550): 43 x 341 air cells inside the
// dead ring and PML, the excitation cell at the frame centre (integer halves,
// as `ceil(frameH / 2) - 1` evaluates on ints) and the listener 5 cells below
// and to the right of it.  No tube, no walls, constant depth.
int prelude_open_space(const vs_tube_spec& s, bool fp64, Prelude& p, std::string& err) {
    phys_tables(s.srate_mul, fp64, p.ph, p.ds32);
    p.num_steps = s.num_steps > 0 ? s.num_steps : float_step_count(s.audio_dur, 1.0f / (float)(44100 * s.srate_mul));
    if (p.num_steps <= 0) { err = "number of steps must be positive"; return VS_E_INVALID; }
    const int pml = s.pml_flag ? 1 : 0;
    p.domainH = 43;
    p.domainW = 341;
    p.H = p.domainH + 2 + 12 * pml;
    p.W = p.domainW + 2 + 12 * pml;
    p.excY = p.H / 2 - 1;
    p.excX = p.W / 2 - 1;
    p.listenerY = p.excY + 5;
    p.listenerX = p.excX + 5;
    p.L = 0;
    p.ts = 1;
    p.te = 0;
    p.midY = p.excY;
    p.mic = 0;
    return VS_OK;
}

int invalid(std::string& err, const char* msg) {
    err = msg;
    return VS_E_INVALID;
}

}  // namespace

void phys_tables(int srate_mul, bool fp64, PhysTables& t, float& ds32) {
    const float c32 = 350.0f, rho32 = 1.14f;
    const int srate = 44100 * srate_mul;
    const float dt32 = 1.0f / (float)srate;
    ds32 = dt32 * c32 * sqrtf(2.0f);
    if (fp64) {
        t.c = 350.0; t.rho = 1.14;
        t.dt = 1.0 / (double)srate;
        t.ds = t.dt * t.c * std::sqrt(2.0);
        const double kappa = t.rho * t.c * t.c;
        t.K1 = (kappa * t.dt) / t.ds;
        t.K2 = kappa * t.dt;
        t.rds = t.rho * t.ds;
    } else {
        const float kappa = rho32 * c32 * c32;
        t.c = c32; t.rho = rho32; t.dt = dt32; t.ds = ds32;
        t.K1 = (float)((kappa * dt32) / ds32);
        t.K2 = (float)(kappa * dt32);
        t.rds = (float)(rho32 * ds32);
    }
    for (int k = 0; k < 16; ++k) { t.beta[k] = 0; t.sigma[k] = 0; }
    t.beta[kAir] = 1; t.beta[kNoPressure] = 1;
    for (int k = 0; k < 6; ++k) t.beta[kPml0 + k] = 1;
    t.sigma[kWall] = t.dt; t.sigma[kExcitation] = t.dt; t.sigma[kHead] = t.dt;
    t.sigma[kDead] = 1000000;
    for (int k = 0; k < 6; ++k) {
        if (fp64) t.sigma[kPml0 + k] = ((double)k / (double)5) * 0.5;
        else t.sigma[kPml0 + k] = ((float)k / (float)5) * 0.5f;
    }
    t.air_normal[0] = 0; t.air_normal[1] = 1; t.air_normal[2] = 1;
    t.air_normal[3] = fp64 ? 0.707106 : (double)0.707106f;
    t.wall_normal[0] = 1; t.wall_normal[1] = 1; t.wall_normal[2] = 1; t.wall_normal[3] = 0;
    t.src_dir[0] = 0; t.src_dir[1] = 0; t.src_dir[2] = 1; t.src_dir[3] = 0;
}

double z_inv_of(float mu, bool fp64, const PhysTables& t) {
    if (fp64) {
        const double m2 = (double)mu * ((2.0 * 0.5 * M_PI) / 1.84);
        const double a = 1.0 / (0.5 + 0.25 * (m2 + (1 / m2)));
        return 1 / (t.rho * t.c * ((1 + std::sqrt(1 - a)) / (1 - std::sqrt(1 - a))));
    }
    const float m2 = mu * ((2.0f * 0.5F * M_PI) / 1.84f);
    const float a = 1.0f / (0.5f + 0.25f * (m2 + (1 / m2)));
    const float rc = (float)t.rho * (float)t.c;
    return 1 / (rc * ((1 + sqrtf(1 - a)) / (1 - sqrtf(1 - a))));
}

int float_step_count(float dur, float dt) {
    int n = 0;
    for (float t = 0; t <= dur - dt; t = t + dt) ++n;
    return n;
}

int prelude(const vs_tube_spec& s, bool fp64, Prelude& p, const MaxChordFn& max_chord, std::string& err) {
    if (s.srate_mul <= 0) return invalid(err, "srate_mul must be positive");
    if (s.fdtd_solver_type != 0 && s.fdtd_solver_type != 1) return invalid(err, "fdtd_solver_type must be 0 or 1");
    if (s.source_type < 0 || s.source_type > 2) return invalid(err, "Invalid source type");
    if (s.open_space) return prelude_open_space(s, fp64, p, err);
    if (s.num_sections <= 0 || !s.length_cm || !s.area_cm2) return invalid(err, "empty area function");
    if (s.cross_sectional_shape < 1 || s.cross_sectional_shape > 3)
        return invalid(err, "cross_sectional_shape must be 1 (circular), 2 (elliptical) or 3 (single-plane symmetric)");
    const int n = s.num_sections;
    for (int i = 0; i < n; ++i)
        if (!(s.area_cm2[i] > 0.0f) || !(s.length_cm[i] > 0.0f))
            return invalid(err, "section area and length must be positive");
    phys_tables(s.srate_mul, fp64, p.ph, p.ds32);
    const float ds = p.ds32;
    p.num_steps = s.num_steps > 0 ? s.num_steps : float_step_count(s.audio_dur, 1.0f / (float)(44100 * s.srate_mul));
    if (p.num_steps <= 0) return invalid(err, "number of steps must be positive");
    // sagittal widths (circle / q:1 ellipse) and minor axes
    std::vector<float> width(n);
    p.minor.resize(n);
    for (int i = 0; i < n; ++i) {
        const float a = s.area_cm2[i] * kCm * kCm;
        if (s.cross_sectional_shape == 3) {
            width[i] = 0.0f;   // maximum chord of the contour, below
            p.minor[i] = 0.0;
        } else if (s.cross_sectional_shape == 1) {
            width[i] = 2.0 * sqrt(a / M_PI);
            p.minor[i] = fp64 ? 2.0 * std::sqrt((double)s.area_cm2[i] * 1e-2 * 1e-2 / M_PI) : (double)width[i];
        } else {
            const float q = s.aspect ? s.aspect[i] : 3.0f;
            if (!(q > 0.0f)) return invalid(err, "ellipse aspect must be positive");
            const float lx = sqrtf(a / (q * 1.0f * M_PI));
            width[i] = lx * (q * 2);
            if (fp64) p.minor[i] = std::sqrt((double)s.area_cm2[i] * 1e-2 * 1e-2 / ((double)q * 1.0 * M_PI)) * 2.0;
            else p.minor[i] = lx * (1.0f * 2);
        }
    }
    if (s.cross_sectional_shape == 3) {
        const int rc = plane_sections(s.area_cm2, n, p.plane, max_chord, err);
        if (rc != VS_OK) return rc;
        for (int i = 0; i < n; ++i) width[i] = p.plane[i].diameter;
    }
    if (s.fdtd_solver_type == 0) {
        // 2D area scaling (TubeGeometry.cu:241-286); pow(x, 2) as x * x
        const float mx = *std::max_element(width.begin(), width.end());
        size_t mi = 0;
        for (size_t i = 0; i < width.size(); ++i) if (width[i] == mx) mi = i;
        std::vector<float> sc(n, 0.0f), m(n, 1.0f);
        sc[mi] = (mx * 0.5 * M_PI) / 1.84f;
        for (int i = 0; i < n; ++i) {
            if ((size_t)i < mi) { const float q = width[i] / width[i + 1]; m[i] = width[i + 1] != 0.0f ? q * q : 1.0f; }
            if ((size_t)i > mi) { const float q = width[i] / width[i - 1]; m[i] = width[i - 1] != 0.0f ? q * q : 1.0f; }
        }
        for (int i = (int)mi - 1; i >= 0; --i) sc[i] = sc[i + 1] * m[i];
        for (int i = (int)mi + 1; i < n; ++i) sc[i] = sc[i - 1] * m[i];
        width = sc;
    }
    p.mic = (int)roundf((3 * kMm) / ds);
    float total = 0.0f;
    p.cum.resize(n);
    for (int i = 0; i < n; ++i) { total = total + s.length_cm[i] * kCm; p.cum[i] = total; }
    p.L = (int)roundf(total / ds);
    if (p.L < 2) return invalid(err, "tube shorter than two cells");
    p.diam.resize(n);
    for (int i = 0; i < n; ++i) {
        int d = (int)roundf(width[i] / ds);
        if (d == 0) d = 1;
        if (d % 2 == 0) {
            const float over = (float)d - width[i] / ds;
            d = (over > 0 || s.cross_sectional_shape >= 3) ? d - 1 : d + 1;
        }
        p.diam[i] = d;
    }
    const int pml = s.pml_flag ? 1 : 0;
    const int maxd = *std::max_element(p.diam.begin(), p.diam.end());
    p.domainW = s.mouth_radiation_cond == 1 ? p.L + 2 : p.L + 1 + p.mic;
    p.domainH = maxd + 2;
    p.W = p.domainW + 2 + 12 * pml;
    p.H = p.domainH + 2 + 12 * pml;
    p.ts = 2 + 6 * pml;
    p.te = p.ts + p.L - 1;
    p.midY = (int)ceilf(p.H / 2.0f) - 1;
    p.listenerY = p.midY;
    p.listenerX = s.mouth_radiation_cond == 1 ? p.te - p.mic : p.te + 3 + p.mic;
    if (p.listenerX > p.W - 2) return invalid(err, "listener falls outside the frame (open mouth without PML)");
    p.excX = p.ts - 1;
    p.excY = p.midY;
    const float mu_def = s.mu3d_default > 0 ? s.mu3d_default : (s.fdtd_solver_type == 1 ? 0.025f : 0.005f);
    p.zsec.resize(n);
    p.dscale.resize(n);
    for (int i = 0; i < n; ++i) {
        const float mu = s.mu3d ? s.mu3d[i] : mu_def;
        if (!(mu > 0.0f)) return invalid(err, "wall admittance must be positive");
        p.zsec[i] = z_inv_of(mu, fp64, p.ph);
        const float sc = s.depth_scale ? s.depth_scale[i] : 1.0f;
        if (!(sc > 0.0f)) return invalid(err, "depth scale must be positive");
        p.dscale[i] = sc;
    }
    return VS_OK;
}

int plane_sections(const float* area_cm2, int n, std::vector<PlaneSection>& out, const MaxChordFn& max_chord,
                   std::string& err) {
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
        const float yc2 = yc1 + a1 + 0.003f * kMm;
        std::vector<float> hy, hz;
        bool ok = false;
        joined_half(a1, b1, a2, b2, yc1, yc2, nh, hy, hz, ok);
        if (!ok) { err = "section " + std::to_string(s) + ": contour shapes do not intersect"; return VS_E_GEOMETRY; }
        float area = 2 * ring_area(hy, hz);
        const float grow = kMm * 0.01f;
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
    const int npt = kPlanePoints;
    const size_t np = (size_t)n * npt;
    std::vector<float> hd(np);
    std::vector<int> hi(np);
    const int rc = max_chord(all_y, all_z, n, npt, hd, hi, err);
    if (rc != VS_OK) return rc;
    if (hd.size() != np || hi.size() != np) { err = "max chord: wrong output size"; return VS_E_INVALID; }
    for (int s = 0; s < n; ++s) {
        PlaneSection& P = out[s];
        const float* d = hd.data() + (size_t)s * npt;
        const int* f = hi.data() + (size_t)s * npt;
        float best = 0.0f;
        int arg = 0;
        for (int i = 0; i < npt; ++i)
            if (d[i] > best) { best = d[i]; arg = i; }
        if (f[arg] < 0 || f[arg] >= npt) { err = "max chord: far index out of range"; return VS_E_INVALID; }
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
    return VS_OK;
}

int enumerate_exact_classes(const uint8_t* cls, int H, int W, uint8_t* xcls, int* keys, int& nk, std::string& err) {
    // a key holds 4 + 1 + 4 + 4 bits
    std::vector<int> id(1 << 13, -1);
    nk = 0;
    for (int r = 1; r < H - 1; ++r)
        for (int col = 1; col < W - 1; ++col) {
            const int q = r * W + col;
            const int key = (cls[q] & kTypeMask) | ((cls[q] & kSegWallBit) ? 16 : 0) | ((cls[q + 1] & kTypeMask) << 5) |
                            ((cls[q - W] & kTypeMask) << 9);
            if (id[key] < 0) {
                if (nk == kExactClasses) {
                    err = "more than " + std::to_string(kExactClasses) + " cell classes";
                    return VS_E_GEOMETRY;
                }
                id[key] = nk;
                keys[nk] = key;
                ++nk;
            }
            xcls[q] = (uint8_t)id[key];
        }
    return VS_OK;
}

int enumerate_fast_classes(const uint16_t* code, const size_t* bounds, const float* luts, int groups, uint8_t* c8,
                           float* tab, int& ncls, std::string& err) {
    if (groups < 1 || bounds[0] != 0) {
        err = "fast classes: cell groups must start at cell 0";
        return VS_E_INVALID;
    }
    for (int g = 0; g < groups; ++g)
        if (bounds[g + 1] < bounds[g]) {
            err = "fast classes: cell group bounds must not decrease";
            return VS_E_INVALID;
        }
    constexpr int kZero = 1 + 8 * 1 + 64 * 1;   // every damping factor zero, no face-form bit
    // a class is the triple of factors (aP, ax, ay) with the two face-form bits
    struct Key {
        float p, x, y;
        int bits;
        bool operator==(const Key& o) const {
            return std::memcmp(&p, &o.p, sizeof p) == 0 && std::memcmp(&x, &o.x, sizeof x) == 0 &&
                   std::memcmp(&y, &o.y, sizeof y) == 0 && bits == o.bits;
        }
    };
    std::vector<Key> keys;
    auto key_of = [](const float* lut, int d) {
        return Key{lut[(d >> 6) & 7], lut[d & 7], lut[(d >> 3) & 7], d & 0x600};
    };
    keys.push_back(key_of(luts, kZero));   // class 0 (a zero factor is the same in every table)
    std::vector<int> cls_of(2048);
    for (int g = 0; g < groups; ++g) {
        const float* lut = luts + 8 * g;
        std::fill(cls_of.begin(), cls_of.end(), -1);
        cls_of[kZero] = 0;
        for (size_t i = bounds[g]; i < bounds[g + 1]; ++i) {
            const int d = (code[i] >> 3) & 0x7ff;
            if (cls_of[d] < 0) {
                const Key k = key_of(lut, d);
                int c = 0;
                while (c < (int)keys.size() && !(keys[c] == k)) ++c;
                if (c == (int)keys.size()) {
                    if (c == kFastClasses) {
                        err = "more than " + std::to_string(kFastClasses - 1) + " cell classes in one batch";
                        return VS_E_INVALID;
                    }
                    keys.push_back(k);
                }
                cls_of[d] = c;
            }
            c8[i] = (uint8_t)cls_of[d];
        }
    }
    ncls = (int)keys.size();
    std::fill(tab, tab + 5 * kFastClasses, 0.0f);
    for (int c = 0; c < ncls; ++c) {
        const Key& k = keys[c];
        const bool s0 = (k.bits & 0x200) != 0, r0 = (k.bits & 0x400) != 0;
        tab[c] = k.p;
        tab[kFastClasses + c] = s0 ? -k.x : k.x;
        tab[2 * kFastClasses + c] = r0 ? -k.y : k.y;
        tab[3 * kFastClasses + c] = s0 ? 0.0f : 1.0f;
        tab[4 * kFastClasses + c] = r0 ? 0.0f : 1.0f;
    }
    return VS_OK;
}

int enumerate_fast_pairs(const uint8_t* c8, const float* tab, const int* H, const int* W, const long long* off,
                         int groups, uint8_t* p8, float* ptab, int& npairs, int* src_col, std::string& err) {
    std::vector<int> id(kFastClasses * kFastClasses, -1);
    std::vector<int> first, second;   // classes of the pair classes, in first-seen order
    auto pair_id = [&](int a, int b) {
        int& p = id[a * kFastClasses + b];
        if (p < 0) {
            p = (int)first.size();
            first.push_back(a);
            second.push_back(b);
        }
        return p;
    };
    pair_id(0, 0);   // pair class 0: two pad cells (every pair past a row's last real cell)
    for (int g = 0; g < groups; ++g) {
        if (H[g] < 3 || W[g] < 3) {
            err = "fast pairs: tube frames must be at least 3 x 3";
            return VS_E_INVALID;
        }
        const int Hi = H[g] - 2, Wi = W[g] - 2;
        src_col[g] = -1;
        auto note_source = [&](int cls, int col) {
            if (tab[4 * kFastClasses + cls] != 0.0f) return;   // r = 1: no source on this face
            if (src_col[g] == -1) src_col[g] = col;
            else if (src_col[g] != col) src_col[g] = -2;
        };
        for (int r = 0; r < Hi; ++r)
            for (int c = 0; c < Wi; c += 2) {
                const long long q = off[g] + (long long)(r + 1) * W[g] + (c + 1);
                const int a = c8[q], b = c + 1 < Wi ? c8[q + 1] : 0;
                if (a >= kFastClasses || b >= kFastClasses) {
                    err = "fast pairs: class byte out of range";
                    return VS_E_INVALID;
                }
                const int p = pair_id(a, b);
                if (p < kFastPairs) p8[q] = (uint8_t)p;
                note_source(a, c);
                if (c + 1 < Wi) note_source(b, c + 1);
            }
    }
    npairs = (int)first.size();
    if (npairs > kFastPairs) return VS_OK;
    std::fill(ptab, ptab + 6 * kFastPairs, 0.0f);
    for (int p = 0; p < npairs; ++p) {
        const int a = first[p], b = second[p];
        ptab[4 * p + 0] = tab[kFastClasses + a];
        ptab[4 * p + 1] = tab[2 * kFastClasses + a];
        ptab[4 * p + 2] = tab[kFastClasses + b];
        ptab[4 * p + 3] = tab[2 * kFastClasses + b];
        ptab[4 * kFastPairs + 2 * p + 0] = tab[a];
        ptab[4 * kFastPairs + 2 * p + 1] = tab[b];
    }
    return VS_OK;
}

}  // namespace vs
