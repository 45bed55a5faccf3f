/*
 * vs_oracle.cpp -- CPU oracle: a plain C++ restatement of the reference's
 * 2.5D FDTD vocal-tract setup path and FDTDEngineCPU time step.
 *
 * TEST INFRASTRUCTURE ONLY (see vs_oracle.h).  Written from the reference
 * source; every block cites the reference file:line it follows.  Compiled with
 * -ffp-contract=off so that every a*b+c rounds twice, as the reference's host
 * code does.
 *
 * Precision model (follows the reference where it is float32):
 *   - geometry integerisation (diameters in cells, tube length in cells, wall
 *     rows, segment lookup) is always float32, exactly as TubeGeometry.cu does;
 *   - depths, coefficients, z_inv, sources and the step run in Real, which is
 *     float for the float32 instantiation (reference semantics, including its
 *     mixed float/double expressions) and double for the FP64 oracle.
 *
 * Parity pins: see tests/test_oracle.py (closed forms, invariants, the
 * z_inv values quoted from the reference, brute-force triples).  Functions
 * without an independent pin are marked "parity unpinned" below.
 */
#include "vs_oracle.h"

/* Planted mistakes for the mutation tests (tests/test_oracle_walls.py builds
 * a copy of this file with -DVSO_MUTATION=n into a temporary directory and
 * checks that the wall pins reject it).  0, the default, plants nothing.
 *   1: the wall-admittance velocity Vb enters with the wrong sign;
 *   2: Vb drops its inward term (xor1 P_right N_in, xor3 P_top N_in);
 *   3: z_inv is taken from whichever cell of the face is a wall, instead of
 *      the current cell only (FDTDEngineCPU.cpp:397-398);
 *   4: the probe reads one element past the end of the pressure array (a
 *      memory error the sanitizer build must report). */
#ifndef VSO_MUTATION
#define VSO_MUTATION 0
#endif

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

/* include/SimulationUnits.h:11-25 -- const float units. */
namespace fu {
const float METER = 1;
const float CENTIMETER = 1e-2 * METER;
const float MILLIMETER = 1e-3 * METER;
const float SECOND = 1;
const float HERTZ = 1 / SECOND;
const float KILOHERTZ = 1e3 * HERTZ;
const float KILOGRAM = 1;
const float GRAM = 1e-3 * KILOGRAM;
const float NEWTON = KILOGRAM * METER / (SECOND * SECOND);
const float DYN = 1e-5 * NEWTON;
}  // namespace fu

/* include/SolverParams.h:21-35 */
enum CellType {
    cell_wall = 0, cell_air, cell_excitation, cell_pml0, cell_pml1, cell_pml2,
    cell_pml3, cell_pml4, cell_pml5, cell_dynamic, cell_dead, cell_noPressure,
    cell_head, NUM_CELL_TYPES
};

/* ------------------------------------------------------------------ */
/* Physics constants: SimulationParams.cu:7-33, FDTDSolver.cu:296-312   */
/* ------------------------------------------------------------------ */
template <typename Real>
struct Phys {
    Real c, rho, kappa, dt, ds, K1, K2;
    int srate;
    Real max_sigma_dt;
    int pml_layers;
    Real beta[NUM_CELL_TYPES];
    Real sigma[NUM_CELL_TYPES];
    Real src_dir[4];
    Real air_normal[4];
    Real wall_normal[4];
};

/* float32 instance: exactly the reference's float expressions. */
Phys<float> phys_f32(int srate_mul) {
    Phys<float> p{};
    p.c = 350.0f * (fu::METER / fu::SECOND);
    p.rho = (1.14f * fu::KILOGRAM) / std::pow(fu::METER, 3);
    int base = (int)(44100 * fu::HERTZ);
    p.kappa = p.rho * p.c * p.c;
    p.srate = base * srate_mul;
    p.dt = 1.0f / static_cast<float>(p.srate);
    p.ds = p.dt * p.c * sqrtf(2.0f);
    p.max_sigma_dt = 0.5;
    p.pml_layers = 6;
    p.K1 = (p.kappa * p.dt) / p.ds;
    p.K2 = p.kappa * p.dt;
    return p;
}

/* float64 instance: the same formulas evaluated in double. */
Phys<double> phys_f64(int srate_mul) {
    Phys<double> p{};
    p.c = 350.0;
    p.rho = 1.14;
    p.kappa = p.rho * p.c * p.c;
    p.srate = 44100 * srate_mul;
    p.dt = 1.0 / static_cast<double>(p.srate);
    p.ds = p.dt * p.c * std::sqrt(2.0);
    p.max_sigma_dt = 0.5;
    p.pml_layers = 6;
    p.K1 = (p.kappa * p.dt) / p.ds;
    p.K2 = p.kappa * p.dt;
    return p;
}

/* SolverParams.cu:5-76 and DeviceSymbols.cu:19-20. */
template <typename Real>
void fill_tables(Phys<Real>& p) {
    for (int k = 0; k < NUM_CELL_TYPES; ++k) { p.beta[k] = 0; p.sigma[k] = 0; }
    p.beta[cell_wall] = 0; p.beta[cell_air] = 1; p.beta[cell_excitation] = 0;
    p.beta[cell_dynamic] = 0; p.beta[cell_dead] = 0; p.beta[cell_noPressure] = 1;
    p.beta[cell_head] = 0;
    for (int k = 0; k < p.pml_layers; ++k) p.beta[cell_pml0 + k] = 1;
    p.sigma[cell_wall] = p.dt; p.sigma[cell_air] = 0; p.sigma[cell_noPressure] = 0;
    p.sigma[cell_excitation] = p.dt; p.sigma[cell_head] = p.dt; p.sigma[cell_dynamic] = 0;
    p.sigma[cell_dead] = 1000000;
    for (int k = 0; k < p.pml_layers; ++k)
        p.sigma[cell_pml0 + k] = ((Real)k / (Real)(p.pml_layers - 1)) * p.max_sigma_dt;
    p.src_dir[0] = 0; p.src_dir[1] = 0; p.src_dir[2] = 1; p.src_dir[3] = 0;
    p.air_normal[0] = 0; p.air_normal[1] = 1.0; p.air_normal[2] = 1.0; p.air_normal[3] = (Real)0.707106;
    p.wall_normal[0] = 1.0; p.wall_normal[1] = 1.0; p.wall_normal[2] = 1.0; p.wall_normal[3] = 0;
}

/* SolverParams.cu:59-66: count of float-accumulated time stamps. */
int float_num_steps(float audio_dur, float dt) {
    float t = 0;
    int n = 0;
    while (t <= audio_dur - dt) { ++n; t = t + dt; }
    return n;
}

/* ------------------------------------------------------------------ */
/* Geometry (float32 integerisation): TubeGeometry.cu:53-179, 182-237,  */
/* 241-286, 316-364, 433-612                                             */
/* ------------------------------------------------------------------ */
/* ------------------------------------------------------------------ */
/* Single-plane-symmetric contour (ShapeGenerator.cpp:35-318,           */
/* ShapeAlignment.cu:11-152, Utils.cpp:59-122).  Boost.Geometry's        */
/* linestring intersection and ring area are restated as the crossing of */
/* two segments and the shoelace sum, both in double rounded to float.   */
/* ------------------------------------------------------------------ */
struct PlaneContour {
    std::vector<float> y, z;
    float minY = 0, minZ = 0, maxZ = 0, diameter = 0, area = 0;
    int steps = 0;
};

std::vector<float> linspace(float a, float b, int n) {
    std::vector<float> v;
    if (n == 0) return v;
    if (n == 1) { v.push_back(a); return v; }
    float step = (b - a) / (n - 1);
    for (int i = 0; i < n; ++i) v.push_back(a + step * i);
    return v;
}

struct Pt { float y, z; };

std::vector<Pt> ellipse_upper(float y0, float y1, float a, float b, int n, float yc) {
    std::vector<float> ys = linspace(y0, y1, n);
    std::vector<float> dd(n);
    for (int i = 0; i < n; ++i) { float t = ys[i] - yc; dd[i] = t * t; }
    float aa = a * a, bb = b * b;
    dd[0] = aa;
    for (int i = 0; i < n / 2; ++i) dd[n - i - 1] = dd[i];
    std::vector<Pt> out(n);
    for (int i = 0; i < n; ++i) out[i] = {ys[i], sqrtf(std::max(0.0f, bb * (1.0f - dd[i] / aa)))};
    return out;
}

/* crossing point of segments p0p1 and q0q1, if any */
bool seg_cross(Pt p0, Pt p1, Pt q0, Pt q1, Pt& out) {
    double rx = (double)p1.y - p0.y, rz = (double)p1.z - p0.z;
    double sx = (double)q1.y - q0.y, sz = (double)q1.z - q0.z;
    double den = rx * sz - rz * sx;
    if (den == 0.0) return false;
    double qx = (double)q0.y - p0.y, qz = (double)q0.z - p0.z;
    double u = (qx * sz - qz * sx) / den;
    double v = (qx * rz - qz * rx) / den;
    if (u < 0.0 || u > 1.0 || v < 0.0 || v > 1.0) return false;
    out = {(float)((double)p0.y + u * rx), (float)((double)p0.z + u * rz)};
    return true;
}

/* first crossing of two polylines in the order of their segments */
bool polyline_cross(const std::vector<Pt>& a, const std::vector<Pt>& b, Pt& out) {
    for (size_t i = 0; i + 1 < a.size(); ++i)
        for (size_t j = 0; j + 1 < b.size(); ++j) {
            if (std::max(b[j].y, b[j + 1].y) < std::min(a[i].y, a[i + 1].y)) continue;
            if (std::min(b[j].y, b[j + 1].y) > std::max(a[i].y, a[i + 1].y)) continue;
            if (seg_cross(a[i], a[i + 1], b[j], b[j + 1], out)) return true;
        }
    return false;
}

float shoelace(const std::vector<Pt>& r) {
    double acc = 0;
    for (size_t i = 0; i < r.size(); ++i) {
        const Pt& p = r[i];
        const Pt& q = r[(i + 1) % r.size()];
        acc += (double)p.y * q.z - (double)q.y * p.z;
    }
    return (float)std::fabs(acc * 0.5);
}

/* lowest and highest crossing of the vertical line y = y0 with a polyline */
bool vertical_chord(const std::vector<float>& y, const std::vector<float>& z, float y0, float& lo, float& hi) {
    lo = std::numeric_limits<float>::infinity();
    hi = -std::numeric_limits<float>::infinity();
    for (size_t k = 0; k + 1 < y.size(); ++k) {
        float ya = y[k], yb = y[k + 1];
        if (y0 < std::min(ya, yb) || y0 > std::max(ya, yb)) continue;
        if (ya == yb) {
            lo = std::min(lo, std::min(z[k], z[k + 1]));
            hi = std::max(hi, std::max(z[k], z[k + 1]));
            continue;
        }
        double t = ((double)y0 - ya) / ((double)yb - ya);
        float zc = (float)((double)z[k] + ((double)z[k + 1] - z[k]) * t);
        lo = std::min(lo, zc);
        hi = std::max(hi, zc);
    }
    return lo <= hi;
}

std::vector<Pt> joined_contour(float a1, float b1, float a2, float b2, float c1, float c2, int nh, bool& ok) {
    std::vector<Pt> e1 = ellipse_upper(c1 - a1, c1 + a1, a1, b1, nh, c1);
    std::vector<Pt> e2 = ellipse_upper(c2 - a2, c2 + a2, a2, b2, nh, c2);
    Pt ip{0, 0};
    ok = polyline_cross(e1, e2, ip);
    std::vector<Pt> out;
    if (!ok) return out;
    std::vector<float> ya = linspace(c1 - a1, ip.y, nh - 1), yb = linspace(ip.y, c2 + a2, nh - 1);
    for (float y : ya) out.push_back({y, sqrtf(std::max(0.0f, (b1 * b1) * (1.0f - ((y - c1) * (y - c1)) / (a1 * a1))))});
    out.push_back(ip);
    for (float y : yb) out.push_back({y, sqrtf(std::max(0.0f, (b2 * b2) * (1.0f - ((y - c2) * (y - c2)) / (a2 * a2))))});
    return out;
}

bool plane_contour(float area_cm2, PlaneContour& pc) {
    const int nh = (1999 + 1) / 2;
    float A = area_cm2 * fu::CENTIMETER * fu::CENTIMETER;
    float m = 3.0f, x = A / (m + 1.0f), A1 = m * x, A2 = 1.0f * x;
    float l1 = sqrtf(A1 / (M_PI * 3.0f * 1.0f)), l2 = sqrtf(A2 / (M_PI * 1.0f * 1.0f));
    float a1 = l1 * 3.0f, b1 = l1 * 1.0f, a2 = l2 * 1.0f, b2 = l2 * 1.0f;
    float c1 = 0.0f, c2 = c1 + a1 + 0.003f * fu::MILLIMETER;
    bool ok = false;
    std::vector<Pt> half = joined_contour(a1, b1, a2, b2, c1, c2, nh, ok);
    if (!ok) { g_err = "contour shapes do not intersect"; return false; }
    float area = 2 * shoelace(half);
    float step = fu::MILLIMETER * 0.01f;
    int steps = 0;
    while (area < A) {
        a1 = a1 + step; b1 = b1 + step; a2 = a2 + step; b2 = b2 + step;
        half = joined_contour(a1, b1, a2, b2, c1, c2, nh, ok);
        if (!ok) { g_err = "contour shapes do not intersect"; return false; }
        area = 2 * shoelace(half);
        ++steps;
    }
    half.front().z = 0.0f;
    half.back().z = 0.0f;
    std::vector<Pt> full = half;
    for (int i = (int)half.size() - 2; i >= 0; --i) full.push_back({half[i].y, -1 * half[i].z});
    /* maximum chord: first point attaining the largest distance, and its first farthest point */
    const int n = (int)full.size();
    float best = 0.0f;
    int bi = 0, bj = 0;
    for (int i = 0; i < n; ++i) {
        float mi = 0.0f;
        int mj = 0;
        for (int j = 0; j < n; ++j) {
            float dy = full[i].y - full[j].y, dz = full[i].z - full[j].z;
            float d = sqrtf(dy * dy + dz * dz);
            if (d > mi) { mi = d; mj = j; }
        }
        if (mi > best) { best = mi; bi = i; bj = mj; }
    }
    float theta = std::atan2(full[bj].z - full[bi].z, full[bj].y - full[bi].y);
    float c = std::cos(-1 * theta), sn = std::sin(-1 * theta);
    for (Pt& p : full) {
        float y = p.y, z = p.z;
        p.y = c * y - sn * z;
        p.z = sn * y + c * z;
    }
    float mnY = full[0].y, mnZ = full[0].z;
    for (const Pt& p : full) { mnY = std::min(mnY, p.y); mnZ = std::min(mnZ, p.z); }
    float shY = (mnY < 0.0f ? -mnY : 0.0f) + 0.05f, shZ = (mnZ < 0.0f ? -mnZ : 0.0f) + 0.05f;
    pc.y.clear(); pc.z.clear();
    for (Pt& p : full) { pc.y.push_back(p.y + shY); pc.z.push_back(p.z + shZ); }
    pc.minY = *std::min_element(pc.y.begin(), pc.y.end());
    pc.minZ = *std::min_element(pc.z.begin(), pc.z.end());
    pc.maxZ = *std::max_element(pc.z.begin(), pc.z.end());
    pc.diameter = best;
    pc.area = area;
    pc.steps = steps;
    return true;
}

struct Geometry {
    int nsec = 0;
    std::vector<PlaneContour> plane;
    std::vector<float> len_m, width_m, minor_m;  /* sagittal width, ellipse minor axis */
    std::vector<int> diam_cells;
    float tube_length = 0;
    int L = 0, mic_cells = 0;
    int domainH = 0, domainW = 0, frameH = 0, frameW = 0;
    int ts = 0, te = 0, midY = 0;
    std::vector<float> cum;
    std::vector<int> upper, lower, seg;
    int listenerX = 0, listenerY = 0, excX = 0, excY = 0;
};

/* TubeGeometry.cu:241-286 (2D area scaling).  pow(x, 2.0f) is evaluated as
 * the correctly rounded square x*x. */
std::vector<float> area_scaling_2d(const std::vector<float>& d) {
    float maxd = *std::max_element(d.begin(), d.end());
    size_t mi = 0;
    for (size_t i = 0; i < d.size(); ++i) if (d[i] == maxd) mi = i;
    std::vector<float> out(d.size(), 0.0f);
    out[mi] = (maxd * 0.5 * M_PI) / 1.84f;
    std::vector<float> m(d.size(), 1.0f);
    for (size_t i = 0; i < d.size(); ++i) {
        if (i < mi) { float r = d[i] / d[i + 1]; m[i] = (d[i + 1] != 0.0f) ? r * r : 1.0f; }
        else if (i > mi) { float r = d[i] / d[i - 1]; m[i] = (d[i - 1] != 0.0f) ? r * r : 1.0f; }
    }
    for (int i = (int)mi - 1; i >= 0; --i) out[i] = out[i + 1] * m[i];
    for (size_t i = mi + 1; i < d.size(); ++i) out[i] = out[i - 1] * m[i];
    return out;
}

bool build_geometry(const vso_spec& s, float ds, Geometry& g) {
    g.nsec = s.num_sections;
    if (g.nsec <= 0 || !s.length_cm || !s.area_cm2) { g_err = "empty area function"; return false; }
    for (int i = 0; i < g.nsec; ++i) {
        if (!(s.area_cm2[i] > 0.0f) || !(s.length_cm[i] > 0.0f)) {
            g_err = "section area and length must be positive"; return false;
        }
    }
    /* Widths: circle TubeGeometry.cu:316-330; ellipse :332-364 (ratio q:1). */
    for (int i = 0; i < g.nsec; ++i) {
        float a_m2 = s.area_cm2[i] * fu::CENTIMETER * fu::CENTIMETER;
        if (s.cross_sectional_shape == 1) {
            float d = 2.0 * sqrt(a_m2 / M_PI);
            g.width_m.push_back(d);
            g.minor_m.push_back(d);
        } else if (s.cross_sectional_shape == 2) {
            const float major = s.aspect ? s.aspect[i] : 3.0f;
            const float minor = 1.0f;
            float lenX = sqrtf(a_m2 / (major * minor * M_PI));
            g.width_m.push_back(lenX * (major * 2));
            g.minor_m.push_back(lenX * (minor * 2));
        } else if (s.cross_sectional_shape == 3) {
            PlaneContour pc;
            if (!plane_contour(s.area_cm2[i], pc)) return false;
            g.width_m.push_back(pc.diameter);
            g.minor_m.push_back(0.0f);
            g.plane.push_back(pc);
        } else {
            g_err = "cross_sectional_shape must be 1 (circular), 2 (elliptical) or 3 (single-plane symmetric)"; return false;
        }
    }
    for (int i = 0; i < g.nsec; ++i) g.len_m.push_back(s.length_cm[i] * fu::CENTIMETER);
    /* mic position TubeGeometry.cu:104-106 */
    float mic = 3 * fu::MILLIMETER;
    g.mic_cells = static_cast<int>(roundf(mic / ds));
    /* setComputationalDomain TubeGeometry.cu:112-179 */
    g.tube_length = std::accumulate(g.len_m.begin(), g.len_m.end(), 0.0f);
    g.L = static_cast<int>(roundf(g.tube_length / ds));
    std::vector<float> width = (s.fdtd_solver_type == 1) ? g.width_m : area_scaling_2d(g.width_m);
    for (float w : width) g.diam_cells.push_back(static_cast<int>(roundf(w / ds)));
    for (size_t i = 0; i < g.diam_cells.size(); ++i) {
        if (g.diam_cells[i] == 0) g.diam_cells[i]++;
        if (g.diam_cells[i] % 2 == 0) {
            float diff = static_cast<float>(g.diam_cells[i]) - (width[i] / ds);
            if (diff > 0 || s.cross_sectional_shape >= 3) g.diam_cells[i]--;
            else g.diam_cells[i]++;
        }
    }
    if (g.L < 2) { g_err = "tube shorter than two cells"; return false; }
    int maxd = *std::max_element(g.diam_cells.begin(), g.diam_cells.end());
    const int pml = s.pml_flag ? 1 : 0;
    if (s.mouth_radiation_cond == 1) g.domainW = g.L + 1 + 1;
    else g.domainW = g.L + 1 + g.mic_cells;
    g.domainH = maxd + 2;
    g.frameW = g.domainW + 2 + 2 * 6 * pml;
    g.frameH = g.domainH + 2 + 2 * 6 * pml;
    /* generateVocalTractWalls TubeGeometry.cu:191-195 */
    g.ts = 0 + 1 + 1 + 6 * pml;
    g.midY = static_cast<int>(ceilf(g.frameH / 2.0f)) - 1;
    g.te = g.ts + g.L - 1;
    /* inclusive scan of segment lengths (TubeGeometry.cu:206-216), sequential */
    g.cum.resize(g.nsec);
    float acc = 0.0f;
    for (int i = 0; i < g.nsec; ++i) { acc = acc + g.len_m[i]; g.cum[i] = acc; }
    /* setVocalTractWalls TubeGeometry.cu:443-471 */
    g.upper.resize(g.L); g.lower.resize(g.L); g.seg.resize(g.L);
    for (int idx = 0; idx < g.L; ++idx) {
        float curr_len = (idx + 1) * ds;
        float half = ds * 0.5f;
        int k = 0;
        while (k < g.nsec - 1 && curr_len > g.cum[k] + half) ++k;
        int radius = (g.diam_cells[k] - 1) / 2;
        g.upper[idx] = g.midY - radius - 1;
        g.lower[idx] = g.midY + radius + 1;
        g.seg[idx] = k;
    }
    /* listener / excitation FDTDSolver.cu:591-611 */
    if (s.mouth_radiation_cond == 1) {
        g.listenerX = g.te - g.mic_cells;
    } else {
        g.listenerX = g.te + 3 + g.mic_cells;
        if (g.listenerX > g.frameW - 2) {
            g_err = "listener falls outside the frame (open mouth without PML)"; return false;
        }
    }
    g.listenerY = g.midY;
    g.excX = g.ts - 1;
    g.excY = g.midY;
    return true;
}

/* Cell types and boundary_segment_type, in the reference's write order:
 * walls + gap fill (TubeGeometry.cu:433-569), excitation column (:571-594),
 * Dirichlet column (:596-612), dead ring (FDTDSolver.cu:816-835), PML rings
 * (FDTDSolver.cu:371-464, 838-857). */
/* openSpaceSimulation, FDTDSolver.cu:479-550: a 43 x 341 domain of air with
 * the excitation cell at (frameH / 2 - 1, frameW / 2 - 1) in integer
 * arithmetic and the listener 5 cells further down and right. */
void build_open_space(const vso_spec& s, Geometry& g) {
    const int pml = s.pml_flag ? 1 : 0;
    g.nsec = 0;
    g.domainH = 43;
    g.domainW = 341;
    g.frameH = g.domainH + 2 + 2 * pml * 6;
    g.frameW = g.domainW + 2 + 2 * pml * 6;
    g.excY = g.frameH / 2 - 1;
    g.excX = g.frameW / 2 - 1;
    g.listenerY = g.excY + 5;
    g.listenerX = g.excX + 5;
    g.L = 0; g.ts = 1; g.te = 0; g.midY = g.excY; g.mic_cells = 0;
}

void build_cells(const vso_spec& s, const Geometry& g, std::vector<int>& type, std::vector<int>& bst) {
    const int H = g.frameH, W = g.frameW;
    type.assign((size_t)H * W, cell_air);
    bst.assign((size_t)H * W, 0);
    if (s.open_space) {
        type[(size_t)g.excY * W + g.excX] = cell_excitation;
    }
    for (int i = 0; i < g.L; ++i) {
        int x = g.ts + i;
        type[(size_t)g.upper[i] * W + x] = cell_wall;
        type[(size_t)g.lower[i] * W + x] = cell_wall;
        bst[(size_t)g.upper[i] * W + x] = g.seg[i] + 1;
        bst[(size_t)g.lower[i] * W + x] = g.seg[i] + 1;
    }
    auto colmax = [&](int col) {
        int m = 0;
        for (int y = 0; y < H; ++y) m = std::max(m, bst[(size_t)y * W + col]);
        return m;
    };
    /* connectSagittalWallsGap: serial, as in the reference */
    if (s.open_space) goto frame;
    {
    int pu = g.upper[0], pl = g.lower[0], px = g.ts;
    for (int i = 1; i <= g.L - 2; ++i) {
        int cu = g.upper[i], cl = g.lower[i], cx = g.ts + i;
        int diff = cu - pu;
        if (diff > 1) {
            int sv = colmax(cx);
            for (int y = pu; y <= cu; ++y) { type[(size_t)y * W + cx] = cell_wall; bst[(size_t)y * W + cx] = sv; }
            for (int y = cl; y <= pl; ++y) { type[(size_t)y * W + cx] = cell_wall; bst[(size_t)y * W + cx] = sv; }
        } else if (diff < -1) {
            int sv = colmax(px);
            for (int y = cu; y <= pu; ++y) { type[(size_t)y * W + px] = cell_wall; bst[(size_t)y * W + px] = sv; }
            for (int y = pl; y <= cl; ++y) { type[(size_t)y * W + px] = cell_wall; bst[(size_t)y * W + px] = sv; }
        }
        pu = cu; pl = cl; px = cx;
    }
    /* setExcitationCells */
    int ex = g.ts - 1;
    for (int y = g.upper[0] + 1; y < g.lower[0]; ++y) type[(size_t)y * W + ex] = cell_excitation;
    type[(size_t)g.upper[0] * W + ex] = cell_wall;
    type[(size_t)g.lower[0] * W + ex] = cell_wall;
    /* setNoPressureCells */
    if (s.mouth_radiation_cond == 1) {
        int mx = g.te + 1;
        for (int y = g.upper[g.L - 1]; y <= g.lower[g.L - 1]; ++y) type[(size_t)y * W + mx] = cell_noPressure;
    }
    }
frame:
    /* setupDeadCells (the reference's idx <= H*W bound touches one element
     * past the end; only in-frame cells are written here) */
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < W; ++c)
            if (r == 0 || r == H - 1 || c == 0 || c == W - 1) type[(size_t)r * W + c] = cell_dead;
    /* PML: horizontal rows then vertical columns, ring k gets pml5 - k */
    if (s.pml_flag == 1) {
        int t = cell_pml5, xs = 1, xe = W - 2;
        for (int k = 0; k < 6; ++k) {
            int ru = 1 + k, rl = H - 1 - 1 - k;
            for (int c = xs; c <= xe; ++c) { type[(size_t)ru * W + c] = t; type[(size_t)rl * W + c] = t; }
            xs++; xe--; t--;
        }
        t = cell_pml5; int ys = 1, ye = H - 2;
        for (int k = 0; k < 6; ++k) {
            int cl = 1 + k, cr = W - 1 - 1 - k;
            for (int r = ys; r <= ye; ++r) { type[(size_t)r * W + cl] = t; type[(size_t)r * W + cr] = t; }
            ys++; ye--; t--;
        }
    }
}

/* z_inv value for a wall admittance mu_3D: FDTDSolver.cu:613-617. */
float z_from_mu_f32(float mu_3D, float rho, float c) {
    float mu_2D = mu_3D * ((2.0f * 0.5F * M_PI) / 1.84f);
    float alpha = 1.0f / (0.5f + 0.25f * (mu_2D + (1 / mu_2D)));
    return 1 / (rho * c * ((1 + std::sqrt(1 - alpha)) / (1 - std::sqrt(1 - alpha))));
}
double z_from_mu_f64(double mu_3D, double rho, double c) {
    double mu_2D = mu_3D * ((2.0 * 0.5 * M_PI) / 1.84);
    double alpha = 1.0 / (0.5 + 0.25 * (mu_2D + (1 / mu_2D)));
    return 1 / (rho * c * ((1 + std::sqrt(1 - alpha)) / (1 - std::sqrt(1 - alpha))));
}

/* ------------------------------------------------------------------ */
/* Two-mass model: TwoMassModel.cpp:4-231 (restated; literal typing kept)*/
/* ------------------------------------------------------------------ */
template <typename Real>
struct TwoMass {
    Real mu, rho, q_factor, gs, m1, m2, d1, d2, k1, k2, kc, etak1, etak2, h1, h2,
        etah1, etah2, Ag0, Ag01, Ag02, lg, z1o, z2o, z1c, z2c, r1_open, r2_open,
        r1_close, r2_close, A1, Ag1, Ag2, ps, p1, f_mass1, f_mass2, p_mass1, p_mass2,
        Rv1, Rv2, Lg1, Lg2, ug_old, ug_curr, ug_next, x1_old, x1_curr, x1_next,
        x2_old, x2_curr, x2_next, a11, a12, a21, a22, b1, b2, s1p, s2p;
    bool abs_double = false;   // fabs resolution of the compiler (see run())

    void init(Real q, Real ps_in) {
        using namespace fu;
        mu = 1.86e-4 * (DYN * SECOND / (CENTIMETER * CENTIMETER));
        rho = 1.14 * (KILOGRAM / (METER * METER * METER));
        q_factor = q;
        gs = 1.0f;
        m1 = (0.125f * GRAM) / q_factor;
        m2 = (0.025f * GRAM) / q_factor;
        d1 = (0.25f * CENTIMETER) / q_factor;
        d2 = (0.05f * CENTIMETER) / q_factor;
        k1 = 80000.0f * (DYN / CENTIMETER) * q_factor;
        k2 = 8000.0f * (DYN / CENTIMETER) * q_factor;
        kc = 25000.0f * (DYN / CENTIMETER) * (q_factor * q_factor);
        etak1 = 100 / (CENTIMETER * CENTIMETER);
        etak2 = 100 / (CENTIMETER * CENTIMETER);
        h1 = 3 * k1;
        h2 = 3 * k2;
        etah1 = 500.0f / (CENTIMETER * CENTIMETER);
        etah2 = 500.0f / (CENTIMETER * CENTIMETER);
        Ag0 = 0.05f * (CENTIMETER * CENTIMETER);
        Ag01 = Ag0; Ag02 = Ag0;
        lg = 1.40f * CENTIMETER;
        z1o = 0.2f; z2o = 0.6f; z1c = 1.1f; z2c = 1.9f;
        r1_open = (2 * z1o * std::sqrt(k1 * m1)) / (gs * gs);
        r2_open = (2 * z2o * std::sqrt(k2 * m2)) / (gs * gs);
        r1_close = (2 * z1c * std::sqrt(k1 * m1)) / (gs * gs);
        r2_close = (2 * z2c * std::sqrt(k2 * m2)) / (gs * gs);
        A1 = 3.0f * (CENTIMETER * CENTIMETER);
        Ag1 = 0.0f; Ag2 = 0.0f;
        ps = ps_in;
        p1 = 0.0f;
        f_mass1 = f_mass2 = p_mass1 = p_mass2 = 0.0f;
        Rv1 = Rv2 = Lg1 = Lg2 = 0.0f;
        ug_old = ug_curr = ug_next = 0.0f;
        x1_old = x1_curr = x1_next = 0.0f;
        x2_old = x2_curr = x2_next = 0.0f;
        a11 = a12 = a21 = a22 = b1 = b2 = s1p = s2p = 0.0f;
    }

    /* runTwoMassModel TwoMassModel.cpp:90-231 */
    void run(int srate_mul) {
        Real srate = 44100 * static_cast<Real>(srate_mul);
        Ag1 = Ag01 + 2 * lg * x1_curr;
        Ag2 = Ag02 + 2 * lg * x2_curr;
        Rv1 = (12 * mu * lg * lg * d1) / (Ag1 * Ag1 * Ag1);
        Lg1 = (rho * d1) / (Ag1);
        Rv2 = (12 * mu * lg * lg * d2) / (Ag2 * Ag2 * Ag2);
        Lg2 = (rho * d2) / (Ag2);
        if (x1_curr > -Ag01 / (2 * lg) && x2_curr > -Ag02 / (2 * lg)) {
            p_mass1 = ps - 1.37 * (rho / 2) * (ug_curr / Ag1) * (ug_curr / Ag1)
                      - 0.5 * (Rv1 * ug_curr + Lg1 * srate * (ug_curr - ug_old));
            f_mass1 = p_mass1 * d1 * lg;
        } else {
            p_mass1 = ps;
            f_mass1 = p_mass1 * d1 * lg;
        }
        if (x1_curr > -Ag01 / (2 * lg)) {
            if (x2_curr > -Ag02 / (2 * lg)) {
                p_mass2 = p_mass1 - (0.5 * (Rv1 + Rv2) * ug_curr + (Lg1 + Lg2) * (ug_curr - ug_old) * srate)
                          - (rho / 2 * ug_curr * ug_curr) * (1 / (Ag2 * Ag2) - 1 / (Ag1 * Ag1));
                f_mass2 = p_mass2 * d2 * lg;
            } else {
                p_mass2 = ps;
                f_mass2 = p_mass2 * d2 * lg;
            }
        } else {
            p_mass2 = 0;
            f_mass2 = p_mass2 * d2 * lg;
        }
        Real h1u, r1u, h2u, r2u;
        if (x1_curr + Ag01 / (2 * lg) < 0) { h1u = h1; r1u = r1_close; }
        else { h1u = 0.0f; r1u = r1_open; }
        if (x2_curr + Ag02 / (2 * lg) < 0) { h2u = h2; r2u = r2_close; }
        else { h2u = 0.0f; r2u = r2_open; }
        a11 = (k1 + h1u + kc) / (srate * srate) + r1u / srate + m1;
        a12 = -kc / (srate * srate);
        a21 = a12;
        a22 = (k2 + h2u + kc) / (srate * srate) + r2u / srate + m2;
        s1p = k1 * etak1 * x1_curr * x1_curr * x1_curr
              + h1u * (Ag01 / (2 * lg) + etah1 * (Ag01 / (2 * lg) + x1_curr) * (Ag01 / (2 * lg) + x1_curr) * (Ag01 / (2 * lg) + x1_curr));
        s2p = k2 * etak2 * x2_curr * x2_curr * x2_curr
              + h2u * (Ag02 / (2 * lg) + etah2 * (Ag02 / (2 * lg) + x2_curr) * (Ag02 / (2 * lg) + x2_curr) * (Ag02 / (2 * lg) + x2_curr));
        b1 = (2 * m1 + r1u / srate) * x1_curr - m1 * x1_old - s1p / (srate * srate) + f_mass1 / (srate * srate);
        b2 = (2 * m2 + r2u / srate) * x2_curr - m2 * x2_old - s2p / (srate * srate) + f_mass2 / (srate * srate);
        Real det = a11 * a22 - a21 * a12;
        if (det == 0) det = 1;
        x1_next = (a22 * b1 - a12 * b2) / det;
        x2_next = (a11 * b2 - a21 * b1) / det;
        x1_old = x1_curr; x1_curr = x1_next;
        x2_old = x2_curr; x2_curr = x2_next;
        Ag1 = Ag01 + 2 * lg * x1_curr;
        Ag2 = Ag02 + 2 * lg * x2_curr;
        if (Ag1 < 0 || Ag2 < 0) {
            ug_next = 0;
        } else {
            // `(rho / 2)*fabs(ug_curr)` (TwoMassModel.cpp:221): MSVC, the
            // reference's compiler, resolves fabs(float) to the float overload,
            // so the product rounds to float before meeting the double 0.37
            // term; glibc's global fabs under g++ is double fabs, which keeps
            // the product exact (abs_double = true reproduces that build)
            const double lead = abs_double ? (double)(rho / 2) * std::fabs((double)ug_curr)
                                           : (double)((rho / 2) * std::fabs(ug_curr));
            Real Rtot = lead * ((0.37 / (Ag1 * Ag1)) + (1 - 2 * (Ag2 / A1) * (1 - Ag2 / A1)) / (Ag2 * Ag2))
                        + (12 * mu * lg * lg * (d1 / (Ag1 * Ag1 * Ag1) + d2 / (Ag2 * Ag2 * Ag2)));
            Real Ltot = rho * (d1 / Ag1 + d2 / Ag2);
            ug_next = ((ps - p1) / srate + Ltot * ug_curr) / (Rtot / srate + Ltot);
        }
        ug_old = ug_curr;
        ug_curr = ug_next;
    }
};

/* ------------------------------------------------------------------ */
/* The oracle tube                                                      */
/* ------------------------------------------------------------------ */
struct TubeBase {
    virtual ~TubeBase() {}
    Geometry g;
    std::vector<int> type, bst;
    int num_steps = 0;
    int dirichlet = 1;
    int source_type = 1;
    int srate_mul = 15;
    int engine_variant = 0;
    int solver_type = 1;
    float ds_f32 = 0;
    virtual void dims(int* out) const = 0;
    virtual void scalars(double* out) const = 0;
    virtual bool field(int f, double* dst, long n) const = 0;
    virtual void run(int first, int n, double* probe, double* src) = 0;
    virtual void reset() = 0;
    virtual void get_state(double*, double*, double*) const = 0;
    virtual void set_state(const double*, const double*, const double*) = 0;
    virtual double run_structured(int n, double* probe) = 0;
    virtual double run_threads(int first, int n, double* probe, int nthreads) = 0;
};

template <typename Real>
struct Tube : TubeBase {
    Phys<Real> ph;
    std::vector<Real> minor_r;        /* ellipse minor axis per section (Real) */
    std::vector<Real> zsec;           /* z_inv per section */
    Real z_default = 0;
    std::vector<Real> dX, dY, dP, zinv;
    std::vector<Real> coef[17];
    std::vector<Real> source;         /* precomputed excitation table */
    std::vector<Real> P, Vx, Vy, Pn, Vxn, Vyn;
    TwoMass<Real> tm;
    Real tm_q = 1, tm_ps = 686.46f;

    int H() const { return g.frameH; }
    int W() const { return g.frameW; }

    bool build(const vso_spec& s);
    bool build_depth(const vso_spec& s);
    void build_coefficients();
    void build_source(const vso_spec& s);

    void dims(int* o) const override {
        o[VSO_FRAME_H] = g.frameH; o[VSO_FRAME_W] = g.frameW;
        o[VSO_DOMAIN_H] = g.domainH; o[VSO_DOMAIN_W] = g.domainW;
        o[VSO_LISTENER_X] = g.listenerX; o[VSO_LISTENER_Y] = g.listenerY;
        o[VSO_EXCITATION_X] = g.excX; o[VSO_EXCITATION_Y] = g.excY;
        o[VSO_NUM_STEPS] = num_steps; o[VSO_TUBE_START_X] = g.ts; o[VSO_TUBE_END_X] = g.te;
        o[VSO_TUBE_LEN_CELLS] = g.L; o[VSO_MID_Y] = g.midY; o[VSO_NUM_SECTIONS] = g.nsec;
        o[VSO_MIC_CELLS] = g.mic_cells;
    }
    void scalars(double* o) const override {
        o[VSO_S_DT] = ph.dt; o[VSO_S_DS] = ph.ds; o[VSO_S_RHO] = ph.rho; o[VSO_S_C] = ph.c;
        o[VSO_S_KAPPA] = ph.kappa; o[VSO_S_K1] = ph.K1; o[VSO_S_K2] = ph.K2;
        o[VSO_S_DS_F32] = ds_f32; o[VSO_S_Z_DEFAULT] = z_default;
    }
    bool field(int f, double* dst, long n) const override;
    void run(int first, int n, double* probe, double* src) override;
    void reset() override {
        std::fill(P.begin(), P.end(), Real(0)); std::fill(Vx.begin(), Vx.end(), Real(0));
        std::fill(Vy.begin(), Vy.end(), Real(0));
        Pn = P; Vxn = Vx; Vyn = Vy;
        tm.init(tm_q, tm_ps);
    }
    void get_state(double* p, double* vx, double* vy) const override {
        for (size_t i = 0; i < P.size(); ++i) { p[i] = P[i]; vx[i] = Vx[i]; vy[i] = Vy[i]; }
    }
    void set_state(const double* p, const double* vx, const double* vy) override {
        for (size_t i = 0; i < P.size(); ++i) { P[i] = (Real)p[i]; Vx[i] = (Real)vx[i]; Vy[i] = (Real)vy[i]; }
        Pn = P; Vxn = Vx; Vyn = Vy;
    }
    double run_structured(int n, double* probe) override;
    double run_threads(int first, int n, double* probe, int nthreads) override;
};

template <typename Real>
bool Tube<Real>::build(const vso_spec& s) {
    dirichlet = s.mouth_radiation_cond == 1 ? 1 : 0;
    source_type = s.source_type;
    srate_mul = s.srate_mul;
    engine_variant = s.engine_variant;
    solver_type = s.fdtd_solver_type;
    if (s.srate_mul <= 0) { g_err = "srate_mul must be positive"; return false; }
    if (s.source_type < 0 || s.source_type > 2) { g_err = "Invalid source type"; return false; }
    if (s.fdtd_solver_type != 0 && s.fdtd_solver_type != 1) { g_err = "fdtd_solver_type must be 0 or 1"; return false; }
    Phys<float> pf = phys_f32(s.srate_mul);
    ds_f32 = pf.ds;
    if constexpr (sizeof(Real) == sizeof(float)) ph = pf;
    else ph = phys_f64(s.srate_mul);
    fill_tables(ph);
    num_steps = s.num_steps > 0 ? s.num_steps : float_num_steps(s.audio_dur, pf.dt);
    if (num_steps <= 0) { g_err = "number of steps must be positive"; return false; }
    if (s.open_space) {
        build_open_space(s, g);
        dirichlet = 0;
    } else if (!build_geometry(s, pf.ds, g)) {
        return false;
    }
    build_cells(s, g, type, bst);
    /* per-section minor axis and z in Real */
    for (int i = 0; i < g.nsec; ++i) {
        if constexpr (sizeof(Real) == sizeof(float)) {
            minor_r.push_back(g.minor_m[i]);
        } else {
            double a_m2 = (double)s.area_cm2[i] * 1e-2 * 1e-2;
            double q = s.cross_sectional_shape == 2 ? (s.aspect ? (double)s.aspect[i] : 3.0) : 1.0;
            double lenX = std::sqrt(a_m2 / (q * 1.0 * M_PI));
            minor_r.push_back(s.cross_sectional_shape == 2 ? lenX * 2.0 : 2.0 * std::sqrt(a_m2 / M_PI));
        }
    }
    float mu_def = s.mu3d_default > 0 ? s.mu3d_default : (s.fdtd_solver_type == 1 ? 0.025f : 0.005f);
    for (int i = 0; i < g.nsec; ++i) {
        float mu = s.mu3d ? s.mu3d[i] : mu_def;
        if (!(mu > 0)) { g_err = "wall admittance must be positive"; return false; }
        if constexpr (sizeof(Real) == sizeof(float)) zsec.push_back(z_from_mu_f32(mu, ph.rho, ph.c));
        else zsec.push_back(z_from_mu_f64((double)mu, ph.rho, ph.c));
    }
    if constexpr (sizeof(Real) == sizeof(float)) z_default = z_from_mu_f32(mu_def, ph.rho, ph.c);
    else z_default = z_from_mu_f64((double)mu_def, ph.rho, ph.c);
    /* setZInverse FDTDSolver.cu:1058-1069 (value per section) */
    zinv.assign(type.size(), Real(0));
    for (size_t i = 0; i < type.size(); ++i)
        if (bst[i] != 0) zinv[i] = zsec[bst[i] - 1];
    /* depth: FDTDSolver.cu:632-641 */
    dX.assign(type.size(), Real(1.0)); dY = dX; dP = dX;
    if (s.open_space) {
        /* constant open-space depth, FDTDSolver.cu:504-514 */
        if (s.fdtd_solver_type == 1) { dX.assign(type.size(), (Real)(1.0 * (Real)0.50f)); dY = dX; dP = dX; }
    } else if (s.fdtd_solver_type == 1 && !build_depth(s)) {
        return false;
    }
    build_coefficients();
    build_source(s);
    tm_q = s.tm_q > 0 ? (Real)s.tm_q : (Real)1.0f;
    tm_ps = s.tm_ps > 0 ? (Real)s.tm_ps : (Real)686.46f;
    P.assign(type.size(), Real(0)); Vx = P; Vy = P;
    reset();
    return true;
}

/* TubeDepth.cu:95-137 (circle), 140-168 (ellipse), 215-265, 313-490, 642-762 */
template <typename Real>
bool Tube<Real>::build_depth(const vso_spec& s) {
    const int H = g.frameH, W = g.frameW;
    const Real open_depth = (Real)0.50f;
    const Real min_depth = (Real)0.001f;
    const Real ds = ph.ds;
    std::vector<Real> X(type.size(), (Real)(1.0 * open_depth)), Y = X;
    for (int j = 0; j < g.L; ++j) {
        int up = g.upper[j], lo = g.lower[j], x = g.ts + j;
        int air = lo - up - 1;
        int half = (int)ceilf(air * 0.5f) - 1;
        Real scale = s.depth_scale ? (Real)s.depth_scale[g.seg[j]] : (Real)1;
        if (s.cross_sectional_shape == 3) {
            /* TubeDepth.cu:493-640: half the vertical chord of the section's
             * contour, positions counted from the lower wall */
            int maxseg = 0;
            for (int y = 0; y < H; ++y) maxseg = std::max(maxseg, bst[(size_t)y * W + x]);
            const PlaneContour& pc = g.plane[maxseg - 1];
            for (int cc = 1; cc <= air; ++cc) {
                float px = (cc - 0.5) * ds_f32;
                float py = cc * ds_f32;
                float zl, zh;
                if (!vertical_chord(pc.y, pc.z, pc.minY + px, zl, zh)) { g_err = "depth: no contour crossing"; return false; }
                Real vx = ((Real)zh - (Real)zl) / (Real)2.0f;
                Real vy = 0;
                if (cc < air) {
                    if (!vertical_chord(pc.y, pc.z, pc.minY + py, zl, zh)) { g_err = "depth: no contour crossing"; return false; }
                    vy = ((Real)zh - (Real)zl) / (Real)2.0f;
                }
                X[(size_t)(lo - cc) * W + x] = vx * scale;
                Y[(size_t)(lo - cc) * W + x] = vy * scale;
            }
            X[(size_t)up * W + x] = 0; X[(size_t)lo * W + x] = 0;
            Y[(size_t)up * W + x] = 0; Y[(size_t)lo * W + x] = 0;
            continue;
        }
        Real sa2 = 0, sb2 = 0, r2 = 0;
        if (s.cross_sectional_shape == 1) {
            Real radius = (static_cast<Real>(air) / 2.0f) * ds;
            r2 = radius * radius;
        } else {
            /* the section of the column is the largest boundary segment id in it (TubeDepth.cu:417-420) */
            int maxseg = 0;
            for (int y = 0; y < H; ++y) maxseg = std::max(maxseg, bst[(size_t)y * W + x]);
            Real a = (air * ds) / 2.0f;
            Real b = minor_r[maxseg - 1] / 2.0f;
            sa2 = a * a; sb2 = b * b;
        }
        for (int k = half; k >= 0; --k) {
            int yu = up + (half - k) + 1;
            int yl = lo - (half - k) - 1;
            Real d = static_cast<Real>(k) * ds;
            Real dl = (static_cast<Real>(k) - 0.5f) * ds;
            Real du = (static_cast<Real>(k) + 0.5f) * ds;
            Real sx, sl, su = 0.0f;
            if (s.cross_sectional_shape == 1) {
                sx = r2 - d * d;
                sl = r2 - dl * dl;
                if (k != half) su = r2 - du * du;
            } else {
                sx = sb2 * (1.0f - ((d * d) / sa2));
                sl = sb2 * (1 - ((dl * dl) / sa2));
                if (k != half) su = sb2 * (1 - ((du * du) / sa2));
            }
            if (sx < 0 || sl < 0 || su < 0) { g_err = "depth sqrt domain error"; return false; }
            Real depx = static_cast<Real>(std::sqrt(sx));
            Real depl = static_cast<Real>(std::sqrt(sl));
            Real depu = static_cast<Real>(std::sqrt(su));
            X[(size_t)yu * W + x] = 2 * depx * scale;
            X[(size_t)yl * W + x] = 2 * depx * scale;
            Y[(size_t)yu * W + x] = 2 * depu * scale;
            Y[(size_t)yl * W + x] = 2 * depl * scale;
        }
        X[(size_t)up * W + x] = 0; X[(size_t)lo * W + x] = 0;
        Y[(size_t)up * W + x] = 0; Y[(size_t)lo * W + x] = 0;
    }
    /* resetOpenSpaceDepth */
    const size_t mid = (size_t)g.midY * W + g.te;
    const Real mx = X[mid], my = Y[mid];
    for (size_t i = 0; i < X.size(); ++i) {
        if (X[i] == open_depth) X[i] = mx;
        if (Y[i] == open_depth) Y[i] = my;
    }
    /* smoothDepthX / smoothDepthY */
    std::vector<Real> Xs(X.size()), Ys(Y.size());
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < W; ++c) {
            size_t i = (size_t)r * W + c;
            Real right = (c < W - 1) ? X[i + 1] : X[mid];
            Xs[i] = 0.5f * (X[i] + right);
            Real above = (r > 0) ? Y[i - W] : Y[mid];
            Ys[i] = 0.5f * (Y[i] + above);
        }
    /* updateDepthP: pivots come from the smoothed maps */
    const Real px = Xs[mid], py = Ys[mid];
    std::vector<Real> Pd(X.size());
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < W; ++c) {
            size_t i = (size_t)r * W + c;
            Real lX = (c > 0) ? Xs[i - 1] : px;
            Real dYv = (r < H - 1) ? Ys[i + W] : py;
            Pd[i] = 0.25f * (Xs[i] + Ys[i] + lX + dYv);
        }
    /* setMinDepthVals */
    for (size_t i = 0; i < X.size(); ++i) {
        if (Pd[i] < min_depth) Pd[i] = min_depth;
        if (Xs[i] < min_depth) Xs[i] = min_depth;
        if (Ys[i] < min_depth) Ys[i] = min_depth;
    }
    dX = Xs; dY = Ys; dP = Pd;
    return true;
}

/* calculateSolverCoefficients FDTDSolver.cu:652-801, kernels :860-1056.
 * Coefficient order: N_out, N_in, minVxBeta, minVyBeta, maxVxSigmaPrimedt,
 * maxVySigmaPrimedt, betaVxSqr_dt_inv_rho_inv_ds, betaVySqr_dt_inv_rho_inv_ds,
 * pressureSigmaPrimedt, areWeNotExcitationVx, areWeNotExcitationVy,
 * excitationWeightVx, excitationWeightVy, xor_val1..4 (FDTDSolver.h:144-160). */
template <typename Real>
void coeff_of_triple(const Phys<Real>& p, int tc, int tr, int tt, Real* o) {
    Real bc = p.beta[tc], br = p.beta[tr], bt = p.beta[tt];
    int lut;
    int ibr = (int)br, ibt = (int)bt;
    if (ibr == cell_air && ibt == cell_air) lut = 0;
    else if (ibr == cell_air && ibt == cell_wall) lut = 1;
    else if (ibr == cell_wall && ibt == cell_air) lut = 2;
    else lut = 3;
    o[0] = p.air_normal[lut] * bc;
    o[1] = p.wall_normal[lut] * (1 - bc);
    Real sc = p.sigma[tc], sr = p.sigma[tr], st = p.sigma[tt];
    int ec = tc == cell_excitation, er = tr == cell_excitation, et = tt == cell_excitation;
    Real mbx = std::min(bc, br), mby = std::min(bc, bt);
    o[2] = mbx; o[3] = mby;
    o[4] = std::max(sc, sr); o[5] = std::max(sc, st);
    Real bx2 = mbx * mbx, by2 = mby * mby;
    o[6] = (bx2 * p.dt) / (p.rho * p.ds);
    o[7] = (by2 * p.dt) / (p.rho * p.ds);
    o[8] = sc;
    o[9] = (Real)((1 - ec) * (1 - er));
    o[10] = (Real)((1 - ec) * (1 - et));
    o[11] = (ec * p.src_dir[2]) + (er * p.src_dir[0]);
    o[12] = (ec * p.src_dir[3]) + (et * p.src_dir[1]);
    o[13] = br * (1 - bc);
    o[14] = bc * (1 - br);
    o[15] = bt * (1 - bc);
    o[16] = bc * (1 - bt);
}

template <typename Real>
void Tube<Real>::build_coefficients() {
    const int H = g.frameH, W = g.frameW;
    const size_t n = (size_t)(H - 2) * (W - 2);
    for (auto& v : coef) v.assign(n, Real(0));
    Real o[17];
    for (int r = 1; r < H - 1; ++r)
        for (int c = 1; c < W - 1; ++c) {
            size_t t = (size_t)(r - 1) * (W - 2) + (c - 1);
            coeff_of_triple(ph, type[(size_t)r * W + c], type[(size_t)r * W + c + 1], type[(size_t)(r - 1) * W + c], o);
            for (int k = 0; k < 17; ++k) coef[k][t] = o[k];
        }
}

/* GlottalExcitations.cu:9-52 */
template <typename Real>
void Tube<Real>::build_source(const vso_spec& s) {
    source.assign(num_steps, Real(0));
    const double dt_d = (double)ph.dt;
    if (s.source_type == 0) {
        Real f = 440 * fu::HERTZ, amp = 25;
        for (int i = 0; i < num_steps; ++i) {
            Real t = i * dt_d;
            source[i] = amp * std::sin(2 * M_PI * f * t);
        }
    } else if (s.source_type == 1) {
        Real f0 = 10.0f * fu::KILOHERTZ;
        Real peak = 0.646f / f0;
        Real width = 0.29f * peak;
        Real amp = 0.025f;
        if (s.gauss_amp > 0) amp = (Real)s.gauss_amp;
        if (s.gauss_t0 > 0) peak = (Real)s.gauss_t0;
        if (s.gauss_width > 0) width = (Real)s.gauss_width;
        for (int i = 0; i < num_steps; ++i) {
            Real t = i * dt_d;
            Real e = (t - peak) / width;
            Real e2 = e * e;
            source[i] = amp * std::exp(-1 * e2);
        }
    }
}

template <typename Real>
bool Tube<Real>::field(int f, double* dst, long n) const {
    const size_t HW = (size_t)g.frameH * g.frameW;
    auto put = [&](const auto& v) {
        if ((long)v.size() > n) return false;
        for (size_t i = 0; i < v.size(); ++i) dst[i] = (double)v[i];
        return true;
    };
    switch (f) {
    case VSO_F_CELL_TYPE: return put(type);
    case VSO_F_BOUNDARY_SEG: return put(bst);
    case VSO_F_DEPTH_P: return put(dP);
    case VSO_F_DEPTH_X: return put(dX);
    case VSO_F_DEPTH_Y: return put(dY);
    case VSO_F_Z_INV: return put(zinv);
    case VSO_F_SOURCE: return put(source);
    case VSO_F_UPPER_WALL: return put(g.upper);
    case VSO_F_LOWER_WALL: return put(g.lower);
    case VSO_F_SEGMENT: return put(g.seg);
    case VSO_F_DIAM_CELLS: return put(g.diam_cells);
    case VSO_F_MINOR_AXIS: return put(minor_r);
    case VSO_F_Z_SECTION: return put(zsec);
    default:
        if (f >= VSO_F_COEFF0 && f < VSO_F_COEFF0 + 17) return put(coef[f - VSO_F_COEFF0]);
    }
    (void)HW;
    return false;
}

/* One time step, flat arrays (FDTDEngineCPU.cpp:257-288, 316-433).
 * The wall-vibration state is identically zero (coupling 0,
 * SimulationParams.cu:25), so -rho_sqrC_dt*w contributes an exact zero and is
 * not swept here; vso_run_structured keeps the sweep. */
template <typename Real>
void Tube<Real>::run(int first, int n, double* probe, double* src) {
    const int H = g.frameH, W = g.frameW, Wi = W - 2;
    const Real K1 = ph.K1;
    const size_t L = (size_t)g.listenerY * W + g.listenerX;
    const size_t E = (size_t)g.excY * W + g.excX;
    for (int s = 0; s < n; ++s) {
        const int t = first + s;
        Real exeV;
        if (source_type == 2) {
            tm.p1 = P[E];
            tm.run(srate_mul);
            exeV = tm.ug_next;
        } else {
            exeV = (t < (int)source.size()) ? source[t] : Real(0);
        }
        for (int r = 1; r < H - 1; ++r)
            for (int c = 1; c < W - 1; ++c) {
                const size_t i = (size_t)r * W + c;
                const size_t k = (size_t)(r - 1) * Wi + (c - 1);
                Real cxvx = Vx[i] * dX[i] - Vx[i - 1] * dX[i - 1];
                Real cyvy = Vy[i] * dY[i] - Vy[i + W] * dY[i + W];
                Real num = (P[i] * dP[i]) - (K1 * (cxvx + cyvy));
                Real den = engine_variant == 0 ? ((1 + coef[8][k]) * dP[i]) : (coef[8][k] * dP[i] + dP[i]);
                Real v = num / den;
                if (dirichlet && type[i] == cell_noPressure) v = 0;
                Pn[i] = v;
            }
        if (probe) probe[t] = (double)Pn[L];
#if VSO_MUTATION == 4
        if (probe) probe[t] = (double)Pn[Pn.size()];
#endif
        if (src) src[t] = (double)exeV;
        for (int r = 1; r < H - 1; ++r)
            for (int c = 1; c < W - 1; ++c) {
                const size_t i = (size_t)r * W + c;
                const size_t k = (size_t)(r - 1) * Wi + (c - 1);
                Real cxp = Pn[i + 1] - Pn[i];
                Real cyp = Pn[i - W] - Pn[i];
                Real vx = (coef[2][k] * Vx[i]) - (coef[6][k] * cxp);
                Real vy = (coef[3][k] * Vy[i]) - (coef[7][k] * cyp);
                vx = vx + (exeV * coef[11][k] * coef[4][k]);
                vy = vy + (exeV * coef[12][k] * coef[5][k]);
                Real bx = (coef[14][k] * Pn[i] * coef[0][k]) - (coef[13][k] * Pn[i + 1] * coef[1][k]);
                Real by = (coef[16][k] * Pn[i] * coef[0][k]) - (coef[15][k] * Pn[i - W] * coef[1][k]);
#if VSO_MUTATION == 2
                bx = (coef[14][k] * Pn[i] * coef[0][k]);
                by = (coef[16][k] * Pn[i] * coef[0][k]);
#endif
#if VSO_MUTATION == 3
                bx = bx * coef[9][k] * std::max(zinv[i], zinv[i + 1]);
                by = by * coef[10][k] * std::max(zinv[i], zinv[i - W]);
#else
                bx = bx * coef[9][k] * zinv[i];
                by = by * coef[10][k] * zinv[i];
#endif
#if VSO_MUTATION == 1
                bx = -bx; by = -by;
#endif
                vx += (coef[4][k] * bx);
                vy += (coef[5][k] * by);
                vx /= (coef[2][k] + coef[4][k]);
                vy /= (coef[3][k] + coef[5][k]);
                Vxn[i] = vx; Vyn[i] = vy;
            }
        /* border ring stays zero: nothing writes it */
        for (int r = 1; r < H - 1; ++r)
            for (int c = 1; c < W - 1; ++c) {
                const size_t i = (size_t)r * W + c;
                P[i] = Pn[i]; Vx[i] = Vxn[i]; Vy[i] = Vyn[i];
            }
    }
}

/* run() with each phase's rows split over nthreads OpenMP threads, for CPU
 * timing on several host cores.  Every cell of a phase reads only the arrays
 * the previous phase wrote, so the results equal run() bit for bit.  Returns
 * the wall seconds of the step loop. */
template <typename Real>
double Tube<Real>::run_threads(int first, int n, double* probe, int nthreads) {
    const int H = g.frameH, W = g.frameW, Wi = W - 2;
    const Real K1 = ph.K1;
    const size_t L = (size_t)g.listenerY * W + g.listenerX;
    const size_t E = (size_t)g.excY * W + g.excX;
    Real exeV = 0;
    auto t0 = std::chrono::high_resolution_clock::now();
#pragma omp parallel num_threads(nthreads)
    for (int s = 0; s < n; ++s) {
        const int t = first + s;
#pragma omp single
        {
            if (source_type == 2) {
                tm.p1 = P[E];
                tm.run(srate_mul);
                exeV = tm.ug_next;
            } else {
                exeV = (t < (int)source.size()) ? source[t] : Real(0);
            }
        }
#pragma omp for schedule(static)
        for (int r = 1; r < H - 1; ++r)
            for (int c = 1; c < W - 1; ++c) {
                const size_t i = (size_t)r * W + c;
                const size_t k = (size_t)(r - 1) * Wi + (c - 1);
                Real cxvx = Vx[i] * dX[i] - Vx[i - 1] * dX[i - 1];
                Real cyvy = Vy[i] * dY[i] - Vy[i + W] * dY[i + W];
                Real num = (P[i] * dP[i]) - (K1 * (cxvx + cyvy));
                Real den = engine_variant == 0 ? ((1 + coef[8][k]) * dP[i]) : (coef[8][k] * dP[i] + dP[i]);
                Real v = num / den;
                if (dirichlet && type[i] == cell_noPressure) v = 0;
                Pn[i] = v;
            }
#pragma omp single nowait
        if (probe) probe[t] = (double)Pn[L];
#pragma omp for schedule(static)
        for (int r = 1; r < H - 1; ++r)
            for (int c = 1; c < W - 1; ++c) {
                const size_t i = (size_t)r * W + c;
                const size_t k = (size_t)(r - 1) * Wi + (c - 1);
                Real cxp = Pn[i + 1] - Pn[i];
                Real cyp = Pn[i - W] - Pn[i];
                Real vx = (coef[2][k] * Vx[i]) - (coef[6][k] * cxp);
                Real vy = (coef[3][k] * Vy[i]) - (coef[7][k] * cyp);
                vx = vx + (exeV * coef[11][k] * coef[4][k]);
                vy = vy + (exeV * coef[12][k] * coef[5][k]);
                Real bx = (coef[14][k] * Pn[i] * coef[0][k]) - (coef[13][k] * Pn[i + 1] * coef[1][k]);
                Real by = (coef[16][k] * Pn[i] * coef[0][k]) - (coef[15][k] * Pn[i - W] * coef[1][k]);
                bx = bx * coef[9][k] * zinv[i];
                by = by * coef[10][k] * zinv[i];
                vx += (coef[4][k] * bx);
                vy += (coef[5][k] * by);
                vx /= (coef[2][k] + coef[4][k]);
                vy /= (coef[3][k] + coef[5][k]);
                Vxn[i] = vx; Vyn[i] = vy;
            }
#pragma omp for schedule(static)
        for (int r = 1; r < H - 1; ++r)
            for (int c = 1; c < W - 1; ++c) {
                const size_t i = (size_t)r * W + c;
                P[i] = Pn[i]; Vx[i] = Vxn[i]; Vy[i] = Vyn[i];
            }
    }
    return std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
}

/* Reference-structured engine for CPU timing: the data structures and sweeps
 * of FDTDEngineCPU.cpp:222-433 (vector<vector<{P,Vx,Vy,type}>>, separate
 * CxVx/CyVy/CxP/CyP/Vb scratch grids, 17 coefficient matrices, the
 * wall-vibration sweep over the full frame and the per-step deep copy). */
template <typename Real>
double Tube<Real>::run_structured(int n, double* probe) {
    struct Cell { Real Pr, Vx, Vy; int cell_type; };
    const int H = g.frameH, W = g.frameW;
    using M = std::vector<std::vector<Real>>;
    std::vector<std::vector<Cell>> PV(H, std::vector<Cell>(W)), PVn;
    M dPm(H, std::vector<Real>(W)), dXm = dPm, dYm = dPm, zm = dPm;
    M w_prev(H, std::vector<Real>(W, 0)), w_next = w_prev, z_prev = w_prev, z_next = w_prev;
    std::vector<M> cm(17, M(H - 2, std::vector<Real>(W - 2)));
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < W; ++c) {
            size_t i = (size_t)r * W + c;
            PV[r][c] = Cell{0, 0, 0, type[i]};
            dPm[r][c] = dP[i]; dXm[r][c] = dX[i]; dYm[r][c] = dY[i]; zm[r][c] = zinv[i];
        }
    for (int k = 0; k < 17; ++k)
        for (int r = 0; r < H - 2; ++r)
            for (int c = 0; c < W - 2; ++c) cm[k][r][c] = coef[k][(size_t)r * (W - 2) + c];
    PVn = PV;
    M CxVx(H - 2, std::vector<Real>(W - 2, 0)), CyVy = CxVx, CxP = CxVx, CyP = CxVx, Vbx = CxVx, Vby = CxVx;
    TwoMass<Real> lm;
    lm.init(tm_q, tm_ps);
    const Real K1 = ph.K1, K2 = ph.K2;
    const Real Mw = 21, Bw = 8000, Kw = 845000, coupling = 0.0f, dt = ph.dt;
    auto t0 = std::chrono::high_resolution_clock::now();
    for (int t = 0; t < n; ++t) {
        Real exeV;
        if (source_type == 2) {
            lm.p1 = PV[g.excY][g.excX].Pr;
            lm.run(srate_mul);
            exeV = lm.ug_next;
        } else {
            exeV = (t < (int)source.size()) ? source[t] : Real(0);
        }
        std::swap(w_prev, w_next);
        std::swap(z_prev, z_next);
        for (int r = 0; r < H - 2; ++r)
            for (int c = 0; c < W - 2; ++c) {
                CxVx[r][c] = PV[r + 1][c + 1].Vx * dXm[r + 1][c + 1] - PV[r + 1][c].Vx * dXm[r + 1][c];
                CyVy[r][c] = PV[r + 1][c + 1].Vy * dYm[r + 1][c + 1] - PV[r + 2][c + 1].Vy * dYm[r + 2][c + 1];
                Real num = ((PV[r + 1][c + 1].Pr * dPm[r + 1][c + 1]) - (K1 * (CxVx[r][c] + CyVy[r][c])) - (K2 * w_prev[r + 1][c + 1]));
                Real den = ((1 + cm[8][r][c]) * dPm[r + 1][c + 1]);
                PVn[r + 1][c + 1].Pr = num / den;
                if (dirichlet && PVn[r + 1][c + 1].cell_type == cell_noPressure) PVn[r + 1][c + 1].Pr = 0;
            }
        if (probe) probe[t] = (double)PVn[g.listenerY][g.listenerX].Pr;
        if (solver_type) {
            for (int r = 0; r < H; ++r)
                for (int c = 0; c < W; ++c) {
                    z_next[r][c] = z_prev[r][c] + w_prev[r][c] * dt;
                    Real num = (Mw * w_prev[r][c]) + (coupling * dt * PVn[r][c].Pr) - (dt * Kw * z_next[r][c]);
                    Real den = Mw + (dt * Bw);
                    w_next[r][c] = num / den;
                }
        }
        for (int r = 0; r < H - 2; ++r)
            for (int c = 0; c < W - 2; ++c) {
                CxP[r][c] = PVn[r + 1][c + 2].Pr - PVn[r + 1][c + 1].Pr;
                CyP[r][c] = PVn[r][c + 1].Pr - PVn[r + 1][c + 1].Pr;
                PVn[r + 1][c + 1].Vx = (cm[2][r][c] * PV[r + 1][c + 1].Vx) - (cm[6][r][c] * CxP[r][c]);
                PVn[r + 1][c + 1].Vy = (cm[3][r][c] * PV[r + 1][c + 1].Vy) - (cm[7][r][c] * CyP[r][c]);
                PVn[r + 1][c + 1].Vx = PVn[r + 1][c + 1].Vx + (exeV * cm[11][r][c] * cm[4][r][c]);
                PVn[r + 1][c + 1].Vy = PVn[r + 1][c + 1].Vy + (exeV * cm[12][r][c] * cm[5][r][c]);
                Vbx[r][c] = (cm[14][r][c] * PVn[r + 1][c + 1].Pr * cm[0][r][c]) - (cm[13][r][c] * PVn[r + 1][c + 2].Pr * cm[1][r][c]);
                Vby[r][c] = (cm[16][r][c] * PVn[r + 1][c + 1].Pr * cm[0][r][c]) - (cm[15][r][c] * PVn[r][c + 1].Pr * cm[1][r][c]);
                Vbx[r][c] = Vbx[r][c] * cm[9][r][c] * zm[r + 1][c + 1];
                Vby[r][c] = Vby[r][c] * cm[10][r][c] * zm[r + 1][c + 1];
                PVn[r + 1][c + 1].Vx += (cm[4][r][c] * Vbx[r][c]);
                PVn[r + 1][c + 1].Vy += (cm[5][r][c] * Vby[r][c]);
                PVn[r + 1][c + 1].Vx /= (cm[2][r][c] + cm[4][r][c]);
                PVn[r + 1][c + 1].Vy /= (cm[3][r][c] + cm[5][r][c]);
            }
        for (int r = 0; r < H; ++r) {
            PVn[r][0].Pr = PVn[r][0].Vx = PVn[r][0].Vy = 0;
            PVn[r][W - 1].Pr = PVn[r][W - 1].Vx = PVn[r][W - 1].Vy = 0;
        }
        for (int c = 0; c < W; ++c) {
            PVn[0][c].Pr = PVn[0][c].Vx = PVn[0][c].Vy = 0;
            PVn[H - 1][c].Pr = PVn[H - 1][c].Vx = PVn[H - 1][c].Vy = 0;
        }
        PV = PVn;
    }
    auto t1 = std::chrono::high_resolution_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // namespace

extern "C" {

const char* vso_last_error(void) { return g_err.c_str(); }

void* vso_create(const vso_spec* spec, int precision) {
    g_err.clear();
    if (!spec) { g_err = "null spec"; return nullptr; }
    try {
        if (precision == 32) {
            auto t = std::make_unique<Tube<float>>();
            if (!t->build(*spec)) return nullptr;
            return static_cast<TubeBase*>(t.release());
        } else if (precision == 64) {
            auto t = std::make_unique<Tube<double>>();
            if (!t->build(*spec)) return nullptr;
            return static_cast<TubeBase*>(t.release());
        }
        g_err = "precision must be 32 or 64";
    } catch (const std::exception& e) {
        g_err = e.what();
    }
    return nullptr;
}

void vso_destroy(void* h) { delete static_cast<TubeBase*>(h); }

int vso_dims(void* h, int* out) {
    if (!h || !out) return -1;
    static_cast<TubeBase*>(h)->dims(out);
    return 0;
}

int vso_scalars(void* h, double* out) {
    if (!h || !out) return -1;
    static_cast<TubeBase*>(h)->scalars(out);
    return 0;
}

int vso_get_field(void* h, int field, double* dst, long n) {
    if (!h || !dst) return -1;
    return static_cast<TubeBase*>(h)->field(field, dst, n) ? 0 : -2;
}

int vso_run(void* h, int first_step, int n_steps, double* probe, double* src_trace) {
    if (!h || n_steps < 0 || first_step < 0) return -1;
    static_cast<TubeBase*>(h)->run(first_step, n_steps, probe, src_trace);
    return 0;
}

int vso_reset(void* h) {
    if (!h) return -1;
    static_cast<TubeBase*>(h)->reset();
    return 0;
}

int vso_get_state(void* h, double* P, double* Vx, double* Vy) {
    if (!h) return -1;
    static_cast<TubeBase*>(h)->get_state(P, Vx, Vy);
    return 0;
}

int vso_set_state(void* h, const double* P, const double* Vx, const double* Vy) {
    if (!h) return -1;
    static_cast<TubeBase*>(h)->set_state(P, Vx, Vy);
    return 0;
}

int vso_run_structured(void* h, int n_steps, double* probe, double* seconds) {
    if (!h || n_steps < 0) return -1;
    double s = static_cast<TubeBase*>(h)->run_structured(n_steps, probe);
    if (seconds) *seconds = s;
    return 0;
}

int vso_run_threads(void* h, int first_step, int n_steps, double* probe, int nthreads, double* seconds) {
    if (!h || n_steps < 0 || first_step < 0 || nthreads < 1) return -1;
    double sec = static_cast<TubeBase*>(h)->run_threads(first_step, n_steps, probe, nthreads);
    if (seconds) *seconds = sec;
    return 0;
}

int vso_float_num_steps(float audio_dur, int srate_mul) {
    return float_num_steps(audio_dur, phys_f32(srate_mul).dt);
}

/* The two-mass recurrence alone, driven as FDTDEngineCPU.cpp:259-267 drives
   it: p1 = the supraglottal pressure of each step, then one run, then the
   volume velocity.  q = 1 and ps = 686.46 are the reference's own parameters
   (TwoMassModel.cpp:10, 61). */
int vso_twomass_run(int precision, double q, double ps, int srate_mul, int abs_double, const double* p1, int n,
                    double* ug) {
    if (n < 0 || (n > 0 && (!p1 || !ug))) return -1;
    if (precision == 32) {
        TwoMass<float> m;
        m.init((float)q, (float)ps);
        m.abs_double = abs_double != 0;
        for (int i = 0; i < n; ++i) { m.p1 = (float)p1[i]; m.run(srate_mul); ug[i] = m.ug_next; }
    } else if (precision == 64) {
        TwoMass<double> m;
        m.init(q, ps);
        m.abs_double = abs_double != 0;
        for (int i = 0; i < n; ++i) { m.p1 = p1[i]; m.run(srate_mul); ug[i] = m.ug_next; }
    } else {
        return -1;
    }
    return 0;
}

int vso_coeff_triple(int precision, int srate_mul, int cur, int right, int top, double* out17) {
    if (cur < 0 || cur >= NUM_CELL_TYPES || right < 0 || right >= NUM_CELL_TYPES || top < 0 || top >= NUM_CELL_TYPES)
        return -1;
    if (precision == 32) {
        Phys<float> p = phys_f32(srate_mul); fill_tables(p);
        float o[17]; coeff_of_triple(p, cur, right, top, o);
        for (int k = 0; k < 17; ++k) out17[k] = o[k];
    } else {
        Phys<double> p = phys_f64(srate_mul); fill_tables(p);
        double o[17]; coeff_of_triple(p, cur, right, top, o);
        for (int k = 0; k < 17; ++k) out17[k] = o[k];
    }
    return 0;
}

int vso_plane_contour(float area_cm2, float* y, float* z, int cap, double* info6) {
    PlaneContour pc;
    if (!plane_contour(area_cm2, pc)) return -1;
    const int n = (int)pc.y.size();
    if (y && z) {
        if (cap < n) return -1;
        for (int i = 0; i < n; ++i) { y[i] = pc.y[i]; z[i] = pc.z[i]; }
    }
    if (info6) {
        info6[0] = pc.minY; info6[1] = pc.minZ; info6[2] = pc.maxZ;
        info6[3] = pc.diameter; info6[4] = pc.area; info6[5] = pc.steps;
    }
    return n;
}

float vso_polygon_area(const float* y, const float* z, int n) {
    std::vector<Pt> r(n);
    for (int i = 0; i < n; ++i) r[i] = {y[i], z[i]};
    return shoelace(r);
}

int vso_vertical_chord(const float* y, const float* z, int n, float y0, float* lo, float* hi) {
    std::vector<float> yy(y, y + n), zz(z, z + n);
    return vertical_chord(yy, zz, y0, *lo, *hi) ? 1 : 0;
}

int vso_segment_cross(const float* p4, const float* q4, float* out2) {
    Pt o{0, 0};
    if (!seg_cross({p4[0], p4[1]}, {p4[2], p4[3]}, {q4[0], q4[1]}, {q4[2], q4[3]}, o)) return 0;
    out2[0] = o.y; out2[1] = o.z;
    return 1;
}

}  // extern "C"
