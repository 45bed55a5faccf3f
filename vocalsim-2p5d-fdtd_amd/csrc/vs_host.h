// vs_host.h -- the library's host-only set-up code, free of HIP: the float32
// geometry prelude of one tube, the single-plane-symmetric contours, and the
// enumeration of the cell classes the step kernels read their coefficients
// from.  libvocalsim.so compiles vs_host.cpp with the host compiler; the
// same file builds on its own (g++, no ROCm), which is how the CPU tests run
// it under AddressSanitizer and UndefinedBehaviorSanitizer.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/vocalsim_abi.h"

namespace vs {

// Cell classes: the reference's CellTypes enum (include/SolverParams.h:21-35).
enum CellType : int {
    kWall = 0, kAir = 1, kExcitation = 2, kPml0 = 3, kPml1, kPml2, kPml3, kPml4, kPml5,
    kDynamic = 9, kDead = 10, kNoPressure = 11, kHead = 12, kNumTypes = 13
};

// Class byte stored per cell in HBM: low nibble = CellType, bit 4 = the cell
// carries a boundary segment id (z_inv applies, FDTDSolver.cu:1068).
constexpr uint8_t kTypeMask = 0x0f;
constexpr uint8_t kSegWallBit = 0x10;

// Physics constants and per-type tables of one tube, stored in double.
// FP32 tubes hold the float32-evaluated values widened exactly.
struct PhysTables {
    double dt, ds, rho, c, K1, K2, rds;   // rds = rho * ds
    double beta[16];
    double sigma[16];
    double air_normal[4];
    double wall_normal[4];
    double src_dir[4];
};

// Single-plane-symmetric cross-section of one tube section: the aligned
// closed contour (3997 points: 1999 on the upper half and their mirror
// images), its bounding values and its maximum chord (the section's sagittal
// width).
constexpr int kPlanePoints = 3997;
struct PlaneSection {
    std::vector<float> y, z;
    float minY = 0, minZ = 0, maxZ = 0;
    float diameter = 0;
    int grow_steps = 0;   // 0.01 mm growth steps taken to reach the area
};

// For `sections` contours of `npt` points each (y and z of section s at
// [s * npt, (s + 1) * npt)), every point's farthest point of the same section
// (the first index attaining the strict maximum) and that distance.  The
// library runs it on the device (k_max_chord); tests pass a CPU search.
// Returns VS_OK or a status code with `err` set.
using MaxChordFn = std::function<int(const std::vector<float>& y, const std::vector<float>& z, int sections, int npt,
                                     std::vector<float>& dist, std::vector<int>& far_idx, std::string& err)>;

// Integer geometry and per-section values of one tube.
struct Prelude {
    int H, W, domainH, domainW, L, ts, te, midY, mic, listenerX, listenerY, excX, excY;
    int num_steps;
    std::vector<float> cum;        // float32 inclusive scan of section lengths [m]
    std::vector<int> diam;         // diameter in cells (odd)
    std::vector<double> minor;     // ellipse minor axis [m] (Real-evaluated, stored in double)
    std::vector<double> zsec;      // z_inv per section
    std::vector<double> dscale;
    std::vector<PlaneSection> plane;   // single-plane-symmetric contours (shape 3)
    PhysTables ph;
    float ds32;
};

// SimulationParams.cu:7-33 in float32, SolverParams.cu:5-57, DeviceSymbols.cu:19-20.
void phys_tables(int srate_mul, bool fp64, PhysTables& t, float& ds32);
// Wall admittance -> z_inv (FDTDSolver.cu:613-617).
double z_inv_of(float mu, bool fp64, const PhysTables& t);
// Time stamps accumulated in float32 (SolverParams.cu:59-66).
int float_step_count(float dur, float dt);

// The prelude of one tube (TubeGeometry.cu:53-179, 241-286;
// FDTDSolver.cu:479-617).  Returns VS_OK, or VS_E_INVALID / VS_E_GEOMETRY /
// the max-chord search's code with `err` set.
int prelude(const vs_tube_spec& s, bool fp64, Prelude& p, const MaxChordFn& max_chord, std::string& err);

// Contours of n single-plane-symmetric sections (ShapeGenerator.cpp:35-318,
// ShapeAlignment.cu:73-152), the maximum chord from `max_chord`.
int plane_sections(const float* area_cm2, int n, std::vector<PlaneSection>& out, const MaxChordFn& max_chord,
                   std::string& err);

// Cell classes of the exact step.  The 17 coefficients of the reference
// (FDTDSolver.cu:860-1056) depend only on the (cur, right, top) cell types,
// so each tube holds them per class: slots 0..16 the coefficients, 17 =
// Dirichlet cell, 18 = wall cell carrying a boundary segment (z_inv
// applies), 19 = flags of the terms the velocity update cannot skip
// (kExactSourceTerms, kExactWallTerms).
constexpr int kExactClasses = 64;
constexpr int kExactSlots = 20;        // 17 coefficients, noPressure, wall, term flags
constexpr int kExactSourceTerms = 1;   // term flags: the class has source coefficients
constexpr int kExactWallTerms = 2;     //             the class carries a wall segment

// For one H x W tube frame of class bytes `cls`: the distinct keys
// cur | segment-wall << 4 | right << 5 | top << 9 of its interior cells in
// first-seen row-major order (keys[0 .. nk)), and each interior cell's index
// into them (xcls; border cells are left as they are).  VS_E_GEOMETRY when
// the frame has more than kExactClasses keys.
int enumerate_exact_classes(const uint8_t* cls, int H, int W, uint8_t* xcls, int* keys, int& nk, std::string& err);

// Cell classes of the FP32 flux-form step: the distinct damping-factor
// triples and face-form bits of the batch become classes 1 .. kFastClasses - 1
// in first-seen order; class 0 is every damping factor zero with no
// face-form bit (the pad cells).  A cell's 16-bit face code holds, in bits
// [3, 14), code_x + 8 code_y + 64 code_p (3-bit damping codes), bit 12
// self-off and bit 13 source face; the cells [bounds[g], bounds[g + 1]) of
// group g (a tube) map damping codes to factors through luts[8 g .. 8 g + 8)
// (the wall damping 1 / (1 + dt) depends on the tube's sample rate), so
// cells of tubes with different tables share a class only where their
// factors are equal.  Per class, `tab` (5 x kFastClasses floats) holds aP,
// ax, ay and two 0/1 factors: s = 1 unless the cell's own pressure is left
// out of its face updates (walls carrying an admittance, excitation) and
// r = 1 unless the x-face carries the source instead of the right
// neighbour's pressure (excitation).  ax and ay carry s = 0, resp. r = 0, in
// their sign bits.  VS_E_INVALID when the batch has more classes or the
// bounds do not run from 0 upwards.
constexpr int kFastClasses = 64;
int enumerate_fast_classes(const uint16_t* code, const size_t* bounds, const float* luts, int groups, uint8_t* c8,
                           float* tab, int& ncls, std::string& err);

// Padded row length of the wide-access on-chip step at K cells per lane (K a
// multiple of 4): the interior width plus at least one zero pad cell,
// rounded up to a multiple of K, so that every cell of a column sits at the
// same position k of its lane's run in every row.
constexpr int fast_wide_wp(int interior_width, int K) { return (interior_width + 1 + K - 1) / K * K; }
// Cells per lane of the wide-access form, smallest first.
constexpr int kFastWideKs[3] = {4, 12, 20};

// Pair classes of the wide-access on-chip step.  Its lanes own runs of an
// even number of consecutive cells of the padded interior j = r * Wp + c
// (r < Hi, c < Wp, Wp even; columns c >= Wi are pad cells of class 0), so
// the cells j, j + 1 with j even form pairs that never cross a row.
// A pair class is the ordered pair (class of j, class of j + 1) in
// first-seen order over the tubes (pair class 0 is two pad cells).  For tube
// g (frame H[g] x W[g] at cell offset off[g] of the class bytes c8, the
// output of enumerate_fast_classes), each pair whose first cell is a real
// cell (c even, c < Wi) gets its pair class at that cell's frame position in
// p8; other positions of p8 are left as they are.  Per pair class p, `ptab`
// (6 x kFastPairs floats) holds the face entry (ax, ay of the first cell, ax,
// ay of the second) at [4p, 4p + 4), with s = 0 and r = 0 in their sign bits
// as in the class table, and the pressure entry (aP of the first cell, aP of
// the second) at [4 kFastPairs + 2p, 4 kFastPairs + 2p + 2).  Sets npairs to
// the number of pair classes; when it exceeds kFastPairs, returns VS_OK with
// p8 and ptab incomplete (the wide-access form is then not used).  For each
// tube, src_col[g] is the interior column of its cells whose x-face carries
// the source (r = 0), -1 when it has none and -2 when they lie in more than
// one column (the wide-access form then is not used either: it applies the
// source term only in the lane-run group of that column).  VS_E_INVALID when
// a tube's frame is smaller than 3 x 3.
constexpr int kFastPairs = 128;
int enumerate_fast_pairs(const uint8_t* c8, const float* tab, const int* H, const int* W, const long long* off,
                         int groups, uint8_t* p8, float* ptab, int& npairs, int* src_col, std::string& err);

}  // namespace vs
