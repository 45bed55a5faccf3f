// vs_common.h -- types shared by the host setup code and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "vs_host.h"

// the two-mass recurrence (vs_twomass.h), callable from host and device code
#define VS_HD __host__ __device__
#include "vs_twomass.h"

namespace vs {

// Status bits raised by device code (read back after setup / each run).
constexpr unsigned kStatusSqrtDomain = 1u;   // negative depth sqrt argument
constexpr unsigned kStatusNonFinite = 2u;    // non-finite probe sample
constexpr unsigned kStatusBarrierTimeout = 4u;   // a cooperative step's grid barrier gave up waiting

// Device descriptor of one tube inside a batch.
struct TubeDesc {
    int H, W, L, nsec;
    int ts, te, midY;
    int listener, excitation;     // flat indices inside the tube frame
    int num_steps;
    int source_type, dirichlet, variant, srate_mul;
    int shape, solver_type;
    float ds_f32;
    int pml;                      // PML rings painted (pml_flag)
    int open_space;               // open-space domain: no tube, point source
    long long cell_off;           // start of the tube's H*W block
    long long col_off;            // start of its W-long column arrays
    long long wall_off;           // start of its L-long tube-column arrays
    long long sec_off;            // start of its section arrays
    long long src_off;            // start of its source table / traces (num_steps)
    PhysTables ph;
};

// Device buffers of a batch (packed over tubes).
template <typename Real>
struct Bufs {
    Real* P; Real* Vx; Real* Vy;
    Real* dP; Real* dX; Real* dY;
    Real* zcol;          // z_inv of the wall cells of each column
    uint8_t* cls;        // class byte per cell
    int* bst;            // boundary segment id per cell (setup / export only)
    int* upper; int* lower; int* seg;   // per tube column
    float* cum;          // per section: float32 cumulative length [m]
    int* diam;           // per section: diameter in cells
    Real* minor;         // per section: ellipse minor axis [m]
    Real* dscale;        // per section: depth scale
    Real* zsec;          // per section: z_inv
    Real* source;        // per step: precomputed excitation
    double* probe;       // per step: listener pressure
    double* strace;      // per step: injected source velocity
    TwoMassState* tm;    // per tube
    double* pexc;        // per tube: P at the excitation cell after the last step
    unsigned* status;    // device status word
    float* poly_y;       // per section: aligned contour (kPlanePoints), single-plane shape only
    float* poly_z;
    float* pbox;         // per section: minY, minZ, maxZ of the aligned contour
    uint8_t* xcls;       // per cell: index of its (cur, right, top, segment-wall) class in its tube
    int* xkey;           // per tube: kExactClasses class keys (cur | wall << 4 | right << 5 | top << 9)
    int* xn;             // per tube: number of classes
    Real* xtab;          // per tube: kExactClasses x kExactSlots coefficients of each class
};

// Maximum chords of single-plane-symmetric contours on the device (the
// max-chord search of plane_sections), on stream `st`; synchronises it.
int max_chord_on_device(hipStream_t st, const std::vector<float>& y, const std::vector<float>& z, int sections,
                        int npt, std::vector<float>& dist, std::vector<int>& far_idx, std::string& err);

// Per-cell coefficients of the FP32 flux-form kernel (vs_fast.hip).
struct FastBufs {
    float* kP; float* gX; float* gY;     // pressure and face coefficients
    float* Uxf; float* Uyf;              // flux state U = V * depth
    uint16_t* code;                      // damping codes + face-form bits
    uint8_t* cls8;                       // class byte (< kFastClasses): damping triple + face-form bits
    float* ctab;                         // per class: aP, ax (sign: !s), ay (sign: !r), s, r (kFastClasses each)
    uint8_t* pcls8;                      // wide form: pair class at the frame position of a pair's first cell
    float* ptab;                         // per pair class: (ax, ay, ax, ay) x kFastPairs, then (aP, aP) x kFastPairs
};

// Buffers of the reverse-time (adjoint) pass (vs_adjoint.hip).  Per-cell
// arrays share the cell layout of Bufs; ck and seg are per-tube slabs of
// whole states (P, Ux, Uy; 3 * H * W values each).
template <typename Real>
struct AdjBufs {
    Real *aP, *kP, *ax, *ay, *cx, *cy;   // flux-form coefficients
    uint8_t *fx, *fy;                    // face kind: 0 ordinary, 1 wall admittance, 2 source
    Real *P, *Ux, *Uy;                   // forward state
    Real *LP, *LX, *LY, *M;              // adjoint state and the pressure adjoint of a step
    double *GkP, *Gx, *Gy;               // accumulated d J / d (kP, cx, cy)
    Real* ck;                            // checkpoints, every S steps
    Real* seg;                           // the S + 1 states of the segment being reversed
    long long *ck_off, *seg_off;         // per tube slab offsets
    double* probe;                       // per step: listener pressure of the flux form
    double* gprobe;                      // per step: d J / d probe
    int S;
};

// On-chip FP32 reverse pass (vs_fast.hip).  With the scaled adjoint
// variables  Kp = kP * M,  Lx = gX * lambda_x,  Ly = gY * lambda_y  the
// reverse step has the stencil of the forward step (pressure-like phase
// reading Lx on the left and Ly below, face phase reading Kp on the right
// and above), so it runs in the same on-chip layout.  Slabs hold, per tube
// (launch order) and per step of a block, fields of K * T floats in
// lane-slot order s = k * T + t (slot k of cell lane t):
//   rec  the forward step's new pressure P' and flux divergence (2 fields)
//   adj  Kp of the reverse step (1 field)
//   start  Lx, Ly as the block's reverse begins (2 fields, per tube)
// The sums kernel replays the Lx, Ly recurrences from start and adj and
// forms the forward factors from rec; the sums over steps of their products
// are the log-derivatives of J with respect to each cell's gX, gY and kP.
struct FastAdj {
    float* rec;
    float* adj;
    float* start;
    float* state;          // per tube 3 * K * T: Kp, Ly, Lx between launches
    float* sink;           // 2 * K * T: stores of waves past every cell of their tube
    long long rec_tube;    // floats per tube of rec (B * 2 * K * T)
    long long adj_tube;    // floats per tube of adj (B * K * T)
    int lanes;             // T: cell lanes per work-group
    const double* gprobe;  // per step: d J / d probe
    double *Lx, *Ly, *Lp;  // d J / d log gX, d log gY, d log kP (cell layout)
};

// Second-order IIR section, a[0] = 1 (vs_audio.hip).
struct AudioFilter {
    double b[3];
    double a[3];
};
// One probe trace to turn into audio: input offset/length in the probe
// array, output offset, decimation factor (srate_mul) and its filter.
struct AudioJob {
    long long in_off, out_off;
    int n, decim;
    AudioFilter flt;
};
AudioFilter butter2_lowpass(double fc, double fs);
hipError_t launch_audio(const double* d_probe, const AudioJob* d_jobs, int njobs, float* d_out, hipStream_t st);

// Kernel launchers (vs_kernels.hip, vs_fast.hip, vs_adjoint.hip).
// Allocation-free FP64 transform (vs_dtft_noalloc): limits of its static
// device scratch array.
constexpr int kNoAllocMaxSamples = 1 << 18;
constexpr int kNoAllocMaxFreqs = 1 << 14;
hipError_t dtft_noalloc(const double* x, int n, double fs, int hann, const double* freqs, int nf, double* re,
                        double* im);
hipError_t measure_lds_read(int width_words, int groups, int iters, float* d_out, hipStream_t st, float* ms);
hipError_t launch_dtft(const double* d_x, long long stride, int nsig, int n, double fs, int hann,
                       const double* d_f, int nf, double* d_re, double* d_im, hipStream_t st);
// On-chip FP32 kernel sizing: cells per lane K (odd, at most kFastMaxK) for the
// batch's largest padded interior Hi * (Wi + 1), zero guard lanes on each
// side of the LDS regions for its widest padded row Wi + 1, and the dynamic
// LDS bytes of a launch.
constexpr int kFastMaxK = 23;
// Wide-access variant (k_fast_wide): K in kFastWideKs (4 mod 8), rows padded
// to a multiple of row_quantum cells (K, or 4; fast_wide_wp, vs_host.h).
// ms: the lane-run group (4 cells) holding the source column of every tube of
// the batch when rows are padded to a multiple of K, -1 when no tube has a
// source face, -2 when every group must handle it.
constexpr int kFastWideMaxK = 20;
size_t fast_wide_lds_bytes(int K, int threads, int guard_floats);
size_t fast_wide_lds_static_bytes();
hipError_t configure_fast_wide(int K, int ms);
hipError_t launch_fast_wide(int K, int ms, const TubeDesc* d_tubes, const int* d_order, int n, Bufs<float> b,
                            FastBufs f, int t0, int nsteps, int threads, int guard_floats, int row_quantum,
                            const int* d_t0, hipStream_t st);
int fast_cells_per_lane(int max_padded_cells, int threads);
int fast_guard_lanes(int K, int max_padded_width);
size_t fast_lds_bytes(int K, int threads, int guard);
size_t fast_lds_static_bytes();
hipError_t launch_fast_setup(const TubeDesc* d_tubes, const TubeDesc* h_tubes, int n, Bufs<float> b,
                             FastBufs f, hipStream_t st);
hipError_t launch_fast_step(int K, const TubeDesc* d_tubes, const int* d_order, int n, Bufs<float> b,
                            FastBufs f, int t0, int nsteps, int threads, int guard, const int* d_t0,
                            hipStream_t st);
// The on-chip step that also writes the forward factors of every step into
// a.rec (slot n - t0 of the segment starting at t0).
hipError_t launch_fast_record(int K, const TubeDesc* d_tubes, const int* d_order, int n, Bufs<float> b,
                              FastBufs f, FastAdj a, int t0, int nsteps, int threads, int guard, hipStream_t st);
// Reverse steps s0 + ns - 1 ... s0 on chip (a.lanes threads per tube), then
// the per-cell sums of the segment into a.Lx, a.Ly, a.Lp.
hipError_t launch_fast_reverse(int K, const TubeDesc* d_tubes, const int* d_order, int n, FastBufs f,
                               FastAdj a, int s0, int ns, int guard, hipStream_t st);
hipError_t launch_fast_adj_sums(int K, const TubeDesc* d_tubes, const int* d_order, int n, FastBufs f, FastAdj a,
                                const float* source, int s0, int ns, int groups, hipStream_t st);
hipError_t launch_bump(int* d_t0, int s, hipStream_t st);
hipError_t configure_fast(int K);
template <typename Real>
hipError_t launch_setup(const TubeDesc* d_tubes, const TubeDesc* h_tubes, int n, Bufs<Real> b,
                        Real* tmpA, Real* tmpB, Real* tmpC, Real* tmpD, hipStream_t st);
template <typename Real>
hipError_t launch_step(const TubeDesc* d_tubes, const int* d_order, int n, Bufs<Real> b,
                       int t0, int nsteps, int threads, int kernel, const int* d_t0, hipStream_t st);
// Exact step of a single tube over `groups` work-groups of one cooperative
// launch, s steps per grid barrier (one: coop_one_cell_per_lane; bar: two
// words of barrier state, zeroed by the launch; P2, Vx2, Vy2: second state
// buffers, cell layout, zero frame ring; tile_bytes: coop_tile_bytes, at
// most kCoopTileMax; spin_limit: polls of a barrier before it gives up).
template <typename Real>
hipError_t launch_step_coop(const TubeDesc* d_tubes, const int* d_order, Bufs<Real> b, int t0, int nsteps,
                            int threads, int groups, int s, bool one, size_t tile_bytes, unsigned* bar, Real* P2,
                            Real* Vx2, Real* Vy2, int spin_limit, hipStream_t st);
constexpr int kCoopSpinDefault = 1 << 22;   // grid-barrier polls before giving up: about a quarter of a second
// LDS tile of one work-group of that step: P, Vx, Vy of H rows by the
// widest column range (own columns plus s on each side).
template <typename Real>
size_t coop_tile_bytes(int H, int W, int groups, int s);
constexpr size_t kCoopTileMax = 144 * 1024;   // dynamic LDS a cooperative step's tile may use
constexpr int kCoopMaxGroups = 64;             // most work-groups of a cooperative step
constexpr int kCoopMaxSteps = 32;              // most steps per grid barrier

// Column partition of that step: work-group g of `groups` owns the frame
// columns [*a, *e) (empty when *a >= *e) of a W-wide frame.
__host__ __device__ inline int coop_chunk(int W, int groups) { return (W - 2 + groups - 1) / groups; }
__host__ __device__ inline void coop_columns(int W, int groups, int g, int* a, int* e) {
    *a = 1 + g * coop_chunk(W, groups);
    const int end = *a + coop_chunk(W, groups);
    *e = end < W - 1 ? end : W - 1;
}
// The work-group owning column c (ring columns go to the nearest owner).
__host__ __device__ inline int coop_owner(int W, int groups, int c) {
    const int cc = c < 1 ? 1 : (c > W - 2 ? W - 2 : c);
    return (cc - 1) / coop_chunk(W, groups);
}
// Does every work-group's first pressure region (own columns, s - 1 more on
// the left and s on the right, within the interior) hold at most one cell
// per lane of the waves after the first of a `threads`-lane work-group.
__host__ __device__ inline bool coop_one_cell_per_lane(int H, int W, int groups, int s, int threads) {
    const int w = coop_chunk(W, groups) + 2 * s - 1;
    return (long long)(H - 2) * (w < W - 2 ? w : W - 2) <= threads - 64;
}
// Does work-group g run the source: its first velocity region (own columns
// and s - 1 more on each side) reaches the excitation column ex or the one
// left of it (the cells with source coefficients), or it owns ex.
__host__ __device__ inline bool coop_runs_source(int W, int groups, int s, int g, int ex) {
    int a, e;
    coop_columns(W, groups, g, &a, &e);
    if (a >= e) return false;
    const int lo = a - (s - 1), hi = e + (s - 1);
    return g == coop_owner(W, groups, ex) || (ex >= lo && ex < hi) || (ex - 1 >= lo && ex - 1 < hi);
}
// A source work-group reads the excitation cell's pressure of sub-steps
// 1 .. s-1 from its own tile, so for s >= 2 every source work-group needs
// the excitation column within [a - 1, e + 2).
__host__ __device__ inline bool coop_source_ok(int W, int groups, int s, int ex) {
    if (s < 2) return true;
    for (int g = 0; g < groups; ++g) {
        if (!coop_runs_source(W, groups, s, g, ex)) continue;
        int a, e;
        coop_columns(W, groups, g, &a, &e);
        if (ex < a - 1 || ex >= e + 2) return false;
    }
    return true;
}
template <typename Real>
hipError_t launch_coeffs(const TubeDesc* d_tubes, const TubeDesc* h_tube, int tube, Bufs<Real> b,
                         Real* out17, hipStream_t st);
template <typename Real>
hipError_t launch_exact_classes(const TubeDesc* d_tubes, int n, Bufs<Real> b, hipStream_t st);
template <typename Real>
hipError_t launch_depth_raw(const TubeDesc* d_tubes, const TubeDesc* h_tubes, int n, Bufs<Real> b, Real* X,
                            Real* Y, long long stride, hipStream_t st);
template <typename Real>
hipError_t launch_adj_coeffs(const TubeDesc* d_tubes, const TubeDesc* h_tubes, int n, Bufs<Real> b,
                             AdjBufs<Real> a, hipStream_t st);
template <typename Real>
hipError_t launch_adj_fwd(const TubeDesc* d_tubes, const int* d_order, int n, Bufs<Real> b, AdjBufs<Real> a,
                          int t0, int ns, int mode, int load, int threads, hipStream_t st);
template <typename Real>
hipError_t launch_adj_rev(const TubeDesc* d_tubes, const int* d_order, int n, Bufs<Real> b, AdjBufs<Real> a,
                          int s0, int ns, int threads, hipStream_t st);

}  // namespace vs
