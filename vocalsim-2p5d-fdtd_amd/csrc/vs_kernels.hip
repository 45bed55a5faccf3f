// vs_kernels.hip -- HIP kernels for gfx950: setup (walls, cell classes,
// depth maps) and the FDTD time step.
//
// Every kernel here rounds each floating-point operation separately
// (-ffp-contract=off for the translation unit) so that the FP64 path
// reproduces the oracle's operation order bit for bit.

#include "vs_common.h"

#include <atomic>

namespace vs {

namespace {

constexpr int kSetupThreads = 256;

__device__ inline float sqrt_r(float x) { return __builtin_sqrtf(x); }
__device__ inline double sqrt_r(double x) { return __builtin_sqrt(x); }

// ---------------------------------------------------------------------------
// Setup 1: wall rows per tube column.  One lane per column: segment lookup
// over the float32 cumulative lengths, then the wall rows mid -/+ (r + 1)
// (TubeGeometry.cu:443-484).  Also fills the per-column z_inv of the walls.
// ---------------------------------------------------------------------------
template <typename Real>
__global__ void k_walls(const TubeDesc* __restrict__ tubes, Bufs<Real> b) {
    const TubeDesc& T = tubes[blockIdx.y];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= T.L) return;
    const float ds = T.ds_f32;
    const float curr_len = (idx + 1) * ds;
    const float half = ds * 0.5f;
    const float* cum = b.cum + T.sec_off;
    int k = 0;
    while (k < T.nsec - 1 && curr_len > cum[k] + half) ++k;
    const int radius = (b.diam[T.sec_off + k] - 1) / 2;
    b.upper[T.wall_off + idx] = T.midY - radius - 1;
    b.lower[T.wall_off + idx] = T.midY + radius + 1;
    b.seg[T.wall_off + idx] = k;
    b.zcol[T.col_off + T.ts + idx] = b.zsec[T.sec_off + k];
}

// ---------------------------------------------------------------------------
// Setup 2: cell class of every cell in closed form.  Reproduces the result
// of the reference's ordered writes: walls and gap fill
// (TubeGeometry.cu:433-569), excitation column (:571-594), Dirichlet column
// (:596-612), dead ring (FDTDSolver.cu:816-835) and PML rings painted
// horizontally then vertically (FDTDSolver.cu:371-464), whose net effect is
// pml5 - min(ring_row, ring_col).
// ---------------------------------------------------------------------------
template <typename Real>
__global__ void k_classify(const TubeDesc* __restrict__ tubes, Bufs<Real> b) {
    const TubeDesc& T = tubes[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int H = T.H, W = T.W;
    if (i >= H * W) return;
    const int r = i / W, c = i - r * W;
    const int* up = b.upper + T.wall_off;
    const int* lo = b.lower + T.wall_off;
    const int* sg = b.seg + T.wall_off;
    const int L = T.L;
    int type = kAir, bst = 0;
    if (T.open_space) {
        // FDTDSolver.cu:519-545: air everywhere, one excitation cell
        if (i == T.excitation) type = kExcitation;
    } else if (c >= T.ts && c <= T.te) {
        const int j = c - T.ts;
        const int uj = up[j], lj = lo[j];
        if (r == uj || r == lj) { type = kWall; bst = sg[j] + 1; }
        // gap fill while this column is the current one
        if (j >= 1 && j <= L - 2) {
            const int d = uj - up[j - 1];
            if (d > 1 && ((r >= up[j - 1] && r <= uj) || (r >= lj && r <= lo[j - 1]))) {
                type = kWall; bst = sg[j] + 1;
            }
        }
        // gap fill while this column is the previous one
        if (j + 1 >= 1 && j + 1 <= L - 2) {
            const int d = up[j + 1] - uj;
            if (d < -1 && ((r >= up[j + 1] && r <= uj) || (r >= lj && r <= lo[j + 1]))) {
                type = kWall; bst = sg[j] + 1;
            }
        }
    } else if (c == T.ts - 1) {
        if (r > up[0] && r < lo[0]) type = kExcitation;
        else if (r == up[0] || r == lo[0]) type = kWall;
    } else if (T.dirichlet && c == T.te + 1) {
        if (r >= up[L - 1] && r <= lo[L - 1]) type = kNoPressure;
    }
    if (r == 0 || r == H - 1 || c == 0 || c == W - 1) {
        type = kDead;
    } else if (T.pml) {
        const int kr = min(r - 1, H - 2 - r), kc = min(c - 1, W - 2 - c);
        const int k = min(kr, kc);
        if (k <= 5) type = kPml5 - k;
    }
    b.cls[T.cell_off + i] = (uint8_t)(type | (bst != 0 ? kSegWallBit : 0));
    b.bst[T.cell_off + i] = bst;
}

// Lowest and highest crossing of the vertical line y = y0 with the closed
// contour (kPlanePoints points); crossing heights are evaluated in double and
// rounded to float, as a float polyline intersection returns them.
__device__ inline bool contour_chord(const float* __restrict__ py, const float* __restrict__ pz, float y0,
                                     float& zlo, float& zhi) {
    zlo = __builtin_inff();
    zhi = -__builtin_inff();
    for (int k = 0; k + 1 < kPlanePoints; ++k) {
        const float ya = py[k], yb = py[k + 1];
        if (!((ya <= y0 && y0 <= yb) || (yb <= y0 && y0 <= ya))) continue;
        const float za = pz[k], zb = pz[k + 1];
        if (ya == yb) {
            zlo = fminf(zlo, fminf(za, zb));
            zhi = fmaxf(zhi, fmaxf(za, zb));
        } else {
            const double t = ((double)y0 - (double)ya) / ((double)yb - (double)ya);
            const float z = (float)((double)za + ((double)zb - (double)za) * t);
            zlo = fminf(zlo, z);
            zhi = fmaxf(zhi, z);
        }
    }
    return zlo <= zhi;
}

// ---------------------------------------------------------------------------
// Setup 3: raw depth maps (TubeDepth.cu:313-490) in closed form per cell.
// Rows above the centre take the "upper" branch, rows below the "lower"
// branch; the centre row keeps the lower branch's depthY, which the
// reference writes last.  Cells outside the tube keep the open-space value
// 0.5 (TubeDepth.cu:108-110).
// ---------------------------------------------------------------------------
template <typename Real>
__global__ void k_depth_raw(const TubeDesc* __restrict__ tubes, Bufs<Real> b, Real* X, Real* Y,
                            long long tmp_stride) {
    const TubeDesc& T = tubes[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int H = T.H, W = T.W;
    if (i >= H * W) return;
    X += (long long)blockIdx.y * tmp_stride;
    Y += (long long)blockIdx.y * tmp_stride;
    const int r = i / W, c = i - r * W;
    Real x = (Real)(1.0 * (Real)0.50f), y = x;
    if (c >= T.ts && c <= T.te) {
        const int j = c - T.ts;
        const int up = b.upper[T.wall_off + j], lo = b.lower[T.wall_off + j];
        if (r == up || r == lo) {
            x = 0; y = 0;
        } else if (r > up && r < lo) {
            const Real ds = (Real)T.ph.ds;
            const int air = lo - up - 1;
            const int half = (int)ceilf(air * 0.5f) - 1;
            const int mid = up + half + 1;
            const int sec = b.seg[T.wall_off + j];
            const Real scale = b.dscale[T.sec_off + sec];
            int k;
            bool upper_side;
            if (r < mid) { k = up + half + 1 - r; upper_side = true; }
            else { k = r - mid; upper_side = false; }
            const Real d = static_cast<Real>(k) * ds;
            const Real dy = upper_side ? (static_cast<Real>(k) + 0.5f) * ds : (static_cast<Real>(k) - 0.5f) * ds;
            Real sx, sy;
            if (T.shape == 3) {
                // single-plane-symmetric contour: half the vertical chord at the
                // face positions counted from the lower wall (TubeDepth.cu:535-626)
                const int cc = lo - r;
                const long long ps = T.sec_off + sec;
                const float* py = b.poly_y + ps * kPlanePoints;
                const float* pz = b.poly_z + ps * kPlanePoints;
                const float minY = b.pbox[3 * ps];
                const float dsf = T.ds_f32;
                const float posx = (float)((cc - 0.5) * (double)dsf);
                const float posy = cc * dsf;
                float zlo, zhi;
                bool okx = contour_chord(py, pz, minY + posx, zlo, zhi);
                const Real hx = okx ? ((Real)zhi - (Real)zlo) / (Real)2.0f : (Real)0;
                bool oky = true;
                Real hy = 0;
                if (cc < air) {
                    oky = contour_chord(py, pz, minY + posy, zlo, zhi);
                    hy = oky ? ((Real)zhi - (Real)zlo) / (Real)2.0f : (Real)0;
                }
                if (!okx || !oky) atomicOr(b.status, kStatusSqrtDomain);
                X[i] = hx * scale;
                Y[i] = hy * scale;
                return;
            }
            if (T.shape == 1) {
                const Real radius = (static_cast<Real>(air) / 2.0f) * ds;
                const Real r2 = radius * radius;
                sx = r2 - d * d;
                sy = (upper_side && k == half) ? (Real)0.0f : r2 - dy * dy;
            } else {
                const Real a = (air * ds) / 2.0f;
                const Real bb = b.minor[T.sec_off + sec] / 2.0f;
                const Real sa2 = a * a, sb2 = bb * bb;
                sx = sb2 * (1.0f - ((d * d) / sa2));
                sy = (upper_side && k == half) ? (Real)0.0f : sb2 * (1 - ((dy * dy) / sa2));
            }
            if (sx < 0 || sy < 0) {
                atomicOr(b.status, kStatusSqrtDomain);
                sx = sx < 0 ? (Real)0 : sx;
                sy = sy < 0 ? (Real)0 : sy;
            }
            x = 2 * sqrt_r(sx) * scale;
            y = 2 * sqrt_r(sy) * scale;
        }
    }
    X[i] = x;
    Y[i] = y;
}

// Setup 4: open-space reset + half-cell smoothing (TubeDepth.cu:642-712).
template <typename Real>
__global__ void k_depth_smooth(const TubeDesc* __restrict__ tubes, const Real* X, const Real* Y,
                               Real* Xs, Real* Ys, long long tmp_stride) {
    const TubeDesc& T = tubes[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int H = T.H, W = T.W;
    if (i >= H * W) return;
    const long long o = (long long)blockIdx.y * tmp_stride;
    X += o; Y += o; Xs += o; Ys += o;
    const int r = i / W, c = i - r * W;
    const int mid = T.midY * W + T.te;
    const Real open = (Real)0.50f;
    const Real mx = X[mid], my = Y[mid];
    Real xc = X[i]; if (xc == open) xc = mx;
    Real yc = Y[i]; if (yc == open) yc = my;
    Real xr = mx;
    if (c < W - 1) { xr = X[i + 1]; if (xr == open) xr = mx; }
    Real ya = my;
    if (r > 0) { ya = Y[i - W]; if (ya == open) ya = my; }
    Xs[i] = 0.5f * (xc + xr);
    Ys[i] = 0.5f * (yc + ya);
}

// Setup 5: depthP average and min-depth clamp (TubeDepth.cu:714-762).
template <typename Real>
__global__ void k_depth_final(const TubeDesc* __restrict__ tubes, Bufs<Real> b, const Real* Xs,
                              const Real* Ys, long long tmp_stride) {
    const TubeDesc& T = tubes[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int H = T.H, W = T.W;
    if (i >= H * W) return;
    const long long o = (long long)blockIdx.y * tmp_stride;
    Xs += o; Ys += o;
    const int r = i / W, c = i - r * W;
    const int mid = T.midY * W + T.te;
    const Real cX = Xs[i], cY = Ys[i];
    const Real lX = (c > 0) ? Xs[i - 1] : Xs[mid];
    const Real dY = (r < H - 1) ? Ys[i + W] : Ys[mid];
    Real p = 0.25f * (cX + cY + lX + dY);
    const Real mn = (Real)0.001f;
    Real x = cX, y = cY;
    if (p < mn) p = mn;
    if (x < mn) x = mn;
    if (y < mn) y = mn;
    if (T.open_space) {   // constant open-space depth, no smoothing (FDTDSolver.cu:504-514)
        const Real od = (Real)(1.0 * (Real)0.50f);
        p = od; x = od; y = od;
    }
    if (T.solver_type == 0) { p = 1.0; x = 1.0; y = 1.0; }   // 2D: unit depth (FDTDSolver.cu:632-635)
    b.dP[T.cell_off + i] = p;
    b.dX[T.cell_off + i] = x;
    b.dY[T.cell_off + i] = y;
}

// ---------------------------------------------------------------------------
// Per-face coefficients of a (cur, right, top) type triple, in the reference's
// arithmetic (FDTDSolver.cu:860-1056).
// ---------------------------------------------------------------------------
template <typename Real>
struct Tables {
    Real beta[16], sigma[16], airN[4], wallN[4], dir[4];
    Real dt, rds, K1;
};

template <typename Real>
__device__ inline void coeffs_of(const Tables<Real>& t, int tc, int tr, int tt, Real* o) {
    const Real bc = t.beta[tc], br = t.beta[tr], bt = t.beta[tt];
    const int ibr = (int)br, ibt = (int)bt;
    const int lut = (ibr == kAir && ibt == kAir) ? 0 : (ibr == kAir && ibt == kWall) ? 1
                  : (ibr == kWall && ibt == kAir) ? 2 : 3;
    const Real sc = t.sigma[tc], sr = t.sigma[tr], st = t.sigma[tt];
    const int ec = tc == kExcitation, er = tr == kExcitation, et = tt == kExcitation;
    const Real mbx = (br < bc) ? br : bc;
    const Real mby = (bt < bc) ? bt : bc;
    o[0] = t.airN[lut] * bc;
    o[1] = t.wallN[lut] * (1 - bc);
    o[2] = mbx;
    o[3] = mby;
    o[4] = (sc < sr) ? sr : sc;
    o[5] = (sc < st) ? st : sc;
    o[6] = (mbx * mbx * t.dt) / t.rds;
    o[7] = (mby * mby * t.dt) / t.rds;
    o[8] = sc;
    o[9] = (Real)((1 - ec) * (1 - er));
    o[10] = (Real)((1 - ec) * (1 - et));
    o[11] = (ec * t.dir[2]) + (er * t.dir[0]);
    o[12] = (ec * t.dir[3]) + (et * t.dir[1]);
    o[13] = br * (1 - bc);
    o[14] = bc * (1 - br);
    o[15] = bt * (1 - bc);
    o[16] = bc * (1 - bt);
}

template <typename Real>
__device__ inline void load_tables(const TubeDesc& T, Tables<Real>& t) {
    for (int k = 0; k < 16; ++k) { t.beta[k] = (Real)T.ph.beta[k]; t.sigma[k] = (Real)T.ph.sigma[k]; }
    for (int k = 0; k < 4; ++k) {
        t.airN[k] = (Real)T.ph.air_normal[k];
        t.wallN[k] = (Real)T.ph.wall_normal[k];
        t.dir[k] = (Real)T.ph.src_dir[k];
    }
    t.dt = (Real)T.ph.dt;
    t.rds = (Real)T.ph.rds;
    t.K1 = (Real)T.ph.K1;
}

// Materialises the 17 interior coefficient grids of one tube (export only:
// the step kernel derives the same values from the class bytes on the fly).
template <typename Real>
__global__ void k_coeffs(const TubeDesc* __restrict__ tubes, int tube, Bufs<Real> b, Real* out) {
    const TubeDesc& T = tubes[tube];
    const int H = T.H, W = T.W, Wi = W - 2;
    const int n = (H - 2) * Wi;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Tables<Real> t;
    load_tables(T, t);
    const int r = 1 + i / Wi, c = 1 + i % Wi;
    const uint8_t* cls = b.cls + T.cell_off;
    const int q = r * W + c;
    Real o[17];
    coeffs_of(t, cls[q] & kTypeMask, cls[q + 1] & kTypeMask, cls[q - W] & kTypeMask, o);
    for (int k = 0; k < 17; ++k) out[(size_t)k * n + i] = o[k];
}

// Coefficients of every class of every tube, in the arithmetic of coeffs_of.
template <typename Real>
__global__ void k_exact_classes(const TubeDesc* __restrict__ tubes, Bufs<Real> b) {
    const int tube = blockIdx.x, c = threadIdx.x;
    if (c >= b.xn[tube]) return;
    const TubeDesc& T = tubes[tube];
    Tables<Real> t;
    load_tables(T, t);
    const int key = b.xkey[tube * kExactClasses + c];
    const int tc = key & 0xf, wall = (key >> 4) & 1, tr = (key >> 5) & 0xf, tt = (key >> 9) & 0xf;
    Real o[17];
    coeffs_of(t, tc, tr, tt, o);
    Real* dst = b.xtab + ((long long)tube * kExactClasses + c) * kExactSlots;
    for (int k = 0; k < 17; ++k) dst[k] = o[k];
    dst[17] = (Real)(tc == kNoPressure ? 1 : 0);
    dst[18] = (Real)wall;
    // terms the velocity update may skip: bit 0 set when the source
    // coefficients (11, 12) are not both zero, bit 1 when the cell carries a
    // wall segment (z_inv applies)
    dst[19] = (Real)(((o[11] != 0 || o[12] != 0) ? kExactSourceTerms : 0) | (wall ? kExactWallTerms : 0));
}

// ---------------------------------------------------------------------------
// Time step, reference operation order: the per-cell updates, the
// one-work-group-per-tube kernel and the cooperative single-tube kernel.
// ---------------------------------------------------------------------------
// One tube's view for the two phases of the exact step.
template <typename Real>
struct ExactView {
    Real* __restrict__ P;
    Real* __restrict__ Vx;
    Real* __restrict__ Vy;
    const Real* __restrict__ dP;
    const Real* __restrict__ dX;
    const Real* __restrict__ dY;
    const Real* __restrict__ zc;
    const uint8_t* __restrict__ xc;
    const Real* ctab;   // this tube's class coefficients (LDS)
    int W, Wi;
    Real K1;
    bool dirichlet, eigen;
};

// New pressure of one cell (FDTDEngineCPU.cpp:316-339) from its pressure
// p, velocities vx, vy, the left cell's vxl and the lower cell's vyb, its
// depths dp, dx, dy, the left cell's dxl and the lower cell's dyb, and its
// class coefficients o.
template <typename Real>
__device__ __forceinline__ Real exact_pressure_v(const ExactView<Real>& x, const Real* o, Real p, Real vx, Real vxl,
                                                 Real vy, Real vyb, Real dp, Real dx, Real dxl, Real dy, Real dyb) {
    const Real cxvx = vx * dx - vxl * dxl;
    const Real cyvy = vy * dy - vyb * dyb;
    const Real num = (p * dp) - (x.K1 * (cxvx + cyvy));
    const Real sp = o[8];   // sigma'dt of the cell's own type
    const Real den = x.eigen ? (sp * dp + dp) : ((1 + sp) * dp);
    Real v = num / den;
    if (x.dirichlet && o[17] != 0) v = 0;
    return v;
}

// New pressure of frame cell q.  The step's input state is read from sP,
// sVx, sVy at index s with row stride sW (the frame itself, or a tile of
// it); depths and classes at frame index q.
template <typename Real>
__device__ __forceinline__ Real exact_pressure_at(const ExactView<Real>& x, const Real* sP, const Real* sVx,
                                                  const Real* sVy, int s, int sW, int q) {
    const int W = x.W;
    return exact_pressure_v(x, x.ctab + x.xc[q] * kExactSlots, sP[s], sVx[s], sVx[s - 1], sVy[s], sVy[s + sW],
                            x.dP[q], x.dX[q], x.dX[q - 1], x.dY[q], x.dY[q + W]);
}

// New pressure of frame cell q from the state in x (P, Vx, Vy of the step's
// input).
template <typename Real>
__device__ __forceinline__ Real exact_pressure_cell(const ExactView<Real>& x, int q) {
    return exact_pressure_at(x, x.P, x.Vx, x.Vy, q, x.W, q);
}

// New velocities of one cell (FDTDEngineCPU.cpp:364-410) from its input
// velocities vxi, vyi, the new pressures pc (own), pr (right) and pt (top),
// the source value ev, its column's z_inv zcol and its class coefficients
// o; stored to vx_out, vy_out.  The source term (ev * o11 * o4) and the
// wall term (o4 * (...) * o9 * z) are added only where the class flags in
// o[19] say they can be non-zero: elsewhere each is a finite value times a
// zero factor, and adding it changes nothing but, at most, the sign of a
// zero velocity, which no later operation can turn into a different
// non-zero value (the step only divides by positive class constants).
template <typename Real>
__device__ __forceinline__ void exact_velocity_v(const Real* o, Real vxi, Real vyi, Real pc, Real pr, Real pt,
                                                 Real ev, Real zcol, Real* vx_out, Real* vy_out) {
    const int terms = (int)o[19];
    const Real cxp = pr - pc;
    const Real cyp = pt - pc;
    Real vx = (o[2] * vxi) - (o[6] * cxp);
    Real vy = (o[3] * vyi) - (o[7] * cyp);
    if (terms & kExactSourceTerms) {
        vx = vx + (ev * o[11] * o[4]);
        vy = vy + (ev * o[12] * o[5]);
    }
    if (terms & kExactWallTerms) {
        const Real z = zcol;
        Real bx = (o[14] * pc * o[0]) - (o[13] * pr * o[1]);
        Real by = (o[16] * pc * o[0]) - (o[15] * pt * o[1]);
        bx = bx * o[9] * z;
        by = by * o[10] * z;
        vx += (o[4] * bx);
        vy += (o[5] * by);
    }
    vx /= (o[2] + o[4]);
    vy /= (o[3] + o[5]);
    *vx_out = vx;
    *vy_out = vy;
}

// New velocities of frame cell q in column c from its input velocities vxi,
// vyi and the new pressures pc, pr, pt.
template <typename Real>
__device__ __forceinline__ void exact_velocity_at(const ExactView<Real>& x, int q, int c, Real vxi, Real vyi, Real pc,
                                                  Real pr, Real pt, Real ev, Real* vx_out, Real* vy_out) {
    const Real* o = x.ctab + x.xc[q] * kExactSlots;
    exact_velocity_v(o, vxi, vyi, pc, pr, pt, ev, (o[18] != 0) ? x.zc[c] : (Real)0, vx_out, vy_out);
}

// The same with the input velocities read from x at q.
template <typename Real>
__device__ __forceinline__ void exact_velocity_cell(const ExactView<Real>& x, int q, int c, Real pc, Real pr, Real pt,
                                                    Real ev, Real* vx_out, Real* vy_out) {
    exact_velocity_at(x, q, c, x.Vx[q], x.Vy[q], pc, pr, pt, ev, vx_out, vy_out);
}

// Pressure phase over the interior cells i = i0, i0 + B, ... < i1
// (row-major, frame cell (r, c) of cell i0 given), updating P in place: a
// cell's pressure update reads only its own P.
template <typename Real>
__device__ __forceinline__ void exact_pressure(const ExactView<Real>& x, int i0, int i1, int B, int r, int c,
                                               int dr, int dc) {
    const int W = x.W, Wi = x.Wi;
    for (int i = i0; i < i1; i += B) {
        const int q = r * W + c;
        x.P[q] = exact_pressure_cell(x, q);
        r += dr; c += dc;
        if (c > Wi) { c -= Wi; ++r; }
    }
}

// Velocity phase over the same cells, in place: a cell's velocity update
// reads only its own V.
template <typename Real>
__device__ __forceinline__ void exact_velocity(const ExactView<Real>& x, Real ev, int i0, int i1, int B, int r,
                                               int c, int dr, int dc) {
    const int W = x.W, Wi = x.Wi;
    for (int i = i0; i < i1; i += B) {
        const int q = r * W + c;
        exact_velocity_cell(x, q, c, x.P[q], x.P[q + 1], x.P[q - W], ev, &x.Vx[q], &x.Vy[q]);
        r += dr; c += dc;
        if (c > Wi) { c -= Wi; ++r; }
    }
}

// One work-group owns one tube for a block of steps; P and V are updated in
// place with a work-group barrier between the pressure and velocity phases
// (a cell's pressure update reads only its own P, a cell's velocity update
// only its own V, so in-place updates equal the reference's double buffer).
// Lane 0 runs the source (table lookup or the two-mass recurrence in double,
// fed by the previous step's P at the excitation cell) and records the probe.
template <typename Real>
__global__ __launch_bounds__(1024) void k_step_exact(const TubeDesc* __restrict__ tubes,
                                                     const int* __restrict__ order, Bufs<Real> b,
                                                     int t0_arg, int nsteps, const int* __restrict__ d_t0) {
    const int tube = order[blockIdx.x];
    // graph replays read the first step from device memory
    const int t0 = d_t0 ? *d_t0 : t0_arg;
    const TubeDesc& T = tubes[tube];
    __shared__ Real ctab[kExactClasses * kExactSlots];   // this tube's class coefficients
    __shared__ Real s_exeV;
    const int tid = threadIdx.x, B = blockDim.x;
    {
        const Real* src = b.xtab + (long long)tube * kExactClasses * kExactSlots;
        const int nv = b.xn[tube] * kExactSlots;
        for (int i = tid; i < nv; i += B) ctab[i] = src[i];
    }
    const int H = T.H, W = T.W, Wi = W - 2, NI = (H - 2) * Wi;
    const int tend = min(t0 + nsteps, T.num_steps);
    Real* __restrict__ P = b.P + T.cell_off;
    Real* __restrict__ Vx = b.Vx + T.cell_off;
    Real* __restrict__ Vy = b.Vy + T.cell_off;
    const Real* __restrict__ dP = b.dP + T.cell_off;
    const Real* __restrict__ dX = b.dX + T.cell_off;
    const Real* __restrict__ dY = b.dY + T.cell_off;
        const Real* __restrict__ zc = b.zcol + T.col_off;
    const Real* __restrict__ src = b.source + T.src_off;
    double* probe = b.probe + T.src_off;
    double* strace = b.strace + T.src_off;
    const bool twomass = T.source_type == 2;
    const bool dirichlet = T.dirichlet != 0;
    const bool eigen = T.variant == 1;
    __shared__ TwoMassState tm;   // lane 0 only; LDS keeps it out of every lane's VGPRs
    double pexc = 0;
    if (tid == 0) {
        pexc = b.pexc[tube];
        if (twomass) tm = b.tm[tube];
    }
    __syncthreads();
    const ExactView<Real> x{P, Vx, Vy, dP, dX, dY, zc, b.xcls + T.cell_off, ctab, W, Wi, (Real)T.ph.K1,
                            dirichlet, eigen};
    // per-thread walk over interior cells i = tid + k * B in row-major order
    const int dr = B / Wi, dc = B - dr * Wi;
    const int r0 = tid / Wi, c0 = tid - r0 * Wi;
    for (int t = t0; t < tend; ++t) {
        if (tid == 0) {
            double e;
            if (twomass) { tm.p1 = pexc; e = twomass_step(tm); }
            else e = (double)src[t];
            s_exeV = (Real)e;
            strace[t] = (double)(Real)e;
        }
        exact_pressure(x, tid, NI, B, r0 + 1, c0 + 1, dr, dc);
        __syncthreads();
        if (tid == 0) {
            probe[t] = (double)P[T.listener];
            pexc = (double)P[T.excitation];
        }
        exact_velocity(x, s_exeV, tid, NI, B, r0 + 1, c0 + 1, dr, dc);
        __syncthreads();
    }
    if (tid == 0) {
        b.pexc[tube] = pexc;
        if (twomass) b.tm[tube] = tm;
    }
}

// Barrier across the work-groups of one cooperative launch.  bar[0] counts
// arrivals, bar[1] is the generation; the last arrival resets the count and
// advances the generation.  Ordering comes from agent-scope fences, one of
// each kind per work-group and barrier (each costs an L2 write-back or
// invalidate): a release fence before arriving, an acquire fence after
// leaving, and the last arrival's acquire before it releases the rest, so
// every work-group's writes before the barrier happen before every read
// after it.  Polling uses relaxed loads.  The wait is bounded: past
// spin_limit polls (or once any work-group has timed out) it raises
// kStatusBarrierTimeout and returns, so every wave exits.
__device__ __forceinline__ void grid_barrier(unsigned* bar, unsigned groups, int spin_limit, unsigned* status) {
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned* count = bar;
        unsigned* gen = bar + 1;
        const unsigned my_gen = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned arrived = __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == groups - 1) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            for (int spins = 0; __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == my_gen;) {
                if (++spins > spin_limit ||
                    (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kStatusBarrierTimeout)) {
                    atomicOr(status, kStatusBarrierTimeout);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
    }
    __syncthreads();
}

// The exact step of one tube spread over gridDim.x work-groups of one
// cooperative launch, s steps per grid barrier.  Work-group g owns the
// frame columns [a, e) of every interior row (coop_columns).  Each period of
// s steps
//  1. copies rows 0..H-1 of the columns [a - s, e + s) of the step's input
//     state into an LDS tile;
//  2. advances the tile s steps in place: sub-step k updates pressures on
//     the columns [a - (s-1-k), e + (s-k)) and then velocities on
//     [a - (s-1-k), e + (s-1-k)).  A pressure reads velocities of its own
//     and the left column, a velocity pressures of its own and the right
//     column, so each region holds every input of the next one; the
//     columns outside [a, e) recompute cells that neighbouring work-groups
//     own, with the same code and inputs, hence the same values;
//  3. stores its own columns into the other state buffer (double buffered
//     as the reference's PV_N / PV_Nplus1) and meets the other work-groups
//     at one grid barrier.
// The per-cell code is that of k_step_exact, so results are bit-identical.
// Only the excitation column and the one left of it carry the source
// (coefficients 11 and 12).  The work-groups whose velocity regions reach
// them (coop_runs_source) run the source on lane 0, the two-mass recurrence
// fed by the excitation cell's new pressure from their own tile (identical
// inputs, identical values); the others' cells have zero source
// coefficients.  The work-group owning the listener column records the
// probe, the one owning the excitation column the source trace and the
// final source state.
// ONE: the first pressure region holds at most one cell per lane of waves
// 1 and up, so each of those lanes keeps one cell for the whole launch,
// with its depths, class and z_inv in registers and its own P, Vx, Vy in
// registers within a period; only the neighbours' values come from the
// tile.  Wave 0 is then the control wave: it runs the two-mass recurrence
// one step ahead, overlapped with the cell waves -- the fold positions of
// step n + 1 during the pressure phase of step n, the flow coefficients and
// (once the new excitation pressure is in the tile) the flow of step n + 1
// during the velocity phase of step n, into the other slot of a
// double-buffered source value.
template <typename Real, bool ONE>
__global__ __launch_bounds__(1024) void k_step_exact_coop(const TubeDesc* __restrict__ tubes,
                                                          const int* __restrict__ order, Bufs<Real> b, int t0,
                                                          int nsteps, int s, unsigned* bar, Real* P2, Real* Vx2,
                                                          Real* Vy2, int spin_limit) {
    extern __shared__ __align__(16) unsigned char coop_tile[];
    const int tube = order[0];
    const TubeDesc& T = tubes[tube];
    __shared__ Real ctab[kExactClasses * kExactSlots];
    __shared__ TwoMassState tm;
    __shared__ TwoMassNext nx;   // ONE: the next step's recurrence before its pressure input
    __shared__ Real s_ev[2];     // source value; ONE: of steps with even / odd k
    // ONE: the period's table source values, and its probe and source-trace
    // samples, kept in LDS and moved once per period: a work-group barrier
    // waits for a wave's outstanding global loads and stores, so none may
    // be issued inside the steps
    __shared__ Real s_srcv[kCoopMaxSteps];
    __shared__ double s_rec[2][kCoopMaxSteps];
    const int tid = threadIdx.x, B = blockDim.x, g = blockIdx.x;
    const int G = (int)gridDim.x;
    {
        const Real* src = b.xtab + (long long)tube * kExactClasses * kExactSlots;
        const int nv = b.xn[tube] * kExactSlots;
        for (int i = tid; i < nv; i += B) ctab[i] = src[i];
    }
    const int H = T.H, W = T.W, Hi = H - 2;
    const int tend = min(t0 + nsteps, T.num_steps);
    int a, e;
    coop_columns(W, G, g, &a, &e);
    const int L = max(0, a - s), R = min(W, e + s), TW = max(0, R - L);
    Real* tP = reinterpret_cast<Real*>(coop_tile);
    Real* tVx = tP + H * TW;
    Real* tVy = tVx + H * TW;
    const long long co = T.cell_off;
    const ExactView<Real> x{b.P + co, b.Vx + co, b.Vy + co, b.dP + co, b.dX + co, b.dY + co, b.zcol + T.col_off,
                            b.xcls + co, ctab, W, W - 2, (Real)T.ph.K1, T.dirichlet != 0, T.variant == 1};
    Real *cP = b.P + co, *cVx = b.Vx + co, *cVy = b.Vy + co;
    Real *nP = P2 + co, *nVx = Vx2 + co, *nVy = Vy2 + co;
    const Real* __restrict__ src = b.source + T.src_off;
    double* probe = b.probe + T.src_off;
    double* strace = b.strace + T.src_off;
    const bool twomass = T.source_type == 2;
    const int ex = T.excitation % W, ey = T.excitation / W;
    const int lx = T.listener % W, ly = T.listener / W;
    const bool source = coop_runs_source(W, G, s, g, ex);
    const bool keeps_source = g == coop_owner(W, G, ex);
    const bool keeps_probe = g == coop_owner(W, G, lx);
    double pexc = 0;   // lane 0 of a source work-group: the source's pressure input
    if (tid == 0 && source) {
        pexc = b.pexc[tube];
        if (twomass) tm = b.tm[tube];
    }
    // ONE: this lane's cell (r1, c1) of the first pressure region, if any
    const int pl0 = max(1, a - (s - 1)), w0 = max(0, min(W - 1, e + s) - pl0);
    const int cell = tid - 64;
    const bool mine = ONE && cell >= 0 && cell < Hi * w0;
    const int r1 = mine ? 1 + cell / w0 : 1, c1 = mine ? pl0 + (cell - (r1 - 1) * w0) : 1;
    const int s1 = r1 * TW + (c1 - L), q1 = r1 * W + c1;
    Real dp1 = 0, dx1 = 0, dxl1 = 0, dy1 = 0, dyb1 = 0, z1 = 0;
    const Real* o1 = ctab;
    if (mine) {
        dp1 = x.dP[q1]; dx1 = x.dX[q1]; dxl1 = x.dX[q1 - 1]; dy1 = x.dY[q1]; dyb1 = x.dY[q1 + W];
        o1 = ctab + x.xc[q1] * kExactSlots;
        z1 = x.zc[c1];
    }
    for (int t = t0; t < tend; t += s) {
        const int sp = min(s, tend - t);
        for (int i = tid; i < H * TW; i += B) {
            const int r = i / TW, c = i - r * TW;
            const int q = r * W + L + c;
            tP[i] = cP[q];
            tVx[i] = cVx[q];
            tVy[i] = cVy[q];
        }
        if (tid == 0 && source && t > t0) pexc = (double)cP[T.excitation];
        if (ONE && source && !twomass && tid < sp) s_srcv[tid] = src[t + tid];
        __syncthreads();
        Real p1 = 0, vx1 = 0, vy1 = 0;   // ONE: this lane's cell
        if (mine) { p1 = tP[s1]; vx1 = tVx[s1]; vy1 = tVy[s1]; }
        for (int k = 0; k < sp; ++k) {
            const int tt = t + k;
            if (tid == 0 && source) {
                if constexpr (ONE) {
                    if (k == 0) {   // later steps of the period were finished one phase ahead
                        double ev;
                        if (twomass) {
                            if (t == t0) { twomass_solve(tm, nx); twomass_flow(tm, nx); }
                            ev = twomass_finish(tm, nx, pexc);
                        } else {
                            ev = (double)s_srcv[0];
                        }
                        s_ev[0] = (Real)ev;
                        s_rec[1][0] = (double)(Real)ev;
                    }
                    if (twomass && tt + 1 < tend) twomass_solve(tm, nx);
                } else {
                    double ev;
                    if (twomass) { tm.p1 = pexc; ev = twomass_step(tm); }
                    else ev = (double)src[tt];
                    s_ev[0] = (Real)ev;
                    if (keeps_source) strace[tt] = (double)(Real)ev;
                }
            }
            const int pl = max(1, a - (sp - 1 - k)), pr = min(W - 1, e + (sp - k));
            if constexpr (ONE) {
                if (mine && c1 >= pl && c1 < pr) {
                    p1 = exact_pressure_v(x, o1, p1, vx1, tVx[s1 - 1], vy1, tVy[s1 + TW], dp1, dx1, dxl1, dy1, dyb1);
                    tP[s1] = p1;
                }
            } else {
                const int pw = pr - pl;
                for (int i = tid; i < Hi * pw; i += B) {
                    const int rr = i / pw, r = 1 + rr, c = pl + (i - rr * pw);
                    const int si = r * TW + (c - L);
                    tP[si] = exact_pressure_at(x, tP, tVx, tVy, si, TW, r * W + c);
                }
            }
            __syncthreads();
            if (tid == 0) {
                if constexpr (ONE) {
                    if (keeps_probe) s_rec[0][k] = (double)tP[ly * TW + (lx - L)];
                    if (source && tt + 1 < tend) {
                        if (twomass) twomass_flow(tm, nx);
                        if (k + 1 < sp) {
                            const double pin = (double)tP[ey * TW + (ex - L)];
                            const double ev = twomass ? twomass_finish(tm, nx, pin) : (double)s_srcv[k + 1];
                            s_ev[(k + 1) & 1] = (Real)ev;
                            s_rec[1][k + 1] = (double)(Real)ev;
                        }
                    }
                } else {
                    if (keeps_probe) probe[tt] = (double)tP[ly * TW + (lx - L)];
                    if (source && k + 1 < sp) pexc = (double)tP[ey * TW + (ex - L)];
                }
            }
            {
                const Real ev = source ? s_ev[ONE ? (k & 1) : 0] : (Real)0;
                const int vl = pl, vr = min(W - 1, e + (sp - 1 - k));
                if constexpr (ONE) {
                    if (mine && c1 >= vl && c1 < vr) {
                        exact_velocity_v(o1, vx1, vy1, p1, tP[s1 + 1], tP[s1 - TW], ev, z1, &vx1, &vy1);
                        tVx[s1] = vx1;
                        tVy[s1] = vy1;
                    }
                } else {
                    const int vw = vr - vl;
                    for (int i = tid; i < Hi * vw; i += B) {
                        const int rr = i / vw, r = 1 + rr, c = vl + (i - rr * vw);
                        const int si = r * TW + (c - L);
                        exact_velocity_at(x, r * W + c, c, tVx[si], tVy[si], tP[si], tP[si + 1], tP[si - TW], ev,
                                          &tVx[si], &tVy[si]);
                    }
                }
            }
            __syncthreads();
        }
        if constexpr (ONE) {
            if (mine && c1 >= a && c1 < e) {
                nP[q1] = p1;
                nVx[q1] = vx1;
                nVy[q1] = vy1;
            }
            if (tid < sp) {
                if (keeps_probe) probe[t + tid] = s_rec[0][tid];
                if (keeps_source) strace[t + tid] = s_rec[1][tid];
            }
        } else {
            const int ow = max(0, e - a);
            for (int i = tid; i < Hi * ow; i += B) {
                const int rr = i / ow, r = 1 + rr, c = a + (i - rr * ow);
                const int si = r * TW + (c - L), q = r * W + c;
                nP[q] = tP[si];
                nVx[q] = tVx[si];
                nVy[q] = tVy[si];
            }
        }
        grid_barrier(bar, (unsigned)G, spin_limit, b.status);
        Real* tp = cP; cP = nP; nP = tp;
        tp = cVx; cVx = nVx; nVx = tp;
        tp = cVy; cVy = nVy; nVy = tp;
    }
    // the final state is in cP; bring it home if that is the second buffer
    if (cP != b.P + co) {
        const int ow = max(0, e - a);
        Real *hP = b.P + co, *hVx = b.Vx + co, *hVy = b.Vy + co;
        for (int i = tid; i < Hi * ow; i += B) {
            const int rr = i / ow, r = 1 + rr, c = a + (i - rr * ow);
            const int q = r * W + c;
            hP[q] = cP[q];
            hVx[q] = cVx[q];
            hVy[q] = cVy[q];
        }
    }
    if (tid == 0 && keeps_source) {
        if (tend > t0) b.pexc[tube] = (double)cP[T.excitation];
        if (twomass) b.tm[tube] = tm;
    }
}

}  // namespace

template <typename Real>
size_t coop_tile_bytes(int H, int W, int groups, int s) {
    return 3 * sizeof(Real) * (size_t)H * (size_t)std::min(W, coop_chunk(W, groups) + 2 * s);
}

template <typename Real>
hipError_t launch_step_coop(const TubeDesc* d_tubes, const int* d_order, Bufs<Real> b, int t0, int nsteps,
                            int threads, int groups, int s, bool one, size_t tile_bytes, unsigned* bar, Real* P2,
                            Real* Vx2, Real* Vy2, int spin_limit, hipStream_t st) {
    static std::atomic<int> configured[2] = {0, 0};   // dynamic LDS limit set, per instantiation
    if (!configured[sizeof(Real) == 8].load()) {
        for (const void* f : {(const void*)k_step_exact_coop<Real, true>, (const void*)k_step_exact_coop<Real, false>}) {
            hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kCoopTileMax);
            if (e != hipSuccess) return e;
        }
        configured[sizeof(Real) == 8].store(1);
    }
    if (tile_bytes > kCoopTileMax || s < 1) return hipErrorInvalidValue;
    const void* fn = one ? (const void*)k_step_exact_coop<Real, true> : (const void*)k_step_exact_coop<Real, false>;
    void* args[] = {(void*)&d_tubes, (void*)&d_order, (void*)&b, (void*)&t0, (void*)&nsteps, (void*)&s,
                    (void*)&bar, (void*)&P2, (void*)&Vx2, (void*)&Vy2, (void*)&spin_limit};
    hipError_t e = hipMemsetAsync(bar, 0, 2 * sizeof(unsigned), st);
    if (e != hipSuccess) return e;
    return hipLaunchCooperativeKernel(fn, dim3(groups), dim3(threads), args, tile_bytes, st);
}

template <typename Real>
hipError_t launch_setup(const TubeDesc* d_tubes, const TubeDesc* h_tubes, int n, Bufs<Real> b,
                        Real* tmpA, Real* tmpB, Real* tmpC, Real* tmpD, hipStream_t st) {
    int maxL = 0, maxHW = 0;
    for (int k = 0; k < n; ++k) {
        maxL = max(maxL, h_tubes[k].L);
        maxHW = max(maxHW, h_tubes[k].H * h_tubes[k].W);
    }
    const long long stride = maxHW;   // raw / smoothed depth scratch: one slice per tube
    dim3 gw((maxL + 63) / 64, n), gc((maxHW + kSetupThreads - 1) / kSetupThreads, n);
    if (maxL > 0) hipLaunchKernelGGL(k_walls<Real>, gw, dim3(64), 0, st, d_tubes, b);
    hipLaunchKernelGGL(k_classify<Real>, gc, dim3(kSetupThreads), 0, st, d_tubes, b);
    hipLaunchKernelGGL(k_depth_raw<Real>, gc, dim3(kSetupThreads), 0, st, d_tubes, b, tmpA, tmpB, stride);
    hipLaunchKernelGGL(k_depth_smooth<Real>, gc, dim3(kSetupThreads), 0, st, d_tubes, tmpA, tmpB, tmpC, tmpD, stride);
    hipLaunchKernelGGL(k_depth_final<Real>, gc, dim3(kSetupThreads), 0, st, d_tubes, b, tmpC, tmpD, stride);
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_exact_classes(const TubeDesc* d_tubes, int n, Bufs<Real> b, hipStream_t st) {
    hipLaunchKernelGGL(k_exact_classes<Real>, dim3(n), dim3(kExactClasses), 0, st, d_tubes, b);
    return hipGetLastError();
}

// Raw (unsmoothed) depth maps only, one stride-long slice per tube.
template <typename Real>
hipError_t launch_depth_raw(const TubeDesc* d_tubes, const TubeDesc* h_tubes, int n, Bufs<Real> b, Real* X,
                            Real* Y, long long stride, hipStream_t st) {
    int maxHW = 0;
    for (int k = 0; k < n; ++k) maxHW = max(maxHW, h_tubes[k].H * h_tubes[k].W);
    dim3 gc((maxHW + kSetupThreads - 1) / kSetupThreads, n);
    hipLaunchKernelGGL(k_depth_raw<Real>, gc, dim3(kSetupThreads), 0, st, d_tubes, b, X, Y, stride);
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_step(const TubeDesc* d_tubes, const int* d_order, int n, Bufs<Real> b,
                       int t0, int nsteps, int threads, int kernel, const int* d_t0, hipStream_t st) {
    (void)kernel;
    hipLaunchKernelGGL(k_step_exact<Real>, dim3(n), dim3(threads), 0, st, d_tubes, d_order, b, t0, nsteps, d_t0);
    return hipGetLastError();
}

namespace {
// Advances the device-resident first step of the next graph replay.
__global__ void k_bump(int* d_t0, int s) {
    if (threadIdx.x == 0) *d_t0 += s;
}
}  // namespace

hipError_t launch_bump(int* d_t0, int s, hipStream_t st) {
    hipLaunchKernelGGL(k_bump, dim3(1), dim3(64), 0, st, d_t0, s);
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_coeffs(const TubeDesc* d_tubes, const TubeDesc* h_tube, int tube, Bufs<Real> b,
                         Real* out17, hipStream_t st) {
    const int n = (h_tube->H - 2) * (h_tube->W - 2);
    hipLaunchKernelGGL(k_coeffs<Real>, dim3((n + 255) / 256), dim3(256), 0, st, d_tubes, tube, b, out17);
    return hipGetLastError();
}

#define VS_INSTANTIATE(Real)                                                                        \
    template hipError_t launch_setup<Real>(const TubeDesc*, const TubeDesc*, int, Bufs<Real>, Real*, \
                                           Real*, Real*, Real*, hipStream_t);                       \
    template hipError_t launch_step<Real>(const TubeDesc*, const int*, int, Bufs<Real>, int, int,    \
                                          int, int, const int*, hipStream_t);                       \
    template hipError_t launch_step_coop<Real>(const TubeDesc*, const int*, Bufs<Real>, int, int, int, \
                                               int, int, bool, size_t, unsigned*, Real*, Real*, Real*, int, \
                                               hipStream_t);                                       \
    template size_t coop_tile_bytes<Real>(int, int, int, int);                                      \
    template hipError_t launch_coeffs<Real>(const TubeDesc*, const TubeDesc*, int, Bufs<Real>,       \
                                            Real*, hipStream_t);                                    \
    template hipError_t launch_depth_raw<Real>(const TubeDesc*, const TubeDesc*, int, Bufs<Real>,    \
                                               Real*, Real*, long long, hipStream_t);                 \
    template hipError_t launch_exact_classes<Real>(const TubeDesc*, int, Bufs<Real>, hipStream_t);

VS_INSTANTIATE(float)
VS_INSTANTIATE(double)

}  // namespace vs
