// vs_adjoint.hip -- reverse-time (adjoint) FDTD for gradients of a probe loss.
//
// Forward model differentiated (flux form, U = V * depth; algebraically the
// reference step FDTDEngineCPU.cpp:316-405):
//   P'_i  = aP_i P_i - kP_i (Ux_i - Ux_{i-1} + Uy_i - Uy_{i+W})
//   Ux'_i = ax_i Ux_i + cx_i (P'_i - P'_{i+1})     fx_i = 0 (ordinary face)
//         = -cx_i P'_{i+1}                        fx_i = 1 (wall with admittance)
//         =  cx_i e_n                             fx_i = 2 (excitation: source)
//   Uy'_i = ay_i Uy_i + cy_i (P'_i - P'_{i-W})     fy_i = 0
//         = -cy_i P'_{i-W}                        fy_i = 1
//   y_n   = P'_L (probe)
// Reverse step n, with LP, LX, LY the adjoints of step n+1's inputs:
//   phase 1: M_i = LP_i + [i = L] g_n + [fx_i = 0] cx_i LX_i - [fx_{i-1} < 2] cx_{i-1} LX_{i-1}
//                 + [fy_i = 0] cy_i LY_i - [fy_{i+W} < 2] cy_{i+W} LY_{i+W}
//            Gx_i += LX_i dUx'_i/dcx_i,  Gy_i += LY_i dUy'_i/dcy_i      (P' = P^{n+1})
//   phase 2: LX_i = [fx_i = 0] ax_i LX_i - kP_i M_i + kP_{i+1} M_{i+1}
//            LY_i = [fy_i = 0] ay_i LY_i - kP_i M_i + kP_{i-W} M_{i-W}
//            LP_i = aP_i M_i,  GkP_i += -M_i div_i(U^n)
// The forward states of a segment are recomputed from checkpoints taken every
// S steps and held in HBM while the segment is reversed.

#include "vs_common.h"

namespace vs {
namespace {

// Per-cell coefficients of the flux form in Real (evaluated in double).
template <typename Real>
__global__ void k_adj_coeffs(const TubeDesc* __restrict__ tubes, Bufs<Real> b, AdjBufs<Real> a) {
    const TubeDesc& T = tubes[blockIdx.y];
    const int H = T.H, W = T.W, Wi = W - 2;
    const int n = (H - 2) * Wi;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = 1 + i / Wi, c = 1 + i % Wi;
    const long long q = T.cell_off + (long long)r * W + c;
    const uint8_t* cls = b.cls + T.cell_off;
    const int qq = r * W + c;
    const int tc = cls[qq] & kTypeMask, tr = cls[qq + 1] & kTypeMask, tt = cls[qq - W] & kTypeMask;
    const bool segwall = (cls[qq] & kSegWallBit) != 0;
    const double dt = T.ph.dt, C = dt / T.ph.rds, K1 = T.ph.K1;
    const double bc = T.ph.beta[tc], br = T.ph.beta[tr], bt = T.ph.beta[tt];
    const double sc = T.ph.sigma[tc], sr = T.ph.sigma[tr], st = T.ph.sigma[tt];
    const double dP = b.dP[q], dX = b.dX[q], dY = b.dY[q];
    const bool nop = T.dirichlet && tc == kNoPressure;
    const double m = nop ? 0.0 : 1.0;
    a.aP[q] = (Real)(m / (1.0 + sc));
    a.kP[q] = (Real)(m * K1 / ((1.0 + sc) * dP));
    const double mbx = br < bc ? br : bc, mby = bt < bc ? bt : bc;
    const double Mx = sc < sr ? sr : sc, My = sc < st ? st : sc;
    uint8_t fx = 0, fy = 0;
    double cx = mbx == 0.0 ? 0.0 : dX * (mbx * mbx * C) / (mbx + Mx);
    double cy = mby == 0.0 ? 0.0 : dY * (mby * mby * C) / (mby + My);
    double ax = mbx == 0.0 ? 0.0 : mbx / (mbx + Mx);
    double ay = mby == 0.0 ? 0.0 : mby / (mby + My);
    if (tc == kExcitation) {
        fx = 2; cx = dX; ax = 0;
        fy = 1; cy = 0; ay = 0;
    } else if (segwall) {
        const int ibr = (int)br, ibt = (int)bt;
        const int lut = (ibr == kAir && ibt == kAir) ? 0 : (ibr == kAir && ibt == kWall) ? 1
                      : (ibr == kWall && ibt == kAir) ? 2 : 3;
        const double nin = T.ph.wall_normal[lut] * (1.0 - bc);
        const double z = b.zcol[T.col_off + c];
        const double nex = tr == kExcitation ? 0.0 : 1.0, ney = tt == kExcitation ? 0.0 : 1.0;
        fx = 1; cx = dX * z * nex * br * nin; ax = 0;
        fy = 1; cy = dY * z * ney * bt * nin; ay = 0;
    }
    a.ax[q] = (Real)ax; a.ay[q] = (Real)ay;
    a.cx[q] = (Real)cx; a.cy[q] = (Real)cy;
    a.fx[q] = fx; a.fy[q] = fy;
}

// Interior walk of one work-group over its tube: i = tid + k*B in row-major order.
struct Walk {
    int r, c, dr, dc, Wi;
    __device__ Walk(int tid, int B, int Wi_) : Wi(Wi_) {
        dr = B / Wi; dc = B - dr * Wi;
        const int r0 = tid / Wi;
        r = r0 + 1; c = tid - r0 * Wi + 1;
    }
    __device__ void next() { r += dr; c += dc; if (c > Wi) { c -= Wi; ++r; } }
};

// Forward flux steps [t0, t0+ns) of one tube per work-group.  mode 1 stores the
// state at every multiple of S into the checkpoint slab; mode 2 stores every
// state of the segment (slot t - t0, including the end state) into the segment
// slab.  load >= 0 first loads checkpoint `load` into the state.
template <typename Real>
__global__ __launch_bounds__(1024) void k_adj_fwd(const TubeDesc* __restrict__ tubes, const int* __restrict__ order,
                                                  Bufs<Real> b, AdjBufs<Real> a, int t0, int ns, int mode, int load) {
    const int tube = order[blockIdx.x];
    const TubeDesc& T = tubes[tube];
    if (t0 >= T.num_steps) return;   // whole work-group: this tube has no such steps (nor checkpoint)
    const int tid = threadIdx.x, B = blockDim.x;
    const int H = T.H, W = T.W, Wi = W - 2, NI = (H - 2) * Wi, HW = H * W;
    const long long co = T.cell_off;
    Real* P = a.P + co;
    Real* Ux = a.Ux + co;
    Real* Uy = a.Uy + co;
    const Real *aP = a.aP + co, *kP = a.kP + co, *ax = a.ax + co, *ay = a.ay + co, *cx = a.cx + co, *cy = a.cy + co;
    const uint8_t *fx = a.fx + co, *fy = a.fy + co;
    Real* ck = a.ck + a.ck_off[tube];
    Real* sg = a.seg + a.seg_off[tube];
    const int S = a.S;
    if (load >= 0) {
        const Real* src = ck + (long long)load * 3 * HW;
        for (int k = tid; k < HW; k += B) { P[k] = src[k]; Ux[k] = src[HW + k]; Uy[k] = src[2 * HW + k]; }
        __syncthreads();
    }
    const int tend = min(t0 + ns, T.num_steps);
    auto store = [&](Real* dst) {
        for (int k = tid; k < HW; k += B) { dst[k] = P[k]; dst[HW + k] = Ux[k]; dst[2 * HW + k] = Uy[k]; }
    };
    for (int t = t0; t < tend; ++t) {
        if (mode == 1 && t % S == 0) store(ck + (long long)(t / S) * 3 * HW);
        if (mode == 2) store(sg + (long long)(t - t0) * 3 * HW);
        __syncthreads();
        const Real e = b.source[T.src_off + t];
        {
            Walk w(tid, B, Wi);
            for (int i = tid; i < NI; i += B, w.next()) {
                const int q = w.r * W + w.c;
                P[q] = aP[q] * P[q] - kP[q] * ((Ux[q] - Ux[q - 1]) + (Uy[q] - Uy[q + W]));
            }
        }
        __syncthreads();
        if (tid == 0) a.probe[T.src_off + t] = (double)P[T.listener];
        {
            Walk w(tid, B, Wi);
            for (int i = tid; i < NI; i += B, w.next()) {
                const int q = w.r * W + w.c;
                const Real pc = P[q], pr = P[q + 1], pt = P[q - W];
                Real ux, uy;
                if (fx[q] == 0) ux = ax[q] * Ux[q] + cx[q] * (pc - pr);
                else if (fx[q] == 1) ux = -cx[q] * pr;
                else ux = cx[q] * e;
                if (fy[q] == 0) uy = ay[q] * Uy[q] + cy[q] * (pc - pt);
                else uy = -cy[q] * pt;
                Ux[q] = ux;
                Uy[q] = uy;
            }
        }
        __syncthreads();
    }
    if (mode == 2 && tend > t0) {
        store(sg + (long long)(tend - t0) * 3 * HW);
        __syncthreads();
    }
}

// Reverse steps n = s0 + ns - 1 ... s0 of one tube per work-group.
template <typename Real>
__global__ __launch_bounds__(1024) void k_adj_rev(const TubeDesc* __restrict__ tubes, const int* __restrict__ order,
                                                  Bufs<Real> b, AdjBufs<Real> a, int s0, int ns) {
    const int tube = order[blockIdx.x];
    const TubeDesc& T = tubes[tube];
    const int tid = threadIdx.x, B = blockDim.x;
    const int H = T.H, W = T.W, Wi = W - 2, NI = (H - 2) * Wi, HW = H * W;
    const long long co = T.cell_off;
    const Real *aP = a.aP + co, *kP = a.kP + co, *ax = a.ax + co, *ay = a.ay + co, *cx = a.cx + co, *cy = a.cy + co;
    const uint8_t *fx = a.fx + co, *fy = a.fy + co;
    Real* LP = a.LP + co;
    Real* LX = a.LX + co;
    Real* LY = a.LY + co;
    Real* M = a.M + co;
    double* GkP = a.GkP + co;
    double* Gx = a.Gx + co;
    double* Gy = a.Gy + co;
    const Real* sg = a.seg + a.seg_off[tube];
    const double* gp = a.gprobe + T.src_off;
    for (int n = s0 + ns - 1; n >= s0; --n) {
        if (n >= T.num_steps) continue;
        const Real* S1 = sg + (long long)(n + 1 - s0) * 3 * HW;   // state n+1
        const Real* S0 = sg + (long long)(n - s0) * 3 * HW;       // state n
        const Real* P1 = S1;
        const Real e = b.source[T.src_off + n];
        {
            Walk w(tid, B, Wi);
            for (int i = tid; i < NI; i += B, w.next()) {
                const int q = w.r * W + w.c;
                Real m = LP[q];
                if (q == T.listener) m += (Real)gp[n];
                const Real lx = LX[q], ly = LY[q];
                if (fx[q] == 0) m += cx[q] * lx;
                if (fx[q - 1] < 2) m -= cx[q - 1] * LX[q - 1];
                if (fy[q] == 0) m += cy[q] * ly;
                if (fy[q + W] < 2) m -= cy[q + W] * LY[q + W];
                M[q] = m;
                const Real pc = P1[q], pr = P1[q + 1], pt = P1[q - W];
                const double dgx = fx[q] == 0 ? (double)(pc - pr) : fx[q] == 1 ? -(double)pr : (double)e;
                const double dgy = fy[q] == 0 ? (double)(pc - pt) : -(double)pt;
                Gx[q] += (double)lx * dgx;
                Gy[q] += (double)ly * dgy;
            }
        }
        __syncthreads();
        {
            Walk w(tid, B, Wi);
            const Real* Ux0 = S0 + HW;
            const Real* Uy0 = S0 + 2 * HW;
            for (int i = tid; i < NI; i += B, w.next()) {
                const int q = w.r * W + w.c;
                const Real m = M[q];
                const Real kpm = kP[q] * m;
                LX[q] = (fx[q] == 0 ? ax[q] * LX[q] : (Real)0) - kpm + kP[q + 1] * M[q + 1];
                LY[q] = (fy[q] == 0 ? ay[q] * LY[q] : (Real)0) - kpm + kP[q - W] * M[q - W];
                LP[q] = aP[q] * m;
                const double dv = (double)((Ux0[q] - Ux0[q - 1]) + (Uy0[q] - Uy0[q + W]));
                GkP[q] -= (double)m * dv;
            }
        }
        __syncthreads();
    }
}

}  // namespace

template <typename Real>
hipError_t launch_adj_coeffs(const TubeDesc* d_tubes, const TubeDesc* h_tubes, int n, Bufs<Real> b, AdjBufs<Real> a,
                             hipStream_t st) {
    int maxHW = 0;
    for (int k = 0; k < n; ++k) maxHW = max(maxHW, h_tubes[k].H * h_tubes[k].W);
    hipLaunchKernelGGL(k_adj_coeffs<Real>, dim3((maxHW + 255) / 256, n), dim3(256), 0, st, d_tubes, b, a);
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_adj_fwd(const TubeDesc* d_tubes, const int* d_order, int n, Bufs<Real> b, AdjBufs<Real> a, int t0,
                          int ns, int mode, int load, int threads, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_fwd<Real>, dim3(n), dim3(threads), 0, st, d_tubes, d_order, b, a, t0, ns, mode, load);
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_adj_rev(const TubeDesc* d_tubes, const int* d_order, int n, Bufs<Real> b, AdjBufs<Real> a, int s0,
                          int ns, int threads, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_rev<Real>, dim3(n), dim3(threads), 0, st, d_tubes, d_order, b, a, s0, ns);
    return hipGetLastError();
}

#define VS_ADJ_INSTANTIATE(Real)                                                                              \
    template hipError_t launch_adj_coeffs<Real>(const TubeDesc*, const TubeDesc*, int, Bufs<Real>, AdjBufs<Real>, \
                                                hipStream_t);                                                 \
    template hipError_t launch_adj_fwd<Real>(const TubeDesc*, const int*, int, Bufs<Real>, AdjBufs<Real>, int,   \
                                             int, int, int, int, hipStream_t);                                \
    template hipError_t launch_adj_rev<Real>(const TubeDesc*, const int*, int, Bufs<Real>, AdjBufs<Real>, int,   \
                                             int, int, hipStream_t);

VS_ADJ_INSTANTIATE(float)
VS_ADJ_INSTANTIATE(double)

}  // namespace vs
