// vs_fast.hip -- FP32 throughput path: one work-group per tube with the
// tube's state held on chip for a whole block of steps.
//
// Layout inside a work-group of T cell threads (plus one control wave).
// The tube interior is indexed with one zero pad column per row,
// j = r * Wp + c (Wp = Wi + 1), and cut into runs of K consecutive cells;
// thread t owns cells j = t*K + k, k < K.
//   LDS   three flat regions, each with G zero guard lanes on both sides:
//         P[j] and Uy[j] (cell j at float j: lane t's run at t*K, so with K
//         odd the 64 lanes of a wave hit distinct banks), and the Ux of each
//         run's last cell (lane t at float t).  The cells below and above,
//         j +- Wp, are then a uniform offset from the lane's run, so their
//         reads pair like the lane's own.  The guards mean every neighbour
//         read of a real cell lands on a real cell, a pad cell or a zero;
//         static tables by cell class: aP, and (ax, ay) with the two
//         face-form flags in their sign bits.
//   VGPRs Ux[K], kP[K], gX[K], gY[K] and the class bytes (4 per VGPR).
//
// Update in flux form (U = V * depth), algebraically equal to the reference
// step (FDTDEngineCPU.cpp:316-405) with every per-cell coefficient
// precomputed at setup:
//   P'  = aP * P - kP * (Ux - Ux_left + Uy - Uy_below)
//   Ux' = ax * Ux + gX * (P' - P'_right)
//   Uy' = ay * Uy + gY * (P' - P'_top)
// aP = m/(1+sigma_c), kP = m*K1/((1+sigma_c)*dP) (m = 0 on Dirichlet cells);
// ax = mbx/(mbx+Mx), gX = dX*Cx/(mbx+Mx) (Cx = mbx^2 dt/(rho ds)), same in y.
// Cells whose face terms differ from that form (walls carrying an admittance,
// whose face velocity is -z*Nin*P_neighbour, and the excitation column, whose
// x-face carries the source velocity) keep it with their own gX, gY and two
// flags of their class: s = 0 drops the cell's own pressure from its faces,
// r = 0 puts the source in place of the right neighbour's pressure.
//
// The same file holds the on-chip reverse pass of this step (the recording
// forward launch, k_fast_reverse and k_fast_adj_sums).

#include "vs_common.h"

#include <atomic>

namespace vs {

namespace {

// Per-cell code written at setup (16 bits): bits [3, 9) hold code_x +
// 8 code_y, bits [9, 12) the pressure damping code; kCodeSelfOff drops the
// cell's own pressure from its face updates (walls carrying an admittance:
// U = -g * P_neighbour); kCodeSrcX makes the x-face carry the source
// (excitation column: Ux = g * exeV).  The host maps the distinct codes of a
// batch to class bytes.
constexpr int kShiftXY = 3;
constexpr int kShiftP = 9;
constexpr uint32_t kCodeSelfOff = 1u << 12;
constexpr uint32_t kCodeSrcX = 1u << 13;
// The step kernel reads a class byte per cell instead: a batch-local class
// (< kFastClasses) of the damping triple and the two face-form bits, indexing
// the tables aP and (ax, ay) (class 0: every factor zero, used for pad cells).
// The sign bit of the ax entry marks s = 0 (the cell's own pressure is left
// out of its face updates), that of ay marks r = 0 (the x-face carries the
// source); the magnitudes are the damping factors.

// ---------------------------------------------------------------------------
// Setup of the per-cell coefficients of the flux form (evaluated in double
// from the FP32 depth maps, then rounded once to float).
// damping code: 0 -> sigma 0, 1 -> zero (Dirichlet cell / blocked face),
//               2 -> sigma dt, 2 + k -> PML ring sigma k/5 * 0.5 (k = 1..5).
// ---------------------------------------------------------------------------
__device__ inline int sigma_code(int type) {
    switch (type) {
    case kWall: case kExcitation: case kHead: return 2;
    case kPml1: return 3; case kPml2: return 4; case kPml3: return 5; case kPml4: return 6; case kPml5: return 7;
    case kDead: return 8;
    default: return 0;   // air, pml0, dynamic, noPressure
    }
}

#ifndef VS_FAST_SPLIT_K   // set by vs_fast_k17.hip, which builds one K of the step kernel only
__global__ void k_fast_setup(const TubeDesc* __restrict__ tubes, Bufs<float> b, FastBufs f) {
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
    // pressure
    const double m = nop ? 0.0 : 1.0;
    f.kP[q] = (float)(m * K1 / ((1.0 + sc) * dP));
    int codeP = nop ? 1 : (sigma_code(tc) == 8 ? 1 : sigma_code(tc));
    // faces
    const double mbx = br < bc ? br : bc, mby = bt < bc ? bt : bc;
    const double Mx = sc < sr ? sr : sc, My = sc < st ? st : sc;
    const int mcx = max(sigma_code(tc), sigma_code(tr)), mcy = max(sigma_code(tc), sigma_code(tt));
    const int codeX = (mbx == 0.0 || mcx == 8) ? 1 : mcx;
    const int codeY = (mby == 0.0 || mcy == 8) ? 1 : mcy;
    float gx = mbx == 0.0 ? 0.0f : (float)(dX * (mbx * mbx * C) / (mbx + Mx));
    float gy = mby == 0.0 ? 0.0f : (float)(dY * (mby * mby * C) / (mby + My));
    uint32_t code = ((uint32_t)(codeX + 8 * codeY) << kShiftXY) | ((uint32_t)codeP << kShiftP);
    // special faces (their damping codes are "zero": beta = 0 on both kinds)
    if (tc == kExcitation) {
        // (exeV * eWx * Mx) / (0 + Mx): the x-face carries the source velocity
        gx = (float)dX;
        gy = 0.0f;
        code |= kCodeSelfOff | kCodeSrcX;
    } else if (segwall) {
        const int ibr = (int)br, ibt = (int)bt;
        const int lut = (ibr == kAir && ibt == kAir) ? 0 : (ibr == kAir && ibt == kWall) ? 1
                      : (ibr == kWall && ibt == kAir) ? 2 : 3;
        const double nin = T.ph.wall_normal[lut] * (1.0 - bc);
        const double z = b.zcol[T.col_off + c];
        const int ec = 0, er = tr == kExcitation, et = tt == kExcitation;
        const double nex = (double)((1 - ec) * (1 - er)), ney = (double)((1 - ec) * (1 - et));
        // (Mx * z * nE * (-xor1 * P_r * Nin)) / (0 + Mx), xor1 = beta_r (beta_c = 0 on walls)
        gx = (float)(dX * z * nex * br * nin);
        gy = (float)(dY * z * ney * bt * nin);
        code |= kCodeSelfOff;
    }
    f.gX[q] = gx;
    f.gY[q] = gy;
    f.code[q] = (uint16_t)code;
}

// Flux <-> velocity conversions of the state (start / end of the fast path).
__global__ void k_vel_to_flux(const TubeDesc* __restrict__ tubes, Bufs<float> b, FastBufs f) {
    const TubeDesc& T = tubes[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T.H * T.W) return;
    const long long q = T.cell_off + i;
    f.Uxf[q] = b.Vx[q] * b.dX[q];
    f.Uyf[q] = b.Vy[q] * b.dY[q];
}

#endif   // VS_FAST_SPLIT_K

template <int K>
__device__ inline uint32_t class_of(const uint32_t* cw, int k) {
    return (cw[k >> 2] >> (8 * (k & 3))) & 0xffu;
}

// Slab stores: a work-group-uniform base plus a per-lane 32-bit byte offset
// (one VGPR, advanced by T floats per slot), so each store is a single
// global store with a scalar base instead of a 64-bit address per slot.
__device__ inline void slab_store(char* base, uint32_t off, float v) {
    *reinterpret_cast<float*>(base + off) = v;
}

// Advances a slab offset through an opaque (non-volatile) asm, so the
// compiler keeps one VGPR instead of K precomputed scalar offsets, without a
// scheduling barrier.
__device__ inline uint32_t next_off(uint32_t vo, uint32_t step) {
    vo += step;
    asm("" : "+v"(vo));
    return vo;
}

// Sign-bit flags of the (ax, ay) class entries.
__device__ __forceinline__ bool flag_off(float e) { return __float_as_int(e) < 0; }

// Walks the padded interior j = r * Wp + c (c < Wp; c == Wp - 1 is the pad
// column) over the work-group: thread i visits j = i, i + B, ... below n,
// with its row and column advanced incrementally.  Consecutive threads visit
// consecutive cells of a row, so the global accesses f makes at
// (r + 1) * W + c + 1 are coalesced.
template <typename F>
__device__ __forceinline__ void for_padded(int n, int Wp, F&& f) {
    const int B = blockDim.x;
    const int Bq = B / Wp, Br = B - Bq * Wp;
    int j = threadIdx.x;
    int r = j / Wp, c = j - r * Wp;
    for (; j < n; j += B) {
        f(j, r, c);
        c += Br;
        r += Bq;
        if (c >= Wp) { c -= Wp; ++r; }
    }
}

// The two phases of one step over a lane's run of K cells.  The pointers are
// __restrict__: in the pressure-like phase only the lane's own P-region is
// written (Uy of this lane and of the lane below are only read), in the face
// phase only the Uy-region (P of this lane, the next and the one above are
// only read), so the compiler may issue the LDS reads of later cells ahead of
// the writes of earlier ones.
template <int K, bool REC>
__device__ __forceinline__ void fast_pressure(float* __restrict__ P, const float* __restrict__ UY,
                                              const float* __restrict__ UYB, const float* __restrict__ ptab,
                                              const float (&Ux)[K], const float (&kP)[K], const uint32_t (&cw)[(K + 3) / 4],
                                              float uxl0, char* dr, uint32_t vo, uint32_t T4, uint32_t KT4) {
#pragma clang fp contract(fast)
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const float p = P[k];
        const float uy = UY[k];
        const float uyb = UYB[k];
        const float left = k > 0 ? Ux[k - 1] : uxl0;
        const float dv = (Ux[k] - left) + (uy - uyb);
        const float aP = ptab[class_of<K>(cw, k)];
        const float pn = aP * p - kP[k] * dv;
        P[k] = pn;
        if (REC) {
            slab_store(dr, vo, pn);
            slab_store(dr + KT4, vo, dv);
            vo = next_off(vo, T4);
        }
    }
}

template <int K>
__device__ __forceinline__ void fast_faces(const float* __restrict__ P, const float* __restrict__ PA,
                                           float* __restrict__ UY, float* __restrict__ UXE,
                                           const float2* __restrict__ ftab, float (&Ux)[K],
                                           const float (&gX)[K], const float (&gY)[K],
                                           const uint32_t (&cw)[(K + 3) / 4], float ev) {
#pragma clang fp contract(fast)
    float pn = P[0];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const float p = pn;
        pn = P[k + 1];   // P[K]: the next lane's first cell
        const float pa = PA[k];
        const float2 c2 = ftab[class_of<K>(cw, k)];   // ax, ay with the s, r flags in their signs
        // walls with admittance (s = 0): U = -g P_neighbour; excitation (s = r = 0): Ux = g exeV
        const float ps = flag_off(c2.x) ? 0.f : p;
        const float prx = flag_off(c2.y) ? -ev : pn;
        const float fx = ps - prx, fy = ps - pa;
        Ux[k] = fabsf(c2.x) * Ux[k] + gX[k] * fx;
        UY[k] = fabsf(c2.y) * UY[k] + gY[k] * fy;
    }
    *UXE = Ux[K - 1];
}

// REC: also write the forward factors of every step (the values that
// multiply gX, gY and kP) into rec, for the reverse pass.
template <int K, bool REC>
__global__ __launch_bounds__(1024) void k_fast_step(const TubeDesc* __restrict__ tubes,
                                                    const int* __restrict__ order, Bufs<float> b,
                                                    FastBufs f, int t0_arg, int nsteps,
                                                    const int* __restrict__ d_t0, int G,
                                                    float* __restrict__ frec, long long slab_tube,
                                                    float* __restrict__ fsink) {
#pragma clang fp contract(fast)
    extern __shared__ float lds[];
    __shared__ float2 ftab[kFastClasses];        // (ax, ay) by class, flags in the signs
    __shared__ float ptab[kFastClasses];         // aP by class
    __shared__ double sexe[2];                   // source value of steps t, t + 1
    __shared__ TwoMassState tm;
    // threads [0, T) own cells; the last wave is the control wave (source,
    // two-mass recurrence, probe).  Keeping the double-precision recurrence
    // out of the cell waves keeps their register budget for the cell arrays.
    const int T = blockDim.x - 64, t = threadIdx.x;
    const int NL = T + 2 * G;                    // lanes with the guards
    float* const Pg = lds + G * K;               // P[j]
    float* const Ug = lds + (NL + G) * K;        // Uy[j]
    float* const Xg = lds + 2 * NL * K + G;      // Ux of lane l's last cell at Xg[l]

    const int tube = order[blockIdx.x];
    const TubeDesc& D = tubes[tube];
    const int H = D.H, W = D.W, Hi = H - 2, Wi = W - 2;
    const int Wp = Wi + 1;
    // graph replays read the first step from device memory
    const int t0 = d_t0 ? *d_t0 : t0_arg;
    const int tend = min(t0 + nsteps, D.num_steps);
    const long long co = D.cell_off;

    // zero guards (LDS starts undefined): G K floats on both sides of the P
    // and Uy regions, G on both sides of the run-end region
    for (int i = t; i < 2 * (2 * G * K + G); i += blockDim.x) {
        const int r = i / (2 * G * K);             // 0: P, 1: Uy, 2: run ends
        const int g = i - r * 2 * G * K;
        if (r < 2) lds[r * NL * K + (g < G * K ? g : g + T * K)] = 0.f;
        else lds[2 * NL * K + (g < G ? g : g + T)] = 0.f;
    }

    // ------------- state and coefficients into the work-group -------------
    // Each field crosses global memory once, coalesced: the work-group copies
    // two fields at a time in padded-index order into the P and Uy regions
    // of LDS, then every cell lane picks up its run j = t K + k.  Slots past
    // the tube's last row (r >= Hi) and pad-column slots are zero.
    const int KT = K * T;
    const int jb = t * K;
    const bool cell_lane = t < T;
    float Ux[K], kP[K], gX[K], gY[K];
    uint32_t cw[(K + 3) / 4];
    auto stage2 = [&](const float* __restrict__ a, const float* __restrict__ c2) {
        for_padded(KT, Wp, [&](int j, int r, int c) {
            float va = 0.f, vb = 0.f;
            if (r < Hi && c < Wi) {
                const long long q = co + (long long)(r + 1) * W + (c + 1);
                va = a[q];
                vb = c2[q];
            }
            Pg[j] = va;
            Ug[j] = vb;
        });
        __syncthreads();
    };
    stage2(f.Uxf, f.kP);
    if (cell_lane) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            Ux[k] = Pg[jb + k];
            kP[k] = Ug[jb + k];
        }
    }
    __syncthreads();
    stage2(f.gX, f.gY);
    if (cell_lane) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            gX[k] = Pg[jb + k];
            gY[k] = Ug[jb + k];
        }
    }
    __syncthreads();
    for_padded(KT, Wp, [&](int j, int r, int c) {
        const uint32_t cd = (r < Hi && c < Wi) ? (uint32_t)f.cls8[co + (long long)(r + 1) * W + (c + 1)] : 0u;
        Ug[j] = __uint_as_float(cd);
    });
    __syncthreads();
    if (cell_lane) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t cd = __float_as_uint(Ug[jb + k]);
            if (k & 3) cw[k >> 2] |= cd << (8 * (k & 3)); else cw[k >> 2] = cd;
        }
    }
    __syncthreads();
    for_padded(KT, Wp, [&](int j, int r, int c) {
        float p = 0.f, uy = 0.f;
        if (r < Hi && c < Wi) {
            const long long q = co + (long long)(r + 1) * W + (c + 1);
            p = b.P[q];
            uy = f.Uyf[q];
        }
        Pg[j] = p;
        Ug[j] = uy;
    });

    if (t >= T) {
        // ------------------------- control wave -------------------------
        const int lane = t - T;
        const bool twomass = D.source_type == 2;
        // The recurrence is a chain of dependent double divisions on one
        // lane; at the highest wave priority its instructions issue ahead of
        // the cell waves sharing its SIMD (with a Gaussian source the wave
        // only moves the source, probe and trace values and keeps priority 0)
        if (twomass) __builtin_amdgcn_s_setprio(3);
        const int jL = (D.listener / W - 1) * Wp + (D.listener % W - 1);
        const int jE = (D.excitation / W - 1) * Wp + (D.excitation % W - 1);
        const float* pL = Pg + jL;
        const float* pE = Pg + jE;
        double pexc = 0.0;
        ptab[lane] = f.ctab[lane];
        ftab[lane] = make_float2(f.ctab[kFastClasses + lane], f.ctab[2 * kFastClasses + lane]);
        // The two-mass steps run one and two steps ahead of the cells.  While
        // the cells do the pressure phase of ts, lane 0 evaluates the flow
        // resistance and inertance of step ts + 1 (twomass_flow) and the part
        // of step ts + 2's fold solve that does not depend on the glottal flow
        // (twomass_pre).  After it, fed by the excitation cell's new pressure,
        // it finishes step ts + 1 (twomass_finish: the source of the cells'
        // next velocity phase) and the solve of ts + 2 (twomass_post).  Each
        // phase thus carries about half of the step's double divisions; the
        // operations and their order are those of twomass_step.
        TwoMassNext nx;
        TwoMassPre pp;
        if (lane == 0) {
            pexc = b.pexc[tube];
            if (twomass) tm = b.tm[tube];
            if (t0 < tend) {
                double e;
                if (twomass) { tm.p1 = pexc; e = twomass_step(tm); }
                else e = (double)b.source[D.src_off + t0];
                sexe[t0 & 1] = (double)(float)e;
                if (twomass && t0 + 1 < tend) twomass_solve(tm, nx);
            }
        }
        __syncthreads();
        for (int ts = t0; ts < tend; ++ts) {
            if (lane == 0 && twomass && ts + 1 < tend) {
                twomass_flow(tm, nx);
                if (ts + 2 < tend) twomass_pre(tm, nx.x1n, nx.x2n, tm.x1_curr, tm.x2_curr, pp);
            }
            __syncthreads();   // pressure phase done
            if (lane == 0) {
                b.probe[D.src_off + ts] = (double)*pL;
                pexc = (double)*pE;
                b.strace[D.src_off + ts] = sexe[ts & 1];
                if (ts + 1 < tend) {
                    double e;
                    if (twomass) {
                        e = twomass_finish(tm, nx, pexc);
                        if (ts + 2 < tend) twomass_post(tm, pp, nx);
                    } else {
                        e = (double)b.source[D.src_off + ts + 1];
                    }
                    sexe[(ts + 1) & 1] = (double)(float)e;
                }
            }
            __syncthreads();   // velocity phase done
        }
        if (lane == 0) {
            b.pexc[tube] = pexc;
            if (twomass) b.tm[tube] = tm;
        }
    } else {
    // --------------------------- cell waves ---------------------------
    Xg[t] = Ux[K - 1];
    __syncthreads();

    for (int ts = t0; ts < tend; ++ts) {
        // Launder the lane index once per phase so the compiler recomputes
        // the run's addresses instead of keeping per-cell copies live across
        // the time loop.
        int tl = t;
        asm volatile("" : "+v"(tl));
#pragma unroll
        for (int k = 0; k < (K + 3) / 4; ++k) asm volatile("" : "+v"(cw[k]));
        // P' and divergence of this step: field f of slot k at float f * KT + k * T + lane;
        // waves wholly past the tube's cells write a small shared sink instead
        // (it stays in L2: the sums never read those slots)
        char* const dr = !REC ? nullptr
                         : __builtin_amdgcn_readfirstlane(tl) * K < Hi * Wp
                             ? reinterpret_cast<char*>(frec + blockIdx.x * slab_tube + (long long)(ts - t0) * 2 * KT)
                             : reinterpret_cast<char*>(fsink);
        const uint32_t T4 = 4u * (uint32_t)T, KT4 = 4u * (uint32_t)KT;
        uint32_t vo = 4u * (uint32_t)tl;
        // ---------------- pressure phase ----------------
        {
            const int j = tl * K;
            // Uy of the cells below (j + Wp); Ux of the previous lane's last cell
            fast_pressure<K, REC>(Pg + j, Ug + j, Ug + j + Wp, ptab, Ux, kP, cw, Xg[tl - 1], dr, vo, T4, KT4);
        }
        __syncthreads();
        const float ev = (float)sexe[ts & 1];
        // ---------------- velocity phase ----------------
        {
            asm volatile("" : "+v"(tl));
#pragma unroll
            for (int k = 0; k < (K + 3) / 4; ++k) asm volatile("" : "+v"(cw[k]));
            const int j = tl * K;
            // P of the cells above (j - Wp)
            fast_faces<K>(Pg + j, Pg + j - Wp, Ug + j, Xg + tl, ftab, Ux, gX, gY, cw, ev);
        }
        __syncthreads();
    }
    }   // cell waves
    // ---------------- write back ----------------
    // P and Uy leave from their LDS regions, coalesced; then the lanes' Ux
    // go through the P region the same way.  V = U / depth as before.
    __syncthreads();
    for_padded(Hi * Wp, Wp, [&](int j, int r, int c) {
        if (c < Wi) {
            const long long q = co + (long long)(r + 1) * W + (c + 1);
            const float uy = Ug[j];
            b.P[q] = Pg[j];
            f.Uyf[q] = uy;
            b.Vy[q] = uy / b.dY[q];
        }
    });
    __syncthreads();
    if (cell_lane) {
#pragma unroll
        for (int k = 0; k < K; ++k) Pg[jb + k] = Ux[k];
    }
    __syncthreads();
    for_padded(Hi * Wp, Wp, [&](int j, int r, int c) {
        if (c < Wi) {
            const long long q = co + (long long)(r + 1) * W + (c + 1);
            const float ux = Pg[j];
            f.Uxf[q] = ux;
            b.Vx[q] = ux / b.dX[q];
        }
    });
}

#ifndef VS_FAST_SPLIT_K
// ---------------------------------------------------------------------------
// Wide-access variant of the step above (forward only), for K = 4 (mod 8).
// Rows are padded to Wp = Wi + 1 rounded up to a multiple of 4, so every
// lane run (t K words) and every row offset is 16-byte aligned: the run's own
// P and Uy, the Uy of the cells below and the P of the cells above move as
// ds_read_b128 / ds_write_b128, and with K / 4 odd the 16 lanes that one
// b128 access serves per pass cover all 64 banks.  A zero guard of GW >= Wp
// floats before the P region (the row above row 0), 4 floats after it (the
// next lane's first P read by the last lane) and GW after the Uy region (the
// row below the last row) keep every neighbour read of a real or pad slot
// inside LDS.
//
// Class tables by cell pair.  K is even and runs start at even j, so the
// cells j, j + 1 (j even) of a run form pairs that never cross a row.  A
// pair's class byte indexes two tables: the face entry (ax, ay of both
// cells, s = 0 and r = 0 in the sign bits, one 16-byte read per pair) and
// the pressure entry (aP of both cells, one 8-byte read per pair).  Per cell
// and step that is 12 bytes of table reads in two accesses per pair instead
// of 20 bytes in four, and the LDS is the pipe this kernel waits on.  The
// face update takes s and r from the sign bits as fast_faces does,
//   fx = (s ? P : 0) - (r ? P_right : -ev),  fy = (s ? P : 0) - P_top,
// and the step is written with the same fused operations as the contracted
// code of k_fast_step, so both kernels give the same values (signed zeros
// aside).
// ---------------------------------------------------------------------------
typedef float f4v __attribute__((ext_vector_type(4)));

// Copies a two-mass state word by word with the 64 lanes of a wave.
__device__ inline void copy_words(float* __restrict__ dst, const float* __restrict__ src, int lane) {
    static_assert(sizeof(TwoMassState) % sizeof(float) == 0, "two-mass state copied in words");
    for (int i = lane; i < (int)(sizeof(TwoMassState) / sizeof(float)); i += 64) dst[i] = src[i];
}

// Pair class of the run's pair h (cells 2h, 2h + 1): byte h of the packed words.
template <int K>
__device__ inline uint32_t pair_of(const uint32_t (&cp)[(K / 2 + 3) / 4], int h) {
    return (cp[h >> 2] >> (8 * (h & 3))) & 0xffu;
}

template <int K>
__device__ __forceinline__ void wide_pressure(float* __restrict__ P, const float* __restrict__ UY,
                                              const float* __restrict__ UYB, const float2* __restrict__ ptab2,
                                              const float (&Ux)[K], const float (&kP)[K],
                                              const uint32_t (&cp)[(K / 2 + 3) / 4], float uxl0) {
#pragma clang fp contract(off)
    f4v* const P4 = reinterpret_cast<f4v*>(P);
    const f4v* const U4 = reinterpret_cast<const f4v*>(UY);
    const f4v* const B4 = reinterpret_cast<const f4v*>(UYB);
#pragma unroll
    for (int m = 0; m < K / 4; ++m) {
        // a group's table reads (aP of its two pairs) go out with its state
        // reads, so the group waits for LDS once
        const float2 a01 = ptab2[pair_of<K>(cp, 2 * m)], a23 = ptab2[pair_of<K>(cp, 2 * m + 1)];
        const f4v p = P4[m], uy = U4[m], uyb = B4[m];
        __builtin_amdgcn_sched_barrier(0);
        const float aP[4] = {a01.x, a01.y, a23.x, a23.y};
        f4v pn;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = 4 * m + i;
            const float left = k > 0 ? Ux[k - 1] : uxl0;
            const float dv = (Ux[k] - left) + (uy[i] - uyb[i]);
            pn[i] = __builtin_fmaf(p[i], aP[i], -(dv * kP[k]));
        }
        P4[m] = pn;
    }
}

// Face update of the run's group m (cells 4m .. 4m + 3).  SRC: the group
// holds the tube's source column, whose cells take -ev in place of the right
// neighbour's pressure (r = 0 in the sign bit of ay); no other cell has r = 0.
template <int K, bool SRC>
__device__ __forceinline__ void wide_face_group(int m, const f4v* __restrict__ P4, const f4v* __restrict__ A4,
                                                f4v* __restrict__ U4, const float* __restrict__ P,
                                                const f4v* __restrict__ ftab4, float (&Ux)[K],
                                                const float (&gX)[K], const float (&gY)[K],
                                                const uint32_t (&cp)[(K / 2 + 3) / 4], float ev, f4v& p) {
#pragma clang fp contract(off)
    // the group's table and state reads go out together (one wait); P of
    // the next four cells, past the run the next lane's first cell
    const f4v e01 = ftab4[pair_of<K>(cp, 2 * m)], e23 = ftab4[pair_of<K>(cp, 2 * m + 1)];
    f4v pnext;
    if (m + 1 < K / 4) pnext = P4[m + 1];
    else pnext[0] = P[K];
    const f4v pa = A4[m], uy = U4[m];
    __builtin_amdgcn_sched_barrier(0);
    const float ax[4] = {e01[0], e01[2], e23[0], e23[2]};   // s = 0 in the sign bit
    const float ay[4] = {e01[1], e01[3], e23[1], e23[3]};   // r = 0 in the sign bit
    f4v un;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int k = 4 * m + i;
        const float pr = i < 3 ? p[i + 1] : pnext[0];
        // walls with admittance (s = 0): U = -g P_neighbour; excitation (s = r = 0): Ux = g exeV
        const float ps = flag_off(ax[i]) ? 0.f : p[i];
        const float prx = SRC && flag_off(ay[i]) ? -ev : pr;
        const float fx = ps - prx, fy = ps - pa[i];
        Ux[k] = __builtin_fmaf(Ux[k], fabsf(ax[i]), gX[k] * fx);
        un[i] = __builtin_fmaf(fabsf(ay[i]), uy[i], gY[k] * fy);
    }
    U4[m] = un;
    p = pnext;
}

// MS: the group that holds the source column of every tube of the batch
// (-1: no tube has a source face, -2: the columns differ, every group
// handles the source).  The condition folds once the loop is unrolled.
template <int K, int MS>
__device__ __forceinline__ void wide_faces(const float* __restrict__ P, const float* __restrict__ PA,
                                           float* __restrict__ UY, float* __restrict__ UXE,
                                           const f4v* __restrict__ ftab4, float (&Ux)[K],
                                           const float (&gX)[K], const float (&gY)[K],
                                           const uint32_t (&cp)[(K / 2 + 3) / 4], float ev) {
    const f4v* const P4 = reinterpret_cast<const f4v*>(P);
    const f4v* const A4 = reinterpret_cast<const f4v*>(PA);
    f4v* const U4 = reinterpret_cast<f4v*>(UY);
    f4v p = P4[0];
#pragma unroll
    for (int m = 0; m < K / 4; ++m) {
        if (MS == -2 || m == MS) wide_face_group<K, true>(m, P4, A4, U4, P, ftab4, Ux, gX, gY, cp, ev, p);
        else wide_face_group<K, false>(m, P4, A4, U4, P, ftab4, Ux, gX, gY, cp, ev, p);
    }
    *UXE = Ux[K - 1];
}

template <int K, int MS>
__global__ __launch_bounds__(1024) void k_fast_wide(const TubeDesc* __restrict__ tubes,
                                                    const int* __restrict__ order, Bufs<float> b,
                                                    FastBufs f, int t0_arg, int nsteps,
                                                    const int* __restrict__ d_t0, int GW, int wq) {
    static_assert(K % 8 == 4, "the wide kernel needs K = 4 (mod 8)");
    extern __shared__ f4v lds4[];
    float* const lds = reinterpret_cast<float*>(lds4);
    __shared__ f4v ftab4[kFastPairs];            // (ax, ay, ax, ay) by pair class, s = 0 / r = 0 in the signs
    __shared__ float2 ptab2[kFastPairs];         // (aP, aP) by pair class
    __shared__ double sexe[2];                   // source value of steps t, t + 1
    __shared__ TwoMassState tm;
    const int T = blockDim.x - 64, t = threadIdx.x;
    const int KT = K * T;
    float* const Pg = lds + GW;                  // P[j], j in [-GW, KT + 4)
    float* const Ug = Pg + KT + 4;               // Uy[j], j in [0, KT + GW)
    float* const Xg = Ug + KT + GW + 4;          // Ux of lane l's last cell at Xg[l], Xg[-1] = 0

    const int tube = order[blockIdx.x];
    const TubeDesc& D = tubes[tube];
    const int H = D.H, W = D.W, Hi = H - 2, Wi = W - 2;
    // rows are padded to a multiple of wq: with wq = K a column sits at the
    // same slot of its lanes' runs in every row (the source column in group
    // MS); wq = 4 runs with MS = -2
    const int Wp = fast_wide_wp(Wi, wq);
    const int t0 = d_t0 ? *d_t0 : t0_arg;
    const int tend = min(t0 + nsteps, D.num_steps);
    const long long co = D.cell_off;

    // zero guards: GW floats before the P region and 4 after it, GW after
    // the Uy region, and Xg[-1]
    for (int i = t; i < 2 * GW + 5; i += blockDim.x) {
        if (i < GW) Pg[i - GW] = 0.f;
        else if (i < GW + 4) Pg[KT + (i - GW)] = 0.f;
        else if (i < 2 * GW + 4) Ug[KT + (i - GW - 4)] = 0.f;
        else Xg[-1] = 0.f;
    }

    const int jb = t * K;
    const bool cell_lane = t < T;
    float Ux[K], kP[K], gX[K], gY[K];
    uint32_t cp[(K / 2 + 3) / 4];   // pair class bytes of the run
    auto stage2 = [&](const float* __restrict__ a, const float* __restrict__ c2) {
        for_padded(KT, Wp, [&](int j, int r, int c) {
            float va = 0.f, vb = 0.f;
            if (r < Hi && c < Wi) {
                const long long q = co + (long long)(r + 1) * W + (c + 1);
                va = a[q];
                vb = c2[q];
            }
            Pg[j] = va;
            Ug[j] = vb;
        });
        __syncthreads();
    };
    stage2(f.Uxf, f.kP);
    if (cell_lane) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            Ux[k] = Pg[jb + k];
            kP[k] = Ug[jb + k];
        }
    }
    __syncthreads();
    stage2(f.gX, f.gY);
    if (cell_lane) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            gX[k] = Pg[jb + k];
            gY[k] = Ug[jb + k];
        }
    }
    __syncthreads();
    // pair classes sit at the frame position of each pair's first cell (even
    // c); pairs of pad cells (c >= Wi) are pair class 0
    for_padded(KT, Wp, [&](int j, int r, int c) {
        const uint32_t cd = (r < Hi && c < Wi) ? (uint32_t)f.pcls8[co + (long long)(r + 1) * W + (c + 1)] : 0u;
        Ug[j] = __uint_as_float(cd);
    });
    __syncthreads();
    if (cell_lane) {
#pragma unroll
        for (int h = 0; h < K / 2; ++h) {
            const uint32_t cd = __float_as_uint(Ug[jb + 2 * h]);
            if (h & 3) cp[h >> 2] |= cd << (8 * (h & 3)); else cp[h >> 2] = cd;
        }
    }
    __syncthreads();
    for_padded(KT, Wp, [&](int j, int r, int c) {
        float p = 0.f, uy = 0.f;
        if (r < Hi && c < Wi) {
            const long long q = co + (long long)(r + 1) * W + (c + 1);
            p = b.P[q];
            uy = f.Uyf[q];
        }
        Pg[j] = p;
        Ug[j] = uy;
    });

    if (t >= T) {
        // ------------------------- control wave -------------------------
        const int lane = t - T;
        const bool twomass = D.source_type == 2;
        // The recurrence is a chain of dependent double divisions on one
        // lane; at the highest wave priority its instructions issue ahead of
        // the cell waves sharing its SIMD (with a Gaussian source the wave
        // only moves the source, probe and trace values and keeps priority 0)
        if (twomass) __builtin_amdgcn_s_setprio(3);
        const int jL = (D.listener / W - 1) * Wp + (D.listener % W - 1);
        const int jE = (D.excitation / W - 1) * Wp + (D.excitation % W - 1);
        const float* pL = Pg + jL;
        const float* pE = Pg + jE;
        double pexc = 0.0;
        for (int q = lane; q < kFastPairs; q += 64) {
            f4v e;
#pragma unroll
            for (int i = 0; i < 4; ++i) e[i] = f.ptab[4 * q + i];
            ftab4[q] = e;
            ptab2[q] = make_float2(f.ptab[4 * kFastPairs + 2 * q], f.ptab[4 * kFastPairs + 2 * q + 1]);
        }
        // The two-mass steps run one and two steps ahead of the cells.  While
        // the cells do the pressure phase of ts, lane 0 evaluates the flow
        // resistance and inertance of step ts + 1 (twomass_flow) and the part
        // of step ts + 2's fold solve that does not depend on the glottal flow
        // (twomass_pre).  After it, fed by the excitation cell's new pressure,
        // it finishes step ts + 1 (twomass_finish: the source of the cells'
        // next velocity phase) and the solve of ts + 2 (twomass_post).  Each
        // phase thus carries about half of the step's double divisions; the
        // operations and their order are those of twomass_step.
        TwoMassNext nx;
        TwoMassPre pp;
        if (twomass) {
            // the wave copies the state into LDS word by word (not through
            // lane 0's registers), then lane 0 reads it
            copy_words(reinterpret_cast<float*>(&tm), reinterpret_cast<const float*>(b.tm + tube), lane);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (lane == 0) {
            pexc = b.pexc[tube];
            if (t0 < tend) {
                double e;
                if (twomass) { tm.p1 = pexc; e = twomass_step(tm); }
                else e = (double)b.source[D.src_off + t0];
                sexe[t0 & 1] = (double)(float)e;
                if (twomass && t0 + 1 < tend) twomass_solve(tm, nx);
            }
        }
        __syncthreads();
        for (int ts = t0; ts < tend; ++ts) {
            if (lane == 0 && twomass && ts + 1 < tend) {
                twomass_flow(tm, nx);
                if (ts + 2 < tend) twomass_pre(tm, nx.x1n, nx.x2n, tm.x1_curr, tm.x2_curr, pp);
            }
            __syncthreads();   // pressure phase done
            if (lane == 0) {
                b.probe[D.src_off + ts] = (double)*pL;
                pexc = (double)*pE;
                b.strace[D.src_off + ts] = sexe[ts & 1];
                if (ts + 1 < tend) {
                    double e;
                    if (twomass) {
                        e = twomass_finish(tm, nx, pexc);
                        if (ts + 2 < tend) twomass_post(tm, pp, nx);
                    } else {
                        e = (double)b.source[D.src_off + ts + 1];
                    }
                    sexe[(ts + 1) & 1] = (double)(float)e;
                }
            }
            __syncthreads();   // velocity phase done
        }
        if (lane == 0) b.pexc[tube] = pexc;
        // lane 0's last writes to tm precede the last work-group barrier
        if (twomass) copy_words(reinterpret_cast<float*>(b.tm + tube), reinterpret_cast<const float*>(&tm), lane);
    } else {
        // --------------------------- cell waves ---------------------------
        Xg[t] = Ux[K - 1];
        __syncthreads();
        for (int ts = t0; ts < tend; ++ts) {
            int tl = t;
            asm volatile("" : "+v"(tl));
#pragma unroll
            for (int k = 0; k < (K / 2 + 3) / 4; ++k) asm volatile("" : "+v"(cp[k]));
            {
                const int j = tl * K;
                wide_pressure<K>(Pg + j, Ug + j, Ug + j + Wp, ptab2, Ux, kP, cp, Xg[tl - 1]);
            }
            __syncthreads();
            const float ev = (float)sexe[ts & 1];
            {
                asm volatile("" : "+v"(tl));
#pragma unroll
                for (int k = 0; k < (K / 2 + 3) / 4; ++k) asm volatile("" : "+v"(cp[k]));
                const int j = tl * K;
                wide_faces<K, MS>(Pg + j, Pg + j - Wp, Ug + j, Xg + tl, ftab4, Ux, gX, gY, cp, ev);
            }
            __syncthreads();
        }
    }
    // ---------------- write back ----------------
    __syncthreads();
    for_padded(Hi * Wp, Wp, [&](int j, int r, int c) {
        if (c < Wi) {
            const long long q = co + (long long)(r + 1) * W + (c + 1);
            const float uy = Ug[j];
            b.P[q] = Pg[j];
            f.Uyf[q] = uy;
            b.Vy[q] = uy / b.dY[q];
        }
    });
    __syncthreads();
    if (cell_lane) {
#pragma unroll
        for (int k = 0; k < K; ++k) Pg[jb + k] = Ux[k];
    }
    __syncthreads();
    for_padded(Hi * Wp, Wp, [&](int j, int r, int c) {
        if (c < Wi) {
            const long long q = co + (long long)(r + 1) * W + (c + 1);
            const float ux = Pg[j];
            f.Uxf[q] = ux;
            b.Vx[q] = ux / b.dX[q];
        }
    });
}
#endif   // VS_FAST_SPLIT_K

// Face update of the scaled adjoint, a L + g (K_neighbour - K_cell), with its
// rounding fixed (a product, a difference, one fused multiply-add) so the
// sums kernel replays it bit for bit.
__device__ __forceinline__ float adj_face(float a, float L, float g, float kn, float kc) {
    return __builtin_fmaf(g, kn - kc, a * L);
}

// The two phases of one reverse step over a lane's run (same aliasing
// structure as fast_pressure / fast_faces: Kp-region written in the first,
// Ly-region in the second).
template <int K>
__device__ __forceinline__ void rev_pressure(float* __restrict__ KP, const float* __restrict__ LY,
                                             const float* __restrict__ LYB, const float* __restrict__ ptab,
                                             const float2* __restrict__ ftab, const float (&LX)[K], const float (&kP)[K],
                                             const uint32_t (&cw)[(K + 3) / 4], float lx0, int kL, float gl,
                                             char* ad, uint32_t vo, uint32_t T4) {
#pragma clang fp contract(fast)
    float left = lx0;   // r Lx of the cell on the left
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t cb = class_of<K>(cw, k);
        const float2 c2 = ftab[cb];   // ax, ay with the s, r flags in their signs
        const float kold = KP[k], ly = LY[k], lx = LX[k];
        const float lyb = LYB[k];
        float br = (flag_off(c2.x) ? 0.f : lx + ly) - left - lyb;
        if (k == kL) br += gl;
        const float kn = ptab[cb] * kold + kP[k] * br;
        left = flag_off(c2.y) ? 0.f : lx;
        KP[k] = kn;
        slab_store(ad, vo, kn);
        vo = next_off(vo, T4);
    }
}

template <int K>
__device__ __forceinline__ void rev_faces(const float* __restrict__ KP, const float* __restrict__ KA,
                                          float* __restrict__ LY, float* __restrict__ LXE,
                                          const float2* __restrict__ ftab, float (&LX)[K],
                                          const float (&gX)[K], const float (&gY)[K],
                                          const uint32_t (&cw)[(K + 3) / 4]) {
#pragma clang fp contract(fast)
    float kn = KP[0];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const float kc = kn;
        kn = KP[k + 1];   // KP[K]: the next lane's first cell
        const float ka = KA[k];
        const float2 c2 = ftab[class_of<K>(cw, k)];
        LX[k] = adj_face(fabsf(c2.x), LX[k], gX[k], kn, kc);
        LY[k] = adj_face(fabsf(c2.y), LY[k], gY[k], ka, kc);
    }
    *LXE = flag_off(ftab[class_of<K>(cw, K - 1)].y) ? 0.f : LX[K - 1];
}

// ---------------------------------------------------------------------------
// Reverse steps of the flux-form step above, one work-group of T = blockDim
// cell lanes per tube, in the same on-chip layout.  Differentiating
//   P'  = aP P - kP (Ux - Ux_left + Uy - Uy_below)
//   Ux' = ax Ux + gX (s P' - r P'_right) + gX (1 - r) ev     s = !SelfOff, r = !SrcX
//   Uy' = ay Uy + gY (s P' - P'_top)
//   probe_n = P'_listener
// and scaling the adjoints by their coefficients (Kp = kP M, Lx = gX lambda_x,
// Ly = gY lambda_y, M the adjoint of P') gives, for reverse step n:
//   Kp' = aP Kp + kP (g_n [listener] + s (Lx + Ly) - r_left Lx_left - Ly_below)
//   Lx' = ax Lx + gX (Kp'_right - Kp')
//   Ly' = ay Ly + gY (Kp'_top - Kp')
// i.e. the forward stencil with Kp in the place of P.  The LDS regions hold
// Kp[j], Ly[j] and r Lx of each run's last cell, laid out as the forward's
// P, Uy and run-end Ux; Lx lives in VGPRs.
// Every step stores Lx, Ly (as they enter the step) and Kp' into a.adj.
// ---------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(1024) void k_fast_reverse(const TubeDesc* __restrict__ tubes,
                                                       const int* __restrict__ order, FastBufs f, FastAdj a,
                                                       int s0, int ns, int G) {
#pragma clang fp contract(fast)
    extern __shared__ float lds[];
    __shared__ float2 ftab[kFastClasses];
    __shared__ float ptab[kFastClasses];
    const int T = blockDim.x, t = threadIdx.x;
    const int KT = K * T;
    const int NL = T + 2 * G;                    // lanes with the guards
    float* const Kg = lds + G * K;               // Kp[j]
    float* const Lg = lds + (NL + G) * K;        // Ly[j]
    float* const Xg = lds + 2 * NL * K + G;      // r Lx of lane l's last cell at Xg[l]

    const int tube = order[blockIdx.x];
    const TubeDesc& D = tubes[tube];
    const int H = D.H, W = D.W, Hi = H - 2, Wi = W - 2;
    const int Wp = Wi + 1;
    const int tend = min(s0 + ns, D.num_steps);
    const long long co = D.cell_off;
    // the listener's lane and slot
    const int jL = (D.listener / W - 1) * Wp + (D.listener % W - 1);
    const int lL = jL / K, kL = jL - (jL / K) * K;
    const double* const gp = a.gprobe + D.src_off;

    for (int i = t; i < 2 * (2 * G * K + G); i += T) {
        const int r = i / (2 * G * K);             // 0: Kp, 1: Ly, 2: run ends
        const int g = i - r * 2 * G * K;
        if (r < 2) lds[r * NL * K + (g < G * K ? g : g + T * K)] = 0.f;
        else lds[2 * NL * K + (g < G ? g : g + T)] = 0.f;
    }
    if (t < 64) {
        ptab[t] = f.ctab[t];
        ftab[t] = make_float2(f.ctab[kFastClasses + t], f.ctab[2 * kFastClasses + t]);
    }
    float LX[K], kP[K], gX[K], gY[K];
    uint32_t cw[(K + 3) / 4];
    float* const st = a.state + (long long)blockIdx.x * 3 * KT + t;
    {
        const int jb = t * K;
        int r = jb / Wp, c = jb - (jb / Wp) * Wp;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            float kp = 0.f, gx = 0.f, gy = 0.f;
            uint32_t cd = 0u;
            if (r < Hi && c < Wi) {
                const long long q = co + (long long)(r + 1) * W + (c + 1);
                kp = f.kP[q]; gx = f.gX[q]; gy = f.gY[q]; cd = f.cls8[q];
            }
            kP[k] = kp; gX[k] = gx; gY[k] = gy;
            if (k & 3) cw[k >> 2] |= cd << (8 * (k & 3)); else cw[k >> 2] = cd;
            const float ly = st[KT + k * T];
            Kg[jb + k] = st[k * T];
            Lg[jb + k] = ly;
            LX[k] = st[2 * KT + k * T];
            // Lx, Ly as this block's reverse begins, for the sums
            a.start[(long long)blockIdx.x * 2 * KT + k * T + t] = LX[k];
            a.start[(long long)blockIdx.x * 2 * KT + KT + k * T + t] = ly;
            if (++c == Wp) { c = 0; ++r; }
        }
        Xg[t] = f.ctab[4 * kFastClasses + class_of<K>(cw, K - 1)] * LX[K - 1];   // r Lx of the run end
    }
    __syncthreads();

    for (int n = tend - 1; n >= s0; --n) {
        int tl = t;
        asm volatile("" : "+v"(tl));
#pragma unroll
        for (int k = 0; k < (K + 3) / 4; ++k) asm volatile("" : "+v"(cw[k]));
        // waves wholly past the tube's cells write the shared sink
        char* const ad = __builtin_amdgcn_readfirstlane(tl) * K < Hi * Wp
                             ? reinterpret_cast<char*>(a.adj + blockIdx.x * a.adj_tube + (long long)(n - s0) * KT)
                             : reinterpret_cast<char*>(a.sink);
        const uint32_t T4 = 4u * (uint32_t)T;
        uint32_t vo = 4u * (uint32_t)tl;
        const float gl = tl == lL ? (float)gp[n] : 0.f;
        // ---------------- pressure-like phase ----------------
        {
            const int j = tl * K;
            // Ly of the cells below (j + Wp); r Lx of the previous lane's last cell
            rev_pressure<K>(Kg + j, Lg + j, Lg + j + Wp, ptab, ftab, LX, kP, cw, Xg[tl - 1], kL, gl, ad, vo, T4);
        }
        __syncthreads();
        // ---------------- face phase ----------------
        {
            asm volatile("" : "+v"(tl));
#pragma unroll
            for (int k = 0; k < (K + 3) / 4; ++k) asm volatile("" : "+v"(cw[k]));
            const int j = tl * K;
            // Kp of the cells above (j - Wp)
            rev_faces<K>(Kg + j, Kg + j - Wp, Lg + j, Xg + tl, ftab, LX, gX, gY, cw);
        }
        __syncthreads();
    }
    {
        int tw = t;
        asm volatile("" : "+v"(tw));
        const int jw = tw * K;
        float* const sw = a.state + (long long)blockIdx.x * 3 * KT + tw;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            sw[k * T] = Kg[jw + k];
            sw[KT + k * T] = Lg[jw + k];
            sw[2 * KT + k * T] = LX[k];
        }
    }
}

#ifndef VS_FAST_SPLIT_K
// Per-cell sums over the steps of one block.  gx work-groups per tube
// stride over its slots s = k * T + t (real cells only), so the slab bases
// are work-group uniform.  From the recorded P' and
// divergence and the reverse pass's Kp, each thread rebuilds, step by step
// backwards:
//   the forward factors  fx = s P' - (r P'_right or -ev),  fy = s P' - P'_top
//   (exactly as the face phase of the forward formed them), and
//   the face adjoints Lx, Ly entering each step, replaying adj_face from the
//   values at the block's start,
// and adds Lx fx, Ly fy, -Kp div (double) to the cell-layout log-derivatives.
// Neighbour values come from the same slab step (L2-resident).
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_fast_adj_sums(const TubeDesc* __restrict__ tubes, const int* __restrict__ order, int ntubes, int K,
                     FastBufs f, FastAdj a, const float* __restrict__ source, int s0, int ns) {
#pragma clang fp contract(fast)
    const int T = a.lanes, KT = K * T;
    // Work-groups are dealt to the 8 XCDs round-robin by linear id; the gx
    // work-groups of tube b get ids w = b % 8 + 8 (x + gx (b / 8)), so all of
    // them share one XCD and its L2, where a thread's right and top
    // neighbour values are the lines its neighbours' threads just loaded.
    const int gx = gridDim.x / (8 * ((ntubes + 7) / 8));
    const int w = blockIdx.x;
    const int b = (w & 7) + 8 * (w / (8 * gx)), x = (w >> 3) % gx;
    if (b >= ntubes) return;
    const TubeDesc& D = tubes[order[b]];
    const int W = D.W, Wp = W - 1, Wi = W - 2, Hi = D.H - 2;
    const int tend = min(s0 + ns, D.num_steps);
    const float* const R = a.rec + b * a.rec_tube;
    const float* const A = a.adj + b * a.adj_tube;
    const float* const st0 = a.start + (long long)b * 2 * KT;
    const float* const src = source + D.src_off;
    for (int s = x * blockDim.x + threadIdx.x; s < KT; s += gx * blockDim.x) {
        const int k = s / T, t = s - (s / T) * T;
        const int j = t * K + k;
        const int r = j / Wp, c = j - (j / Wp) * Wp;
        if (r >= Hi || c >= Wi) continue;
        const long long q = D.cell_off + (long long)(r + 1) * W + (c + 1);
        // neighbour slots: right j + 1 (a real or pad cell of this row), top j - Wp
        const int jr = j + 1, jt = j - Wp;
        const int sr = (jr % K) * T + jr / K;
        const bool has_top = jt >= 0;
        const int stp = has_top ? (jt % K) * T + jt / K : s;
        const uint32_t cb = f.cls8[q];
        const float axe = f.ctab[kFastClasses + cb], aye = f.ctab[2 * kFastClasses + cb];
        const float ax = fabsf(axe), ay = fabsf(aye);
        const bool soff = flag_off(axe), roff = flag_off(aye);
        const float gx = f.gX[q], gy = f.gY[q];
        float lx = st0[s], ly = st0[KT + s];
        double sx = 0.0, sy = 0.0, sp = 0.0;
        // step rows walked backwards from uniform bases; the per-thread
        // offsets (s, sr, stp) stay fixed, and the top loads are always in
        // bounds (stp = s without a top neighbour)
        // the step rows are loaded two steps ahead of their sums (the source
        // value included, so that no later load forces a wait for all of
        // them): the loop is unrolled by three over the rows r0, r1, r2, so a
        // row keeps its registers from its load to its use, with the rows of
        // the next two steps in flight.  The step index is clamped at the
        // block's first step (every load in bounds).
        struct Row { float p, pr, ptr, dv, kc, kr, ktr, ev; };
        auto load = [&](int n) {
            const int m = max(n, s0);
            const long long i = (long long)(m - s0);
            const float* Rn = R + i * 2 * KT;
            const float* An = A + i * KT;
            return Row{Rn[s], Rn[sr], Rn[stp], Rn[KT + s], An[s], An[sr], An[stp], src[m]};
        };
        auto sum = [&](const Row& c) {
            const float pt = has_top ? c.ptr : 0.f, kt = has_top ? c.ktr : 0.f;
            const float ps = soff ? 0.f : c.p;
            const float prx = roff ? -c.ev : c.pr;
            const float fx = ps - prx, fy = ps - pt;
            sx += (double)lx * (double)fx;
            sy += (double)ly * (double)fy;
            sp += (double)c.kc * (double)c.dv;
            lx = adj_face(ax, lx, gx, c.kr, c.kc);
            ly = adj_face(ay, ly, gy, kt, c.kc);
        };
        // (a scheduling barrier after each row's loads keeps the compiler
        // from sinking them next to their use)
        Row r0 = load(tend - 1), r1 = load(tend - 2), r2;
        for (int n = tend - 1;; n -= 3) {
            r2 = load(n - 2);
            __builtin_amdgcn_sched_barrier(0);
            sum(r0);
            if (n - 1 < s0) break;
            r0 = load(n - 3);
            __builtin_amdgcn_sched_barrier(0);
            sum(r1);
            if (n - 2 < s0) break;
            r1 = load(n - 4);
            __builtin_amdgcn_sched_barrier(0);
            sum(r2);
            if (n - 3 < s0) break;
        }
        a.Lx[q] += sx;
        a.Ly[q] += sy;
        a.Lp[q] -= sp;
    }
}

#endif   // VS_FAST_SPLIT_K

// Raises a kernel's dynamic LDS limit to the on-chip kernels' share of the
// 160 KB (once per kernel; configure_fast does it ahead of graph capture).
hipError_t allow_fast_lds(const void* fn, std::atomic<int>& done) {
    if (done.load()) return hipSuccess;
    const int dyn = 160 * 1024 - (int)fast_lds_static_bytes();
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
    if (e == hipSuccess) done.store(1);
    return e;
}

template <int K, bool REC>
hipError_t configure_step() {
    static std::atomic<int> done{0};
    return allow_fast_lds((const void*)k_fast_step<K, REC>, done);
}

template <int K>
hipError_t configure_rev() {
    static std::atomic<int> done{0};
    return allow_fast_lds((const void*)k_fast_reverse<K>, done);
}

template <int K>
hipError_t configure_k() {
    hipError_t e = configure_step<K, false>();
    if (e == hipSuccess) e = configure_step<K, true>();
    if (e == hipSuccess) e = configure_rev<K>();
    return e;
}

template <int K>
hipError_t launch_k(const TubeDesc* d_tubes, const int* d_order, int n, Bufs<float> b, FastBufs f,
                    int t0, int nsteps, int threads, int guard, const int* d_t0, hipStream_t st) {
    const size_t lds = fast_lds_bytes(K, threads, guard);
    hipError_t e = configure_step<K, false>();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_fast_step<K, false>), dim3(n), dim3(threads), lds, st, d_tubes, d_order, b, f, t0,
                       nsteps, d_t0, guard, (float*)nullptr, 0LL, (float*)nullptr);
    return hipGetLastError();
}

template <int K>
hipError_t launch_rec_k(const TubeDesc* d_tubes, const int* d_order, int n, Bufs<float> b, FastBufs f,
                        FastAdj a, int t0, int nsteps, int threads, int guard, hipStream_t st) {
    const size_t lds = fast_lds_bytes(K, threads, guard);
    hipError_t e = configure_step<K, true>();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_fast_step<K, true>), dim3(n), dim3(threads), lds, st, d_tubes, d_order, b, f, t0,
                       nsteps, (const int*)nullptr, guard, a.rec, a.rec_tube, a.sink);
    return hipGetLastError();
}

#ifndef VS_FAST_SPLIT_K
template <int K, int MS>
hipError_t configure_wide_km() {
    static std::atomic<int> done{0};
    if (done.load()) return hipSuccess;
    const int dyn = 160 * 1024 - (int)fast_wide_lds_static_bytes();
    const hipError_t e = hipFuncSetAttribute((const void*)k_fast_wide<K, MS>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             dyn);
    if (e == hipSuccess) done.store(1);
    return e;
}

template <int K, int MS>
hipError_t launch_wide_km(const TubeDesc* d_tubes, const int* d_order, int n, Bufs<float> b, FastBufs f, int t0,
                          int nsteps, int threads, int GW, int wq, const int* d_t0, hipStream_t st) {
    // a source group needs rows padded to a multiple of K
    if (MS != -2 && wq != K) return hipErrorInvalidValue;
    if (wq != K && wq != 4) return hipErrorInvalidValue;
    const size_t lds = fast_wide_lds_bytes(K, threads, GW);
    hipError_t e = configure_wide_km<K, MS>();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_fast_wide<K, MS>), dim3(n), dim3(threads), lds, st, d_tubes, d_order, b, f, t0, nsteps, d_t0,
                       GW, wq);
    return hipGetLastError();
}

// The instantiation for a source group ms in [-2, K / 4) chosen at run time.
template <int K, int MS = K / 4 - 1>
hipError_t configure_wide_k(int ms) {
    if (ms == MS) return configure_wide_km<K, MS>();
    if constexpr (MS > -2) return configure_wide_k<K, MS - 1>(ms);
    else return hipErrorInvalidValue;
}

template <int K, int MS = K / 4 - 1>
hipError_t launch_wide_k(int ms, const TubeDesc* d_tubes, const int* d_order, int n, Bufs<float> b, FastBufs f,
                         int t0, int nsteps, int threads, int GW, int wq, const int* d_t0, hipStream_t st) {
    if (ms == MS) return launch_wide_km<K, MS>(d_tubes, d_order, n, b, f, t0, nsteps, threads, GW, wq, d_t0, st);
    if constexpr (MS > -2)
        return launch_wide_k<K, MS - 1>(ms, d_tubes, d_order, n, b, f, t0, nsteps, threads, GW, wq, d_t0, st);
    else return hipErrorInvalidValue;
}
#endif

template <int K>
hipError_t launch_rev_k(const TubeDesc* d_tubes, const int* d_order, int n, FastBufs f, FastAdj a, int s0,
                        int ns, int guard, hipStream_t st) {
    const size_t lds = fast_lds_bytes(K, a.lanes + 64, guard);
    hipError_t e = configure_rev<K>();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fast_reverse<K>, dim3(n), dim3(a.lanes), lds, st, d_tubes, d_order, f, a, s0, ns, guard);
    return hipGetLastError();
}

}  // namespace

// K = kFastSplitK: the forward and recording step kernels of that K are
// built in their own translation unit (vs_fast_k17.hip) under another
// instruction scheduler.  Under the default one the forward kernel of this K
// alone spills; the other spills at K = 23.  Its reverse kernel stays here.
// The kernels exist for odd K up to kFastMaxK (fast_cells_per_lane never
// picks an even K).
constexpr int kFastSplitK = 17;
hipError_t split_configure_step(bool rec);
hipError_t split_launch_step(const TubeDesc* d_tubes, const int* d_order, int n, Bufs<float> b, FastBufs f,
                             int t0, int nsteps, int threads, int guard, const int* d_t0, hipStream_t st);
hipError_t split_launch_record(const TubeDesc* d_tubes, const int* d_order, int n, Bufs<float> b, FastBufs f,
                               FastAdj a, int t0, int nsteps, int threads, int guard, hipStream_t st);

#ifndef VS_FAST_SPLIT_K
size_t fast_lds_bytes(int K, int threads, int guard) {
    const int T = threads - 64;   // cell threads
    return (size_t)(T + 2 * guard) * (2 * K + 1) * sizeof(float);
}

// Odd K keeps the lane stride of the P and Uy regions odd, so the 64 lanes
// of a wave hit distinct LDS banks (an even stride shares banks: K = 24 ran
// 3.3 times slower than K = 23).
int fast_cells_per_lane(int max_padded_cells, int threads) {
    const int T = threads - 64;
    return ((max_padded_cells + T - 1) / T) | 1;
}

int fast_guard_lanes(int K, int max_padded_width) { return (max_padded_width + K - 1) / K + 1; }

size_t fast_wide_lds_bytes(int K, int threads, int guard_floats) {
    const int T = threads - 64;
    return (size_t)(2 * K * T + 2 * guard_floats + 8 + T) * sizeof(float);
}

size_t fast_wide_lds_static_bytes() {
    return sizeof(TwoMassState) + kFastPairs * (4 * sizeof(float) + 2 * sizeof(float)) + 2 * sizeof(double) + 64;
}

hipError_t configure_fast_wide(int K, int ms) {
    switch (K) {
    case 4: return configure_wide_k<4>(ms);
    case 12: return configure_wide_k<12>(ms);
    case 20: return configure_wide_k<20>(ms);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fast_wide(int K, int ms, const TubeDesc* d_tubes, const int* d_order, int n, Bufs<float> b,
                            FastBufs f, int t0, int nsteps, int threads, int guard_floats, int row_quantum,
                            const int* d_t0, hipStream_t st) {
    const int q = row_quantum;
    switch (K) {
    case 4: return launch_wide_k<4>(ms, d_tubes, d_order, n, b, f, t0, nsteps, threads, guard_floats, q, d_t0, st);
    case 12: return launch_wide_k<12>(ms, d_tubes, d_order, n, b, f, t0, nsteps, threads, guard_floats, q, d_t0, st);
    case 20: return launch_wide_k<20>(ms, d_tubes, d_order, n, b, f, t0, nsteps, threads, guard_floats, q, d_t0, st);
    default: return hipErrorInvalidValue;
    }
}

size_t fast_lds_static_bytes() {
    return sizeof(TwoMassState) + kFastClasses * (sizeof(float2) + sizeof(float)) + 2 * sizeof(double) + 64;
}

hipError_t launch_fast_setup(const TubeDesc* d_tubes, const TubeDesc* h_tubes, int n, Bufs<float> b,
                             FastBufs f, hipStream_t st) {
    int maxHW = 0;
    for (int k = 0; k < n; ++k) maxHW = max(maxHW, h_tubes[k].H * h_tubes[k].W);
    dim3 g((maxHW + 255) / 256, n);
    hipLaunchKernelGGL(k_fast_setup, g, dim3(256), 0, st, d_tubes, b, f);
    hipLaunchKernelGGL(k_vel_to_flux, g, dim3(256), 0, st, d_tubes, b, f);
    return hipGetLastError();
}

hipError_t configure_fast(int K) {
    switch (K) {
#define VS_K(k) case k: return configure_k<k>();
        VS_K(1) VS_K(3) VS_K(5) VS_K(7) VS_K(9) VS_K(11) VS_K(13) VS_K(15) VS_K(19) VS_K(21)
        VS_K(23)
#undef VS_K
    case kFastSplitK: {
        hipError_t e = split_configure_step(false);
        if (e == hipSuccess) e = split_configure_step(true);
        if (e == hipSuccess) e = configure_rev<kFastSplitK>();
        return e;
    }
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fast_step(int K, const TubeDesc* d_tubes, const int* d_order, int n, Bufs<float> b,
                            FastBufs f, int t0, int nsteps, int threads, int guard, const int* d_t0,
                            hipStream_t st) {
    switch (K) {
#define VS_K(k) case k: return launch_k<k>(d_tubes, d_order, n, b, f, t0, nsteps, threads, guard, d_t0, st);
        VS_K(1) VS_K(3) VS_K(5) VS_K(7) VS_K(9) VS_K(11) VS_K(13) VS_K(15) VS_K(19) VS_K(21)
        VS_K(23)
#undef VS_K
    case kFastSplitK:
        return split_launch_step(d_tubes, d_order, n, b, f, t0, nsteps, threads, guard, d_t0, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fast_record(int K, const TubeDesc* d_tubes, const int* d_order, int n, Bufs<float> b,
                              FastBufs f, FastAdj a, int t0, int nsteps, int threads, int guard, hipStream_t st) {
    if (threads - 64 != a.lanes) return hipErrorInvalidValue;
    switch (K) {
#define VS_K(k) case k: return launch_rec_k<k>(d_tubes, d_order, n, b, f, a, t0, nsteps, threads, guard, st);
        VS_K(1) VS_K(3) VS_K(5) VS_K(7) VS_K(9) VS_K(11) VS_K(13) VS_K(15) VS_K(19) VS_K(21)
        VS_K(23)
#undef VS_K
    case kFastSplitK:
        return split_launch_record(d_tubes, d_order, n, b, f, a, t0, nsteps, threads, guard, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fast_reverse(int K, const TubeDesc* d_tubes, const int* d_order, int n, FastBufs f,
                               FastAdj a, int s0, int ns, int guard, hipStream_t st) {
    switch (K) {
#define VS_K(k) case k: return launch_rev_k<k>(d_tubes, d_order, n, f, a, s0, ns, guard, st);
        VS_K(1) VS_K(3) VS_K(5) VS_K(7) VS_K(9) VS_K(11) VS_K(13) VS_K(15) VS_K(17) VS_K(19)
        VS_K(21) VS_K(23)
#undef VS_K
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fast_adj_sums(int K, const TubeDesc* d_tubes, const int* d_order, int n, FastBufs f, FastAdj a,
                                const float* source, int s0, int ns, int groups, hipStream_t st) {
    const int gx = std::max(1, groups / n);
    hipLaunchKernelGGL(k_fast_adj_sums, dim3(8 * ((n + 7) / 8) * gx), dim3(1024), 0, st, d_tubes, d_order, n, K, f, a,
                       source, s0, ns);
    return hipGetLastError();
}


#else   // VS_FAST_SPLIT_K: this unit provides the split K only

static_assert(VS_FAST_SPLIT_K == kFastSplitK, "vs_fast_k17.hip builds kFastSplitK");

hipError_t split_configure_step(bool rec) {
    return rec ? configure_step<VS_FAST_SPLIT_K, true>() : configure_step<VS_FAST_SPLIT_K, false>();
}

hipError_t split_launch_step(const TubeDesc* d_tubes, const int* d_order, int n, Bufs<float> b, FastBufs f,
                             int t0, int nsteps, int threads, int guard, const int* d_t0, hipStream_t st) {
    return launch_k<VS_FAST_SPLIT_K>(d_tubes, d_order, n, b, f, t0, nsteps, threads, guard, d_t0, st);
}

hipError_t split_launch_record(const TubeDesc* d_tubes, const int* d_order, int n, Bufs<float> b, FastBufs f,
                               FastAdj a, int t0, int nsteps, int threads, int guard, hipStream_t st) {
    return launch_rec_k<VS_FAST_SPLIT_K>(d_tubes, d_order, n, b, f, a, t0, nsteps, threads, guard, st);
}

#endif   // VS_FAST_SPLIT_K

}  // namespace vs
