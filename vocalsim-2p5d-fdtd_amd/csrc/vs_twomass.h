// vs_twomass.h -- the two-mass vocal-fold model (Ishizaka-Flanagan,
// TwoMassModel.cpp:4-231) in double: its state, parameter set-up and step.
// Free of HIP: device code includes it through vs_common.h, which defines
// VS_HD as __host__ __device__; host-only builds (the CPU tests) include it
// directly.
#pragma once

#ifndef VS_HD
#define VS_HD
#endif

namespace vs {

// Two-mass vocal-fold state (Ishizaka-Flanagan, TwoMassModel.cpp:4-231),
// evaluated in double on the device: one lane per tube per step.
struct TwoMassState {
    double mu, rho, q, gs, m1, m2, d1, d2, k1, k2, kc, etak1, etak2, h1, h2,
        etah1, etah2, Ag01, Ag02, lg, r1_open, r2_open, r1_close, r2_close, A1,
        ps, p1, ug_old, ug_curr, ug_next, x1_old, x1_curr, x2_old, x2_curr;
    // Sub-expressions of the step that do not depend on the state, or only
    // on whether a fold is closed (suffix c) or open (suffix o), evaluated
    // once by twomass_init with the step's own operations: sr = 44100 *
    // srate_mul, sr2 = sr * sr, thr = -Ag0 / (2 lg), g = Ag0 / (2 lg),
    // a11, a22 and a12 of the 2x2 system, cb = 2 m + r / sr.
    double sr, sr2, thr1, thr2, g1, g2, a11c, a11o, a22c, a22o, a12, cb1c, cb1o, cb2c, cb2o;
};

// Parameter set-up (TwoMassModel.cpp:4-87).  The float sub-expressions of the
// reference (units are float32 constants) are kept as float.
inline void twomass_init(TwoMassState& s, double q, double ps, int srate_mul) {
#pragma clang fp contract(off)
    const float CM = 1e-2f, GRAM = 1e-3f, DYN = 1e-5f, KG = 1.0f, M = 1.0f, SEC = 1.0f;
    s.mu = 1.86e-4 * (DYN * SEC / (CM * CM));
    s.rho = 1.14 * (KG / (M * M * M));
    s.q = q;
    s.gs = 1.0f;
    s.m1 = (0.125f * GRAM) / s.q;
    s.m2 = (0.025f * GRAM) / s.q;
    s.d1 = (0.25f * CM) / s.q;
    s.d2 = (0.05f * CM) / s.q;
    s.k1 = 80000.0f * (DYN / CM) * s.q;
    s.k2 = 8000.0f * (DYN / CM) * s.q;
    s.kc = 25000.0f * (DYN / CM) * (s.q * s.q);
    s.etak1 = 100 / (CM * CM);
    s.etak2 = 100 / (CM * CM);
    s.h1 = 3 * s.k1;
    s.h2 = 3 * s.k2;
    s.etah1 = 500.0f / (CM * CM);
    s.etah2 = 500.0f / (CM * CM);
    s.Ag01 = 0.05f * (CM * CM);
    s.Ag02 = s.Ag01;
    s.lg = 1.40f * CM;
    const double z1o = 0.2f, z2o = 0.6f, z1c = 1.1f, z2c = 1.9f;
    s.r1_open = (2 * z1o * __builtin_sqrt(s.k1 * s.m1)) / (s.gs * s.gs);
    s.r2_open = (2 * z2o * __builtin_sqrt(s.k2 * s.m2)) / (s.gs * s.gs);
    s.r1_close = (2 * z1c * __builtin_sqrt(s.k1 * s.m1)) / (s.gs * s.gs);
    s.r2_close = (2 * z2c * __builtin_sqrt(s.k2 * s.m2)) / (s.gs * s.gs);
    s.A1 = 3.0f * (CM * CM);
    s.ps = ps;
    s.p1 = 0;
    s.ug_old = s.ug_curr = s.ug_next = 0;
    s.x1_old = s.x1_curr = s.x2_old = s.x2_curr = 0;
    // TwoMassModel.cpp:90-231, the terms of the step that are constants
    const double sr = 44100 * static_cast<double>(srate_mul), lg = s.lg;
    s.sr = sr;
    s.sr2 = sr * sr;
    s.thr1 = -s.Ag01 / (2 * lg);
    s.thr2 = -s.Ag02 / (2 * lg);
    s.g1 = s.Ag01 / (2 * lg);
    s.g2 = s.Ag02 / (2 * lg);
    const double h1c = s.h1, h1o = 0.0f, h2c = s.h2, h2o = 0.0f;
    s.a11c = (s.k1 + h1c + s.kc) / (sr * sr) + s.r1_close / sr + s.m1;
    s.a11o = (s.k1 + h1o + s.kc) / (sr * sr) + s.r1_open / sr + s.m1;
    s.a22c = (s.k2 + h2c + s.kc) / (sr * sr) + s.r2_close / sr + s.m2;
    s.a22o = (s.k2 + h2o + s.kc) / (sr * sr) + s.r2_open / sr + s.m2;
    s.a12 = -s.kc / (sr * sr);
    s.cb1c = 2 * s.m1 + s.r1_close / sr;
    s.cb1o = 2 * s.m1 + s.r1_open / sr;
    s.cb2c = 2 * s.m2 + s.r2_close / sr;
    s.cb2o = 2 * s.m2 + s.r2_open / sr;
}

// One step of the recurrence (TwoMassModel.cpp:90-231), in three parts so
// that a caller can compute the first two before the pressure input p1 is
// known: twomass_solve (the new fold positions), twomass_flow (flow
// resistance and inertance at those positions) and twomass_finish (the new
// glottal flow from p1, committing the state).  twomass_step runs the three
// in a row.  Every operation is rounded separately (no contraction), in the
// reference's order, so host and device agree bitwise.
struct TwoMassNext {
    double x1n, x2n, Rtot, Ltot;
    bool closed;
};

// The fold solve split by its dependence on the glottal flow.  twomass_pre
// evaluates what depends only on the fold positions of the step (x1, x2) and
// the step before (x1o, x2o): glottal areas, viscous resistances and
// inertances, the opening and closure tests, the spring terms, the 2x2 matrix
// with its determinant and the right-hand sides up to their pressure-force
// terms.  twomass_post finishes the solve from the state's flows (ug_curr,
// ug_old).  Every sub-expression is the same operation on the same operands
// as in the one-piece solve, so pre then post round identically; a caller can
// run pre a phase earlier, before the previous step's flow is known.
struct TwoMassPre {
    double Ag1, Rv1, Lg1sr, rvs, lgs, inv, a11, a12, a22, det, b1p, b2p;
    bool open1, open2;
};

VS_HD inline void twomass_pre(const TwoMassState& s, double x1, double x2, double x1o, double x2o, TwoMassPre& p) {
#pragma clang fp contract(off)
    const double sr = s.sr;
    const double lg = s.lg, rho = s.rho, mu = s.mu, d1 = s.d1, d2 = s.d2;
    const double Ag1 = s.Ag01 + 2 * lg * x1;
    const double Ag2 = s.Ag02 + 2 * lg * x2;
    const double Rv1 = (12 * mu * lg * lg * d1) / (Ag1 * Ag1 * Ag1);
    const double Lg1 = (rho * d1) / (Ag1);
    const double Rv2 = (12 * mu * lg * lg * d2) / (Ag2 * Ag2 * Ag2);
    const double Lg2 = (rho * d2) / (Ag2);
    p.Ag1 = Ag1;
    p.Rv1 = Rv1;
    p.Lg1sr = Lg1 * sr;
    p.rvs = 0.5 * (Rv1 + Rv2);
    p.lgs = Lg1 + Lg2;
    p.inv = 1 / (Ag2 * Ag2) - 1 / (Ag1 * Ag1);
    p.open1 = x1 > s.thr1;
    p.open2 = x2 > s.thr2;
    const bool closed1 = x1 + s.g1 < 0, closed2 = x2 + s.g2 < 0;
    const double h1u = closed1 ? s.h1 : 0.0f, h2u = closed2 ? s.h2 : 0.0f;
    const double a11 = closed1 ? s.a11c : s.a11o;
    const double a12 = s.a12;
    const double a21 = a12;
    const double a22 = closed2 ? s.a22c : s.a22o;
    const double g1 = s.g1, g2 = s.g2;
    const double s1p = s.k1 * s.etak1 * x1 * x1 * x1 + h1u * (g1 + s.etah1 * (g1 + x1) * (g1 + x1) * (g1 + x1));
    const double s2p = s.k2 * s.etak2 * x2 * x2 * x2 + h2u * (g2 + s.etah2 * (g2 + x2) * (g2 + x2) * (g2 + x2));
    p.b1p = (closed1 ? s.cb1c : s.cb1o) * x1 - s.m1 * x1o - s1p / s.sr2;
    p.b2p = (closed2 ? s.cb2c : s.cb2o) * x2 - s.m2 * x2o - s2p / s.sr2;
    double det = a11 * a22 - a21 * a12;
    if (det == 0) det = 1;
    p.a11 = a11;
    p.a12 = a12;
    p.a22 = a22;
    p.det = det;
}

VS_HD inline void twomass_post(const TwoMassState& s, const TwoMassPre& p, TwoMassNext& n) {
#pragma clang fp contract(off)
    const double lg = s.lg, rho = s.rho, d1 = s.d1, d2 = s.d2;
    const double ug = s.ug_curr, Ag1 = p.Ag1;
    double pm1;
    if (p.open1 && p.open2)
        pm1 = s.ps - 1.37 * (rho / 2) * (ug / Ag1) * (ug / Ag1) - 0.5 * (p.Rv1 * ug + p.Lg1sr * (ug - s.ug_old));
    else
        pm1 = s.ps;
    const double f1 = pm1 * d1 * lg;
    double pm2;
    if (p.open1) {
        if (p.open2)
            pm2 = pm1 - (p.rvs * ug + p.lgs * (ug - s.ug_old) * s.sr) - (rho / 2 * ug * ug) * p.inv;
        else
            pm2 = s.ps;
    } else {
        pm2 = 0;
    }
    const double f2 = pm2 * d2 * lg;
    const double b1 = p.b1p + f1 / s.sr2;
    const double b2 = p.b2p + f2 / s.sr2;
    const double a21 = p.a12;
    n.x1n = (p.a22 * b1 - p.a12 * b2) / p.det;
    n.x2n = (p.a11 * b2 - a21 * b1) / p.det;
}

VS_HD inline void twomass_solve(const TwoMassState& s, TwoMassNext& n) {
    TwoMassPre p;
    twomass_pre(s, s.x1_curr, s.x2_curr, s.x1_old, s.x2_old, p);
    twomass_post(s, p, n);
}

VS_HD inline void twomass_flow(const TwoMassState& s, TwoMassNext& n) {
#pragma clang fp contract(off)
    const double lg = s.lg, rho = s.rho, mu = s.mu, d1 = s.d1, d2 = s.d2;
    const double ug = s.ug_curr;
    const double Ag1 = s.Ag01 + 2 * lg * n.x1n;
    const double Ag2 = s.Ag02 + 2 * lg * n.x2n;
    n.closed = Ag1 < 0 || Ag2 < 0;
    n.Rtot = n.Ltot = 0;
    if (!n.closed) {
        n.Rtot = (rho / 2) * __builtin_fabs(ug) * ((0.37 / (Ag1 * Ag1)) + (1 - 2 * (Ag2 / s.A1) * (1 - Ag2 / s.A1)) / (Ag2 * Ag2))
                 + (12 * mu * lg * lg * (d1 / (Ag1 * Ag1 * Ag1) + d2 / (Ag2 * Ag2 * Ag2)));
        n.Ltot = rho * (d1 / Ag1 + d2 / Ag2);
    }
}

VS_HD inline double twomass_finish(TwoMassState& s, const TwoMassNext& n, double p1) {
#pragma clang fp contract(off)
    const double sr = s.sr;
    const double ug = s.ug_curr;
    s.p1 = p1;
    s.x1_old = s.x1_curr; s.x1_curr = n.x1n;
    s.x2_old = s.x2_curr; s.x2_curr = n.x2n;
    const double ugn = n.closed ? 0 : ((s.ps - s.p1) / sr + n.Ltot * ug) / (n.Rtot / sr + n.Ltot);
    s.ug_old = ug;
    s.ug_curr = ugn;
    s.ug_next = ugn;
    return ugn;
}

// One step with the pressure input already in s.p1.
VS_HD inline double twomass_step(TwoMassState& s) {
    TwoMassNext n;
    twomass_solve(s, n);
    twomass_flow(s, n);
    return twomass_finish(s, n, s.p1);
}

}  // namespace vs
