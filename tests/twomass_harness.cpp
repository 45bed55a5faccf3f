// twomass_harness.cpp -- the library's two-mass recurrence
// (vocalsim-2p5d-fdtd_amd/csrc/vs_twomass.h) built for the CPU, run in the
// two orders the library uses: one twomass_step per step (the exact step
// kernels), and the order of the on-chip kernels' control wave, where the
// flow terms and the flow-independent part of the fold solve run a phase
// ahead (twomass_flow, twomass_pre) and the rest after the pressure input
// arrives (twomass_finish, twomass_post), restarted with a whole
// twomass_step at the start of every launch of `launch` steps.

#include "../vocalsim-2p5d-fdtd_amd/csrc/vs_twomass.h"

extern "C" {

// Glottal flow ug[i] of n steps, p1[i] being the pressure input of step i.
// schedule 0: twomass_step per step; 1: the control wave's order.
int tm_run(double q, double ps, int srate_mul, int schedule, int launch, const double* p1, int n, double* ug,
           double* x1) {
    if (n < 0 || launch < 1 || (schedule != 0 && schedule != 1)) return -1;
    vs::TwoMassState tm;
    vs::twomass_init(tm, q, ps, srate_mul);
    if (schedule == 0) {
        for (int i = 0; i < n; ++i) {
            tm.p1 = p1[i];
            ug[i] = vs::twomass_step(tm);
            x1[i] = tm.x1_curr;
        }
        return 0;
    }
    for (int t0 = 0; t0 < n; t0 += launch) {
        const int tend = t0 + launch < n ? t0 + launch : n;
        vs::TwoMassNext nx;
        vs::TwoMassPre pp;
        tm.p1 = p1[t0];
        ug[t0] = vs::twomass_step(tm);
        x1[t0] = tm.x1_curr;
        if (t0 + 1 < tend) vs::twomass_solve(tm, nx);
        for (int ts = t0; ts < tend; ++ts) {
            if (ts + 1 < tend) {
                vs::twomass_flow(tm, nx);
                if (ts + 2 < tend) vs::twomass_pre(tm, nx.x1n, nx.x2n, tm.x1_curr, tm.x2_curr, pp);
            }
            // the pressure phase ends here: the input of step ts + 1 is known
            if (ts + 1 < tend) {
                ug[ts + 1] = vs::twomass_finish(tm, nx, p1[ts + 1]);
                x1[ts + 1] = tm.x1_curr;
                if (ts + 2 < tend) vs::twomass_post(tm, pp, nx);
            }
        }
    }
    return 0;
}

}  // extern "C"
