/*
 * TwoMassModel.h -- drop-in for the reference's Ishizaka-Flanagan two-mass
 * vocal-fold model (include/TwoMassModel.h:122-126): set-up, supraglottal
 * pressure input, one step, volume velocity.  Host-side; the solver runs the
 * same recurrence on the GPU, one lane per tube.
 */
#ifndef VOCALSIM_TWOMASSMODEL_H_
#define VOCALSIM_TWOMASSMODEL_H_

class TwoMassModel {
private:
    float mu = 0, rho = 0, q_factor = 1, gs = 1;
    float m1 = 0, m2 = 0, d1 = 0, d2 = 0, k1 = 0, k2 = 0, kc = 0, etak1 = 0, etak2 = 0;
    float h1 = 0, h2 = 0, etah1 = 0, etah2 = 0, Ag0 = 0, Ag01 = 0, Ag02 = 0, lg = 0;
    float r1_open = 0, r2_open = 0, r1_close = 0, r2_close = 0, A1 = 0;
    float ps = 0, p1 = 0;
    float ug_old = 0, ug_curr = 0, ug_next = 0;
    float x1_old = 0, x1_curr = 0, x2_old = 0, x2_curr = 0;

public:
    void setTwoMassModelParams();
    /* Extension: subglottal pressure [Pa] and pitch factor q (reference: 686.46, 1). */
    void setTwoMassModelParams(float subglottal_pressure, float q);
    void runTwoMassModel(int srate_mul);
    void setSupraGlottalPressure(float pressure) { p1 = pressure; }
    float getVolumeVelocity() { return ug_next; }
};

#endif
