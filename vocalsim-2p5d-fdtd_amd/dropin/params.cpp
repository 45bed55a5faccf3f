// params.cpp -- SimulationParams, SolverParams, GlottalExcitations and
// TwoMassModel of the C++ drop-in API.  Float arithmetic as the reference's
// host code evaluates it.
#include <cmath>
#include <stdexcept>
#include <vector>

#include "vocalsim/GlottalExcitations.h"
#include "vocalsim/SimulationParams.h"
#include "vocalsim/SolverParams.h"
#include "vocalsim/TwoMassModel.h"

// SimulationParams.cu:7-33
SimulationParams::SimulationParams(int srate_mul_) {
    sound_speed = 350.0f * (METER / SECOND);
    rho = (1.14f * KILOGRAM) / std::pow(METER, 3);
    base_sample_srate = 44100 * HERTZ;
    kappa = rho * sound_speed * sound_speed;
    srate = base_sample_srate * srate_mul_;
    dt = 1.0f / static_cast<float>(srate);
    ds = dt * sound_speed * sqrtf(2.0f);
    max_sigma_dt = 0.5;
    pml_layers = 6;
    wall_pressure_coupling_coeff = 0.0f;
    M0 = 21 * (KILOGRAM / (METER * METER));
    B0 = 8000 * (KILOGRAM / (METER * METER * SECOND));
    K0 = 845000 * (KILOGRAM / (METER * METER * SECOND * SECOND));
    Mw = M0; Bw = B0; Kw = K0;
    srate_mul = srate_mul_;
}

// SimulationParams.cu:36-57
SimulationParams::SimulationParams(const int fdtd_solver_type_, const int simulation_type_, int tube_geometry_,
                                   int vowel_type_, int cross_sectional_shape_, const int mouth_radiation_cond_,
                                   const int srate_mul_, const float audio_dur_, const int source_type_,
                                   const int pml_flag_)
    : fdtd_solver_type(fdtd_solver_type_), simulation_type(simulation_type_), tube_geometry(tube_geometry_),
      vowel_type(vowel_type_), cross_sectional_shape(cross_sectional_shape_),
      mouth_radiation_cond(mouth_radiation_cond_), srate_mul(srate_mul_), audio_dur(audio_dur_),
      source_type(source_type_), pml_flag(pml_flag_) {}

int SimulationParams::getFDTDSolverType() { return fdtd_solver_type; }
int SimulationParams::getSimulationType() { return simulation_type; }
int SimulationParams::getTubeGeometry() { return tube_geometry; }
int SimulationParams::getVowelType() { return vowel_type; }
int SimulationParams::getCrossSectionalShape() { return cross_sectional_shape; }
int SimulationParams::getMouthRadiationCond() { return mouth_radiation_cond; }
int SimulationParams::getSrateMul() { return srate_mul; }
float SimulationParams::getAudioDur() { return audio_dur; }
int SimulationParams::getSourceType() { return source_type; }
int SimulationParams::getPMLFlag() { return pml_flag; }

// SolverParams.cu:5-76
void SolverParams::setBeta(int n_pml) {
    for (float& b : beta) b = 0;
    beta[cell_air] = 1;
    beta[cell_noPressure] = 1;
    for (int k = 0; k < n_pml; ++k) beta[cell_pml0 + k] = 1;
}

void SolverParams::setSigmaPrimeDt(float max_sigma_dt, float dt, int n_pml) {
    for (float& s : sigma_prime_dt) s = 0;
    sigma_prime_dt[cell_wall] = dt;
    sigma_prime_dt[cell_excitation] = dt;
    sigma_prime_dt[cell_head] = dt;
    sigma_prime_dt[cell_dead] = 1000000;
    for (int k = 0; k < n_pml; ++k)
        sigma_prime_dt[cell_pml0 + k] = ((float)k / (float)(n_pml - 1)) * max_sigma_dt;
}

void SolverParams::setTimeStep(float audio_time, float dt) {
    time_step.clear();
    for (float t = 0; t <= audio_time - dt; t = t + dt) time_step.push_back(t);
}

void SolverParams::setSourceDirection() {
    source_direction[0] = 0; source_direction[1] = 0; source_direction[2] = 1; source_direction[3] = 0;
}

float* SolverParams::getBeta() { return beta; }
float* SolverParams::getSigmaPrimeDt() { return sigma_prime_dt; }
std::vector<float> SolverParams::getTimeStep() { return time_step; }
float* SolverParams::getSourceDirection() { return source_direction; }

// GlottalExcitations.cu:9-52
void GlottalExcitations::setExcitationVelocity(int source_type, int num_steps, double dt) {
    host_excitation_velocity.assign(num_steps, 0.0f);
    if (source_type == 0) {
        const float f = 440 * HERTZ, amp = 25;
        for (int i = 0; i < num_steps; ++i) {
            const float t = i * dt;
            host_excitation_velocity[i] = amp * std::sin(2 * M_PI * f * t);
        }
    } else if (source_type == 1) {
        const float f0 = 10.0f * KILOHERTZ;
        const float peak = 0.646f / f0;
        const float width = 0.29f * peak;
        const float amp = 0.025f;
        for (int i = 0; i < num_steps; ++i) {
            const float t = i * dt;
            const float e = (t - peak) / width;
            host_excitation_velocity[i] = amp * expf(-1 * (e * e));
        }
    } else if (source_type != 2) {
        throw std::invalid_argument("Invalid source type");
    }
}

std::vector<float> GlottalExcitations::getExcitationVelocity() { return host_excitation_velocity; }

// TwoMassModel.cpp:4-87
void TwoMassModel::setTwoMassModelParams() { setTwoMassModelParams(686.46f, 1.0f); }

void TwoMassModel::setTwoMassModelParams(float subglottal_pressure, float q) {
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
    const float z1o = 0.2f, z2o = 0.6f, z1c = 1.1f, z2c = 1.9f;
    r1_open = (2 * z1o * sqrtf(k1 * m1)) / (gs * gs);
    r2_open = (2 * z2o * sqrtf(k2 * m2)) / (gs * gs);
    r1_close = (2 * z1c * sqrtf(k1 * m1)) / (gs * gs);
    r2_close = (2 * z2c * sqrtf(k2 * m2)) / (gs * gs);
    A1 = 3.0f * (CENTIMETER * CENTIMETER);
    ps = subglottal_pressure;
    p1 = 0.0f;
    ug_old = ug_curr = ug_next = 0.0f;
    x1_old = x1_curr = x2_old = x2_curr = 0.0f;
}

// TwoMassModel.cpp:90-231: glottal areas -> forces -> 2x2 solve -> flow.
void TwoMassModel::runTwoMassModel(int srate_mul) {
    const float sr = 44100 * static_cast<float>(srate_mul);
    float Ag1 = Ag01 + 2 * lg * x1_curr;
    float Ag2 = Ag02 + 2 * lg * x2_curr;
    const float Rv1 = (12 * mu * lg * lg * d1) / (Ag1 * Ag1 * Ag1);
    const float Lg1 = (rho * d1) / (Ag1);
    const float Rv2 = (12 * mu * lg * lg * d2) / (Ag2 * Ag2 * Ag2);
    const float Lg2 = (rho * d2) / (Ag2);
    const bool open1 = x1_curr > -Ag01 / (2 * lg), open2 = x2_curr > -Ag02 / (2 * lg);
    float pm1 = ps;
    if (open1 && open2)
        pm1 = ps - 1.37 * (rho / 2) * (ug_curr / Ag1) * (ug_curr / Ag1) - 0.5 * (Rv1 * ug_curr + Lg1 * sr * (ug_curr - ug_old));
    const float f1 = pm1 * d1 * lg;
    float pm2 = 0;
    if (open1)
        pm2 = open2 ? (float)(pm1 - (0.5 * (Rv1 + Rv2) * ug_curr + (Lg1 + Lg2) * (ug_curr - ug_old) * sr)
                              - (rho / 2 * ug_curr * ug_curr) * (1 / (Ag2 * Ag2) - 1 / (Ag1 * Ag1)))
                    : ps;
    const float f2 = pm2 * d2 * lg;
    const bool closed1 = x1_curr + Ag01 / (2 * lg) < 0, closed2 = x2_curr + Ag02 / (2 * lg) < 0;
    const float h1u = closed1 ? h1 : 0.0f, r1u = closed1 ? r1_close : r1_open;
    const float h2u = closed2 ? h2 : 0.0f, r2u = closed2 ? r2_close : r2_open;
    const float a11 = (k1 + h1u + kc) / (sr * sr) + r1u / sr + m1;
    const float a12 = -kc / (sr * sr), a21 = a12;
    const float a22 = (k2 + h2u + kc) / (sr * sr) + r2u / sr + m2;
    const float g1 = Ag01 / (2 * lg), g2 = Ag02 / (2 * lg);
    const float s1p = k1 * etak1 * x1_curr * x1_curr * x1_curr + h1u * (g1 + etah1 * (g1 + x1_curr) * (g1 + x1_curr) * (g1 + x1_curr));
    const float s2p = k2 * etak2 * x2_curr * x2_curr * x2_curr + h2u * (g2 + etah2 * (g2 + x2_curr) * (g2 + x2_curr) * (g2 + x2_curr));
    const float b1 = (2 * m1 + r1u / sr) * x1_curr - m1 * x1_old - s1p / (sr * sr) + f1 / (sr * sr);
    const float b2 = (2 * m2 + r2u / sr) * x2_curr - m2 * x2_old - s2p / (sr * sr) + f2 / (sr * sr);
    float det = a11 * a22 - a21 * a12;
    if (det == 0) det = 1;
    const float x1n = (a22 * b1 - a12 * b2) / det, x2n = (a11 * b2 - a21 * b1) / det;
    x1_old = x1_curr; x1_curr = x1n;
    x2_old = x2_curr; x2_curr = x2n;
    Ag1 = Ag01 + 2 * lg * x1_curr;
    Ag2 = Ag02 + 2 * lg * x2_curr;
    if (Ag1 < 0 || Ag2 < 0) {
        ug_next = 0;
    } else {
        const float Rtot = (rho / 2) * std::fabs(ug_curr) * ((0.37 / (Ag1 * Ag1)) + (1 - 2 * (Ag2 / A1) * (1 - Ag2 / A1)) / (Ag2 * Ag2))
                           + (12 * mu * lg * lg * (d1 / (Ag1 * Ag1 * Ag1) + d2 / (Ag2 * Ag2 * Ag2)));
        const float Ltot = rho * (d1 / Ag1 + d2 / Ag2);
        ug_next = ((ps - p1) / sr + Ltot * ug_curr) / (Rtot / sr + Ltot);
    }
    ug_old = ug_curr;
    ug_curr = ug_next;
}
