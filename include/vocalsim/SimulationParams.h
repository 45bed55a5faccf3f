/*
 * SimulationParams.h -- drop-in for the reference's SimulationParams
 * (include/SimulationParams.h:26-98): the physics constants of
 * SimulationParams(int srate_mul) and the ten user flags of the second
 * constructor, with the same names, types and getters.
 */
#ifndef VOCALSIM_SIMULATIONPARAMS_H_
#define VOCALSIM_SIMULATIONPARAMS_H_

#include "SimulationUnits.h"

class SimulationParams {
private:
    int fdtd_solver_type = 1;
    int simulation_type = 1;
    int tube_geometry = 1;
    int vowel_type = 1;
    int cross_sectional_shape = 1;
    int mouth_radiation_cond = 1;
    int srate_mul = 15;
    float audio_dur = 0.1f;
    int source_type = 1;
    int pml_flag = 0;

public:
    float sound_speed = 0;
    float rho = 0;
    int base_sample_srate = 0;
    float kappa = 0;
    int srate = 0;
    float max_sigma_dt = 0;
    int pml_layers = 0;
    float wall_pressure_coupling_coeff = 0;
    float M0 = 0, B0 = 0, K0 = 0;
    float Mw = 0, Bw = 0, Kw = 0;
    float ds = 0;
    float dt = 0;

    explicit SimulationParams(int srate_mul);
    SimulationParams(const int fdtd_solver_type, const int simulation_type, int tube_geometry, int vowel_type,
                     int cross_sectional_shape, const int mouth_radiation_cond, const int srate_mul,
                     const float audio_dur, const int source_type, const int pml_flag);

    int getFDTDSolverType();
    int getSimulationType();
    int getTubeGeometry();
    int getVowelType();
    int getCrossSectionalShape();
    int getMouthRadiationCond();
    int getSrateMul();
    float getAudioDur();
    int getSourceType();
    int getPMLFlag();
};

#endif
