/*
 * SolverParams.h -- drop-in for the reference's SolverParams and CellTypes
 * (include/SolverParams.h:21-56): per-type beta and sigma'dt tables, the
 * float-accumulated time vector and the source direction.
 */
#ifndef VOCALSIM_SOLVERPARAMS_H_
#define VOCALSIM_SOLVERPARAMS_H_

#include <vector>

#define NUM_CELL_TYPES 13

enum CellTypes {
    cell_wall, cell_air, cell_excitation, cell_pml0, cell_pml1, cell_pml2, cell_pml3, cell_pml4,
    cell_pml5, cell_dynamic, cell_dead, cell_noPressure, cell_head,
};

class SolverParams {
private:
    std::vector<float> time_step;
    float source_direction[4] = {0, 0, 0, 0};
    float beta[NUM_CELL_TYPES] = {};
    float sigma_prime_dt[NUM_CELL_TYPES] = {};

public:
    void setBeta(int num_pml_layers);
    void setSigmaPrimeDt(float max_sigma_dt, float dt, int num_pml_layers);
    void setTimeStep(float audio_dur, float dt);
    void setSourceDirection();
    float* getBeta();
    float* getSigmaPrimeDt();
    std::vector<float> getTimeStep();
    float* getSourceDirection();
};

#endif
