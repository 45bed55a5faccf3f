/*
 * FDTDEngineCPU.h -- drop-in for the reference's time-step engine
 * (include/FDTDEngineCPU.h:16-126).  The class keeps the reference's name and
 * interface so existing drivers compile unchanged, but startSolverEngine()
 * advances the solver on the MI355X through the C ABI (vs_run); nothing falls
 * back to the CPU.  record_pressure holds the mouth-probe waveform; the
 * probe is written one value per line with 6 significant digits, as
 * savePressureSamples does (FDTDEngineCPU.cpp:435-450).
 */
#ifndef VOCALSIM_FDTDENGINECPU_H_
#define VOCALSIM_FDTDENGINECPU_H_

#include <string>
#include <vector>

#include "FDTDSolver.h"

class FDTDEngineCPU {
public:
    int frameH;
    int frameW;
    int listenerX;
    int listenerY;
    std::vector<float> record_pressure;
    std::vector<double> record_pressure_f64;   // extension: full-precision probe
    std::string output_path = "../../results/recorded-pressure.txt";
    std::string validation_dir = "../../data/";
    bool write_output = true;                  // extension: skip the text file
    double elapsed_seconds = 0;

    explicit FDTDEngineCPU(FDTDSolver& fdtdSolver, bool validationFlag = false);
    void startSolverEngine();

protected:
    FDTDSolver& _fdtdSolver;
    int engine_variant = 0;
    void savePressureSamples();
    void validateSolverCoefficients();
};

#endif
