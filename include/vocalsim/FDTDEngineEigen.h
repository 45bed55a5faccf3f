/*
 * FDTDEngineEigen.h -- drop-in for the reference's Eigen engine
 * (include/FDTDEngineEigen.h:94-96): same maths with the pressure denominator
 * rounded as sigma*dP + dP (FDTDEngineEigen.cpp:251); runs on the GPU.
 */
#ifndef VOCALSIM_FDTDENGINEEIGEN_H_
#define VOCALSIM_FDTDENGINEEIGEN_H_

#include "FDTDEngineCPU.h"

class FDTDEngineEigen : public FDTDEngineCPU {
public:
    FDTDEngineEigen(FDTDEngineCPU& fdtdEngineCPU, FDTDSolver& fdtdSolver, bool validationFlag = false);
    void startSolverEngine();
};

#endif
