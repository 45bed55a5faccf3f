/* VocalSim.h -- umbrella header of the C++ drop-in API. */
#ifndef VOCALSIM_VOCALSIM_H_
#define VOCALSIM_VOCALSIM_H_
#include "SimulationUnits.h"
#include "SimulationParams.h"
#include "SolverParams.h"
#include "GlottalExcitations.h"
#include "TwoMassModel.h"
#include "FDTDSolver.h"
#include "TubeGeometry.h"
#include "TubeDepth.h"
#include "FDTDEngineCPU.h"
#include "FDTDEngineEigen.h"
#endif
