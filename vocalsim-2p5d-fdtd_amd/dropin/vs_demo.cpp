// vs_demo.cpp -- a driver written against the C++ drop-in API exactly as the
// reference's main() drives its classes (src/main.cu:65-146): flags ->
// SimulationParams -> FDTDSolver::setupWaveSolver -> FDTDEngineCPU (or
// FDTDEngineEigen) -> startSolverEngine.  Flags come from the command line:
//   vs_demo solver_type simulation_type tube_geometry vowel cross_section
//           mouth srate_mul audio_dur source pml engine(0 cpu / 1 eigen)
//           precision(32/64) num_steps(0 = from audio_dur) output_path
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "vocalsim/VocalSim.h"

int main(int argc, char** argv) {
    int a[10] = {1, 1, 1, 1, 1, 1, 15, 0, 1, 1};
    float dur = 0.02f;
    int engine = 0, precision = 64, steps = 0;
    std::string out = "recorded-pressure.txt";
    for (int i = 1; i < argc && i <= 13; ++i) {
        if (i == 8) dur = std::atof(argv[i]);
        else if (i <= 10) a[i - 1] = std::atoi(argv[i]);
        else if (i == 11) engine = std::atoi(argv[i]);
        else if (i == 12) precision = std::atoi(argv[i]);
        else if (i == 13) steps = std::atoi(argv[i]);
    }
    if (argc > 14) out = argv[14];
    try {
        SimulationParams params(a[0], a[1], a[2], a[3], a[4], a[5], a[6], dur, a[8], a[9]);
        FDTDSolver solver(params);
        solver.setPrecision(precision);
        if (steps > 0) solver.setNumSteps(steps);
        solver.setupWaveSolver();
        FDTDEngineCPU cpu(solver, false);
        cpu.output_path = out;
        if (engine == 1) {
            FDTDEngineEigen eigen(cpu, solver, false);
            eigen.startSolverEngine();
        } else {
            cpu.startSolverEngine();
        }
        std::printf("frame %d x %d, listener (%d, %d), steps %d\n", solver.frameH, solver.frameW, solver.listenerY,
                    solver.listenerX, solver.num_steps);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "vs_demo: %s\n", e.what());
        return 1;
    }
    return 0;
}
