/*
 * TestTwoMassModel.h -- drop-in for the reference's two-mass test driver
 * (include/TestTwoMassModel.h:9-54, included by src/main.cu:15): one second
 * of the two-mass model at fs = 44100 * 15 with no supraglottal pressure
 * feedback, written as "<t in ms>\t<U_g / max U_g>" lines to
 * ../../results/norm_ug.txt, after which the program ends with EXIT_SUCCESS,
 * as the reference's function does.
 *
 * writeTwoMassNormUg() is the same run written to a path of the caller's
 * choice; it returns the number of samples written, or -1 when the file
 * cannot be opened, and leaves the program running.
 */
#ifndef VOCALSIM_TESTTWOMASSMODEL_H_
#define VOCALSIM_TESTTWOMASSMODEL_H_

#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "SolverParams.h"
#include "TwoMassModel.h"

inline int writeTwoMassNormUg(const char* path, int srate_mul = 15, float audio_time = 1.0f) {
    TwoMassModel model;
    model.setTwoMassModelParams();
    const float dt = 1.0f / static_cast<float>(44100 * srate_mul);
    SolverParams timing;
    timing.setTimeStep(audio_time, dt);
    const std::vector<float> t = timing.getTimeStep();
    std::vector<float> ug(t.size(), 0.0f);
    for (size_t i = 0; i < t.size(); ++i) {
        model.runTwoMassModel(srate_mul);
        ug[i] = model.getVolumeVelocity();
    }
    const float peak = ug.empty() ? 1.0f : *std::max_element(ug.begin(), ug.end());
    std::ofstream out(path);
    if (!out) {
        std::cerr << "ERROR: could not open output file.\n";
        return -1;
    }
    for (size_t i = 0; i < t.size(); ++i) out << t[i] * 1000.0f << '\t' << ug[i] / peak << '\n';
    return static_cast<int>(t.size());
}

inline void testTwoMassModel() {
    const int n = writeTwoMassNormUg("../../results/norm_ug.txt");
    if (n < 0) return;
    std::cout << "Completed writing " << n << " samples to norm_ug.txt\n";
    std::exit(EXIT_SUCCESS);
}

#endif
