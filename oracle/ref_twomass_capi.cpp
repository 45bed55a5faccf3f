// C binding of the reference's own two-mass vocal-fold model, for the tests
// only.  `make -C oracle ref` compiles it together with the reference source
// src/TwoMassModel.cpp where it lies (include/TwoMassModel.h and
// include/SimulationUnits.h are its only headers) into oracle/_ref/.  Nothing
// of the reference is copied here: this file only drives the class the way
// FDTDEngineCPU.cpp:246-267 does.
#include "TwoMassModel.h"

extern "C" int ref_twomass_run(int srate_mul, const float* p1, int n, float* ug) {
    if (n < 0 || (n > 0 && (!p1 || !ug))) return -1;
    TwoMassModel model;
    model.setTwoMassModelParams();                // FDTDEngineCPU.cpp:247-248
    for (int i = 0; i < n; ++i) {
        model.setSupraGlottalPressure(p1[i]);     // FDTDEngineCPU.cpp:262-263
        model.runTwoMassModel(srate_mul);         // :266
        ug[i] = model.getVolumeVelocity();        // :267
    }
    return 0;
}
