/*
 * GlottalExcitations.h -- drop-in for the reference's precomputed source
 * table (include/GlottalExcitations.h): sine (type 0), Gaussian pulse
 * (type 1) and two-mass (type 2, table left at zero).  Throws
 * std::invalid_argument on an unknown type where the reference exits.
 */
#ifndef VOCALSIM_GLOTTALEXCITATIONS_H_
#define VOCALSIM_GLOTTALEXCITATIONS_H_

#include <vector>

class GlottalExcitations {
private:
    std::vector<float> host_excitation_velocity;

public:
    void setExcitationVelocity(int source_type, int num_steps, double dt);
    std::vector<float> getExcitationVelocity();
};

#endif
