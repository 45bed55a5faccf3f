/*
 * TubeDepth.h -- drop-in for the reference's TubeDepth (include/TubeDepth.h):
 * the 2.5D depth maps are built by the GPU setup kernels (circle and ellipse
 * chords, open-space reset, half-cell smoothing, depthP average, 1 mm clamp);
 * this class exposes them with the reference's names.
 */
#ifndef VOCALSIM_TUBEDEPTH_H_
#define VOCALSIM_TUBEDEPTH_H_

#include <vector>

class TubeGeometry;

class TubeDepth {
public:
    float min_depth = 0.001f;
    float open_space_depth = 0.50f;

    TubeDepth(TubeGeometry& tubeGeometry, std::vector<float>& depthX, std::vector<float>& depthY,
              std::vector<float>& depthP);
    void setTubeDepth();
    void generateDepthP();
    void setMinDepth();

private:
    TubeGeometry& _tubeGeometry;
    std::vector<float>& depthX;
    std::vector<float>& depthY;
    std::vector<float>& depthP;
};

#endif
