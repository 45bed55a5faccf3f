/*
 * TubeGeometry.h -- drop-in for the reference's TubeGeometry
 * (include/TubeGeometry.h:35-102).  The constructor reads the area-function
 * file the flags select (the data/tube-geometry text files, "<length cm> <area cm^2>"
 * per line, with the reference's file choice: vowel 1/2/3 -> /a/ /i/ /u/,
 * other vowel numbers -> cylindrical tube, tube_geometry == 0 -> no file).
 * The wall, excitation and mouth cells are built by the GPU setup kernels
 * when the solver is set up; generate*() are kept for source compatibility.
 */
#ifndef VOCALSIM_TUBEGEOMETRY_H_
#define VOCALSIM_TUBEGEOMETRY_H_

#include <string>
#include <vector>

#define CIRCULAR 1
#define ELLIPTICAL 2
#define SINGLE_PLANE_SYMMETRIC 3

class FDTDSolver;

class TubeGeometry {
public:
    int tube_start_posX = 0, tube_start_posY = 0, tube_end_posX = 0, tube_end_posY = 0;
    float mic_position = 0;
    int mic_position_cells = 0;
    int* upper_wall_pos = nullptr;   // device pointers, valid after setup
    int* lower_wall_pos = nullptr;

    explicit TubeGeometry(FDTDSolver& fdtdSolver);

    void setComputationalDomain();
    void generateVocalTractWalls();
    void generateExcitationCells();
    void generateMouthEndCells();

    int getFrameH();
    int getFrameW();
    int getVocalTractShape();
    float getSpatialResolution();
    int getTubeLengthCells();
    int* getBoundarySegmentType();
    std::vector<float> getMinorAxisLength();
    std::vector<float> getTubeSegmentLengthInCM();
    std::vector<float> getCrossSectionalAreaInCM2() const { return cross_sectional_area_incm2; }
    const std::string& fileName() const { return file_name; }

    /* Reads "<segment_len_cm> <area_cm2>" pairs (TubeGeometry.cu:83-88). */
    static bool readAreaFile(const std::string& path, std::vector<float>& len_cm, std::vector<float>& area_cm2);
    /* File the reference's flags select (TubeGeometry.cu:58-71), "" if none. */
    static std::string selectFile(int tube_geometry, int vowel_type);

private:
    std::vector<float> tube_segment_length_incm;
    std::vector<float> cross_sectional_area_incm2;
    std::string file_name;
    FDTDSolver& _fdtdSolver;
    void pull();
};

#endif
