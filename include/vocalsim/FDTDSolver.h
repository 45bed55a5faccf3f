/*
 * FDTDSolver.h -- drop-in for the reference's FDTDSolver
 * (include/FDTDSolver.h:35-177): same constructor, setupWaveSolver(), public
 * fields and coefficient getters.  The setup runs on the GPU through the C
 * ABI (vs_create); the getters return device pointers the solver owns, valid
 * until the solver is destroyed (the reference's engine frees them, this
 * build never transfers ownership).  Host-only header: no HIP or Thrust types;
 * the reference's thrust::device_vector accessors return host copies.
 */
#ifndef VOCALSIM_FDTDSOLVER_H_
#define VOCALSIM_FDTDSOLVER_H_

#include <string>
#include <vector>

#include "SimulationParams.h"
#include "SolverParams.h"

#define SIMULATION_2D 0
#define SIMULATION_2_5D 1

struct vs_handle;
class TubeGeometry;

class FDTDSolver {
public:
    int fdtd_solver_type;
    int simulation_type;
    int tube_geometry;
    int vowel_type;
    int cross_sectional_shape;
    int mouth_radiation_condition;
    float audio_dur;
    int source_type;
    int pml_flag;
    int srate_mul;

    float kappa = 0;
    float rho_sqrC_dt_inv_ds = 0;
    float rho_sqrC_dt = 0;
    float open_space_depth = 0.5f;
    int num_pml_layers = 6;

    float sound_speed = 0;
    float rho = 0;
    float wall_pressure_coupling_coeff = 0;
    float Mw = 0, Bw = 0, Kw = 0;
    float ds = 0;
    float dt = 0;

    int frameH = 0, frameW = 0, domainH = 0, domainW = 0;
    int listenerY = 0, listenerX = 0, excitationY = 0, excitationX = 0;

    float max_sigma_dt = 0;
    float* beta = nullptr;
    float* sigma_prime_dt = nullptr;
    float* source_direction = nullptr;

    std::vector<float> time_step;
    int num_steps = 0;

    struct GridCellComponents {
        float Pr;
        float Vx;
        float Vy;
        int cell_type;
    };

    explicit FDTDSolver(SimulationParams simulationParams);
    ~FDTDSolver();
    FDTDSolver(const FDTDSolver&) = delete;
    FDTDSolver& operator=(const FDTDSolver&) = delete;

    void setupWaveSolver();

    // Coefficient grids, (frameH-2) x (frameW-2), device float* (FP32 values).
    float* getN_out() const;
    float* getN_in() const;
    float* getMinVxBeta() const;
    float* getMinVyBeta() const;
    float* getMaxVxSigmaPrimedt() const;
    float* getMaxVySigmaPrimedt() const;
    float* getBetaVxSqr_dt_inv_rho_inv_ds() const;
    float* getBetaVySqr_dt_inv_rho_inv_ds() const;
    float* getPressureSigmaPrimedt() const;
    float* getAreWeNotExcitationVx() const;
    float* getAreWeNotExcitationVy() const;
    float* getExcitationWeightVx() const;
    float* getExcitationWeightVy() const;
    float* getXorVal1() const;
    float* getXorVal2() const;
    float* getXorVal3() const;
    float* getXorVal4() const;
    int* getBoundarySegmentType();
    float* getZ_inv() const;

    // For debugging purposes (FDTDSolver.cu:88-284): CSV dumps into debug_dir.
    // Pointers may be device pointers handed out by the getters or host memory.
    void printInMatrixFormat(float* data, std::string file_name, int num_rows, int num_cols);
    void printInMatrixFormat(int* data, std::string file_name, int num_rows, int num_cols);
    void printGridCellComponents();
    void printNeighboringCellsData(std::string file_name, float* data_cur, float* data_right, float* data_top);
    void printNeighboringCellsData(std::string file_name, int* data_cur, int* data_right, int* data_top);
    void printDepthComponents(float* data, std::string file_name);
    std::string debug_dir = "../../data/";
    // setupWaveSolver() writes cell_type_vals, beta_vals and sigma_prime_dt_vals
    // as the reference's calculateSolverCoefficients does (FDTDSolver.cu:689-712)
    bool write_setup_csv = true;

    // Host copies of the reference's device vectors.
    std::vector<float> getDepthP();
    std::vector<float> getDepthX();
    std::vector<float> getDepthY();
    std::vector<GridCellComponents> getPV_N();
    std::vector<float> getExcitationVelocity();

    // ---- extensions (inputs the reference hard-codes), set before setupWaveSolver ----
    void setPrecision(int bits);                 // 64 (default) or 32
    void setAreaFunction(const std::vector<float>& length_cm, const std::vector<float>& area_cm2);
    void setSectionAspect(const std::vector<float>& major_to_minor);
    void setSectionDepthScale(const std::vector<float>& scale);
    void setWallAdmittance(float mu3d);
    void setGaussianPulse(double amplitude, double t0, double width);
    void setTwoMassParams(double subglottal_pressure, double q);
    void setNumSteps(int steps);                 // overrides the float-accumulated count
    static void setDataDirectory(const std::string& dir);
    static std::string dataDirectory();

    vs_handle* handle() const { return h_; }
    int precision() const { return precision_; }

private:
    friend class TubeGeometry;
    friend class FDTDEngineCPU;
    friend class FDTDEngineEigen;
    vs_handle* h_ = nullptr;
    int precision_ = 64;
    std::vector<float> len_cm_, area_cm2_, aspect_, dscale_;
    float mu3d_ = 0;
    double g_amp_ = 0, g_t0_ = 0, g_w_ = 0, tm_ps_ = 0, tm_q_ = 0;
    int explicit_steps_ = 0;
    float beta_[NUM_CELL_TYPES] = {};
    float sigma_[NUM_CELL_TYPES] = {};
    float srcdir_[4] = {};
    mutable std::vector<float*> fptr_cache_;
    float* coeff_ptr(int k) const;
    void load_area_function();
    void write_setup_dumps();
};

#endif
