// solver.cpp -- FDTDSolver, TubeGeometry, TubeDepth and the engines of the
// C++ drop-in API, implemented over the C ABI (include/vocalsim_abi.h).
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "vocalsim_abi.h"
#include "vocalsim/FDTDEngineCPU.h"
#include "vocalsim/FDTDEngineEigen.h"
#include "vocalsim/FDTDSolver.h"
#include "vocalsim/GlottalExcitations.h"
#include "vocalsim/TubeDepth.h"
#include "vocalsim/TubeGeometry.h"

namespace {
std::string g_data_dir;

void check(int rc, const char* what) {
    if (rc != VS_OK) throw std::runtime_error(std::string(what) + ": " + vs_last_error());
}
}  // namespace

// ------------------------------------------------------------------ FDTDSolver
FDTDSolver::FDTDSolver(SimulationParams p)
    : fdtd_solver_type(p.getFDTDSolverType()), simulation_type(p.getSimulationType()),
      tube_geometry(p.getTubeGeometry()), vowel_type(p.getVowelType()),
      cross_sectional_shape(p.getCrossSectionalShape()), mouth_radiation_condition(p.getMouthRadiationCond()),
      audio_dur(p.getAudioDur()), source_type(p.getSourceType()), pml_flag(p.getPMLFlag()),
      srate_mul(p.getSrateMul()) {}

FDTDSolver::~FDTDSolver() {
    if (h_) vs_destroy(h_);
}

void FDTDSolver::setDataDirectory(const std::string& dir) { g_data_dir = dir; }

std::string FDTDSolver::dataDirectory() {
    if (!g_data_dir.empty()) return g_data_dir;
    if (const char* env = std::getenv("VOCALSIM_DATA_DIR")) return env;
    return "../../data/tube-geometry/";   // the reference's relative path (TubeGeometry.cu:73)
}

void FDTDSolver::setPrecision(int bits) {
    if (bits != 32 && bits != 64) throw std::invalid_argument("precision must be 32 or 64");
    precision_ = bits;
}
void FDTDSolver::setAreaFunction(const std::vector<float>& l, const std::vector<float>& a) {
    if (l.size() != a.size() || l.empty()) throw std::invalid_argument("area function needs matching, non-empty arrays");
    len_cm_ = l; area_cm2_ = a;
}
void FDTDSolver::setSectionAspect(const std::vector<float>& q) { aspect_ = q; }
void FDTDSolver::setSectionDepthScale(const std::vector<float>& s) { dscale_ = s; }
void FDTDSolver::setWallAdmittance(float mu) { mu3d_ = mu; }
void FDTDSolver::setGaussianPulse(double a, double t0, double w) { g_amp_ = a; g_t0_ = t0; g_w_ = w; }
void FDTDSolver::setTwoMassParams(double ps, double q) { tm_ps_ = ps; tm_q_ = q; }
void FDTDSolver::setNumSteps(int n) { explicit_steps_ = n; }

void FDTDSolver::load_area_function() {
    if (!len_cm_.empty()) return;
    const std::string name = TubeGeometry::selectFile(tube_geometry, vowel_type);
    if (name.empty() || !TubeGeometry::readAreaFile(dataDirectory() + name, len_cm_, area_cm2_))
        throw std::runtime_error("Error: Unable to open file " + name);
}

// FDTDSolver.cu:289-476: constants, tables, step count, then the geometry,
// depth, PML and coefficient setup (here: the GPU setup kernels of vs_create).
void FDTDSolver::setupWaveSolver() {
    SimulationParams sp(srate_mul);
    sound_speed = sp.sound_speed; rho = sp.rho; ds = sp.ds; dt = sp.dt;
    wall_pressure_coupling_coeff = sp.wall_pressure_coupling_coeff;
    Mw = sp.Mw; Bw = sp.Bw; Kw = sp.Kw;
    kappa = sp.kappa;
    max_sigma_dt = sp.max_sigma_dt;
    num_pml_layers = sp.pml_layers;
    rho_sqrC_dt_inv_ds = (kappa * sp.dt) / sp.ds;
    rho_sqrC_dt = kappa * sp.dt;
    SolverParams solver;
    solver.setBeta(num_pml_layers);
    solver.setSigmaPrimeDt(max_sigma_dt, sp.dt, num_pml_layers);
    solver.setSourceDirection();
    for (int k = 0; k < NUM_CELL_TYPES; ++k) { beta_[k] = solver.getBeta()[k]; sigma_[k] = solver.getSigmaPrimeDt()[k]; }
    for (int k = 0; k < 4; ++k) srcdir_[k] = solver.getSourceDirection()[k];
    beta = beta_; sigma_prime_dt = sigma_; source_direction = srcdir_;
    solver.setTimeStep(audio_dur, sp.dt);
    time_step = solver.getTimeStep();
    num_steps = explicit_steps_ > 0 ? explicit_steps_ : (int)time_step.size();
    // simulation_type 0: open-space domain (FDTDSolver.cu:479-550), no area function
    if (simulation_type != 0) load_area_function();

    vs_tube_spec s{};
    s.fdtd_solver_type = fdtd_solver_type;
    s.cross_sectional_shape = cross_sectional_shape;
    s.mouth_radiation_cond = mouth_radiation_condition;
    s.srate_mul = srate_mul;
    s.pml_flag = pml_flag;
    s.source_type = source_type;
    s.num_steps = num_steps;
    s.audio_dur = audio_dur;
    s.num_sections = (int)len_cm_.size();
    s.length_cm = len_cm_.data();
    s.area_cm2 = area_cm2_.data();
    s.aspect = aspect_.size() == len_cm_.size() ? aspect_.data() : nullptr;
    s.depth_scale = dscale_.size() == len_cm_.size() ? dscale_.data() : nullptr;
    s.mu3d = nullptr;
    s.mu3d_default = mu3d_;
    s.gauss_amp = g_amp_; s.gauss_t0 = g_t0_; s.gauss_width = g_w_;
    s.tm_ps = tm_ps_; s.tm_q = tm_q_;
    s.engine_variant = 0;
    s.open_space = simulation_type == 0 ? 1 : 0;
    if (h_) { vs_destroy(h_); h_ = nullptr; }
    fptr_cache_.clear();
    check(vs_create(&s, 1, precision_, &h_), "vs_create");
    vs_dims d{};
    check(vs_get_dims(h_, 0, &d), "vs_get_dims");
    frameH = d.frameH; frameW = d.frameW; domainH = d.domainH; domainW = d.domainW;
    listenerX = d.listenerX; listenerY = d.listenerY;
    excitationX = d.excitationX; excitationY = d.excitationY;
    std::cout << "Simulation frame height = " << frameH << std::endl;
    std::cout << "Simulation frame width = " << frameW << std::endl;
    if (write_setup_csv) write_setup_dumps();
}

float* FDTDSolver::coeff_ptr(int k) const {
    if (!h_) throw std::runtime_error("setupWaveSolver() has not run");
    if (fptr_cache_.empty()) fptr_cache_.assign(19, nullptr);
    if (!fptr_cache_[k]) {
        const int field = k < 17 ? VS_F_COEFF0 + k : VS_F_Z_INV;
        void* p = nullptr;
        check(vs_field_ptr(h_, 0, field | VS_F_AS_F32, &p, nullptr), "vs_field_ptr");
        fptr_cache_[k] = static_cast<float*>(p);
    }
    return fptr_cache_[k];
}

float* FDTDSolver::getN_out() const { return coeff_ptr(0); }
float* FDTDSolver::getN_in() const { return coeff_ptr(1); }
float* FDTDSolver::getMinVxBeta() const { return coeff_ptr(2); }
float* FDTDSolver::getMinVyBeta() const { return coeff_ptr(3); }
float* FDTDSolver::getMaxVxSigmaPrimedt() const { return coeff_ptr(4); }
float* FDTDSolver::getMaxVySigmaPrimedt() const { return coeff_ptr(5); }
float* FDTDSolver::getBetaVxSqr_dt_inv_rho_inv_ds() const { return coeff_ptr(6); }
float* FDTDSolver::getBetaVySqr_dt_inv_rho_inv_ds() const { return coeff_ptr(7); }
float* FDTDSolver::getPressureSigmaPrimedt() const { return coeff_ptr(8); }
float* FDTDSolver::getAreWeNotExcitationVx() const { return coeff_ptr(9); }
float* FDTDSolver::getAreWeNotExcitationVy() const { return coeff_ptr(10); }
float* FDTDSolver::getExcitationWeightVx() const { return coeff_ptr(11); }
float* FDTDSolver::getExcitationWeightVy() const { return coeff_ptr(12); }
float* FDTDSolver::getXorVal1() const { return coeff_ptr(13); }
float* FDTDSolver::getXorVal2() const { return coeff_ptr(14); }
float* FDTDSolver::getXorVal3() const { return coeff_ptr(15); }
float* FDTDSolver::getXorVal4() const { return coeff_ptr(16); }
float* FDTDSolver::getZ_inv() const { return coeff_ptr(17); }

int* FDTDSolver::getBoundarySegmentType() {
    if (!h_) throw std::runtime_error("setupWaveSolver() has not run");
    void* p = nullptr;
    check(vs_field_ptr(h_, 0, VS_F_BOUNDARY_SEG, &p, nullptr), "vs_field_ptr");
    return static_cast<int*>(p);
}

namespace {
std::vector<float> real_field(vs_handle* h, int precision, int field, size_t n) {
    std::vector<float> out(n);
    if (precision == 64) {
        std::vector<double> tmp(n);
        check(vs_get_field(h, 0, field, tmp.data(), n * sizeof(double)), "vs_get_field");
        for (size_t i = 0; i < n; ++i) out[i] = (float)tmp[i];
    } else {
        check(vs_get_field(h, 0, field, out.data(), n * sizeof(float)), "vs_get_field");
    }
    return out;
}
}  // namespace

std::vector<float> FDTDSolver::getDepthP() { return real_field(h_, precision_, VS_F_DEPTH_P, (size_t)frameH * frameW); }
std::vector<float> FDTDSolver::getDepthX() { return real_field(h_, precision_, VS_F_DEPTH_X, (size_t)frameH * frameW); }
std::vector<float> FDTDSolver::getDepthY() { return real_field(h_, precision_, VS_F_DEPTH_Y, (size_t)frameH * frameW); }
std::vector<float> FDTDSolver::getExcitationVelocity() { return real_field(h_, precision_, VS_F_SOURCE, (size_t)num_steps); }

std::vector<FDTDSolver::GridCellComponents> FDTDSolver::getPV_N() {
    const size_t n = (size_t)frameH * frameW;
    std::vector<float> P = real_field(h_, precision_, VS_F_PRESSURE, n);
    std::vector<float> Vx = real_field(h_, precision_, VS_F_VX, n);
    std::vector<float> Vy = real_field(h_, precision_, VS_F_VY, n);
    std::vector<int> t(n);
    check(vs_get_field(h_, 0, VS_F_CELL_TYPE, t.data(), n * sizeof(int)), "vs_get_field");
    std::vector<GridCellComponents> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = GridCellComponents{P[i], Vx[i], Vy[i], t[i]};
    return out;
}

// ------------------------------------------------------- debug CSV dumps
// FDTDSolver.cu:88-284.  Values use the default stream format, as there.
namespace {
template <typename T>
std::vector<T> host_copy(const T* p, size_t n) {
    std::vector<T> v(n);
    check(vs_copy_to_host(v.data(), p, n * sizeof(T)), "vs_copy_to_host");
    return v;
}
}  // namespace

void FDTDSolver::printInMatrixFormat(float* data, std::string file_name, int num_rows, int num_cols) {
    const std::vector<float> v = host_copy(data, (size_t)num_rows * num_cols);
    std::ofstream out(debug_dir + file_name + ".csv", std::ios::out | std::ios::trunc);
    if (!out) { std::cerr << "Failed to open " << file_name + ".csv" << " for writing." << std::endl; return; }
    for (int i = 0; i < num_rows; ++i) {
        for (int j = 0; j < num_cols; ++j) {
            out << v[(size_t)i * num_cols + j];
            if (j < frameW - 1) out << ",";   // the reference tests against frameW, not num_cols
        }
        out << "\n";
    }
}

void FDTDSolver::printInMatrixFormat(int* data, std::string file_name, int num_rows, int num_cols) {
    const std::vector<int> v = host_copy(data, (size_t)num_rows * num_cols);
    std::ofstream out(debug_dir + file_name + ".csv", std::ios::out | std::ios::trunc);
    if (!out) { std::cerr << "Failed to open " << file_name + ".csv" << " for writing." << std::endl; return; }
    for (int i = 0; i < num_rows; ++i) {
        for (int j = 0; j < num_cols; ++j) {
            out << v[(size_t)i * num_cols + j];
            if (j < frameW - 1) out << ",";
        }
        out << "\n";
    }
}

void FDTDSolver::printGridCellComponents() {
    const std::vector<GridCellComponents> pv = getPV_N();
    auto write = [&](auto get, const std::string& name) {
        std::ofstream out(debug_dir + name, std::ios::out | std::ios::trunc);
        if (!out) { std::cerr << "Failed to open " << debug_dir + name << "\n"; return; }
        for (int i = 0; i < frameH; ++i) {
            for (int j = 0; j < frameW; ++j) {
                out << get(pv[(size_t)i * frameW + j]);
                if (j + 1 < frameW) out << ',';
            }
            out << '\n';
        }
    };
    write([](const GridCellComponents& c) { return c.Pr; }, "Pressure.csv");
    write([](const GridCellComponents& c) { return c.Vx; }, "Vx.csv");
    write([](const GridCellComponents& c) { return c.Vy; }, "Vy.csv");
    write([](const GridCellComponents& c) { return c.cell_type; }, "cell_types.csv");
}

void FDTDSolver::printDepthComponents(float* data, std::string file_name) {
    const std::vector<float> v = host_copy(data, (size_t)frameH * frameW);
    std::ofstream out(debug_dir + file_name + ".csv", std::ios::out | std::ios::trunc);
    if (!out) { std::cerr << "Failed to open " << file_name + ".csv" << " for writing." << std::endl; return; }
    for (int i = 0; i < frameH; ++i) {
        for (int j = 0; j < frameW; ++j) {
            out << v[(size_t)i * frameW + j];
            if (j + 1 < frameW) out << ",";
        }
        out << "\n";
    }
}

namespace {
template <typename T>
void neighbor_csv(const std::string& path, const char* header, const std::vector<T>& a, const std::vector<T>& b,
                  const std::vector<T>& c) {
    std::ofstream out(path, std::ios::out | std::ios::trunc);
    if (!out) { std::cerr << "Failed to open " << path << " for writing." << std::endl; return; }
    out << header;
    for (size_t i = 0; i < a.size(); ++i) out << a[i] << "," << b[i] << "," << c[i] << "\n";
}
}  // namespace

void FDTDSolver::printNeighboringCellsData(std::string file_name, float* cur, float* right, float* top) {
    const size_t n = (size_t)(frameH - 2) * (frameW - 2);
    // the float overload's header carries the reference's spelling
    neighbor_csv(debug_dir + file_name + ".csv", "cuurent, right, top\n", host_copy(cur, n), host_copy(right, n),
                 host_copy(top, n));
}

void FDTDSolver::printNeighboringCellsData(std::string file_name, int* cur, int* right, int* top) {
    const size_t n = (size_t)(frameH - 2) * (frameW - 2);
    neighbor_csv(debug_dir + file_name + ".csv", "current, right, top\n", host_copy(cur, n), host_copy(right, n),
                 host_copy(top, n));
}

// The three dumps of calculateSolverCoefficients (FDTDSolver.cu:664-712):
// cell types, beta and sigma'dt of each interior cell and its right and top
// neighbours, interior cells in row-major order.
void FDTDSolver::write_setup_dumps() {
    const size_t HW = (size_t)frameH * frameW, n = (size_t)(frameH - 2) * (frameW - 2);
    std::vector<int> t(HW);
    check(vs_get_field(h_, 0, VS_F_CELL_TYPE, t.data(), HW * sizeof(int)), "vs_get_field");
    std::vector<int> tc(n), tr(n), tt(n);
    std::vector<float> bc(n), br(n), bt(n), sc(n), sr(n), st(n);
    for (size_t i = 0; i < n; ++i) {
        const size_t r = 1 + i / (frameW - 2), c = 1 + i % (frameW - 2);
        tc[i] = t[r * frameW + c];
        tr[i] = t[r * frameW + c + 1];
        tt[i] = t[(r - 1) * frameW + c];
        bc[i] = beta_[tc[i]]; br[i] = beta_[tr[i]]; bt[i] = beta_[tt[i]];
        sc[i] = sigma_[tc[i]]; sr[i] = sigma_[tr[i]]; st[i] = sigma_[tt[i]];
    }
    neighbor_csv(debug_dir + "cell_type_vals.csv", "current, right, top\n", tc, tr, tt);
    neighbor_csv(debug_dir + "beta_vals.csv", "cuurent, right, top\n", bc, br, bt);
    neighbor_csv(debug_dir + "sigma_prime_dt_vals.csv", "cuurent, right, top\n", sc, sr, st);
}

// ---------------------------------------------------------------- TubeGeometry
std::string TubeGeometry::selectFile(int tube_geometry, int vowel_type) {
    // TubeGeometry.cu:58-71: the trailing else binds to the innermost if, so a
    // vocal-tract run with an unknown vowel falls back to the cylinder and a
    // regular-tube run selects no file at all.
    if (!tube_geometry) return "";
    if (vowel_type == 1) return "vocaltract_a.txt";
    if (vowel_type == 2) return "vocaltract_i.txt";
    if (vowel_type == 3) return "vocaltract_u.txt";
    return "cylindrical-tube.txt";
}

bool TubeGeometry::readAreaFile(const std::string& path, std::vector<float>& len_cm, std::vector<float>& area_cm2) {
    std::ifstream f(path);
    if (!f.is_open()) return false;
    len_cm.clear(); area_cm2.clear();
    float l, a;
    while (f >> l >> a) { len_cm.push_back(l); area_cm2.push_back(a); }
    return !len_cm.empty();
}

TubeGeometry::TubeGeometry(FDTDSolver& s) : _fdtdSolver(s) {
    file_name = selectFile(s.tube_geometry, s.vowel_type);
    if (!s.len_cm_.empty()) {
        tube_segment_length_incm = s.len_cm_;
        cross_sectional_area_incm2 = s.area_cm2_;
    } else if (file_name.empty() || !readAreaFile(FDTDSolver::dataDirectory() + file_name, tube_segment_length_incm,
                                                    cross_sectional_area_incm2)) {
        std::cerr << "Error: Unable to open file " << file_name << std::endl;
    }
    mic_position = 3 * MILLIMETER;
    mic_position_cells = static_cast<int>(roundf(mic_position / SimulationParams(s.srate_mul).ds));
    if (s.h_) pull();
}

void TubeGeometry::pull() {
    vs_dims d{};
    check(vs_get_dims(_fdtdSolver.h_, 0, &d), "vs_get_dims");
    tube_start_posX = d.tube_start_x; tube_end_posX = d.tube_end_x;
    tube_start_posY = tube_end_posY = d.midY;
    mic_position_cells = d.mic_cells;
    void* p = nullptr;
    check(vs_field_ptr(_fdtdSolver.h_, 0, VS_F_UPPER_WALL, &p, nullptr), "vs_field_ptr");
    upper_wall_pos = static_cast<int*>(p);
    check(vs_field_ptr(_fdtdSolver.h_, 0, VS_F_LOWER_WALL, &p, nullptr), "vs_field_ptr");
    lower_wall_pos = static_cast<int*>(p);
}

void TubeGeometry::setComputationalDomain() {
    if (!_fdtdSolver.h_) _fdtdSolver.setupWaveSolver();
    pull();
}
void TubeGeometry::generateVocalTractWalls() { setComputationalDomain(); }
void TubeGeometry::generateExcitationCells() { setComputationalDomain(); }
void TubeGeometry::generateMouthEndCells() { setComputationalDomain(); }
int TubeGeometry::getFrameH() { return _fdtdSolver.frameH; }
int TubeGeometry::getFrameW() { return _fdtdSolver.frameW; }
int TubeGeometry::getVocalTractShape() { return _fdtdSolver.cross_sectional_shape; }
float TubeGeometry::getSpatialResolution() { return _fdtdSolver.ds; }
int TubeGeometry::getTubeLengthCells() {
    vs_dims d{};
    check(vs_get_dims(_fdtdSolver.h_, 0, &d), "vs_get_dims");
    return d.tube_len_cells;
}
int* TubeGeometry::getBoundarySegmentType() { return _fdtdSolver.getBoundarySegmentType(); }
std::vector<float> TubeGeometry::getTubeSegmentLengthInCM() { return tube_segment_length_incm; }
std::vector<float> TubeGeometry::getMinorAxisLength() {
    // ellipse minor axis per section (TubeGeometry.cu:353-359), ratio 3:1 unless set
    std::vector<float> out;
    const std::vector<float>& q = _fdtdSolver.aspect_;
    for (size_t i = 0; i < cross_sectional_area_incm2.size(); ++i) {
        const float a = cross_sectional_area_incm2[i] * CENTIMETER * CENTIMETER;
        const float major = q.size() == cross_sectional_area_incm2.size() ? q[i] : 3.0f;
        const float lx = sqrtf(a / (major * 1.0f * M_PI));
        out.push_back(lx * (1.0f * 2));
    }
    return out;
}

// ------------------------------------------------------------------- TubeDepth
TubeDepth::TubeDepth(TubeGeometry& g, std::vector<float>& x, std::vector<float>& y, std::vector<float>& p)
    : _tubeGeometry(g), depthX(x), depthY(y), depthP(p) {}
void TubeDepth::setTubeDepth() { _tubeGeometry.setComputationalDomain(); }
void TubeDepth::generateDepthP() {}
void TubeDepth::setMinDepth() {}

// --------------------------------------------------------------------- engines
FDTDEngineCPU::FDTDEngineCPU(FDTDSolver& s, bool validationFlag)
    : frameH(s.frameH), frameW(s.frameW), listenerX(s.listenerX), listenerY(s.listenerY), _fdtdSolver(s) {
    if (!s.h_) throw std::runtime_error("setupWaveSolver() has not run");
    if (validationFlag) validateSolverCoefficients();
}

// FDTDEngineCPU.cpp:222-304: the time loop, here one vs_run on the GPU.
void FDTDEngineCPU::startSolverEngine() {
    vs_handle* h = _fdtdSolver.h_;
    check(vs_set_option(h, VS_OPT_ENGINE_VARIANT, engine_variant), "vs_set_option");
    int done = 0;
    check(vs_steps_done(h, &done), "vs_steps_done");
    if (done) check(vs_reset(h), "vs_reset");
    const auto t0 = std::chrono::high_resolution_clock::now();
    check(vs_run(h, _fdtdSolver.num_steps), "vs_run");
    check(vs_sync(h), "vs_sync");
    const auto t1 = std::chrono::high_resolution_clock::now();
    elapsed_seconds = std::chrono::duration<double>(t1 - t0).count();
    std::cout << "======================================" << std::endl;
    std::cout << "Elapsed time [in seconds]: " << elapsed_seconds << " s" << std::endl;
    std::cout << "======================================" << std::endl;
    record_pressure_f64.assign(_fdtdSolver.num_steps, 0.0);
    check(vs_get_probe(h, 0, record_pressure_f64.data(), record_pressure_f64.size()), "vs_get_probe");
    record_pressure.assign(record_pressure_f64.begin(), record_pressure_f64.end());
    if (write_output) savePressureSamples();
}

// FDTDEngineCPU.cpp:435-450: one value per line, default stream format.
void FDTDEngineCPU::savePressureSamples() {
    std::ofstream out(output_path);
    if (!out) throw std::runtime_error("Error: could not open " + output_path + " for writing");
    for (size_t i = 0; i < record_pressure.size(); ++i) out << record_pressure[i] << "\n";
}

// FDTDEngineCPU.cpp:455-497: coefficient grids as CSV, one row per line,
// every value followed by ','.  z_inv is written over the same interior-sized
// window as the other grids, as the reference does.
void FDTDEngineCPU::validateSolverCoefficients() {
    static const char* names[17] = {"N_out", "N_in", "minVxBeta", "minVyBeta", "maxVxSigmaPrimedt",
                                    "maxVySigmaPrimedt", "betaVxSqr_dt_inv_rho_inv_ds",
                                    "betaVySqr_dt_inv_rho_inv_ds", "pressureSigmaPrimedt", "areWeNotExcitationVx",
                                    "areWeNotExcitationVy", "excitationWeightVx", "excitationWeightVy", "xor_val1",
                                    "xor_val2", "xor_val3", "xor_val4"};
    const int R = frameH - 2, C = frameW - 2;
    auto dump = [&](const std::vector<float>& v, int pitch, const std::string& name) {
        std::ofstream f(validation_dir + name + ".csv");
        for (int r = 0; r < R; ++r) {
            for (int c = 0; c < C; ++c) f << v[(size_t)r * pitch + c] << ',';
            f << std::endl;
        }
    };
    for (int k = 0; k < 17; ++k)
        dump(real_field(_fdtdSolver.h_, _fdtdSolver.precision_, VS_F_COEFF0 + k, (size_t)R * C), C, names[k]);
    dump(real_field(_fdtdSolver.h_, _fdtdSolver.precision_, VS_F_Z_INV, (size_t)frameH * frameW), frameW, "z_inv");
}

FDTDEngineEigen::FDTDEngineEigen(FDTDEngineCPU& cpu, FDTDSolver& s, bool validationFlag)
    : FDTDEngineCPU(s, validationFlag) {
    output_path = cpu.output_path;
    write_output = cpu.write_output;
    engine_variant = 1;
}

void FDTDEngineEigen::startSolverEngine() { FDTDEngineCPU::startSolverEngine(); }
