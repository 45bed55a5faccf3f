/*
 * vs_oracle.h -- C API of the CPU oracle for the 2.5D FDTD vocal-tract solver.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain C++ restatement of the
 * reference's setup path and FDTDEngineCPU time step, written from the
 * reference source (citations in vs_oracle.cpp).  Only tests/, the smoke()
 * check and the cpu_baseline leg of bench.py may load it, and only as the
 * checker or the timed CPU baseline.  The product library never links it.
 */
#ifndef VS_ORACLE_H_
#define VS_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Tube description handed to the oracle.  Field meaning mirrors the
 * reference's SimulationParams flags plus the inputs the configs add. */
typedef struct vso_spec {
    int fdtd_solver_type;       /* 0: 2D, 1: 2.5D */
    int cross_sectional_shape;  /* 1: circular, 2: elliptical, 3: single-plane symmetric */
    int mouth_radiation_cond;   /* 1: Dirichlet mouth, 0: open */
    int srate_mul;              /* fs = 44100 * srate_mul */
    int pml_flag;               /* 0/1 */
    int source_type;            /* 0: sine, 1: Gaussian, 2: two-mass */
    int num_steps;              /* > 0: explicit; <= 0: reference float accumulation of audio_dur */
    float audio_dur;            /* seconds */
    int num_sections;
    const float* length_cm;     /* per-section length [cm] */
    const float* area_cm2;      /* per-section area [cm^2] */
    const float* aspect;        /* ellipse major:minor ratio per section; NULL -> 3 */
    const float* depth_scale;   /* per-section depth multiplier; NULL -> 1 */
    const float* mu3d;          /* per-section wall admittance; NULL -> mu3d_default */
    float mu3d_default;         /* <= 0 -> reference default (0.025 in 2.5D, 0.005 in 2D) */
    double gauss_amp;           /* <= 0 -> reference Gaussian (0.025, t0 = 64.6 us, w = 0.29 t0) */
    double gauss_t0;
    double gauss_width;
    double tm_ps;               /* two-mass subglottal pressure [Pa]; <= 0 -> 686.46 */
    double tm_q;                /* two-mass pitch factor; <= 0 -> 1 */
    int engine_variant;         /* 0: FDTDEngineCPU op order, 1: FDTDEngineEigen op order */
    int open_space;             /* 1: open-space domain (FDTDSolver.cu:479-550); sections ignored */
} vso_spec;

enum vso_dim {
    VSO_FRAME_H = 0, VSO_FRAME_W, VSO_DOMAIN_H, VSO_DOMAIN_W,
    VSO_LISTENER_X, VSO_LISTENER_Y, VSO_EXCITATION_X, VSO_EXCITATION_Y,
    VSO_NUM_STEPS, VSO_TUBE_START_X, VSO_TUBE_END_X, VSO_TUBE_LEN_CELLS,
    VSO_MID_Y, VSO_NUM_SECTIONS, VSO_MIC_CELLS, VSO_NUM_DIMS
};

enum vso_field {
    VSO_F_CELL_TYPE = 0,   /* H*W, cell type enum value */
    VSO_F_BOUNDARY_SEG,    /* H*W, boundary_segment_type */
    VSO_F_DEPTH_P,         /* H*W */
    VSO_F_DEPTH_X,         /* H*W */
    VSO_F_DEPTH_Y,         /* H*W */
    VSO_F_Z_INV,           /* H*W */
    VSO_F_SOURCE,          /* num_steps: precomputed excitation table */
    VSO_F_UPPER_WALL,      /* L */
    VSO_F_LOWER_WALL,      /* L */
    VSO_F_SEGMENT,         /* L: 0-based section index of each tube column */
    VSO_F_DIAM_CELLS,      /* num_sections */
    VSO_F_MINOR_AXIS,      /* num_sections: ellipse minor axis [m] */
    VSO_F_Z_SECTION,       /* num_sections: z_inv value per section */
    VSO_F_COEFF0 = 100     /* + k, k in [0,17): interior coefficient grid k, (H-2)*(W-2) */
};

/* Scalars (index into vso_scalars output). */
enum vso_scalar {
    VSO_S_DT = 0, VSO_S_DS, VSO_S_RHO, VSO_S_C, VSO_S_KAPPA,
    VSO_S_K1, VSO_S_K2, VSO_S_DS_F32, VSO_S_Z_DEFAULT, VSO_NUM_SCALARS
};

/* precision: 32 or 64.  Returns NULL on error (message via vso_last_error). */
void* vso_create(const vso_spec* spec, int precision);
void  vso_destroy(void* h);
const char* vso_last_error(void);
int   vso_dims(void* h, int* out /* VSO_NUM_DIMS ints */);
int   vso_scalars(void* h, double* out /* VSO_NUM_SCALARS */);
int   vso_get_field(void* h, int field, double* dst, long n);

/* Run n_steps from the current state, appending probe samples (and the
 * injected source velocity per step) at offset first_step.  Returns 0. */
int   vso_run(void* h, int first_step, int n_steps, double* probe, double* src_trace);
/* Reset state (P, V, two-mass) to zero. */
int   vso_reset(void* h);
/* Snapshot of the state: P, Vx, Vy each H*W. */
int   vso_get_state(void* h, double* P, double* Vx, double* Vy);
int   vso_set_state(void* h, const double* P, const double* Vx, const double* Vy);

/* Reference-structured engine (vector<vector<cell>> layout, per-step deep
 * copy, wall-vibration sweep) for CPU timing.  Writes wall seconds of the
 * step loop to *seconds. */
int   vso_run_structured(void* h, int n_steps, double* probe, double* seconds);

/* vso_run with each phase's rows split over nthreads OpenMP threads (CPU
 * timing on several cores; bit-identical to vso_run).  Writes wall seconds of
 * the step loop to *seconds. */
int   vso_run_threads(void* h, int first_step, int n_steps, double* probe, int nthreads, double* seconds);

/* Reference float32 time-step count for a duration (SolverParams.cu:59-66). */
int   vso_float_num_steps(float audio_dur, int srate_mul);

/* The two-mass recurrence alone (TwoMassModel.cpp:90-231), fed p1[i] before
 * step i as FDTDEngineCPU.cpp:259-267 does; writes the volume velocity of
 * every step.  abs_double selects g++'s resolution of fabs(float) (double)
 * instead of MSVC's (float), the reference's compiler. */
int   vso_twomass_run(int precision, double q, double ps, int srate_mul, int abs_double, const double* p1, int n,
                      double* ug);
/* 17 solver coefficients of one (cur, right, top) type triple. */
int   vso_coeff_triple(int precision, int srate_mul, int cur, int right, int top, double* out17);

/* Single-plane-symmetric contour of one section (ShapeGenerator.cpp:35-197,
 * ShapeAlignment.cu:11-113): writes the aligned closed contour (returns its
 * point count, or -1) and info6 = {minY, minZ, maxZ, max chord, contour area,
 * growth steps}.  y/z may be NULL to query the count. */
int   vso_plane_contour(float area_cm2, float* y, float* z, int cap, double* info6);
/* Geometry helpers of that path: shoelace |area| of a closed ring, lowest and
 * highest crossing of a vertical line with a polyline (returns 1 if any), and
 * the crossing of two segments p = (y0,z0,y1,z1), q likewise (returns 1 if any). */
float vso_polygon_area(const float* y, const float* z, int n);
int   vso_vertical_chord(const float* y, const float* z, int n, float y0, float* lo, float* hi);
int   vso_segment_cross(const float* p4, const float* q4, float* out2);

#ifdef __cplusplus
}
#endif
#endif
