/*
 * vocalsim_abi.h -- C ABI of the MI355X-native 2.5D FDTD vocal-tract solver.
 *
 * Plain C: POD structs, raw pointers and sizes, int status codes.  No HIP or
 * torch types appear in any signature; device pointers are `void*`.
 *
 * One handle owns a batch of n >= 1 tubes (n = 1 is the plain single-tube
 * solver).  The handle owns all device memory; pointers handed out stay valid
 * until vs_destroy.  Every call returns VS_OK (0) or a negative VS_E* code and
 * sets a thread-local message readable with vs_last_error().  Nothing calls
 * exit() or traps.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   vs_create            FDTDSolver::FDTDSolver + setupWaveSolver()
 *                        (src/FDTDSolver.cu:71-83, 289-476) together with the
 *                        TubeGeometry / TubeDepth setup they drive
 *                        (src/TubeGeometry.cu:53-237, src/TubeDepth.cu:95-265)
 *   vs_get_dims          public fields frameH/frameW/domainH/domainW,
 *                        listenerX/Y, excitationX/Y, num_steps
 *                        (include/FDTDSolver.h:39-81)
 *   vs_get_field         coefficient getters getN_out() ... getZ_inv() and
 *                        getDepthP/X/Y(), getPV_N() (include/FDTDSolver.h:143-170)
 *   vs_run               FDTDEngineCPU::startSolverEngine time loop
 *                        (src/FDTDEngineCPU.cpp:222-304), FDTDEngineEigen
 *                        (src/FDTDEngineEigen.cpp:121-183)
 *   vs_get_probe         record_pressure / savePressureSamples
 *                        (src/FDTDEngineCPU.cpp:342, 435-450)
 *   vs_destroy           FDTDEngineCPU::freeCUDAMemory (src/FDTDEngineCPU.cpp:199-219)
 *                        and the solver's device vectors
 */
#ifndef VOCALSIM_ABI_H_
#define VOCALSIM_ABI_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VS_ABI_VERSION 2

enum vs_status {
    VS_OK = 0,
    VS_E_INVALID = -1,     /* bad argument or spec */
    VS_E_GEOMETRY = -2,    /* geometry cannot be built (e.g. depth sqrt domain) */
    VS_E_DEVICE = -3,      /* HIP runtime error */
    VS_E_RANGE = -4,       /* index / size out of range */
    VS_E_STATE = -5,       /* call not valid in the current state */
    VS_E_NUMERIC = -6,     /* non-finite value detected on the device */
    VS_E_HOST_MEMORY = -7  /* host memory allocation failed */
};

enum vs_precision { VS_FP32 = 32, VS_FP64 = 64 };

/* Tube description: the reference's SimulationParams flags
 * (include/SimulationParams.h:74-98) plus the inputs the workloads add.
 * Section arrays are float32, as the reference reads them
 * (src/TubeGeometry.cu:83-88). */
typedef struct vs_tube_spec {
    int fdtd_solver_type;       /* 0: 2D, 1: 2.5D                          */
    int cross_sectional_shape;  /* 1: circular, 2: elliptical, 3: single-plane symmetric */
    int mouth_radiation_cond;   /* 1: Dirichlet mouth, 0: open              */
    int srate_mul;              /* fs = 44100 * srate_mul                   */
    int pml_flag;               /* 0 / 1                                    */
    int source_type;            /* 0: sine, 1: Gaussian, 2: two-mass        */
    int num_steps;              /* > 0 explicit; <= 0 float accumulation of audio_dur */
    float audio_dur;            /* seconds                                  */
    int num_sections;
    const float* length_cm;     /* [num_sections]                           */
    const float* area_cm2;      /* [num_sections]                           */
    const float* aspect;        /* ellipse major:minor per section, NULL -> 3 */
    const float* depth_scale;   /* per-section depth multiplier, NULL -> 1  */
    const float* mu3d;          /* per-section wall admittance, NULL -> default */
    float mu3d_default;         /* <= 0 -> 0.025 (2.5D) / 0.005 (2D)        */
    double gauss_amp;           /* <= 0 -> reference pulse (0.025, 64.6 us, 0.29*t0) */
    double gauss_t0;
    double gauss_width;
    double tm_ps;               /* two-mass subglottal pressure, <= 0 -> 686.46 Pa */
    double tm_q;                /* two-mass pitch factor, <= 0 -> 1         */
    int engine_variant;         /* 0: FDTDEngineCPU rounding, 1: FDTDEngineEigen */
    int open_space;             /* 1: open-space domain (simulation_type 0, FDTDSolver.cu:479-550):
                                 * 43 x 341 air cells, point source at the centre,
                                 * listener 5 cells down and right; the area function is ignored */
} vs_tube_spec;

typedef struct vs_dims {
    int frameH, frameW, domainH, domainW;
    int listenerX, listenerY, excitationX, excitationY;
    int num_steps, tube_start_x, tube_end_x, tube_len_cells;
    int midY, num_sections, mic_cells;
} vs_dims;

/* Field ids for vs_get_field / vs_field_ptr. */
enum vs_field {
    VS_F_CELL_TYPE = 0,   /* int32 [H*W]  CellTypes enum value                     */
    VS_F_BOUNDARY_SEG,    /* int32 [H*W]  boundary_segment_type                    */
    VS_F_DEPTH_P,         /* Real  [H*W]                                           */
    VS_F_DEPTH_X,         /* Real  [H*W]                                           */
    VS_F_DEPTH_Y,         /* Real  [H*W]                                           */
    VS_F_Z_INV,           /* Real  [H*W]                                           */
    VS_F_SOURCE,          /* Real  [num_steps] precomputed excitation table        */
    VS_F_UPPER_WALL,      /* int32 [L]                                             */
    VS_F_LOWER_WALL,      /* int32 [L]                                             */
    VS_F_SEGMENT,         /* int32 [L]   section index of each tube column         */
    VS_F_PRESSURE,        /* Real  [H*W] current P                                 */
    VS_F_VX,              /* Real  [H*W] current Vx                                */
    VS_F_VY,              /* Real  [H*W] current Vy                                */
    VS_F_CELL_CLASS,      /* uint8 [H*W] packed class byte (type | segment-wall bit) */
    VS_F_AS_F32 = 0x1000, /* flag for vs_field_ptr: float32 copy of a Real field (FP64 handles) */
    VS_F_COEFF0 = 100     /* + k, k in [0,17): Real [(H-2)*(W-2)] coefficient grid k
                             in getter order N_out, N_in, minVxBeta, minVyBeta,
                             maxVxSigmaPrimedt, maxVySigmaPrimedt,
                             betaVxSqr_dt_inv_rho_inv_ds, betaVySqr_dt_inv_rho_inv_ds,
                             pressureSigmaPrimedt, areWeNotExcitationVx/Vy,
                             excitationWeightVx/Vy, xor_val1..4                     */
};

/* Tunables for vs_set_option. */
enum vs_option {
    VS_OPT_STEPS_PER_LAUNCH = 0, /* steps advanced by one launch of the step kernel (default 256) */
    VS_OPT_THREADS = 1,          /* work-group size of the step kernel (64..1024) */
    VS_OPT_USE_GRAPH = 2,        /* 1: replay step blocks from a captured hipGraph */
    VS_OPT_KERNEL = 3,           /* step kernel family, see vs_kernel */
    VS_OPT_TIMING = 4,           /* 1: time every launch with HIP events */
    VS_OPT_ENGINE_VARIANT = 5,   /* 0: FDTDEngineCPU rounding, 1: FDTDEngineEigen (all tubes) */
    VS_OPT_ADJOINT_KERNEL = 6,   /* FP32 reverse-time pass: 0 (default) on-chip kernels when every
                                    tube fits one work-group, 1: per-step global-memory kernels.
                                    FP64 always uses the per-step kernels. */
    VS_OPT_ADJOINT_BLOCK = 7,    /* on-chip reverse pass: most steps recorded per block (0, the
                                    default: as many of the checkpoint spacing as fit in half the
                                    free device memory); shorter blocks trade one extra forward
                                    pass for memory */
    VS_OPT_EXACT_GROUPS = 8,     /* exact step of a single-tube handle: 1024-thread
                                    work-groups of one cooperative launch (VS_OPT_THREADS
                                    does not apply), the tube's columns split across them
                                    (bit-identical results).  0, the default: the fewest (at
                                    most 48) whose columns plus a halo of 2 x
                                    VS_OPT_EXACT_STEPS hold one cell per thread (one
                                    work-group when the tube itself does); 1: one
                                    work-group; 2..64 */
    VS_OPT_EXACT_STEPS = 9,      /* steps of that cooperative launch between two grid
                                    barriers: each work-group recomputes halo columns to run
                                    them without exchanging data.  0, the default: up to 8,
                                    as many as 48 work-groups allow at one cell per thread
                                    (with VS_OPT_EXACT_GROUPS set: as many as that count
                                    allows, else 1); 1..32 (lowered where the LDS tile or
                                    the source placement needs it) */
    VS_OPT_BARRIER_SPIN = 10,    /* polls a grid barrier of that cooperative launch makes
                                    before it gives up: the launch then finishes without
                                    synchronising and vs_sync reports VS_E_DEVICE.  0, the
                                    default: about a quarter of a second; >= 1: that many */
    VS_OPT_FAST_WIDE = 11,       /* FP32 on-chip kernel layout: 1 (default) the wide-access form
                                    (rows padded to 4 cells, 4 (mod 8) cells per lane, 16-byte LDS
                                    accesses) whenever the batch fits it at the work-group size,
                                    else the odd-run form; 0: always the odd-run form.  Both
                                    give the same values (signed zeros aside). */
    VS_OPT_FAST_LAYOUT_K = 12,   /* read-only (vs_get_option): cells per lane of the on-chip
                                    kernel the current work-group size runs, 4 (mod 8) for the
                                    wide-access form, odd for the odd-run form, 0 when neither
                                    fits the batch */
    VS_OPT_FAST_PAIR_CLASSES = 13, /* read-only (vs_get_option): pair classes of the batch (the
                                    ordered class pairs of neighbouring cells the wide-access form
                                    reads its coefficients by), 0 before the on-chip kernel's
                                    set-up has run; above 128 the batch runs the odd-run form */
    VS_OPT_FAST_ROW_QUANTUM = 14 /* read-only (vs_get_option): the wide-access form pads rows to a
                                    multiple of this many cells: its cells per lane (each column at
                                    one slot of its lanes' runs, the source face handled in one
                                    group of four cells), or 4 when such rows do not fit; 0 when
                                    the wide-access form does not run */
};

enum vs_kernel {
    VS_K_EXACT = 0,   /* reference operation order (bit-exact vs the oracle's instantiation) */
    VS_K_FAST = 1     /* FP32 reassociated coefficient form (batched throughput path) */
};

typedef struct vs_handle vs_handle;

const char* vs_last_error(void);
int vs_abi_version(void);

/* Builds n tubes on the current HIP device: host float32 integerisation
 * followed by the setup kernels (walls, classification, depth maps, z_inv,
 * source tables).  precision: VS_FP32 or VS_FP64. */
int vs_create(const vs_tube_spec* specs, int n_tubes, int precision, vs_handle** out);
int vs_destroy(vs_handle* h);

int vs_num_tubes(const vs_handle* h, int* n);
int vs_get_dims(const vs_handle* h, int tube, vs_dims* out);

/* Copies a field of one tube to host memory. bytes must be at least the
 * field size; returns VS_E_RANGE otherwise. */
int vs_get_field(vs_handle* h, int tube, int field, void* dst, size_t bytes);

/* Advances every tube by n_steps (tubes stop at their own num_steps).
 * Asynchronous on the handle's stream; vs_sync waits. */
int vs_run(vs_handle* h, int n_steps);
int vs_sync(vs_handle* h);
/* Steps already advanced. */
int vs_steps_done(const vs_handle* h, int* steps);
/* Zeros P, V and the source state; probe history restarts at step 0. */
int vs_reset(vs_handle* h);

/* Probe waveform (P at the listener after each step's pressure phase) and
 * the injected source velocity of each step, as float64. */
int vs_get_probe(vs_handle* h, int tube, double* dst, size_t n);
int vs_get_source_trace(vs_handle* h, int tube, double* dst, size_t n);

int vs_set_option(vs_handle* h, int option, int value);
int vs_get_option(const vs_handle* h, int option, int* value);

/* Device pointer of a field (same ids as vs_get_field, tube's slice), and the
 * handle's stream as an opaque pointer (hipStream_t).  Coefficient grids,
 * cell types and z_inv are materialised on first request into buffers the
 * handle owns (the step kernels derive them on the fly). */
int vs_field_ptr(vs_handle* h, int tube, int field, void** dptr, size_t* bytes);
int vs_stream(vs_handle* h, void** stream);

/* Timing of the last vs_run with HIP events on the handle's stream: total
 * milliseconds, number of step-kernel launches and (with VS_OPT_TIMING) the
 * mean duration of one launch. */
int vs_last_run_timing(const vs_handle* h, double* total_ms, int* launches, double* mean_launch_ms);

/* Autotuner (no reference counterpart; the reference has no GPU time step).
 * Times the step kernel over the work-group shapes that fit the batch (for
 * the on-chip FP32 kernel: every multiple of 64 threads whose cells-per-lane
 * count K and LDS footprint fit; for the exact kernel 256..1024 threads) and
 * steps per launch {128, 256, 512}, each with a short warm-up and a timed
 * block measured with HIP events, then keeps the fastest and resets the
 * state (vs_reset).  Results are cached per (batch shape, precision, kernel)
 * for the life of the process.  Requires steps_done == 0.  Writes a JSON
 * report of the candidates into report (if non-NULL). */
int vs_autotune(vs_handle* h, int timed_steps, char* report, size_t report_len);

/* FP64 discrete-time Fourier transform X(f) = sum_n x[n] w[n] e^{-i 2 pi f n / fs}
 * at nf arbitrary frequencies [Hz] (w = Hann window when hann != 0, else 1).
 * Host arrays in and out; computed on the GPU.  Replaces the np.fft.fft
 * spectrum of results/formants_bandwidths.py:34-36 and
 * results/plt_transfer_function.py:21-23 with an evaluation at any frequency,
 * which the sub-bin formant estimates need. */
int vs_dtft(const double* x, int n, double fs, int hann, const double* freqs, int nf, double* re, double* im);
/* vs_dtft without any device allocation (no counterpart in the reference:
 * for callers that cannot allocate device memory, e.g. code under graph
 * capture or in a serving loop that forbids hipMalloc).  The device copies of
 * x, freqs, re and im live in a static device array of the library, one per
 * GPU and held until the process ends (2.375 MiB): n <= 262144 samples and
 * nf <= 16384 frequencies, VS_E_RANGE beyond.  Runs on the current device's
 * null stream and returns when the results are in re and im.  Calls on one
 * GPU share that array: the library serialises its own calls per device, and
 * calls from other processes on the same GPU race. */
int vs_dtft_noalloc(const double* x, int n, double fs, int hann, const double* freqs, int nf, double* re,
                    double* im);
/* Same transform of a tube's device-resident probe waveform (kind 0) or
 * injected source velocity (kind 1) over the steps advanced so far. */
int vs_probe_dtft(vs_handle* h, int tube, int kind, int hann, const double* freqs, int nf, double* re, double* im);

/* Copies bytes from a device pointer handed out by vs_field_ptr (or from host
 * memory) into host memory: the cudaMemcpy of the reference's debug printers
 * (src/FDTDSolver.cu:95, 125, 201, 232-234). */
int vs_copy_to_host(void* dst, const void* src, size_t bytes);

/* Interior cell updates of one step summed over the batch:
 * sum (frameH-2)(frameW-2)  (FDTDEngineCPU.cpp:316-317). */
int vs_cells_per_step(const vs_handle* h, long long* cells);

/* LDS read bandwidth of the current device in bytes per second (no reference
 * counterpart: the peak against which the on-chip step kernel's LDS traffic
 * is rated).  A streaming kernel, 8 work-groups of 1024 threads per compute
 * unit, each lane reading lane-consecutive vectors of width_bytes (4, 8 or
 * 16) from a 32 KiB region; best of three timed launches after a warm-up. */
int vs_measure_lds_bandwidth(int width_bytes, double* bytes_per_s);

/* Audio export.  Replaces results/generate_audio.py:9-42: the probe of a
 * tube (tube >= 0) or of every tube concatenated in tube order (tube = -1),
 * over the steps advanced so far, normalised by its max |p|, clipped to
 * [-1, 1] and, when fs = 44100 * srate_mul differs from 44100, low-passed by
 * the second-order Butterworth at 5050 Hz of scipy.signal.butter(2, ...)
 * applied as scipy.signal.lfilter does, then decimated by srate_mul
 * (samples 0, srate_mul, 2 srate_mul, ...).  Computed on the GPU from the
 * device-resident probe; float32 samples at 44.1 kHz.  *n_out receives the
 * sample count; out may be NULL to query it. */
int vs_probe_audio(vs_handle* h, int tube, float* out, size_t cap, size_t* n_out);
/* Coefficients b[3], a[3] (a[0] = 1) of that low-pass for a cutoff and a
 * sample rate (host only, no device needed). */
int vs_butter2_lowpass(double cutoff_hz, double fs, double* b3, double* a3);
/* Writes mono float32 samples as a WAV file (IEEE float, 18-byte fmt chunk
 * and a fact chunk, the layout scipy.io.wavfile.write produces; host only). */
int vs_write_wav(const char* path, const float* samples, size_t n, int rate);

/* Reverse-time (adjoint) gradients.  No reference counterpart: the reference
 * only runs the forward time loop (FDTDEngineCPU.cpp:316-410); these entry
 * points differentiate that loop, written in flux form U = V * depth, which
 * is algebraically the same step.
 *
 * vs_adjoint_forward runs every tube of the handle from the zero state over
 * its num_steps in the handle's precision, keeps a state checkpoint
 * (P, Ux, Uy) every checkpoint_every steps in device memory, and writes the
 * probe traces (tube-major, num_steps values per tube) into probe when probe
 * is non-NULL.  probe may be host memory or device memory of the handle's
 * device (the copy's direction is taken from the pointer); it is complete
 * when the call returns.  It leaves the handle's ordinary state (vs_run) untouched.
 * Table sources only (source_type 0 or 1): a two-mass source feeds back the
 * pressure and is rejected with VS_E_INVALID.
 *
 * vs_adjoint_backward, after vs_adjoint_forward, evaluates the gradient of
 * J = sum_tubes sum_n gprobe[n] * probe[n] (gprobe tube-major like probe)
 * with respect to each section's area_cm2, depth_scale and wall admittance
 * mu3d (outputs section-major per tube, tubes concatenated; n_sections_total
 * = sum of num_sections).  gprobe may be host memory or device memory of
 * the handle's device; a device gprobe must be complete before the call (the
 * library reads it on its own stream).  The three gradient outputs are host
 * memory.  Segments of checkpoint_every steps are recomputed
 * from their checkpoint and reversed step by step on the GPU; the per-cell
 * coefficient gradients are then chained on the host through the depth
 * clamps, the depth-P average, the half-cell smoothing, the ellipse depth
 * profile and the admittance map.  The integer geometry (diameters in cells)
 * is held fixed: its dependence on the area is piecewise constant.  A
 * circular cross-section's depth does not read the area, so its area
 * gradient is 0; so are the area and depth-scale gradients of the 2D solver.
 * Memory: checkpoints ceil(num_steps / checkpoint_every) * 3 * H * W values
 * and a segment slab (checkpoint_every + 1) * 3 * H * W values per tube. */
int vs_adjoint_forward(vs_handle* h, int checkpoint_every, double* probe, size_t probe_len);
int vs_adjoint_backward(vs_handle* h, const double* gprobe, size_t gprobe_len, double* g_area,
                        double* g_depth_scale, double* g_mu3d, size_t n_sections_total);
/* Wall-clock split of the last vs_adjoint_backward: the reverse pass on the
 * device including its copies back (reverse_ms) and the host chain from the
 * per-cell log-derivatives to the per-section inputs (chain_ms). */
int vs_adjoint_timing(const vs_handle* h, double* reverse_ms, double* chain_ms);
/* Frees the device memory the library keeps after vs_destroy (no
 * counterpart in the reference: for callers that create and destroy handles
 * in a loop, such as an optimisation loop that builds a handle per forward
 * pass of the gradient).  A handle takes its field memory in chunks of
 * 4 to 64 MiB (sized from its batch), and the on-chip FP32 reverse pass its
 * record slabs (several GB for a large batch).  When the handle is
 * destroyed, every such piece of 4 MiB or more is kept, per device and up to
 * 64 GiB in all (the oldest are freed beyond that), and a later request of
 * between half a kept piece's size and its size takes it instead of
 * allocating.  Kept memory
 * counts as free when a handle sizes its record slabs, and a failed
 * allocation frees the device's kept memory before it is retried once.
 * The pinned host blocks vs_adjoint_backward copies its per-cell results
 * into (tens of MB, up to 4 GiB kept) are freed here too.  Call this when
 * the work is done and the memory is wanted elsewhere. */
int vs_release_cached_memory(void);

#ifdef __cplusplus
}
#endif
#endif
