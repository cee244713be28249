/*
 * kwave_hip.h -- C ABI of libkwavehip, the MI355X (gfx950) implementation of KWaveSwift's
 * k-space pseudospectral first-order time loop `kspaceFirstOrder` (1D/2D/3D).
 *
 * Boundary. The reference's public entry point is
 *   kspaceFirstOrder(grid:medium:source:sensor:options:) -> SimulationOutput
 *   (Sources/KWave/Solver/KSpaceFirstOrder.swift:79-107), dispatched on `grid.dim` at :94.
 * A Swift shim selected by `options.device == .hip` (DeviceKind, Sources/KWave/KWave.swift:12-22)
 * flattens the KWaveGrid / KWaveMedium / KWaveSource / KWaveSensor / SimulationOptions values
 * into `kwh_sim_desc`, calls kwh_sim_create + kwh_sim_run, and wraps the outputs read with
 * kwh_sim_read_output back into SimulationOutput (KSpaceFirstOrder.swift:47-73).
 *
 * Conventions.
 *  - All entry points are thread-compatible: one kwh_sim per host thread. Separate sims may run
 *    concurrently (each owns a HIP stream).
 *  - Input pointers are borrowed for the duration of the call only; everything is copied.
 *  - Outputs go to caller-allocated buffers.
 *  - Every call returns kwh_status; on failure kwh_last_error() holds a message (thread-local,
 *    valid until the next kwh_* call on the same thread).
 *  - Arrays are C-order (row-major) over [nx, ny, nz], matching the reference's MLXArray layout.
 *  - Sensor / source point lists are ascending C-order flat indices, as produced by
 *    flatNonzeroIndices (KSpaceFirstOrder.swift:778-782).
 */
#ifndef KWAVE_HIP_H
#define KWAVE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KWH_ABI_VERSION 2

typedef enum {
    KWH_OK = 0,
    KWH_EINVAL = -1,        /* invalid argument / descriptor (reference: precondition traps) */
    KWH_ENOMEM = -2,        /* device allocation failed */
    KWH_EDEVICE = -3,       /* HIP runtime error or no device */
    KWH_ENONFINITE = -4,    /* non-finite pressure (kwh_sim_set_nonfinite_check) */
    KWH_EUNSUPPORTED = -5,  /* configuration outside what this library implements */
    KWH_ESTATE = -6         /* call out of order (e.g. reading an output before running) */
} kwh_status;

/* SourceMode (Sources/KWave/Source/KWaveSource.swift:4-8). */
typedef enum { KWH_SRC_DIRICHLET = 0, KWH_SRC_ADDITIVE = 1, KWH_SRC_ADDITIVE_NO_CORRECTION = 2 } kwh_source_mode;

/* AbsorptionMode (Sources/KWave/Medium/KWaveMedium.swift:8-17). */
typedef enum { KWH_ALPHA_NONE = 0, KWH_ALPHA_POWER_LAW = 1, KWH_ALPHA_NO_DISPERSION = 2, KWH_ALPHA_STOKES = 3 } kwh_alpha_mode;

/* DTypePrecision (Sources/KWave/KWave.swift:4-9); honoured on this path. */
typedef enum { KWH_FLOAT32 = 0, KWH_FLOAT64 = 1 } kwh_dtype;

/* DirectivityPattern (Sources/KWave/Sensor/KWaveSensor.swift:24-29). */
typedef enum { KWH_DIR_PRESSURE = 0, KWH_DIR_GRADIENT = 1 } kwh_directivity_pattern;

/* RecordField bits (Sources/KWave/Sensor/KWaveSensor.swift:8-20). */
enum {
    KWH_REC_P = 1 << 0, KWH_REC_P_MAX = 1 << 1, KWH_REC_P_MIN = 1 << 2, KWH_REC_P_RMS = 1 << 3,
    KWH_REC_P_FINAL = 1 << 4, KWH_REC_UX = 1 << 5, KWH_REC_UY = 1 << 6, KWH_REC_UZ = 1 << 7,
    KWH_REC_U_MAX = 1 << 8, KWH_REC_U_RMS = 1 << 9, KWH_REC_U_FINAL = 1 << 10,
    KWH_REC_I_AVG = 1 << 11, KWH_REC_I_MAX = 1 << 12
};

/* Output fields of SimulationOutput (KSpaceFirstOrder.swift:47-73). Series are [nSensor, nt],
 * aggregates [nSensor], finals grid-shaped. Element type is the sim's dtype. */
typedef enum {
    KWH_OUT_P = 0, KWH_OUT_P_MAX, KWH_OUT_P_MIN, KWH_OUT_P_RMS, KWH_OUT_P_FINAL,
    KWH_OUT_UX, KWH_OUT_UY, KWH_OUT_UZ,
    KWH_OUT_UX_MAX, KWH_OUT_UY_MAX, KWH_OUT_UZ_MAX,
    KWH_OUT_UX_RMS, KWH_OUT_UY_RMS, KWH_OUT_UZ_RMS,
    KWH_OUT_UX_FINAL, KWH_OUT_UY_FINAL, KWH_OUT_UZ_FINAL,
    KWH_OUT_IX_AVG, KWH_OUT_IY_AVG, KWH_OUT_IZ_AVG,
    KWH_OUT_COUNT
} kwh_output;

/* A scalar-or-grid float field: count == 1 is a scalar, count == nx*ny*nz a C-order map.
 * data == NULL means "absent" (the reference's optional). */
typedef struct {
    const float* data;
    int64_t count;
} kwh_field;

/* One time-varying source field (SourceOp, KSpaceFirstOrder.swift:129-201).
 * signal is [rows][len], rows == 1 (broadcast) or rows == n_points.
 * Delayed copies (a transducer drive, KWaveTransducer.swift:185-258): with delay != NULL the
 * signal is one row and point k plays scale[k] * signal[t - delay[k]] (zero outside the row;
 * scale NULL = 1); the per-point rows, len + max(delay) samples long, are built on the device. */
typedef struct {
    const int64_t* index;     /* ascending flat indices of mask nonzeros; NULL = absent */
    int64_t n_points;
    const float* signal;      /* NULL = absent */
    int32_t rows;
    int32_t len;
    int32_t mode;             /* kwh_source_mode */
    const int32_t* delay;     /* [n_points] samples >= 0, or NULL */
    const float* scale;       /* [n_points], or NULL */
} kwh_source_field;

/* Flattened KWaveGrid + KWaveMedium + KWaveSource + KWaveSensor + SimulationOptions. */
typedef struct {
    int32_t abi_version;      /* must be KWH_ABI_VERSION */
    /* KWaveGrid (Sources/KWave/Grid/KWaveGrid.swift:9-143) */
    int32_t dim;              /* 1, 2 or 3 */
    int32_t n[3];             /* nx, ny, nz (unused entries 1) */
    double d[3];              /* dx, dy, dz */
    double dt;
    int32_t nt;
    int32_t dtype;            /* kwh_dtype */
    /* SimulationOptions (KSpaceFirstOrder.swift:17-40) */
    int32_t pml_size;         /* PMLSize.uniform(n) */
    double pml_alpha;
    int32_t smooth_p0, smooth_c0, smooth_rho0;
    /* KWaveMedium (Sources/KWave/Medium/KWaveMedium.swift:21-76) */
    kwh_field sound_speed;
    kwh_field density;
    kwh_field alpha_coeff;    /* dB/(MHz^y cm); absent = lossless */
    double alpha_power;       /* y; the caller substitutes 1.5 when alphaPower is nil */
    int32_t alpha_mode;       /* kwh_alpha_mode */
    kwh_field bona;           /* B/A; absent or all-zero = linear */
    /* KWaveSource (Sources/KWave/Source/KWaveSource.swift:11-71) */
    kwh_field p0;
    kwh_source_field p_source;
    kwh_source_field u_source[3];   /* ux, uy, uz share the same mask in the reference */
    /* KWaveSensor (Sources/KWave/Sensor/KWaveSensor.swift:32-62) */
    const int64_t* sensor_index;    /* binary mask as flat indices; NULL if none / Cartesian */
    int64_t n_sensor;
    const float* sensor_cartesian;  /* [dim][n_cartesian] coordinates in metres; NULL if none */
    int64_t n_cartesian;
    uint32_t record;                /* RecordField bits */
    const float* directivity_angle; /* grid-shaped angles (2D binary masks only); NULL = off */
    double directivity_size;        /* <= 0: default 10*max(dx,dy) */
    int32_t directivity_pattern;    /* kwh_directivity_pattern */
} kwh_sim_desc;

typedef struct kwh_ctx kwh_ctx;
typedef struct kwh_sim kwh_sim;
typedef struct kwh_batch kwh_batch;

/* Called once per step, in order, between step chunks (up to 64 steps per device sync; a chunk
 * ends at every step whose pressure is delivered). t is the step index, p_host the pressure field
 * after step t (sim dtype, grid-shaped) when t % cb_interval == 0, else NULL; cb_interval <= 0
 * never delivers the field (progress only). Replaces the reference's progress / fieldMonitor
 * closures (KSpaceFirstOrder.swift:30-37, called at :337-340). p_host is valid during the call. */
typedef void (*kwh_step_cb)(int32_t t, int32_t nt, const void* p_host, void* user);

/* Per-kernel timing record filled by kwh_sim_kernel_stats. */
typedef struct {
    char name[48];
    int64_t launches;
    double total_ms;          /* summed HIP-event durations over the timed launches */
    double bytes_per_launch;  /* algorithmic bytes moved per launch, averaged over the timed launches */
} kwh_kernel_stat;

/* ABI version compiled into the library. */
int kwh_abi_version(void);

/* Thread-local message for the last failing call on this thread. */
const char* kwh_last_error(void);

/* Open device `device_ordinal` (HIP ordinal). */
int kwh_ctx_create(int device_ordinal, kwh_ctx** out);
void kwh_ctx_destroy(kwh_ctx* ctx);

/* Host-only validation of a descriptor (no device access): every check kwh_sim_create performs
 * before touching the GPU, with the same error codes (the reference's precondition traps,
 * KSpaceFirstOrder.swift:210-212, :221-222, Recording.swift:76-77). */
int kwh_sim_validate(const kwh_sim_desc* desc);

/* Validate the descriptor, copy all inputs to the device and precompute the operators
 * (setup of kspaceFirstOrder1D/2D/3D, KSpaceFirstOrder.swift:203-308, 351-491, 557-703). */
int kwh_sim_create(kwh_ctx* ctx, const kwh_sim_desc* desc, kwh_sim** out);

/* The axisymmetric solver kspaceFirstOrderAS(grid:medium:source:sensor:options:)
 * (Sources/KWave/Solver/KSpaceFirstOrderAS.swift:17-23) on the same descriptor: a 2D grid with x
 * axial and y radial (y index 0 on the axis), WSWA-FFT radial symmetry on a 4x radially expanded
 * grid, a one-sided radial PML, sources p0 / p (three modes) / ux, uy (three modes), Stokes
 * absorption, B/A; the pressure is recorded (p and its aggregates, pFinal). The handle is used
 * with kwh_sim_run / kwh_sim_read_output like any other; step graphs, checkpoints, the adjoint and
 * the autotuner are not available for it (KWH_EUNSUPPORTED). */
int kwh_sim_create_axisymmetric(kwh_ctx* ctx, const kwh_sim_desc* desc, kwh_sim** out);

/* Advance steps [t_begin, t_end) (the hot loop, KSpaceFirstOrder.swift:309-342, 492-546,
 * 704-751). Runs must be contiguous: t_begin equals the number of steps already run.
 * cb may be NULL. Returns after the device has finished the requested steps. */
int kwh_sim_run(kwh_sim* sim, int32_t t_begin, int32_t t_end, kwh_step_cb cb, int32_t cb_interval,
                void* user);

/* Enqueue steps [t_begin, t_end) without waiting (for timing harnesses); kwh_sim_sync waits. */
int kwh_sim_launch(kwh_sim* sim, int32_t t_begin, int32_t t_end);
int kwh_sim_sync(kwh_sim* sim);

/* Shape of an output after the run (finalizeRecording, Recording.swift:129-168).
 * Returns KWH_EINVAL if the field was not recorded. With a sensor mask that has no points, the
 * recorded sensor fields are empty: series [0, steps run], aggregates [0] (read them with any
 * buffer; nothing is copied). */
int kwh_sim_output_shape(kwh_sim* sim, int32_t field, int64_t shape[3], int32_t* ndim);

/* Copy an output (kwh_output) into dst (element type = sim dtype). */
int kwh_sim_read_output(kwh_sim* sim, int32_t field, void* dst, size_t dst_bytes);

/* kwh_sim_read_output without device allocations. No counterpart in the reference: it is for
 * callers that cannot allocate device memory between set-up and teardown (serving loops that read
 * results after every run, harnesses that forbid hipMalloc). Same arguments, same results (bit
 * for bit) as kwh_sim_read_output; the staging that call allocates per read (sensor-major series
 * rows, pRms / uRms, iAvg) lives instead in a static 8 MiB device array per GPU, resolved once per
 * device by kwh_ctx_create, and the outputs are produced in chunks of sensors that fit it. Limit:
 * one sensor's recorded series (steps run x element size) must fit the 8 MiB, else
 * KWH_EUNSUPPORTED. Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (sims on several host threads) race and must be serialised by the caller. */
int kwh_sim_read_output_noalloc(kwh_sim* sim, int32_t field, void* dst, size_t dst_bytes);

/* Number of steps completed so far. */
int kwh_sim_steps_done(kwh_sim* sim, int32_t* steps);

/* Save / restore the time-loop state (p, u, rho, sensor accumulators) for chunked runs.
 * kwh_sim_state_bytes gives the buffer size. */
int kwh_sim_state_bytes(kwh_sim* sim, size_t* bytes);
int kwh_sim_save_state(kwh_sim* sim, void* dst, size_t bytes);
int kwh_sim_load_state(kwh_sim* sim, const void* src, size_t bytes);

/* Kernel timing with HIP events around every launch (off by default; adds event overhead). */
int kwh_sim_set_kernel_timing(kwh_sim* sim, int enable);
/* Event timing of one kernel only (a name kwh_sim_kernel_stats reports, e.g. "z_density"):
 * two events per launch of that kernel, none elsewhere. Resets the statistics. */
int kwh_sim_set_kernel_timing_only(kwh_sim* sim, const char* kernel_name);
int kwh_sim_kernel_stats(kwh_sim* sim, kwh_kernel_stat* stats, int32_t capacity, int32_t* count);

/* Optional blow-up detection (no counterpart in the reference, whose solver only traps on its
 * preconditions, KSpaceFirstOrder.swift:564-576): with every > 0, kwh_sim_run scans the pressure
 * field on the device after every `every`-th step and after the last step of each call, and
 * returns KWH_ENONFINITE as soon as a scan finds a NaN or an infinity (typically a time step above
 * the CFL stability limit). Without a callback the run stops at the end of the first chunk of
 * `every` steps whose field is not finite; with one, after the callback of the flagged step. The
 * steps run so far stay done (kwh_sim_steps_done) and the outputs stay readable. 0 (the default)
 * turns the check off; the scan reads the pressure once (N values) per check. */
int kwh_sim_set_nonfinite_check(kwh_sim* sim, int32_t every);

/* Replay the time step from captured hipGraphs (1 = on): one graph per set of active
 * k-space-corrected sources, captured on first use; results are bitwise those of eager runs. */
int kwh_sim_set_graph(kwh_sim* sim, int enable);

/* Plane mode (no counterpart in the reference, whose MLX transforms are whole-array calls,
 * KSpaceFirstOrder.swift:647-686): 3-D grids whose (y, z) plane is 32 x 32, 64 x 64, 96 x 96 or 128 x 128 run the y
 * and z transforms of each field in one kernel per x plane (the plane held in LDS), so a step
 * reads and writes each field once for the y and z axes together instead of once per axis. It
 * covers every medium and source kind and pressure recording, and is ON BY DEFAULT
 * wherever it applies (environment KWH_PLANE=0 turns the default off); checkpointing and the
 * adjoint switch the sim to the split passes while they are set. enable = 0 selects the split
 * passes; enable = 1 returns KWH_EUNSUPPORTED for configurations plane mode does not cover
 * (collocated velocity output, checkpointing, other plane shapes). May be
 * switched between steps. Results agree with the split passes to rounding (the transforms are
 * factored differently); a sim is deterministic in either mode. */
int kwh_sim_set_plane(kwh_sim* sim, int enable);

/* Receive combination on the device (combineSensorData, KWaveTransducer.swift:263-289):
 * out[e][t] = scale * sum over sensor points k with element[k] == e of
 * series[k][t + delay[k]] (terms with t + delay[k] beyond the recorded steps are dropped).
 * field is a recorded series (KWH_OUT_P / _UX / _UY / _UZ); element[] in [0, n_elements),
 * delay[] >= 0, both [n_sensor]. dst is [n_elements][steps recorded], sim dtype. */
int kwh_sim_element_traces(kwh_sim* sim, int32_t field, const int32_t* element, const int32_t* delay,
                           int32_t n_elements, double scale, void* dst, size_t dst_bytes);

/* Time-reversal reconstruction (timeReversal, TimeReversal.swift:20-51): the grid, medium and
 * options of desc, its binary sensor mask as a Dirichlet pressure source playing `recorded`
 * ([n_sensor][n_recorded_steps], float, host) reversed in time (on the device); runs the loop and
 * writes max(0, compensation_factor * pFinal) to p0_out (grid-shaped, desc dtype). Every source
 * and recording field of desc other than the sensor mask is ignored. */
int kwh_time_reversal(kwh_ctx* ctx, const kwh_sim_desc* desc, const float* recorded, int32_t n_recorded_steps,
                      double compensation_factor, void* p0_out);

/* Autotuner (no counterpart in the reference): before the first step, times one full time step
 * under each launch-schedule candidate (tile columns of the strided / middle passes, row pairs
 * per workgroup of the row kernels, LDS budget of the generic FFT passes) on this sim's buffers
 * and keeps the fastest. The choice is cached per device (gfx architecture and CU count), grid
 * shape and precision: in memory for the process and in a text file (KWH_AUTOTUNE_CACHE, a path
 * or "off"; default $XDG_CACHE_HOME/kwavehip/autotune.txt or $HOME/.cache/kwavehip/autotune.txt),
 * so later sims and later processes on the same kind of GPU take it without timing. Schedules
 * change launch geometry only: results are bitwise those of the default schedule. report (may be
 * NULL) receives the candidates' step times or the cached entry. The state is reset afterwards. */
int kwh_sim_autotune(kwh_sim* sim, char* report, size_t report_bytes);

/* Adjoint (reverse-time) gradient; no counterpart in the reference.
 * kwh_sim_set_checkpointing: save (p, u, rho) after every `every`-th step of later runs
 * (0 = off), so a following kwh_sim_vjp needs no forward pass of its own.
 * kwh_sim_vjp: for the p0 problem on the lossless linear loop (no time-varying sources, no
 * smoothing, pressure series recorded), given dL/dy [n_sensor][nt] of the recorded pressure
 * (sim dtype), writes dL/dp0 and dL/dc (grid-shaped, sim dtype). The gradient is the partial
 * derivative at fixed c_ref = max(c) (kappa and the PML held fixed). If the sim has not just
 * completed nt steps with checkpoints every `checkpoint_every` steps, the forward pass is run
 * first. The reverse sweep recomputes each checkpoint segment; afterwards the sim is back at its
 * end-of-run state, outputs unchanged. */
int kwh_sim_set_checkpointing(kwh_sim* sim, int32_t every);
int kwh_sim_vjp(kwh_sim* sim, const void* dl_dp, int32_t checkpoint_every, void* grad_p0, void* grad_c0);

/* Batched independent simulations of uneven sizes (per-scan-line transmits) on one device.
 * kwh_batch_create validates and sets up n sims (descs[i] as for kwh_sim_create).
 * kwh_batch_run advances every sim to its nt. mode is a set of KWH_BATCH_* flags:
 *   0                  each sim enqueues its steps eagerly on its own stream;
 *   KWH_BATCH_GRAPH    each sim replays its captured step graph on its own stream, the sims
 *                      interleaved step by step;
 *   KWH_BATCH_FUSED    one launch per pass of the step covers every sim whose step has the same
 *                      launch sequence and launch geometry (the sims of one grid shape); each
 *                      grid shape runs on a stream of its own, overlapping the others; with
 *                      KWH_BATCH_GRAPH each shape's batched step is captured once and replayed;
 *   KWH_BATCH_PLANE    switch every sim that plane mode covers into it (kwh_sim_set_plane; it is
 *                      the default where it applies, the flag restores it for sims switched
 *                      off); without the flag every sim keeps its own mode.
 * Every mode gives results bitwise equal to running the sims one by one in the same plane mode. Outputs are read per sim
 * (kwh_batch_read_output, or the borrowed handle from kwh_batch_sim with the kwh_sim_* readers). */
#define KWH_BATCH_GRAPH 1
#define KWH_BATCH_FUSED 2
#define KWH_BATCH_PLANE 4
int kwh_batch_create(kwh_ctx* ctx, const kwh_sim_desc* descs, int32_t n, kwh_batch** out);
int kwh_batch_run(kwh_batch* batch, int32_t mode);
int kwh_batch_sim(kwh_batch* batch, int32_t i, kwh_sim** out);
int kwh_batch_read_output(kwh_batch* batch, int32_t i, int32_t field, void* dst, size_t dst_bytes);
void kwh_batch_destroy(kwh_batch* batch);

void kwh_sim_destroy(kwh_sim* sim);

/* ---- k-Wave HDF5 input / output -------------------------------------------------------------
 * The k-Wave C++ file format the reference writes and reads through CHDF5 (KWaveCpp.writeInput2D,
 * readPFinal2D: Sources/KWave/IO/KWaveInput.swift:12-99; HDF5File: Sources/KWave/IO/HDF5.swift:
 * 26-190), so that `kwave-cli run --input sim.h5 --output results.h5` (REQUIREMENTS.md 8.3) runs
 * on the device. Host-only calls: libhdf5 (>= 1.10) is loaded with dlopen on first use
 * (KWH_HDF5_LIBRARY, the usual sonames, /opt/conda/lib); without it every call returns
 * KWH_EUNSUPPORTED. Layout: uint64 [1,1,1] integer scalars and flags, float32 [1,1,1] real
 * scalars, float32 [Nz,Ny,Nx] fields (x fastest), uint64 [1,1,Ns] sensor_mask_index (1-based,
 * column-major), root attributes file_type / major_version "1" / minor_version "2". */
typedef struct kwh_h5_input kwh_h5_input;

/* KWH_OK when libhdf5 can be loaded, else KWH_EUNSUPPORTED with the reason. */
int kwh_h5_available(void);

/* Write a k-Wave input file for a p0 problem (writeInput2D, KWaveInput.swift:12-99, extended to
 * 3D grids, sound-speed / density maps and B/A): grid, Nt, dt, one PML size and alpha on every
 * axis, c0 / c_ref = max(c0) / rho0 and its staggered copies (scalars or fields), BonA, the p0
 * field as given (the smoothing flags are not stored: the file holds the final p0) and the binary
 * sensor mask. Time-varying sources, absorption, Cartesian sensors and directivity are
 * KWH_EUNSUPPORTED. */
int kwh_h5_write_input(const char* path, const kwh_sim_desc* desc);

/* Read a k-Wave input file (2D or 3D, p0 source, index sensor mask, lossless, optional B/A) into
 * a handle whose descriptor (kwh_h5_input_desc) points into the handle's buffers: float32, no
 * smoothing, record = pFinal; copy it and set record / dtype before kwh_sim_create. The sensor
 * points are in ascending C-order like every kwh_sim_desc; the handle remembers the file's
 * order for kwh_h5_write_output. Files with other sources or flags set: KWH_EUNSUPPORTED. */
int kwh_h5_read_input(const char* path, kwh_h5_input** out);
const kwh_sim_desc* kwh_h5_input_desc(kwh_h5_input* in);
void kwh_h5_input_destroy(kwh_h5_input* in);

/* Write a k-Wave output file from a sim that has run: p_final as [Nz,Ny,Nx] (readPFinal2D,
 * KWaveInput.swift:101-111), and the recorded p / p_max / p_min / p_rms as [1,Nt,Ns] / [1,1,Ns]
 * with the sensor points in the input file's order when `in` is given (else ascending C-order).
 * Values in the sim's precision (float32 sims write float32). */
int kwh_h5_write_output(const char* path, kwh_sim* sim, const kwh_h5_input* in);

/* Read any numeric dataset as float64 (HDF5File.readFloatDataset, HDF5.swift:50-71): dims[ndim]
 * as stored; dst == NULL queries the shape only. */
int kwh_h5_read_dataset(const char* path, const char* name, int64_t dims[8], int32_t* ndim, double* dst,
                        size_t dst_count);

/* Read a root string attribute (HDF5File.readRootStringAttribute, HDF5.swift:149-161). */
int kwh_h5_read_root_attribute(const char* path, const char* name, char* dst, size_t dst_bytes);

#ifdef __cplusplus
}
#endif

#endif /* KWAVE_HIP_H */
