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

#define KWH_ABI_VERSION 1

typedef enum {
    KWH_OK = 0,
    KWH_EINVAL = -1,        /* invalid argument / descriptor (reference: precondition traps) */
    KWH_ENOMEM = -2,        /* device allocation failed */
    KWH_EDEVICE = -3,       /* HIP runtime error or no device */
    KWH_ENONFINITE = -4,    /* non-finite pressure detected by the optional check */
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
 * signal is [rows][len], rows == 1 (broadcast) or rows == n_points. */
typedef struct {
    const int64_t* index;     /* ascending flat indices of mask nonzeros; NULL = absent */
    int64_t n_points;
    const float* signal;      /* NULL = absent */
    int32_t rows;
    int32_t len;
    int32_t mode;             /* kwh_source_mode */
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

/* Called between step chunks. t is the last completed step index, p_host the pressure field
 * copied to host when the callback is due (every cb_interval steps), else NULL. Replaces the
 * reference's progress / fieldMonitor closures (KSpaceFirstOrder.swift:30-37). */
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

/* Advance steps [t_begin, t_end) (the hot loop, KSpaceFirstOrder.swift:309-342, 492-546,
 * 704-751). Runs must be contiguous: t_begin equals the number of steps already run.
 * cb may be NULL. Returns after the device has finished the requested steps. */
int kwh_sim_run(kwh_sim* sim, int32_t t_begin, int32_t t_end, kwh_step_cb cb, int32_t cb_interval,
                void* user);

/* Enqueue steps [t_begin, t_end) without waiting (for timing harnesses); kwh_sim_sync waits. */
int kwh_sim_launch(kwh_sim* sim, int32_t t_begin, int32_t t_end);
int kwh_sim_sync(kwh_sim* sim);

/* Shape of an output after the run (finalizeRecording, Recording.swift:129-168).
 * Returns KWH_EINVAL if the field was not recorded. */
int kwh_sim_output_shape(kwh_sim* sim, int32_t field, int64_t shape[3], int32_t* ndim);

/* Copy an output (kwh_output) into dst (element type = sim dtype). */
int kwh_sim_read_output(kwh_sim* sim, int32_t field, void* dst, size_t dst_bytes);

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

/* Capture one time step into a hipGraph and replay it for subsequent steps (1 = on). */
int kwh_sim_set_graph(kwh_sim* sim, int enable);

void kwh_sim_destroy(kwh_sim* sim);

#ifdef __cplusplus
}
#endif

#endif /* KWAVE_HIP_H */
