/*
 * kwh_cli.c -- k-Wave input files on the device from the command line, over the C ABI of
 * libkwavehip (include/kwave_hip.h). The commands are those of the reference's planned CLI
 * (REQUIREMENTS.md 8.3: `kwave-cli run --input sim.h5 --output results.h5`, `validate`, `info`)
 * on the file format of KWaveCpp.writeInput2D / readPFinal2D (Sources/KWave/IO/KWaveInput.swift).
 *
 *   kwh_cli info --input IN.h5
 *       grid, time axis, PML, media (scalar or map), sensor points, as one JSON line.
 *   kwh_cli validate --input IN.h5
 *       reads the file and runs the solver's host-side checks (no device); exit status 0 if valid.
 *   kwh_cli run --input IN.h5 --output OUT.h5 [--p] [--p_max] [--p_min] [--p_rms] [--float64]
 *               [--device N]
 *       runs every step on the device and writes p_final (always) and the requested sensor
 *       fields; prints one JSON line with the step rate.
 */
#define _POSIX_C_SOURCE 199309L   /* clock_gettime under -std=c11 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "kwave_hip.h"

static int fail(const char* what, int code) {
    fprintf(stderr, "kwh_cli: %s failed (%d): %s\n", what, code, kwh_last_error());
    return code == KWH_EUNSUPPORTED ? 3 : 1;
}

static int usage(void) {
    fprintf(stderr,
            "usage: kwh_cli info --input IN.h5\n"
            "       kwh_cli validate --input IN.h5\n"
            "       kwh_cli run --input IN.h5 --output OUT.h5 [--p] [--p_max] [--p_min] [--p_rms] [--float64] "
            "[--device N]\n");
    return 2;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char** argv) {
    if (argc < 2) return usage();
    const char* cmd = argv[1];
    const char *in_path = NULL, *out_path = NULL;
    uint32_t record = KWH_REC_P_FINAL;
    int dtype = KWH_FLOAT32, device = 0;
    for (int i = 2; i < argc; ++i) {
        if (!strcmp(argv[i], "--input") && i + 1 < argc) in_path = argv[++i];
        else if (!strcmp(argv[i], "--output") && i + 1 < argc) out_path = argv[++i];
        else if (!strcmp(argv[i], "--p")) record |= KWH_REC_P;
        else if (!strcmp(argv[i], "--p_max")) record |= KWH_REC_P_MAX;
        else if (!strcmp(argv[i], "--p_min")) record |= KWH_REC_P_MIN;
        else if (!strcmp(argv[i], "--p_rms")) record |= KWH_REC_P_RMS;
        else if (!strcmp(argv[i], "--p_final")) record |= KWH_REC_P_FINAL;
        else if (!strcmp(argv[i], "--float64")) dtype = KWH_FLOAT64;
        else if (!strcmp(argv[i], "--device") && i + 1 < argc) device = atoi(argv[++i]);
        else return usage();
    }
    if (!in_path) return usage();

    kwh_h5_input* in = NULL;
    int code = kwh_h5_read_input(in_path, &in);
    if (code != KWH_OK) return fail("reading the input file", code);
    kwh_sim_desc d = *kwh_h5_input_desc(in);
    d.record = record;
    d.dtype = dtype;

    int rc = 0;
    if (!strcmp(cmd, "info")) {
        printf("{\"dim\": %d, \"n\": [%d, %d, %d], \"d\": [%g, %g, %g], \"dt\": %g, \"nt\": %d, \"pml_size\": %d, "
               "\"pml_alpha\": %g, \"sound_speed\": \"%s\", \"density\": \"%s\", \"nonlinear\": %s, "
               "\"n_sensor\": %lld}\n",
               d.dim, d.n[0], d.n[1], d.n[2], d.d[0], d.d[1], d.d[2], d.dt, d.nt, d.pml_size, d.pml_alpha,
               d.sound_speed.count == 1 ? "scalar" : "map", d.density.count == 1 ? "scalar" : "map",
               d.bona.data ? "true" : "false", (long long)d.n_sensor);
    } else if (!strcmp(cmd, "validate")) {
        code = kwh_sim_validate(&d);
        if (code != KWH_OK) rc = fail("validation", code);
        else printf("{\"valid\": true}\n");
    } else if (!strcmp(cmd, "run")) {
        if (!out_path) {
            kwh_h5_input_destroy(in);
            return usage();
        }
        kwh_ctx* ctx = NULL;
        kwh_sim* sim = NULL;
        double t_run = 0.0;
        if ((code = kwh_ctx_create(device, &ctx)) != KWH_OK) rc = fail("kwh_ctx_create", code);
        else if ((code = kwh_sim_create(ctx, &d, &sim)) != KWH_OK) rc = fail("kwh_sim_create", code);
        else {
            const double t0 = now_s();
            if ((code = kwh_sim_run(sim, 0, d.nt, NULL, 0, NULL)) != KWH_OK) rc = fail("kwh_sim_run", code);
            t_run = now_s() - t0;
            if (!rc && (code = kwh_h5_write_output(out_path, sim, in)) != KWH_OK) rc = fail("writing the output file", code);
        }
        if (!rc) {
            const double cells = (double)d.n[0] * d.n[1] * (d.dim == 3 ? d.n[2] : 1);
            printf("{\"output\": \"%s\", \"steps\": %d, \"run_s\": %.6f, \"Mcell_steps_per_s\": %.1f}\n", out_path, d.nt,
                   t_run, cells * d.nt / (t_run > 0 ? t_run : 1e-9) / 1e6);
        }
        if (sim) kwh_sim_destroy(sim);
        if (ctx) kwh_ctx_destroy(ctx);
    } else {
        rc = usage();
    }
    kwh_h5_input_destroy(in);
    return rc;
}
