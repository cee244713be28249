/*
 * kwh_ivp.c -- a plain C caller of libkwavehip (include/kwave_hip.h), with no Python in between:
 * the sequence a Swift or C++ host follows behind `kspaceFirstOrder` (KSpaceFirstOrder.swift:79-107,
 * outputs :47-73) on a 3-D initial-value problem.
 *
 *   kwh_ivp --validate
 *       host-only checks (no device): ABI version, a valid descriptor passes kwh_sim_validate, one
 *       with nt = 0 returns KWH_EINVAL with a message.
 *   kwh_ivp N NT OUT [f32|f64]
 *       N^3 grid, dx = 1e-4 m, c = 1500 m/s, rho = 1000 kg/m^3, PML 10, smoothP0 off,
 *       dt = 0.3 dx / c; p0 = 1 Pa on the cells with (i-N/2)^2 + (j-N/2)^2 + (k-N/2)^2 <= (N/8)^2;
 *       sensor plane x = 3N/4 recording p and pFinal. Writes OUT as
 *       [int64 n_sensor][int64 nt][p series n_sensor*nt][pFinal N^3] in the sim's element type
 *       and prints one JSON line: the rate of steps 2..NT-1 (one kwh_sim_run call, no callback)
 *       and the number of callbacks of steps 0 and 1 that carried the pressure field.
 */
#define _POSIX_C_SOURCE 199309L   /* clock_gettime under -std=c11 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "kwave_hip.h"

static int fail(const char* what, int code) {
    fprintf(stderr, "kwh_ivp: %s failed (%d): %s\n", what, code, kwh_last_error());
    return 1;
}

static void base_desc(kwh_sim_desc* d, int n, int nt, int dtype) {
    memset(d, 0, sizeof *d);
    d->abi_version = KWH_ABI_VERSION;
    d->dim = 3;
    d->n[0] = d->n[1] = d->n[2] = n;
    d->d[0] = d->d[1] = d->d[2] = 1e-4;
    d->dt = 0.3 * 1e-4 / 1500.0;
    d->nt = nt;
    d->dtype = dtype;
    d->pml_size = 10;
    d->pml_alpha = 2.0;
    d->alpha_power = 1.5;
    d->alpha_mode = KWH_ALPHA_NONE;
    d->record = KWH_REC_P | KWH_REC_P_FINAL;
}

static int validate_only(void) {
    static const float c = 1500.0f, rho = 1000.0f;
    static const float p0 = 1.0f;
    kwh_sim_desc d;
    int code;
    if (kwh_abi_version() != KWH_ABI_VERSION) {
        fprintf(stderr, "kwh_ivp: library ABI %d, header %d\n", kwh_abi_version(), KWH_ABI_VERSION);
        return 1;
    }
    base_desc(&d, 32, 10, KWH_FLOAT32);
    d.sound_speed.data = &c;
    d.sound_speed.count = 1;
    d.density.data = &rho;
    d.density.count = 1;
    float* p0map = (float*)calloc((size_t)32 * 32 * 32, sizeof(float));
    if (!p0map) return 1;
    p0map[16 * 32 * 32 + 16 * 32 + 16] = p0;
    d.p0.data = p0map;
    d.p0.count = 32 * 32 * 32;
    code = kwh_sim_validate(&d);
    if (code != KWH_OK) {
        free(p0map);
        return fail("kwh_sim_validate (valid descriptor)", code);
    }
    d.nt = 0;   /* time not set: the reference traps in its solver (makeTime first) */
    code = kwh_sim_validate(&d);
    free(p0map);
    if (code != KWH_EINVAL || kwh_last_error()[0] == '\0') {
        fprintf(stderr, "kwh_ivp: nt = 0 returned %d (message '%s'), expected KWH_EINVAL\n", code,
                kwh_last_error());
        return 1;
    }
    printf("{\"validate\": \"ok\", \"abi\": %d}\n", kwh_abi_version());
    fprintf(stderr, "kwh_ivp: nt = 0 rejected: %s\n", kwh_last_error());
    return 0;
}

/* progress callback (the reference's progress / fieldMonitor closures): counts the calls that
 * carried a pressure field */
static void count_steps(int32_t t, int32_t nt, const void* p_host, void* user) {
    (void)t;
    (void)nt;
    if (p_host) ++*(int*)user;
}

int main(int argc, char** argv) {
    if (argc == 2 && strcmp(argv[1], "--validate") == 0) return validate_only();
    if (argc < 4) {
        fprintf(stderr, "usage: kwh_ivp --validate | kwh_ivp N NT OUT [f32|f64]\n");
        return 2;
    }
    const int n = atoi(argv[1]), nt = atoi(argv[2]);
    const int dtype = (argc > 4 && strcmp(argv[4], "f64") == 0) ? KWH_FLOAT64 : KWH_FLOAT32;
    if (n < 16 || n > 1024 || nt < 1) {
        fprintf(stderr, "kwh_ivp: need 16 <= N <= 1024 and NT >= 1\n");
        return 2;
    }
    const int64_t cells = (int64_t)n * n * n;
    static const float c = 1500.0f, rho = 1000.0f;
    float* p0 = (float*)calloc((size_t)cells, sizeof(float));
    int64_t* sensor = (int64_t*)malloc(sizeof(int64_t) * (size_t)n * n);
    if (!p0 || !sensor) return 1;
    const int h = n / 2, r = n / 8, xs = 3 * n / 4;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                const int di = i - h, dj = j - h, dk = k - h;
                if (di * di + dj * dj + dk * dk <= r * r) p0[((int64_t)i * n + j) * n + k] = 1.0f;
            }
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) sensor[(int64_t)j * n + k] = ((int64_t)xs * n + j) * n + k;

    kwh_sim_desc d;
    base_desc(&d, n, nt, dtype);
    d.sound_speed.data = &c;
    d.sound_speed.count = 1;
    d.density.data = &rho;
    d.density.count = 1;
    d.p0.data = p0;
    d.p0.count = cells;
    d.sensor_index = sensor;
    d.n_sensor = (int64_t)n * n;

    kwh_ctx* ctx = NULL;
    kwh_sim* sim = NULL;
    int code = kwh_ctx_create(0, &ctx);
    if (code != KWH_OK) return fail("kwh_ctx_create", code);
    code = kwh_sim_create(ctx, &d, &sim);
    free(p0);   /* inputs are copied by kwh_sim_create */
    free(sensor);
    if (code != KWH_OK) return fail("kwh_sim_create", code);

    /* the first steps with a callback after every step (p copied to the host each time), the
     * rest in one call without one: runs continue where the previous call stopped */
    int callbacks = 0;
    const int head = nt < 2 ? nt : 2;
    code = kwh_sim_run(sim, 0, head, count_steps, 1, &callbacks);
    if (code != KWH_OK) return fail("kwh_sim_run (with callback)", code);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    code = head < nt ? kwh_sim_run(sim, head, nt, NULL, 0, NULL) : KWH_OK;
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (code != KWH_OK) return fail("kwh_sim_run", code);
    const double sec = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);

    int64_t shp[3];
    int32_t nd = 0;
    code = kwh_sim_output_shape(sim, KWH_OUT_P, shp, &nd);
    if (code != KWH_OK) return fail("kwh_sim_output_shape", code);
    if (nd != 2 || shp[0] != (int64_t)n * n || shp[1] != nt) {
        fprintf(stderr, "kwh_ivp: unexpected p shape (%d: %lld x %lld)\n", nd, (long long)shp[0], (long long)shp[1]);
        return 1;
    }
    const size_t es = dtype == KWH_FLOAT64 ? 8 : 4;
    const size_t series_bytes = (size_t)shp[0] * (size_t)shp[1] * es, final_bytes = (size_t)cells * es;
    void* series = malloc(series_bytes);
    void* pfinal = malloc(final_bytes);
    if (!series || !pfinal) return 1;
    code = kwh_sim_read_output(sim, KWH_OUT_P, series, series_bytes);
    if (code != KWH_OK) return fail("kwh_sim_read_output(p)", code);
    code = kwh_sim_read_output(sim, KWH_OUT_P_FINAL, pfinal, final_bytes);
    if (code != KWH_OK) return fail("kwh_sim_read_output(pFinal)", code);
    kwh_sim_destroy(sim);
    kwh_ctx_destroy(ctx);

    FILE* f = fopen(argv[3], "wb");
    if (!f) {
        fprintf(stderr, "kwh_ivp: cannot open %s\n", argv[3]);
        return 1;
    }
    const int64_t hdr[2] = {shp[0], shp[1]};
    const int ok = fwrite(hdr, sizeof hdr, 1, f) == 1 && fwrite(series, series_bytes, 1, f) == 1 &&
                   fwrite(pfinal, final_bytes, 1, f) == 1;
    fclose(f);
    free(series);
    free(pfinal);
    if (!ok) {
        fprintf(stderr, "kwh_ivp: short write to %s\n", argv[3]);
        return 1;
    }
    printf("{\"grid\": %d, \"nt\": %d, \"dtype\": \"%s\", \"run_s\": %.4f, \"Mcell_steps_per_s\": %.1f, "
           "\"callbacks\": %d}\n",
           n, nt, dtype == KWH_FLOAT64 ? "f64" : "f32", sec, nt > head ? (double)cells * (nt - head) / sec / 1e6 : 0.0, callbacks);
    return 0;
}
