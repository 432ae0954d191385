/*
 * A C99 caller of the C-ABI (include/bcsim_c.h), written the way a C program or
 * a foreign-function binding drives the "gpu" algorithm: create, configure,
 * simulate_lines into a host buffer, read a debug key, destroy.
 *
 * Usage: c_caller <out.bin>
 * Simulates a small fixed-scatterer sector scan and writes the inputs and the
 * IQ frame to out.bin for a test to re-evaluate with the FP64 oracle:
 *   int32 L, S, N
 *   float origins[L][3], dirs[L][3], lats[L][3], times[L]
 *   float xyza[N][4], excitation[55]
 *   complex64 iq[L][S]
 * Exit status 0 on success; a failed call prints bcsim_last_error().
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "bcsim_c.h"

enum { L = 16, N = 4000, NE = 55 };

static uint32_t g_state = 2024u;

static float uniform(float a, float b) {
    g_state = g_state * 1664525u + 1013904223u;
    return a + (b - a) * (float)(g_state >> 8) * (1.0f / 16777216.0f);
}

static int check(bcsim_status st, const char* what) {
    if (st != BCSIM_OK) {
        fprintf(stderr, "%s failed (%d): %s\n", what, st, bcsim_last_error());
        return 1;
    }
    return 0;
}

int main(int argc, char** argv) {
    static float origins[L * 3], dirs[L * 3], lats[L * 3], times[L], xyza[N * 4], exc[NE];
    bcsim_sim* s = NULL;
    float* iq = NULL;
    FILE* f = NULL;
    int lines = 0, samples = 0, i, rc = 1;
    const float fs = 50e6f, fc = 2.5e6f, line_length = 0.04f;
    double memset_ms[4];
    size_t count = 0, total = 0;

    if (argc < 2) {
        fprintf(stderr, "usage: c_caller <out.bin>\n");
        return 2;
    }
    for (i = 0; i < NE; i++) {
        const float t = (float)(i - NE / 2) / fs;
        exc[i] = expf(-(t * t) / (2.0f * 1.6e-7f * 1.6e-7f)) * cosf(6.2831853f * fc * t);
    }
    for (i = 0; i < L; i++) {
        const double a = -0.4 + 0.8 * i / (L - 1);
        origins[3 * i] = origins[3 * i + 1] = origins[3 * i + 2] = 0.0f;
        dirs[3 * i] = (float)sin(a);
        dirs[3 * i + 1] = 0.0f;
        dirs[3 * i + 2] = (float)cos(a);
        lats[3 * i] = (float)cos(a);
        lats[3 * i + 1] = 0.0f;
        lats[3 * i + 2] = (float)-sin(a);
        times[i] = 0.0f;
    }
    for (i = 0; i < N; i++) {
        xyza[4 * i] = uniform(-0.015f, 0.015f);
        xyza[4 * i + 1] = uniform(-0.003f, 0.003f);
        xyza[4 * i + 2] = uniform(0.005f, 0.04f);
        xyza[4 * i + 3] = uniform(-1.0f, 1.0f);
    }

    if (check(bcsim_create("gpu", &s), "bcsim_create")) return 1;
    if (check(bcsim_set_parameter(s, "radial_decimation", "2"), "bcsim_set_parameter") ||
        check(bcsim_set_parameter(s, "store_kernel_details", "on"), "bcsim_set_parameter") ||
        check(bcsim_set_excitation(s, exc, NE, NE / 2, fs, fc), "bcsim_set_excitation") ||
        check(bcsim_set_scan_sequence(s, L, origins, dirs, lats, times, line_length, 0), "bcsim_set_scan_sequence") ||
        check(bcsim_set_fixed_scatterers(s, xyza, N), "bcsim_set_fixed_scatterers") ||
        check(bcsim_set_analytical_profile(s, 1e-3f, 1e-3f), "bcsim_set_analytical_profile") ||
        check(bcsim_get_output_shape(s, &lines, &samples), "bcsim_get_output_shape"))
        goto done;
    total = (size_t)lines * (size_t)samples;
    iq = (float*)malloc(total * 2 * sizeof(float));
    if (!iq) goto done;
    if (check(bcsim_simulate_lines(s, iq, total, &lines, &samples), "bcsim_simulate_lines") ||
        check(bcsim_get_debug_data(s, "kernel_memset_ms", memset_ms, 4, &count), "bcsim_get_debug_data"))
        goto done;
    f = fopen(argv[1], "wb");
    if (!f) goto done;
    {
        const int32_t hdr[3] = {lines, samples, N};
        if (fwrite(hdr, sizeof hdr, 1, f) != 1 || fwrite(origins, sizeof origins, 1, f) != 1 ||
            fwrite(dirs, sizeof dirs, 1, f) != 1 || fwrite(lats, sizeof lats, 1, f) != 1 ||
            fwrite(times, sizeof times, 1, f) != 1 || fwrite(xyza, sizeof xyza, 1, f) != 1 ||
            fwrite(exc, sizeof exc, 1, f) != 1 || fwrite(iq, 2 * sizeof(float), total, f) != total)
            goto done;
    }
    printf("c_caller: %d lines x %d samples, %zu memset timing(s)\n", lines, samples, count);
    rc = 0;
done:
    if (f) fclose(f);
    free(iq);
    if (s && check(bcsim_destroy(s), "bcsim_destroy")) rc = 1;
    return rc;
}
