/* dropin_tick.c -- the reference README's frame loop (README.md:19-110),
 * compiled unchanged against the drop-in header and linked with
 * libpolesitter_hip.so.  Host arrays, as in the README.
 *
 *   dropin_tick N THETA OUT.bin [FRAMES]
 *
 * Inputs are tests/accuracy.c's fill_uniform (LCG seed 7331, cube
 * [-50,50)^3, masses in [0.5,2)).  The first tick's sorted arrays and forces
 * are written to OUT.bin (x, y, z, mass, fx, fy, fz as float32, id, codes as
 * uint32, N values each) for comparison with the oracle; then FRAMES more
 * frames of the README loop run (id reset, prepare, calc_forces, velocity
 * gather by id, semi-implicit Euler). */
#define PS_MULTITHREADING
#define POLESITTER_IMPLEMENTATION
#include "polesitter.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint32_t lcg_next(uint32_t* s) {
    *s = (*s * 1664525U) + 1013904223U;
    return *s;
}

static float rand_unit(uint32_t* s) {
    return (float)(lcg_next(s) >> 8) / 16777215.0F;
}

int main(int argc, char** argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s N THETA OUT.bin [FRAMES]\n", argv[0]);
        return 2;
    }
    size_t n      = (size_t)strtoull(argv[1], NULL, 10);
    float  theta  = (float)atof(argv[2]);
    int    frames = argc > 4 ? atoi(argv[4]) : 3;

    float*    x   = malloc(n * sizeof(float));
    float*    y   = malloc(n * sizeof(float));
    float*    z   = malloc(n * sizeof(float));
    float*    m   = malloc(n * sizeof(float));
    float*    fx  = calloc(n, sizeof(float));
    float*    fy  = calloc(n, sizeof(float));
    float*    fz  = calloc(n, sizeof(float));
    float*    vx  = calloc(n, sizeof(float));
    float*    vy  = calloc(n, sizeof(float));
    float*    vz  = calloc(n, sizeof(float));
    float*    tmp = malloc(3 * n * sizeof(float));
    uint32_t* ids = malloc(n * sizeof(uint32_t));
    uint32_t* mc  = malloc(n * sizeof(uint32_t));
    if (!x || !y || !z || !m || !fx || !fy || !fz || !vx || !vy || !vz || !tmp || !ids || !mc) {
        return 3;
    }

    uint32_t st = 7331U;
    for (size_t i = 0; i < n; ++i) { /* tests/accuracy.c:111-121 */
        x[i] = (rand_unit(&st) * 100.0F) - 50.0F;
        y[i] = (rand_unit(&st) * 100.0F) - 50.0F;
        z[i] = (rand_unit(&st) * 100.0F) - 50.0F;
        m[i] = 0.5F + (rand_unit(&st) * 1.5F);
    }

    size_t        size = (size_t)(16u << 20) + 24 * n + 200 * n;
    void*         buf  = malloc(size);
    ps_context_t* ctx  = NULL;
    ps_config_t   cfg  = {buf, size, n, theta, 4};
    ps_result_t   rc   = ps_init(&ctx, &cfg);
    if (rc != PS_OK) {
        fprintf(stderr, "ps_init failed: %d\n", (int)rc);
        return 4;
    }
    ps_particle_arrs_t arrs = {x, y, z, m, fx, fy, fz, ids, n};

    float dt = 0.016F;
    for (int f = 0; f <= frames; ++f) {
        for (size_t i = 0; i < n; ++i) {
            ids[i] = (uint32_t)i;
            fx[i] = 0.0F; fy[i] = 0.0F; fz[i] = 0.0F;
        }
        float min_b, max_b, range;
        rc = ps_prepare_particles(&arrs, mc, &min_b, &max_b, &range);
        if (rc != PS_OK) { fprintf(stderr, "prepare failed: %d\n", (int)rc); return 5; }
        float root_c  = min_b + (range / 2.0F);
        float root_hw = range / 2.0F;
        rc = ps_calc_forces(ctx, &arrs, mc, root_c, root_c, root_c, root_hw);
        if (rc != PS_OK) { fprintf(stderr, "calc_forces failed: %d\n", (int)rc); return 6; }

        if (f == 0) {
            FILE* o = fopen(argv[3], "wb");
            if (!o) return 7;
            fwrite(x, 4, n, o); fwrite(y, 4, n, o); fwrite(z, 4, n, o); fwrite(m, 4, n, o);
            fwrite(fx, 4, n, o); fwrite(fy, 4, n, o); fwrite(fz, 4, n, o);
            fwrite(ids, 4, n, o); fwrite(mc, 4, n, o);
            fclose(o);
        }
        /* gather the external velocity arrays by the new order */
        for (size_t i = 0; i < n; ++i) {
            uint32_t k = ids[i];
            tmp[i] = vx[k]; tmp[n + i] = vy[k]; tmp[2 * n + i] = vz[k];
        }
        memcpy(vx, tmp, n * 4); memcpy(vy, tmp + n, n * 4); memcpy(vz, tmp + 2 * n, n * 4);
        for (size_t i = 0; i < n; ++i) { /* semi-implicit Euler */
            if (m[i] <= 0.0F) continue;
            vx[i] += (fx[i] / m[i]) * dt;
            vy[i] += (fy[i] / m[i]) * dt;
            vz[i] += (fz[i] / m[i]) * dt;
            x[i] += vx[i] * dt;
            y[i] += vy[i] * dt;
            z[i] += vz[i] * dt;
        }
    }
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) sum += (double)fx[i] + (double)fy[i] + (double)fz[i];
    printf("dropin_tick n=%zu frames=%d force_sum=%.9g\n", n, frames, sum);
    ps_destroy(ctx);
    free(buf);
    return 0;
}
