/* demo_headless.c -- the frame loop of the reference's raylib demos
 * (demo/unary.c:117-184, demo/binary.c:135-212) without the window, compiled
 * against the drop-in header and linked with libpolesitter_hip.so.  Host
 * arrays and the demos' own calls, including their use of the reference's
 * internals: ps_impl_arena_clear(&ctx->arena) before every frame.
 *
 *   demo_unary  OUT.bin [FRAMES]   40,000 particles, one black hole pinned at
 *                                  the origin, PS_MULTITHREADING, 4 threads
 *   demo_binary OUT.bin [FRAMES]   10,000 particles around two black holes,
 *                                  single-threaded traversal, 1 thread
 *
 * raylib's GetRandomValue(lo, hi) is replaced by rand() % (hi - lo + 1) + lo
 * after srand(1), and GetFrameTime() by a fixed 1/60 s.  OUT.bin receives
 * the first frame's inputs (x, y, z, mass as float32, before the tick) and
 * then its sorted arrays and forces (x, y, z, mass, fx, fy, fz as float32,
 * id, codes as uint32) for comparison with the oracle; the last line printed
 * summarises the final state. */
#ifdef PS_DEMO_UNARY
#define PS_MULTITHREADING
#endif
#define POLESITTER_IMPLEMENTATION
#include "polesitter.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#ifdef PS_DEMO_UNARY
#define PARTICLE_COUNT 40000
#define THREADS        4
#else
#define PARTICLE_COUNT 10000
#define THREADS        1
#endif
#define BLACKHOLE_MASS 1000.0F
#define ARENA_SIZE     (1024ULL * 1024ULL * 256ULL)
#define THETA          2.0F
#define DEG_TO_RAD     0.017453292519943295F

static float    vx[PARTICLE_COUNT], vy[PARTICLE_COUNT], vz[PARTICLE_COUNT];
static float    px[PARTICLE_COUNT], py[PARTICLE_COUNT], pz[PARTICLE_COUNT];
static float    mass[PARTICLE_COUNT];
static float    fx[PARTICLE_COUNT], fy[PARTICLE_COUNT], fz[PARTICLE_COUNT];
static float    tmp_vx[PARTICLE_COUNT], tmp_vy[PARTICLE_COUNT], tmp_vz[PARTICLE_COUNT];
static uint32_t morton_codes[PARTICLE_COUNT];
static uint32_t id[PARTICLE_COUNT];

static int random_value(int lo, int hi) { return (rand() % (hi - lo + 1)) + lo; }

#ifdef PS_DEMO_UNARY
static void init_particles(void) { /* demo/unary.c:71-98 */
    for (int i = 0; i < PARTICLE_COUNT; i++) {
        float angle  = (float)random_value(0, 360) * DEG_TO_RAD;
        float radius = (float)random_value(1, 1000) / 1000.0F;
        radius       = radius * radius * 30.0F;
        px[i] = cosf(angle) * radius;
        py[i] = ((float)random_value(-100, 100) / 100.0F) * (2.0F / (radius + 1.0F));
        pz[i] = sinf(angle) * radius;
        mass[i] = 0.1F;
        float v_mag = sqrtf(BLACKHOLE_MASS / (radius + 1.0F));
        vx[i] = -sinf(angle) * v_mag;
        vy[i] = 0.0F;
        vz[i] = cosf(angle) * v_mag;
    }
    px[0] = py[0] = pz[0] = 0.0F;
    vx[0] = vy[0] = vz[0] = 0.0F;
    mass[0] = BLACKHOLE_MASS;
}
#else
static void init_particles(void) { /* demo/binary.c:71-113 */
    px[0] = -15.0F; py[0] = 0.0F; pz[0] = 0.0F;
    vx[0] = 0.0F; vy[0] = 0.0F; vz[0] = 3.0F;
    mass[0] = BLACKHOLE_MASS;
    px[1] = 15.0F; py[1] = 0.0F; pz[1] = 0.0F;
    vx[1] = 0.0F; vy[1] = 0.0F; vz[1] = -3.0F;
    mass[1] = BLACKHOLE_MASS;
    int half_p = PARTICLE_COUNT / 2;
    for (int i = 2; i < PARTICLE_COUNT; i++) {
        int   bh     = (i < half_p) ? 0 : 1;
        float angle  = (float)random_value(0, 360) * DEG_TO_RAD;
        float radius = (float)random_value(1, 1000) / 1000.0F;
        radius       = radius * radius * 15.0F;
        px[i] = px[bh] + cosf(angle) * radius;
        py[i] = py[bh] + ((float)random_value(-100, 100) / 100.0F) * (1.0F / (radius + 1.0F));
        pz[i] = pz[bh] + sinf(angle) * radius;
        mass[i] = 0.1F;
        float v_mag = sqrtf(BLACKHOLE_MASS / (radius + 1.0F));
        vx[i] = vx[bh] - sinf(angle) * v_mag;
        vy[i] = vy[bh];
        vz[i] = vz[bh] + cosf(angle) * v_mag;
    }
}
#endif

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s OUT.bin [FRAMES]\n", argv[0]);
        return 2;
    }
    int frames = argc > 2 ? atoi(argv[2]) : 60;
    void* buffer = malloc(ARENA_SIZE);
    if (!buffer) return 3;
    ps_context_t* ctx = NULL;
    ps_config_t   cfg = {buffer, ARENA_SIZE, PARTICLE_COUNT, THETA, THREADS};
    if (ps_init(&ctx, &cfg) != PS_OK) return 4;
    ps_particle_arrs_t arrs = {px, py, pz, mass, fx, fy, fz, id, PARTICLE_COUNT};
    srand(1);
    init_particles();

    FILE* out = fopen(argv[1], "wb");
    if (!out) return 7;
    fwrite(px, 4, PARTICLE_COUNT, out); fwrite(py, 4, PARTICLE_COUNT, out);
    fwrite(pz, 4, PARTICLE_COUNT, out); fwrite(mass, 4, PARTICLE_COUNT, out);

    const float dt = 1.0F / 60.0F;
    for (int f = 0; f < frames; ++f) {
        ps_impl_arena_clear(&ctx->arena); /* demo/unary.c:120, demo/binary.c:158 */
        for (int i = 0; i < PARTICLE_COUNT; ++i) {
            id[i] = (uint32_t)i;
            fx[i] = fy[i] = fz[i] = 0.0F;
        }
        float min_b, max_b, range;
        if (ps_prepare_particles(&arrs, morton_codes, &min_b, &max_b, &range) != PS_OK) return 5;
        float root_c  = min_b + (range / 2.0F);
        float root_hw = range / 2.0F;
        if (ps_calc_forces(ctx, &arrs, morton_codes, root_c, root_c, root_c, root_hw) != PS_OK)
            return 6;
        if (f == 0) {
            fwrite(px, 4, PARTICLE_COUNT, out); fwrite(py, 4, PARTICLE_COUNT, out);
            fwrite(pz, 4, PARTICLE_COUNT, out); fwrite(mass, 4, PARTICLE_COUNT, out);
            fwrite(fx, 4, PARTICLE_COUNT, out); fwrite(fy, 4, PARTICLE_COUNT, out);
            fwrite(fz, 4, PARTICLE_COUNT, out); fwrite(id, 4, PARTICLE_COUNT, out);
            fwrite(morton_codes, 4, PARTICLE_COUNT, out);
            fclose(out);
        }
        for (int i = 0; i < PARTICLE_COUNT; ++i) { /* velocities follow the sort */
            uint32_t k = id[i];
            tmp_vx[i] = vx[k]; tmp_vy[i] = vy[k]; tmp_vz[i] = vz[k];
        }
        for (int i = 0; i < PARTICLE_COUNT; ++i) {
            vx[i] = tmp_vx[i]; vy[i] = tmp_vy[i]; vz[i] = tmp_vz[i];
        }
        for (int i = 0; i < PARTICLE_COUNT; i++) {
#ifdef PS_DEMO_UNARY
            if (mass[i] == BLACKHOLE_MASS) { /* pinned at the origin */
                vx[i] = vy[i] = vz[i] = 0.0F;
                px[i] = py[i] = pz[i] = 0.0F;
                continue;
            }
#endif
            if (mass[i] <= 0.0F) continue;
            vx[i] += (fx[i] / mass[i]) * dt;
            vy[i] += (fy[i] / mass[i]) * dt;
            vz[i] += (fz[i] / mass[i]) * dt;
            px[i] += vx[i] * dt;
            py[i] += vy[i] * dt;
            pz[i] += vz[i] * dt;
        }
    }
    double rad = 0.0, vmax = 0.0;
    int    finite = 1, holes = 0;
    for (int i = 0; i < PARTICLE_COUNT; ++i) {
        double r = sqrt((double)px[i] * px[i] + (double)py[i] * py[i] + (double)pz[i] * pz[i]);
        double v = sqrt((double)vx[i] * vx[i] + (double)vy[i] * vy[i] + (double)vz[i] * vz[i]);
        if (!isfinite(r) || !isfinite(v)) finite = 0;
        rad += r;
        if (v > vmax) vmax = v;
        if (mass[i] == BLACKHOLE_MASS) holes++;
    }
    printf("demo n=%d frames=%d finite=%d black_holes=%d mean_radius=%.6g max_speed=%.6g\n",
           PARTICLE_COUNT, frames, finite, holes, rad / PARTICLE_COUNT, vmax);
    ps_destroy(ctx);
    free(buffer);
    return finite ? 0 : 8;
}
