/*
 * ref_driver.c -- thin C-ABI around the reference's own src/polesitter.h,
 * compiled by oracle/Makefile into oracle/_ref/ (never into the product).
 *
 * TEST INFRASTRUCTURE ONLY.  It lets the tests and the golden-fixture script
 * run the genuine reference tick on chosen inputs, dump its octree in the
 * canonical level order, and time its CPU path for the baseline.  The
 * reference header is included from the reference tree by -I path; no
 * reference source lives in this repository.
 */
#define _POSIX_C_SOURCE 199309L
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define POLESITTER_IMPLEMENTATION
#include "polesitter.h"

typedef struct {
    uint32_t  cap;
    uint32_t  n_nodes;
    uint8_t*  depth;
    uint8_t*  leaf;
    uint32_t* first;
    uint32_t* cnt;
    uint32_t* parent;
    float*    geom; /* 4 per node: x, y, z, half_width */
    float*    mp;   /* 4 per node */
    float*    loc;  /* 4 per node */
} ref_dump_t;

static void* ref_buffer(size_t size) {
    /* zero-filled: the reference leaves its spinlock uninitialised when it
     * starts no worker thread (H:750-758) */
    void* b = malloc(size);
    if (b) memset(b, 0, size);
    return b;
}

int ref_prepare(size_t n, float* x, float* y, float* z, uint32_t* codes,
                float* out3) {
    ps_particle_arrs_t arrs = {x, y, z, NULL, NULL, NULL, NULL, NULL, n};
    return (int)ps_prepare_particles(&arrs, codes, &out3[0], &out3[1],
                                     &out3[2]);
}

/* breadth-first walk in child-slot order == (depth, Morton prefix) order */
static int dump_tree(ps_context_t* ctx, ref_dump_t* d) {
    if (!ctx->root) return 0;
    size_t      cap   = d->cap;
    ps_node_t** queue = (ps_node_t**)malloc((cap + 1) * sizeof(ps_node_t*));
    uint32_t*   par   = (uint32_t*)malloc((cap + 1) * sizeof(uint32_t));
    uint8_t*    dep   = (uint8_t*)malloc(cap + 1);
    size_t      head = 0, tail = 0;
    queue[tail] = ctx->root; par[tail] = 0xFFFFFFFFu; dep[tail] = 0; ++tail;
    while (head < tail) {
        ps_node_t* n = queue[head];
        size_t     k = head++;
        if (k >= cap) { free(queue); free(par); free(dep); return -1; }
        int leaf = (int)(n->data.leaf.is_leaf & 1);
        d->depth[k]  = dep[k];
        d->leaf[k]   = (uint8_t)leaf;
        d->parent[k] = par[k];
        d->geom[4 * k + 0] = n->x;
        d->geom[4 * k + 1] = n->y;
        d->geom[4 * k + 2] = n->z;
        d->geom[4 * k + 3] = n->half_width;
        const float* l = ps_impl_get_local(ctx, n);
        for (int c = 0; c < 4; ++c) {
            d->mp[4 * k + c]  = n->multipole[c];
            d->loc[4 * k + c] = l[c];
        }
        if (leaf) {
            d->first[k] = n->data.leaf.first_particle_idx;
            d->cnt[k]   = n->data.leaf.particle_cnt;
        } else {
            d->first[k] = 0xFFFFFFFFu;
            d->cnt[k]   = 0;
            for (int s = 0; s < 8; ++s) {
                ps_node_t* c = ps_impl_get_node(ctx, n->data.children_offs[s]);
                if (!c) continue;
                if (tail >= cap) { free(queue); free(par); free(dep); return -1; }
                queue[tail] = c; par[tail] = (uint32_t)k;
                dep[tail] = (uint8_t)(dep[k] + 1); ++tail;
            }
        }
    }
    d->n_nodes = (uint32_t)tail;
    free(queue); free(par); free(dep);
    return (int)tail;
}

/* One public-API tick exactly as tests/accuracy.c:486-503 does it:
 * ps_init, ps_prepare_particles, ps_calc_forces(root = min + range/2),
 * then (optionally) dump the octree.  Returns the last ps_result_t, or the
 * node count when a dump was asked for. */
int ref_tick(size_t n, float* x, float* y, float* z, float* m, uint32_t* id,
             uint32_t* codes, float* fx, float* fy, float* fz, float theta,
             size_t thrd_cnt, size_t buff_size, float* out3, ref_dump_t* d) {
    void* buff = ref_buffer(buff_size);
    if (!buff) return -100;
    ps_context_t* ctx = NULL;
    ps_config_t   cfg = {buff, buff_size, n, theta, thrd_cnt};
    int           res = (int)ps_init(&ctx, &cfg);
    if (res != PS_OK) { free(buff); return res; }
    ps_particle_arrs_t arrs = {x, y, z, m, fx, fy, fz, id, n};
    res = (int)ps_prepare_particles(&arrs, codes, &out3[0], &out3[1], &out3[2]);
    if (res == PS_OK) {
        float rc = out3[0] + (out3[2] * 0.5F);
        res = (int)ps_calc_forces(ctx, &arrs, codes, rc, rc, rc, out3[2] * 0.5F);
    }
    if (res == PS_OK && d) {
        res = dump_tree(ctx, d);
    }
    ps_destroy(ctx);
    free(buff);
    return res;
}

/* calc_forces only, with caller-chosen codes and root cell */
int ref_calc(size_t n, float* x, float* y, float* z, float* m, uint32_t* id,
             uint32_t* codes, float* fx, float* fy, float* fz, float cx,
             float cy, float cz, float hw, float theta, size_t thrd_cnt,
             size_t buff_size, ref_dump_t* d) {
    void* buff = ref_buffer(buff_size);
    if (!buff) return -100;
    ps_context_t* ctx = NULL;
    ps_config_t   cfg = {buff, buff_size, n, theta, thrd_cnt};
    int           res = (int)ps_init(&ctx, &cfg);
    if (res != PS_OK) { free(buff); return res; }
    ps_particle_arrs_t arrs = {x, y, z, m, fx, fy, fz, id, n};
    res = (int)ps_calc_forces(ctx, &arrs, codes, cx, cy, cz, hw);
    if (res == PS_OK && d) res = dump_tree(ctx, d);
    ps_destroy(ctx);
    free(buff);
    return res;
}

static double wall(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + (double)ts.tv_nsec * 1e-9;
}

/* CPU-baseline timing in the window of tests/benchmark.c:175-184:
 * ps_prepare_particles + ps_calc_forces per tick, ps_init outside.  Ids are
 * reset before each tick (README.md:96-100), outside the window. */
int ref_time_ticks(size_t n, float* x, float* y, float* z, float* m,
                   uint32_t* id, uint32_t* codes, float* fx, float* fy,
                   float* fz, float theta, size_t thrd_cnt, size_t buff_size,
                   int n_ticks, double* out_secs) {
    void* buff = ref_buffer(buff_size);
    if (!buff) return -100;
    ps_context_t* ctx = NULL;
    ps_config_t   cfg = {buff, buff_size, n, theta, thrd_cnt};
    int           res = (int)ps_init(&ctx, &cfg);
    if (res != PS_OK) { free(buff); return res; }
    ps_particle_arrs_t arrs = {x, y, z, m, fx, fy, fz, id, n};
    for (int t = 0; t < n_ticks && res == PS_OK; ++t) {
        for (size_t i = 0; i < n; ++i) id[i] = (uint32_t)i;
        double t0 = wall();
        float  lo, hi, range;
        res = (int)ps_prepare_particles(&arrs, codes, &lo, &hi, &range);
        if (res == PS_OK) {
            float rc = lo + (range / 2.0F);
            res = (int)ps_calc_forces(ctx, &arrs, codes, rc, rc, rc, range / 2.0F);
        }
        out_secs[t] = wall() - t0;
    }
    ps_destroy(ctx);
    free(buff);
    return res;
}
