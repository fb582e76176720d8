/*
    polesitter (MI355X) -- FMM N-body forces on AMD gfx950, C99 single header.

    Drop-in for the reference's src/polesitter.h (cited below as H:line):
    the same four functions, types, field orders and result codes, so callers
    that use positional initialisers (ps_config_t cfg = {buf, size, n, theta,
    thrd}) compile unchanged.

        #define PS_MULTITHREADING          // optional, selects the traversal
        #define POLESITTER_IMPLEMENTATION  // in exactly one C or C++ file
        #include "polesitter.h"
        ... link with -lpolesitter_hip (libpolesitter_hip.so, gfx950)

    What differs from the CPU reference:
      * The force computation runs on the current HIP device.  Particle
        arrays may live in device memory (used in place, no copies) or in
        host memory (staged through device memory and copied back).
      * PS_MULTITHREADING no longer starts threads; it selects the same
        traversal roots the reference's threaded build uses (targets at
        depth 2 against the root, H:1772-1801), which decide the interaction
        lists.  Without it the single (root, root) traversal is used
        (H:1888).  thrd_cnt is accepted and ignored.
      * The caller's buffer holds the context (as in H:1705) and keeps the
        reference's size contract (radix scratch 24 B per particle,
        H:1717-1723; node capacity (size - ctx - scratch) / 80, H:1742-1743).
        Device memory for that many particles and nodes is allocated by
        ps_init and released by ps_destroy.
      * Where the reference is silent the drop-in reports an error:
        cnt > max_particles or Morton codes wider than 30 bits -> PS_EINVAL,
        a tree larger than the node capacity -> PS_EOOM (the reference
        truncates the tree, H:930-933), device failures -> PS_EDEVICE.
*/
#ifndef POLESITTER_H
#define POLESITTER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* H:294-300, plus PS_EDEVICE */
typedef enum {
    PS_OK      = 0,
    PS_EOOM    = -1, /* out of memory */
    PS_EINVAL  = -2, /* invalid argument */
    PS_ETHRD   = -3, /* thread creation failed (never returned here) */
    PS_EALLOC  = -4, /* allocation failed */
    PS_EDEVICE = -5  /* HIP device error */
} ps_result_t;

typedef struct ps_context ps_context_t;

/* H:304-310 */
typedef struct {
    void*  buff;          /* host memory provided by the caller */
    size_t buff_size;     /* total size in bytes */
    size_t max_particles; /* maximum capacity for the solver */
    float  theta;         /* MAC threshold, 0.0F defaults to 1.0F */
    size_t thrd_cnt;      /* accepted, unused on the GPU */
} ps_config_t;

/* H:312-322 */
typedef struct {
    float*    x;
    float*    y;
    float*    z;
    float*    mass;
    float*    fx;
    float*    fy;
    float*    fz;
    uint32_t* id;
    size_t    cnt;
} ps_particle_arrs_t;

/* H:325 */
ps_result_t ps_init(ps_context_t** out_ctx, const ps_config_t* cfg);

/* H:328-330: sorts x, y, z, mass, id and morton_codes in place by code,
 * builds the octree under the given root cell and writes fx, fy, fz. */
ps_result_t ps_calc_forces(ps_context_t* ctx, const ps_particle_arrs_t* arrs,
                           uint32_t* morton_codes, float root_cx, float root_cy,
                           float root_cz, float root_hw);

/* H:334-336 */
ps_result_t ps_prepare_particles(ps_particle_arrs_t* arrs,
                                 uint32_t* out_morton_codes, float* out_min_b,
                                 float* out_max_b, float* out_range);

/* H:339 */
ps_result_t ps_destroy(ps_context_t* ctx);

/* ---- extensions (no reference counterpart) ------------------------------ */

/* ps_prepare_particles with an explicit Morton quantisation mode:
 * 0 = the reference's scalar truncation (default), 1 = its AVX2 rounding. */
ps_result_t ps_prepare_particles_ex(ps_particle_arrs_t* arrs,
                                    uint32_t* out_morton_codes,
                                    float* out_min_b, float* out_max_b,
                                    float* out_range, int morton_mode);

/* engine options, see include/polesitter_hip.h (PS_HIP_OPT_*) */
ps_result_t ps_set_option(ps_context_t* ctx, int key, double value);

/* the engine behind a context, for the ps_hip_* entry points */
struct ps_hip_engine;
struct ps_hip_engine* ps_get_engine(ps_context_t* ctx);

/* ps_config_t with caller-owned device memory appended.  ps_config_t itself
 * stays the reference's five fields: appending to it would make every
 * positional initialiser of the reference's callers (tests/main.c:82) miss
 * fields, which -Wextra (CMakeLists.txt:37-45) reports.  With dev_buff set
 * (device memory of at least ps_device_scratch_bytes(max_particles) bytes)
 * ps_init_ex makes no allocation at all: the device buffers of every tick
 * are carved from dev_buff and the solver's host state lives in buff after
 * the context.  NULL: the same as ps_init. */
typedef struct {
    ps_config_t base;
    void*       dev_buff;
    size_t      dev_buff_size;
} ps_config_ex_t;

ps_result_t ps_init_ex(ps_context_t** out_ctx, const ps_config_ex_t* cfg);
size_t      ps_device_scratch_bytes(size_t max_particles);

/* ps_calc_forces enqueued on a HIP stream (NULL: the default stream) without
 * waiting; device arrays only.  ps_calc_forces_wait completes it and returns
 * what ps_calc_forces would have (see ps_hip_calc_forces_async). */
ps_result_t ps_calc_forces_async(ps_context_t* ctx, const ps_particle_arrs_t* arrs,
                                 uint32_t* morton_codes, float root_cx,
                                 float root_cy, float root_cz, float root_hw,
                                 void* stream);
ps_result_t ps_calc_forces_wait(ps_context_t* ctx);

#ifdef __cplusplus
}
#endif

#endif /* POLESITTER_H */

#ifdef POLESITTER_IMPLEMENTATION
#ifndef POLESITTER_IMPLEMENTATION_DONE
#define POLESITTER_IMPLEMENTATION_DONE

#include "polesitter_hip.h"

#ifndef PS_MORTON_MODE
#define PS_MORTON_MODE PS_HIP_MORTON_TRUNC
#endif

/* Compatibility with callers of the reference's internals: its demos clear
 * the node arena before every frame, ps_impl_arena_clear(&ctx->arena)
 * (demo/unary.c:120, demo/binary.c:158; H:878-880).  The device rebuilds
 * its tree in every tick, so clearing only resets this offset. */
typedef struct {
    size_t offset;
} ps_impl_arena_t;

static inline void ps_impl_arena_clear(ps_impl_arena_t* arena) {
    if (arena) {
        arena->offset = 0;
    }
}

struct ps_context {
    struct ps_hip_engine* engine;
    float                 theta;
    size_t                max_particles;
    size_t                max_nodes;
    ps_impl_arena_t       arena;
};

static size_t ps_impl_align_forward(size_t v, size_t align) {
    return (v + (align - 1)) & ~(align - 1);
}

static ps_result_t ps_impl_result(int rc) {
    switch (rc) {
    case PS_HIP_OK: return PS_OK;
    case PS_HIP_EOOM: return PS_EOOM;
    case PS_HIP_EINVAL: return PS_EINVAL;
    default: return PS_EDEVICE;
    }
}

static ps_result_t ps_impl_init(ps_context_t** out_ctx, const ps_config_t* cfg,
                                void* dev_buff, size_t dev_buff_size) {
    /* validation and buffer contract of H:1700-1743 */
    if (!cfg || !cfg->buff || cfg->buff_size < sizeof(ps_context_t) || !out_ctx) {
        return PS_EINVAL;
    }
    ps_context_t* ctx = (ps_context_t*)cfg->buff;
    ctx->engine        = NULL;
    ctx->arena.offset  = 0;
    ctx->theta         = (cfg->theta > 0.0F) ? cfg->theta : 1.0F;
    ctx->max_particles = cfg->max_particles;

    size_t radix_tmp_size = 6 * cfg->max_particles * sizeof(float);
    size_t arena_start    = ps_impl_align_forward(sizeof(ps_context_t), 64);
    if (arena_start + radix_tmp_size > cfg->buff_size) {
        return PS_EOOM;
    }
    size_t usable  = cfg->buff_size - arena_start - radix_tmp_size;
    ctx->max_nodes = usable / (64 + 16);

#ifdef PS_MULTITHREADING
    int traversal = PS_HIP_TRAV_MT;
#else
    int traversal = PS_HIP_TRAV_ST;
#endif
    /* with caller device memory, the engine's host state goes into the
     * part of buff the reference uses for its node arena */
    void* host_mem = dev_buff ? (void*)((char*)cfg->buff + arena_start) : NULL;
    int   rc = ps_hip_engine_create_in(&ctx->engine, cfg->max_particles, ctx->max_nodes,
                                       ctx->theta, traversal, host_mem,
                                       dev_buff ? usable : 0, dev_buff, dev_buff_size);
    if (rc != PS_HIP_OK) {
        ctx->engine = NULL;
        if (dev_buff && rc == PS_HIP_EOOM) {
            return PS_EOOM; /* dev_buff smaller than ps_device_scratch_bytes */
        }
        return rc == PS_HIP_EOOM ? PS_EALLOC : ps_impl_result(rc);
    }
    *out_ctx = ctx;
    return PS_OK;
}

ps_result_t ps_init(ps_context_t** out_ctx, const ps_config_t* cfg) {
    return ps_impl_init(out_ctx, cfg, NULL, 0);
}

ps_result_t ps_init_ex(ps_context_t** out_ctx, const ps_config_ex_t* cfg) {
    if (!cfg) {
        return PS_EINVAL;
    }
    return ps_impl_init(out_ctx, &cfg->base, cfg->dev_buff, cfg->dev_buff_size);
}

size_t ps_device_scratch_bytes(size_t max_particles) {
    return ps_hip_engine_scratch_bytes(max_particles);
}

ps_result_t ps_calc_forces_async(ps_context_t* ctx, const ps_particle_arrs_t* arrs,
                                 uint32_t* morton_codes, float root_cx,
                                 float root_cy, float root_cz, float root_hw,
                                 void* stream) {
    if (!ctx || !arrs || !ctx->engine) {
        return PS_EINVAL;
    }
    if (arrs->cnt == 0) {
        return PS_OK;
    }
    if (!morton_codes || arrs->cnt > ctx->max_particles) {
        return PS_EINVAL;
    }
    return ps_impl_result(ps_hip_calc_forces_async(
        ctx->engine, arrs->x, arrs->y, arrs->z, arrs->mass, arrs->fx, arrs->fy,
        arrs->fz, arrs->id, arrs->cnt, morton_codes, root_cx, root_cy, root_cz,
        root_hw, stream));
}

ps_result_t ps_calc_forces_wait(ps_context_t* ctx) {
    if (!ctx || !ctx->engine) {
        return PS_EINVAL;
    }
    return ps_impl_result(ps_hip_calc_forces_wait(ctx->engine));
}

ps_result_t ps_calc_forces(ps_context_t* ctx, const ps_particle_arrs_t* arrs,
                           uint32_t* morton_codes, float root_cx, float root_cy,
                           float root_cz, float root_hw) {
    if (!ctx || !arrs || !ctx->engine) { /* H:1843 */
        return PS_EINVAL;
    }
    if (arrs->cnt == 0) {
        return PS_OK;
    }
    if (!morton_codes || arrs->cnt > ctx->max_particles) {
        return PS_EINVAL;
    }
    return ps_impl_result(ps_hip_calc_forces(
        ctx->engine, arrs->x, arrs->y, arrs->z, arrs->mass, arrs->fx, arrs->fy,
        arrs->fz, arrs->id, arrs->cnt, morton_codes, root_cx, root_cy, root_cz,
        root_hw));
}

ps_result_t ps_prepare_particles_ex(ps_particle_arrs_t* arrs,
                                    uint32_t* out_morton_codes,
                                    float* out_min_b, float* out_max_b,
                                    float* out_range, int morton_mode) {
    if (!arrs || !out_morton_codes || arrs->cnt == 0) { /* H:1914 */
        return PS_EINVAL;
    }
    return ps_impl_result(ps_hip_prepare(arrs->x, arrs->y, arrs->z, arrs->cnt,
                                         out_morton_codes, out_min_b,
                                         out_max_b, out_range, morton_mode));
}

ps_result_t ps_prepare_particles(ps_particle_arrs_t* arrs,
                                 uint32_t* out_morton_codes, float* out_min_b,
                                 float* out_max_b, float* out_range) {
    return ps_prepare_particles_ex(arrs, out_morton_codes, out_min_b, out_max_b,
                                   out_range, PS_MORTON_MODE);
}

ps_result_t ps_destroy(ps_context_t* ctx) {
    if (!ctx) { /* H:2173 */
        return PS_EINVAL;
    }
    if (ctx->engine) {
        ps_hip_engine_destroy(ctx->engine);
        ctx->engine = NULL;
    }
    return PS_OK;
}

ps_result_t ps_set_option(ps_context_t* ctx, int key, double value) {
    if (!ctx || !ctx->engine) {
        return PS_EINVAL;
    }
    if (key == PS_HIP_OPT_THETA) {
        ctx->theta = value > 0.0 ? (float)value : 1.0F;
    }
    return ps_impl_result(ps_hip_engine_set_option(ctx->engine, key, value));
}

struct ps_hip_engine* ps_get_engine(ps_context_t* ctx) {
    return ctx ? ctx->engine : NULL;
}

#endif /* POLESITTER_IMPLEMENTATION_DONE */
#endif /* POLESITTER_IMPLEMENTATION */
