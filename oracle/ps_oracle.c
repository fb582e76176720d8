/*
 * ps_oracle.c -- CPU restatement of the polesitter FMM tick (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker.  The product (the HIP library)
 * never links or calls it.
 *
 * Build with -O2 -ffp-contract=off -fno-fast-math (oracle/Makefile): every
 * float expression below keeps the reference's operation order, so the
 * results are bit-identical to a strict, non-AVX2 build of
 * src/polesitter.h.  Citations are H:line into src/polesitter.h.
 */
#include "ps_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Morton helpers, H:777-799                                           */
/* ------------------------------------------------------------------ */

uint32_t or_expand_bits(uint32_t v) {
    /* spread the 10 low bits of v to every third bit (bit k -> bit 3k) */
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

uint32_t or_morton_encode(uint32_t x, uint32_t y, uint32_t z) {
    /* x -> bit 0, y -> bit 1, z -> bit 2 of every 3-bit group, H:797-798 */
    return or_expand_bits(x) | (or_expand_bits(y) << 1) |
           (or_expand_bits(z) << 2);
}

static int clamp1023(int q) {
    if (q < 0) {
        q = 0;
    }
    if (q > 1023) {
        q = 1023;
    }
    return q;
}

/* ------------------------------------------------------------------ */
/* ps_prepare_particles, H:1911-2170                                   */
/* ------------------------------------------------------------------ */

int or_prepare(size_t n, const float* x, const float* y, const float* z,
               uint32_t* codes, int mode, float* out_min, float* out_max,
               float* out_range) {
    if (!x || !y || !z || !codes || n == 0) {
        return -2; /* H:1914 */
    }

    /* joint min/max over x, y and z: one cube, H:1918-2006 */
    float lo = FLT_MAX;
    float hi = -FLT_MAX;
    for (size_t i = 0; i < n; ++i) {
        if (x[i] < lo) lo = x[i];
        if (y[i] < lo) lo = y[i];
        if (z[i] < lo) lo = z[i];
        if (x[i] > hi) hi = x[i];
        if (y[i] > hi) hi = y[i];
        if (z[i] > hi) hi = z[i];
    }

    float range = hi - lo; /* H:2008-2011 */
    if (range < 0.001F) {
        range = 0.001F;
    }

    size_t i = 0;
    if (mode == OR_MORTON_AVX2) {
        /* 8-wide blocks: (v-min)*fl(1023/range), clamped, then rounded to
         * nearest-even by the int conversion, H:2015-2072 */
        float scale = 1023.0F / range;
        i           = n - (n % 8u); /* whole 8-blocks: 8*floor(n/8) */
        for (size_t k = 0; k < i; ++k) {
            float v[3] = {(x[k] - lo) * scale, (y[k] - lo) * scale,
                          (z[k] - lo) * scale};
            uint32_t q[3];
            for (int a = 0; a < 3; ++a) {
                float t = v[a] < 1023.0F ? v[a] : 1023.0F; /* min(v,1023) */
                t       = 0.0F > t ? 0.0F : t;             /* max(0,v) */
                q[a]    = (uint32_t)(int)nearbyintf(t);
            }
            codes[k] = or_morton_encode(q[0], q[1], q[2]);
        }
    }

    /* scalar path: trunc(((v-min)/range)*1023), clamped, H:2132-2157 */
    for (; i < n; ++i) {
        int qx = (int)(((x[i] - lo) / range) * 1023.0F);
        int qy = (int)(((y[i] - lo) / range) * 1023.0F);
        int qz = (int)(((z[i] - lo) / range) * 1023.0F);
        codes[i] = or_morton_encode((uint32_t)clamp1023(qx),
                                    (uint32_t)clamp1023(qy),
                                    (uint32_t)clamp1023(qz));
    }

    if (out_min) *out_min = lo;
    if (out_max) *out_max = hi;
    if (out_range) *out_range = range;
    return 0;
}

/* ------------------------------------------------------------------ */
/* stable sort by key, H:1588-1693                                     */
/* ------------------------------------------------------------------ */

void or_sort(size_t n, uint32_t* codes, uint32_t* id, float* x, float* y,
             float* z, float* m) {
    /* Any stable sort yields the reference's permutation; this one is a
     * byte-wise LSD counting sort, like the reference's 4 x 8-bit passes. */
    if (n == 0) {
        return;
    }
    uint32_t* k2 = (uint32_t*)malloc(n * sizeof(uint32_t));
    uint32_t* i2 = (uint32_t*)malloc(n * sizeof(uint32_t));
    float*    x2 = (float*)malloc(n * sizeof(float));
    float*    y2 = (float*)malloc(n * sizeof(float));
    float*    z2 = (float*)malloc(n * sizeof(float));
    float*    m2 = (float*)malloc(n * sizeof(float));

    uint32_t* ks = codes; uint32_t* is = id;
    float* xs = x; float* ys = y; float* zs = z; float* ms = m;
    uint32_t* kd = k2; uint32_t* idd = i2;
    float* xd = x2; float* yd = y2; float* zd = z2; float* md = m2;

    for (int pass = 0; pass < 4; ++pass) {
        size_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        int shift = pass * 8;
        for (size_t i = 0; i < n; ++i) {
            cnt[((ks[i] >> shift) & 0xFFu) + 1]++;
        }
        for (int b = 0; b < 256; ++b) {
            cnt[b + 1] += cnt[b];
        }
        for (size_t i = 0; i < n; ++i) {
            size_t p = cnt[(ks[i] >> shift) & 0xFFu]++;
            kd[p] = ks[i]; idd[p] = is[i];
            xd[p] = xs[i]; yd[p] = ys[i]; zd[p] = zs[i]; md[p] = ms[i];
        }
        uint32_t* tk = ks; ks = kd; kd = tk;
        uint32_t* ti = is; is = idd; idd = ti;
        float* t;
        t = xs; xs = xd; xd = t;
        t = ys; ys = yd; yd = t;
        t = zs; zs = zd; zd = t;
        t = ms; ms = md; md = t;
    }
    /* four passes: the data is back in the caller's arrays (H:1688-1690) */
    free(k2); free(i2); free(x2); free(y2); free(z2); free(m2);
}

/* ------------------------------------------------------------------ */
/* octree build, H:882-967                                             */
/* ------------------------------------------------------------------ */

typedef struct {
    float    x, y, z, hw;
    uint32_t first, end, depth;
    int      leaf;
    int32_t  kid[8];
} dfs_node_t;

typedef struct {
    dfs_node_t* v;
    size_t      len, cap;
} dfs_pool_t;

static int32_t dfs_alloc(dfs_pool_t* p) {
    if (p->len == p->cap) {
        size_t nc = p->cap ? p->cap * 2 : 1024;
        p->v      = (dfs_node_t*)realloc(p->v, nc * sizeof(dfs_node_t));
        p->cap    = nc;
    }
    memset(&p->v[p->len], 0, sizeof(dfs_node_t));
    for (int i = 0; i < 8; ++i) {
        p->v[p->len].kid[i] = -1;
    }
    return (int32_t)p->len++;
}

/* binary search for the first index in [start,end) whose octant at this
 * shift is >= target_oct, H:886-903 */
static size_t find_split(const uint32_t* codes, size_t start, size_t end,
                         int shift, uint32_t target_oct) {
    size_t left = start, right = end;
    while (left < right) {
        size_t   mid = left + ((right - left) / 2);
        uint32_t oct = (codes[mid] >> shift) & 0x7u;
        if (oct < target_oct) {
            left = mid + 1;
        } else {
            right = mid;
        }
    }
    return left;
}

static void build_rec(dfs_pool_t* p, int32_t node, uint32_t depth,
                      size_t start, size_t end, const uint32_t* codes) {
    dfs_node_t* nd = &p->v[node];
    nd->first      = (uint32_t)start;
    nd->end        = (uint32_t)end;
    nd->depth      = depth;
    if (depth == OR_MAX_DEPTH || end - start <= 1) { /* H:911-916 */
        nd->leaf = 1;
        return;
    }
    int    shift = (int)(27 - (depth * 3)); /* H:919 */
    size_t curr  = start;
    for (uint32_t oct = 0; oct < 8; ++oct) {
        size_t oct_end = find_split(codes, curr, end, shift, oct + 1);
        if (oct_end > curr) {
            int32_t c = dfs_alloc(p);
            nd        = &p->v[node]; /* the pool may have moved */
            float hw  = nd->hw * 0.5F; /* H:935-939 */
            p->v[c].x  = nd->x + ((oct & 1) ? hw : -hw);
            p->v[c].y  = nd->y + ((oct & 2) ? hw : -hw);
            p->v[c].z  = nd->z + ((oct & 4) ? hw : -hw);
            p->v[c].hw = hw;
            nd->kid[oct] = c;
            build_rec(p, c, depth + 1, curr, oct_end, codes);
            nd = &p->v[node];
        }
        curr = oct_end;
        if (curr == end) {
            break;
        }
    }
}

long or_build(size_t n, const uint32_t* codes, float cx, float cy, float cz,
              float hw, or_nodes_t* nodes) {
    dfs_pool_t p = {0, 0, 0};
    int32_t    r = dfs_alloc(&p);
    p.v[r].x     = cx;
    p.v[r].y     = cy;
    p.v[r].z     = cz;
    p.v[r].hw    = hw;
    build_rec(&p, r, 0, 0, n, codes);

    if (p.len > nodes->cap) {
        free(p.v);
        return -1;
    }

    /* breadth-first renumbering with children in slot order gives the
     * canonical (depth, Morton prefix) order */
    int32_t*  queue = (int32_t*)malloc(p.len * sizeof(int32_t));
    uint32_t* canon = (uint32_t*)malloc(p.len * sizeof(uint32_t));
    size_t    head = 0, tail = 0;
    queue[tail++] = r;
    while (head < tail) {
        int32_t v  = queue[head];
        canon[v]   = (uint32_t)head;
        ++head;
        for (int s = 0; s < 8; ++s) {
            if (p.v[v].kid[s] >= 0) {
                queue[tail++] = p.v[v].kid[s];
            }
        }
    }

    for (int d = 0; d < 12; ++d) {
        nodes->level_start[d] = (uint32_t)p.len;
    }
    uint32_t maxd = 0;
    for (size_t k = p.len; k-- > 0;) {
        const dfs_node_t* v = &p.v[queue[k]];
        nodes->level_start[v->depth] = (uint32_t)k;
    }
    for (int d = 10; d >= 0; --d) { /* empty levels start where the next does */
        if (nodes->level_start[d] > nodes->level_start[d + 1]) {
            nodes->level_start[d] = nodes->level_start[d + 1];
        }
    }

    for (size_t k = 0; k < p.len; ++k) {
        const dfs_node_t* v = &p.v[queue[k]];
        nodes->x[k]     = v->x;
        nodes->y[k]     = v->y;
        nodes->z[k]     = v->z;
        nodes->hw[k]    = v->hw;
        nodes->first[k] = v->first;
        nodes->end[k]   = v->end;
        nodes->depth[k] = (uint8_t)v->depth;
        nodes->parent[k] = OR_NONE;
        if (v->depth > maxd) maxd = v->depth;
        uint32_t mask = 0, fc = OR_NONE;
        for (int s = 0; s < 8; ++s) {
            if (v->kid[s] >= 0) {
                mask |= 1u << s;
                if (fc == OR_NONE) fc = canon[v->kid[s]];
            }
        }
        nodes->mask[k]  = v->leaf ? 0u : mask;
        nodes->child[k] = v->leaf ? OR_NONE : fc;
        for (int c = 0; c < 4; ++c) {
            nodes->mp[4 * k + c]  = 0.0F;
            nodes->loc[4 * k + c] = 0.0F;
        }
    }
    for (size_t k = 0; k < p.len; ++k) {
        if (nodes->mask[k]) {
            uint32_t c = nodes->child[k];
            for (uint32_t j = 0; j < (uint32_t)__builtin_popcount(nodes->mask[k]); ++j) {
                nodes->parent[c + j] = (uint32_t)k;
            }
        }
    }
    nodes->n_nodes   = (uint32_t)p.len;
    nodes->max_depth = maxd;

    free(queue);
    free(canon);
    free(p.v);
    return (long)nodes->n_nodes;
}

static inline int is_leaf(const or_nodes_t* t, uint32_t n) {
    return t->mask[n] == 0;
}

/* child in slot s, or OR_NONE */
static inline uint32_t child_at(const or_nodes_t* t, uint32_t n, int s) {
    uint32_t m = t->mask[n];
    if (!(m & (1u << s))) {
        return OR_NONE;
    }
    return t->child[n] + (uint32_t)__builtin_popcount(m & ((1u << s) - 1u));
}

/* ------------------------------------------------------------------ */
/* upward pass, H:969-1046                                             */
/* ------------------------------------------------------------------ */

static void up_rec(or_nodes_t* t, uint32_t n, const float* x, const float* y,
                   const float* z, const float* m) {
    float* mp = &t->mp[4 * n];
    if (is_leaf(t, n)) { /* P2M about the cell centre, H:978-1004 */
        float m0 = 0.0F, mx = 0.0F, my = 0.0F, mz = 0.0F;
        for (uint32_t i = t->first[n]; i < t->end[n]; ++i) {
            float mass = m[i];
            float dx   = x[i] - t->x[n];
            float dy   = y[i] - t->y[n];
            float dz   = z[i] - t->z[n];
            m0 += mass;
            mx += mass * dx;
            my += mass * dy;
            mz += mass * dz;
        }
        mp[0] = m0; mp[1] = mx; mp[2] = my; mp[3] = mz;
        return;
    }
    /* M2M in slot order, H:1007-1045 */
    float p0 = 0.0F, px = 0.0F, py = 0.0F, pz = 0.0F;
    for (int s = 0; s < 8; ++s) {
        uint32_t c = child_at(t, n, s);
        if (c == OR_NONE) continue;
        up_rec(t, c, x, y, z, m);
        float dx = t->x[c] - t->x[n];
        float dy = t->y[c] - t->y[n];
        float dz = t->z[c] - t->z[n];
        const float* cm = &t->mp[4 * c];
        p0 += cm[0];
        px += cm[1] + (cm[0] * dx);
        py += cm[2] + (cm[0] * dy);
        pz += cm[3] + (cm[0] * dz);
    }
    mp[0] = p0; mp[1] = px; mp[2] = py; mp[3] = pz;
}

void or_upward(or_nodes_t* nodes, const float* x, const float* y,
               const float* z, const float* m) {
    if (nodes->n_nodes) {
        up_rec(nodes, 0, x, y, z, m);
    }
}

/* ------------------------------------------------------------------ */
/* MAC and M2L, H:1056-1092                                            */
/* ------------------------------------------------------------------ */

int or_mac_far(float tx, float ty, float tz, float thw, float sx, float sy,
               float sz, float shw, float theta) {
    float dx      = sx - tx;
    float dy      = sy - ty;
    float dz      = sz - tz;
    float dist_sq = (dx * dx) + (dy * dy) + (dz * dz);
    float hw_sum  = thw + shw;
    return dist_sq > (theta * theta) * (hw_sum * hw_sum);
}

void or_m2l(float tx, float ty, float tz, float sx, float sy, float sz,
            const float* smp, float* loc3) {
    float dx           = sx - tx;
    float dy           = sy - ty;
    float dz           = sz - tz;
    float dist_sq      = (dx * dx) + (dy * dy) + (dz * dz);
    float soft_dist_sq = dist_sq + 0.1F;
    float dist         = sqrtf(soft_dist_sq);
    float inv_r3       = 1.0F / (dist * soft_dist_sq);
    float inv_r5       = inv_r3 / soft_dist_sq;
    float m0 = smp[0], mx = smp[1], my = smp[2], mz = smp[3];
    float fmx = m0 * dx * inv_r3;
    float fmy = m0 * dy * inv_r3;
    float fmz = m0 * dz * inv_r3;
    float m_dot_r      = (mx * dx) + (my * dy) + (mz * dz);
    float dipole_coeff = 3.0F * m_dot_r * inv_r5;
    float fdx = (mx * inv_r3) - (dx * dipole_coeff);
    float fdy = (my * inv_r3) - (dy * dipole_coeff);
    float fdz = (mz * inv_r3) - (dz * dipole_coeff);
    loc3[0] += fmx + fdx;
    loc3[1] += fmy + fdy;
    loc3[2] += fmz + fdz;
}

static inline int mac_far_nodes(const or_nodes_t* t, uint32_t a, uint32_t b,
                                float theta) {
    return or_mac_far(t->x[a], t->y[a], t->z[a], t->hw[a], t->x[b], t->y[b],
                      t->z[b], t->hw[b], theta);
}

/* ------------------------------------------------------------------ */
/* interaction pass, H:1050-1146                                       */
/* ------------------------------------------------------------------ */

static void inter_rec(or_nodes_t* t, uint32_t tg, uint32_t src, float theta,
                      or_counts_t* c) {
    c->visited_pairs++;
    if (mac_far_nodes(t, tg, src, theta)) {
        or_m2l(t->x[tg], t->y[tg], t->z[tg], t->x[src], t->y[src], t->z[src],
               &t->mp[4 * src], &t->loc[4 * tg + 1]);
        c->m2l_pairs++;
        return;
    }
    int tl = is_leaf(t, tg), sl = is_leaf(t, src);
    if (tl && sl) {
        return; /* left to P2P, H:1098-1101 */
    }
    if (tl) { /* open the source, H:1103-1113 */
        for (int j = 0; j < 8; ++j) {
            uint32_t s = child_at(t, src, j);
            if (s != OR_NONE) inter_rec(t, tg, s, theta, c);
        }
    } else if (sl) { /* open the target, H:1114-1124 */
        for (int i = 0; i < 8; ++i) {
            uint32_t g = child_at(t, tg, i);
            if (g != OR_NONE) inter_rec(t, g, src, theta, c);
        }
    } else { /* all child pairs, H:1125-1145 */
        for (int i = 0; i < 8; ++i) {
            uint32_t g = child_at(t, tg, i);
            if (g == OR_NONE) continue;
            for (int j = 0; j < 8; ++j) {
                uint32_t s = child_at(t, src, j);
                if (s != OR_NONE) inter_rec(t, g, s, theta, c);
            }
        }
    }
}

/* job roots of H:1772-1801: targets at depth 2, or shallower leaves */
typedef void (*job_fn)(or_nodes_t*, uint32_t tg, uint32_t src, void* arg);

static void dispatch_rec(or_nodes_t* t, uint32_t tg, int depth, job_fn fn,
                         void* arg) {
    if (depth == 2 || is_leaf(t, tg)) {
        fn(t, tg, 0, arg);
        return;
    }
    for (int i = 0; i < 8; ++i) {
        uint32_t g = child_at(t, tg, i);
        if (g != OR_NONE) dispatch_rec(t, g, depth + 1, fn, arg);
    }
}

typedef struct {
    float        theta;
    or_counts_t* c;
    const float *x, *y, *z, *m;
    float *fx, *fy, *fz;
} pass_arg_t;

static void inter_job(or_nodes_t* t, uint32_t tg, uint32_t src, void* a) {
    pass_arg_t* pa = (pass_arg_t*)a;
    inter_rec(t, tg, src, pa->theta, pa->c);
}

void or_interaction(or_nodes_t* nodes, float theta, int trav,
                    or_counts_t* counts) {
    if (!nodes->n_nodes) return;
    or_counts_t dummy;
    pass_arg_t  pa;
    memset(&pa, 0, sizeof(pa));
    pa.theta = theta;
    pa.c     = counts ? counts : &dummy;
    if (trav == OR_TRAV_MT) {
        dispatch_rec(nodes, 0, 0, inter_job, &pa);
    } else {
        inter_rec(nodes, 0, 0, theta, pa.c);
    }
}

/* ------------------------------------------------------------------ */
/* downward pass, H:1152-1254 / H:1804-1836                            */
/* ------------------------------------------------------------------ */

static void down_rec(or_nodes_t* t, uint32_t n, const float* m, float* fx,
                     float* fy, float* fz) {
    const float* l = &t->loc[4 * n];
    if (is_leaf(t, n)) { /* L2P, H:1224-1233 */
        for (uint32_t i = t->first[n]; i < t->end[n]; ++i) {
            float mass = m[i];
            fx[i] += mass * l[1];
            fy[i] += mass * l[2];
            fz[i] += mass * l[3];
        }
        return;
    }
    for (int s = 0; s < 8; ++s) { /* L2L, H:1238-1253 */
        uint32_t c = child_at(t, n, s);
        if (c == OR_NONE) continue;
        float* cl = &t->loc[4 * c];
        cl[1] += l[1];
        cl[2] += l[2];
        cl[3] += l[3];
        down_rec(t, c, m, fx, fy, fz);
    }
}

void or_downward(or_nodes_t* nodes, const float* m, float* fx, float* fy,
                 float* fz) {
    if (nodes->n_nodes) {
        down_rec(nodes, 0, m, fx, fy, fz);
    }
}

/* ------------------------------------------------------------------ */
/* P2P pass, H:1258-1494 (scalar inner loop H:1422-1448)               */
/* ------------------------------------------------------------------ */

static void p2p_leaf_pair(or_nodes_t* t, uint32_t tg, uint32_t src,
                          pass_arg_t* pa) {
    const float *sx = pa->x, *sy = pa->y, *sz = pa->z, *sm = pa->m;
    for (uint32_t ti = t->first[tg]; ti < t->end[tg]; ++ti) {
        float t_x = sx[ti], t_y = sy[ti], t_z = sz[ti], t_mass = sm[ti];
        float f_x = 0.0F, f_y = 0.0F, f_z = 0.0F;
        for (uint32_t j = t->first[src]; j < t->end[src]; ++j) {
            float p_dx = sx[j] - t_x;
            float p_dy = sy[j] - t_y;
            float p_dz = sz[j] - t_z;
            float p_dist_sq =
                (p_dx * p_dx) + (p_dy * p_dy) + (p_dz * p_dz) + 0.1F;
            float inv_dist  = 1.0F / sqrtf(p_dist_sq);
            float inv_dist3 = inv_dist * inv_dist * inv_dist;
            float force     = sm[j] * inv_dist3;
            f_x += p_dx * force;
            f_y += p_dy * force;
            f_z += p_dz * force;
        }
        pa->fx[ti] += t_mass * f_x;
        pa->fy[ti] += t_mass * f_y;
        pa->fz[ti] += t_mass * f_z;
    }
    pa->c->p2p_leaf_pairs++;
    pa->c->p2p_pairs += (uint64_t)(t->end[tg] - t->first[tg]) *
                        (uint64_t)(t->end[src] - t->first[src]);
}

static void p2p_rec(or_nodes_t* t, uint32_t tg, uint32_t src, pass_arg_t* pa) {
    if (mac_far_nodes(t, tg, src, pa->theta)) {
        return;
    }
    int tl = is_leaf(t, tg), sl = is_leaf(t, src);
    if (tl && sl) {
        p2p_leaf_pair(t, tg, src, pa);
        return;
    }
    if (tl) {
        for (int j = 0; j < 8; ++j) {
            uint32_t s = child_at(t, src, j);
            if (s != OR_NONE) p2p_rec(t, tg, s, pa);
        }
    } else if (sl) {
        for (int i = 0; i < 8; ++i) {
            uint32_t g = child_at(t, tg, i);
            if (g != OR_NONE) p2p_rec(t, g, src, pa);
        }
    } else {
        for (int i = 0; i < 8; ++i) {
            uint32_t g = child_at(t, tg, i);
            if (g == OR_NONE) continue;
            for (int j = 0; j < 8; ++j) {
                uint32_t s = child_at(t, src, j);
                if (s != OR_NONE) p2p_rec(t, g, s, pa);
            }
        }
    }
}

static void p2p_job(or_nodes_t* t, uint32_t tg, uint32_t src, void* a) {
    p2p_rec(t, tg, src, (pass_arg_t*)a);
}

void or_p2p(or_nodes_t* nodes, float theta, int trav, const float* x,
            const float* y, const float* z, const float* m, float* fx,
            float* fy, float* fz, or_counts_t* counts) {
    if (!nodes->n_nodes) return;
    or_counts_t dummy;
    pass_arg_t  pa;
    pa.theta = theta;
    pa.c     = counts ? counts : &dummy;
    pa.x = x; pa.y = y; pa.z = z; pa.m = m;
    pa.fx = fx; pa.fy = fy; pa.fz = fz;
    if (trav == OR_TRAV_MT) {
        dispatch_rec(nodes, 0, 0, p2p_job, &pa);
    } else {
        p2p_rec(nodes, 0, 0, &pa);
    }
}

/* ------------------------------------------------------------------ */
/* ps_calc_forces, H:1840-1909                                         */
/* ------------------------------------------------------------------ */

int or_calc_forces(size_t n, float* x, float* y, float* z, float* m,
                   uint32_t* id, uint32_t* codes, float* fx, float* fy,
                   float* fz, float cx, float cy, float cz, float hw,
                   float theta, int trav, or_nodes_t* nodes,
                   or_counts_t* counts) {
    if (!x || !y || !z || !m || !id || !codes || !fx || !fy || !fz ||
        !nodes) {
        return -2;
    }
    if (theta <= 0.0F) {
        theta = 1.0F; /* ps_init's default, H:1707 */
    }
    or_counts_t local;
    if (!counts) counts = &local;
    memset(counts, 0, sizeof(*counts));

    or_sort(n, codes, id, x, y, z, m); /* H:1851 */
    for (size_t i = 0; i < n; ++i) {   /* H:1868-1872 */
        fx[i] = 0.0F; fy[i] = 0.0F; fz[i] = 0.0F;
    }
    if (or_build(n, codes, cx, cy, cz, hw, nodes) < 0) {
        return -1;
    }
    or_upward(nodes, x, y, z, m);
    or_interaction(nodes, theta, trav, counts);
    or_downward(nodes, m, fx, fy, fz);
    or_p2p(nodes, theta, trav, x, y, z, m, fx, fy, fz, counts);
    return 0;
}

/* ------------------------------------------------------------------ */
/* FP64 direct sum, tests/accuracy.c:42-72                             */
/* ------------------------------------------------------------------ */

void or_direct_f64(size_t n, const float* x, const float* y, const float* z,
                   const float* m, size_t n_t, const uint32_t* targets,
                   double* fx, double* fy, double* fz) {
    const double eps2 = (double)0.1F; /* SOFTENING_SQ is the float 0.1F */
    for (size_t k = 0; k < n_t; ++k) {
        size_t i  = targets ? targets[k] : k;
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (size_t j = 0; j < n; ++j) {
            if (i == j) continue;
            double dx     = (double)x[j] - (double)x[i];
            double dy     = (double)y[j] - (double)y[i];
            double dz     = (double)z[j] - (double)z[i];
            double r2     = (dx * dx) + (dy * dy) + (dz * dz) + eps2;
            double inv_r  = 1.0 / sqrt(r2);
            double inv_r3 = inv_r * inv_r * inv_r;
            double scale  = (double)m[i] * (double)m[j] * inv_r3;
            ax += dx * scale;
            ay += dy * scale;
            az += dz * scale;
        }
        fx[k] = ax;
        fy[k] = ay;
        fz[k] = az;
    }
}
