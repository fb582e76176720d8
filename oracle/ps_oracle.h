/*
 * ps_oracle.h -- CPU restatement of the polesitter FMM tick.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; it is the checker that the HIP path is compared against
 * (tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py).
 *
 * Every function restates one piece of the reference header
 * src/polesitter.h (cited as H:line) with the same IEEE single-precision
 * operation order as a strict build of the reference
 * (-O2 -ffp-contract=off -fno-fast-math, no AVX2: the scalar code paths).
 * Built with those flags, its outputs are bit-identical to that build of the
 * reference on the same inputs; tests/test_oracle_golden.py pins this against
 * fixtures written by the compiled reference.
 *
 * Nodes are stored in canonical level order: depth-major, and inside one
 * depth in ascending Morton-prefix order (children of one parent are
 * contiguous, in octant order).  The GPU tree uses the same order, so node
 * arrays compare index by index.
 */
#ifndef PS_ORACLE_H
#define PS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#define OR_MAX_DEPTH 10
#define OR_NONE      0xFFFFFFFFu

/* traversal roots, H:1772-1801 (MT) and H:1888,1905 (ST) */
#define OR_TRAV_MT 0 /* jobs (target at depth 2 or shallower leaf, root) */
#define OR_TRAV_ST 1 /* one pair (root, root) */

/* Morton quantisation, H:2132-2157 (scalar) vs H:2040-2072 (AVX2) */
#define OR_MORTON_TRUNC 0
#define OR_MORTON_AVX2  1

typedef struct {
    uint32_t  cap;            /* capacity of every per-node array */
    uint32_t  n_nodes;        /* out */
    uint32_t  max_depth;      /* out */
    uint32_t  level_start[12];/* out: depth d occupies [ls[d], ls[d+1]) */
    float*    x;              /* node centre, H:935-939 */
    float*    y;
    float*    z;
    float*    hw;             /* half width */
    uint32_t* first;          /* first particle of the node's range */
    uint32_t* end;            /* one past the last particle */
    uint32_t* child;          /* first child (internal) or OR_NONE (leaf) */
    uint32_t* mask;           /* occupied child slots, 0 for a leaf */
    uint32_t* parent;         /* OR_NONE for the root */
    uint8_t*  depth;
    float*    mp;             /* 4 per node: m0, px, py, pz (H:1000-1003) */
    float*    loc;            /* 4 per node: 0, Ex, Ey, Ez (H:1089-1092) */
} or_nodes_t;

typedef struct {
    uint64_t m2l_pairs;       /* far pairs evaluated by M2L */
    uint64_t p2p_leaf_pairs;  /* near leaf x leaf pairs */
    uint64_t p2p_pairs;       /* sum over leaf pairs of n_t * n_s */
    uint64_t visited_pairs;   /* every (target, source) the M2L pass tests */
} or_counts_t;

/* H:1911-2170.  Returns 0, or -2 (EINVAL) like the reference. */
int or_prepare(size_t n, const float* x, const float* y, const float* z,
               uint32_t* codes, int mode, float* out_min, float* out_max,
               float* out_range);

/* H:777-799 */
uint32_t or_expand_bits(uint32_t v);
uint32_t or_morton_encode(uint32_t x, uint32_t y, uint32_t z);

/* H:1588-1693: stable sort by the full 32-bit key, payload moved along. */
void or_sort(size_t n, uint32_t* codes, uint32_t* id, float* x, float* y,
             float* z, float* m);

/* H:882-967.  Returns the node count, or -1 when nodes->cap is too small. */
long or_build(size_t n, const uint32_t* codes, float cx, float cy, float cz,
              float hw, or_nodes_t* nodes);

/* H:969-1046 */
void or_upward(or_nodes_t* nodes, const float* x, const float* y,
               const float* z, const float* m);

/* H:1048-1146 plus the dispatch of H:1772-1801 / H:1888 */
void or_interaction(or_nodes_t* nodes, float theta, int trav,
                    or_counts_t* counts);

/* H:1148-1254 plus H:1804-1836: L2L, then L2P as the first write of f. */
void or_downward(or_nodes_t* nodes, const float* m, float* fx, float* fy,
                 float* fz);

/* H:1256-1494 (scalar P2P, H:1422-1448) */
void or_p2p(or_nodes_t* nodes, float theta, int trav, const float* x,
            const float* y, const float* z, const float* m, float* fx,
            float* fy, float* fz, or_counts_t* counts);

/* H:1840-1909: the whole ps_calc_forces on caller arrays, in place.
 * Returns 0, -1 (node capacity) or -2 (EINVAL). */
int or_calc_forces(size_t n, float* x, float* y, float* z, float* m,
                   uint32_t* id, uint32_t* codes, float* fx, float* fy,
                   float* fz, float cx, float cy, float cz, float hw,
                   float theta, int trav, or_nodes_t* nodes,
                   or_counts_t* counts);

/* tests/accuracy.c:42-72: FP64 direct sum (self term skipped, eps^2 = 0.1)
 * for the listed target indices. */
void or_direct_f64(size_t n, const float* x, const float* y, const float* z,
                   const float* m, size_t n_t, const uint32_t* targets,
                   double* fx, double* fy, double* fz);

/* MAC of H:1056-1064 / H:1265-1273 on explicit cells (1 = far). */
int or_mac_far(float tx, float ty, float tz, float thw, float sx, float sy,
               float sz, float shw, float theta);

/* M2L of H:1065-1092 for one pair, accumulated into loc3. */
void or_m2l(float tx, float ty, float tz, float sx, float sy, float sz,
            const float* smp, float* loc3);

#endif
