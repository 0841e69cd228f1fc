/*
 * oracle.c — CPU restatement of the reference path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Only tests/, the smoke test and
 * bench.py's CPU-baseline leg load this library; the product never does.
 *
 * Citations are file:line in the reference repository.  Build flags:
 * -O2 -ffp-contract=off (oracle/Makefile) so each float op rounds once in
 * source order.
 */
#include "oracle.h"
/* sin, cos and pow: OpenCL leaves their exact results to the runtime; the
 * oracle uses the library's one definition of each (double evaluation,
 * one rounding), pinned against libm in tests/test_fmath.py. */
#include "../yune_amd/csrc/yn_fmath.h"

#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

/* Planted mistakes, for tests only: built with -DOR_PLANT=k the oracle
 * carries mistake k (a dropped term, a wrong weight, a swapped operand; the
 * list is in tests/test_oracle_planted.py), so a test can show that the
 * checks which pin the oracle fail on it.  The default build, 0, has none. */
#ifndef OR_PLANT
#define OR_PLANT 0
#endif

#define PI      3.14159265359f      /* udpt.cl:3 */
#define INV_PI  0.31830988618f      /* udpt.cl:4 */
#define EPSILON 0.0001f             /* udpt.cl:5 */

/* ------------------------------------------------------------------ */
/* float4 arithmetic with OpenCL semantics (all four lanes)            */
/* ------------------------------------------------------------------ */
typedef struct { float x, y, z, w; } v4;

static inline v4 V(float x, float y, float z, float w) { v4 r = {x, y, z, w}; return r; }
static inline v4 F4(yn_float4 a) { return V(a.x, a.y, a.z, a.w); }
static inline v4 add(v4 a, v4 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static inline v4 sub(v4 a, v4 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
static inline v4 mul(v4 a, v4 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
static inline v4 muls(v4 a, float s) { return V(a.x * s, a.y * s, a.z * s, a.w * s); }
static inline v4 divs(v4 a, float s) { return V(a.x / s, a.y / s, a.z / s, a.w / s); }
static inline v4 neg(v4 a) { return V(-a.x, -a.y, -a.z, -a.w); }
static inline v4 bcast(float s) { return V(s, s, s, s); }
/* OpenCL dot/length/normalize/cross given one fixed definition. */
static inline float dot(v4 a, v4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
static inline float length(v4 a) { return sqrtf(dot(a, a)); }
static inline float length3(v4 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline v4 normalize(v4 a) { float inv = 1.0f / sqrtf(dot(a, a)); return muls(a, inv); }
static inline v4 cross(v4 a, v4 b)
{
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.0f);
}
static inline int any_nan(v4 a) { return isnan(a.x) || isnan(a.y) || isnan(a.z) || isnan(a.w); }

/* ------------------------------------------------------------------ */
/* RNG                                                                  */
/* ------------------------------------------------------------------ */
uint32_t or_wang_hash(uint32_t seed)        /* udpt.cl:1131-1139 */
{
    seed = (seed ^ 61u) ^ (seed >> 16);
    seed *= 9u;
    seed = seed ^ (seed >> 4);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15);
    return seed;
}

uint32_t or_xor_shift(uint32_t seed)        /* udpt.cl:1141-1147 */
{
    seed ^= seed << 13;
    seed ^= seed >> 17;
    seed ^= seed << 5;
    return seed;
}

/* `seed / (float) UINT_MAX`: (float)UINT_MAX rounds to 2^32. */
float or_u2f(uint32_t seed) { return (float)seed / 4294967296.0f; }

static inline float next_rand(uint32_t* seed) { *seed = or_xor_shift(*seed); return or_u2f(*seed); }

/* udpt.cl:175-180 */
uint32_t or_pixel_seed(uint32_t rand, int x, int y, int width)
{
    uint32_t seed = (uint32_t)((y + 1) * width + (x + 1));
    seed = rand * seed;
    seed = or_wang_hash(seed);
    if (seed == 0) seed = or_wang_hash(seed);
    return seed;
}

/* std::mt19937 (RendererCore.cpp:61-63, 268: uniform_int_distribution over
 * the full 32-bit range returns the raw engine output). */
void or_mt19937(uint32_t seed, int n, uint32_t* out)
{
    uint32_t mt[624];
    int idx = 624;
    mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
    for (int k = 0; k < n; ++k) {
        if (idx >= 624) {
            for (int i = 0; i < 624; ++i) {
                uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % 624] & 0x7fffffffu);
                mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            idx = 0;
        }
        uint32_t y = mt[idx++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        out[k] = y;
    }
}

/* ------------------------------------------------------------------ */
/* Camera: Camera::setViewMatrix (Camera.cpp:105-117) + setBuffer      */
/* (66-82); view_plane_dist = 1/tan(y_FOV*3.14/360) in double (:49).   */
/* glm::normalize of a vec4 with w = 0 is v * inversesqrt(dot(v,v)).   */
/* ------------------------------------------------------------------ */
static void glm_normalize3(const float in[3], float out[3])
{
    float d = in[0] * in[0] + in[1] * in[1] + in[2] * in[2];
    float inv = 1.0f / sqrtf(d);
    out[0] = in[0] * inv; out[1] = in[1] * inv; out[2] = in[2] * inv;
}

void or_camera(const float side[3], const float up[3], const float look_at[3],
               const float eye[3], float yfov_deg, yn_cam* c)
{
    float s[3], u[3], l[3];
    glm_normalize3(side, s);
    glm_normalize3(up, u);
    glm_normalize3(look_at, l);
    /* view2world = mat4(side, up, -look_at, eye) as columns; rows packed. */
    c->r1.x = s[0]; c->r1.y = u[0]; c->r1.z = -l[0]; c->r1.w = eye[0];
    c->r2.x = s[1]; c->r2.y = u[1]; c->r2.z = -l[1]; c->r2.w = eye[1];
    c->r3.x = s[2]; c->r3.y = u[2]; c->r3.z = -l[2]; c->r3.w = eye[2];
    c->r4.x = 0.0f; c->r4.y = 0.0f; c->r4.z = 0.0f; c->r4.w = 1.0f;
    c->view_plane_dist = (float)(1.0 / tan((double)yfov_deg * 3.14 / 360.0));
    c->pad[0] = c->pad[1] = c->pad[2] = 0.0f;
}

/* ------------------------------------------------------------------ */
/* SAH BVH, restating BVH.cpp / TriangleCPU.cpp / BVHNodeCPU.cpp        */
/* ------------------------------------------------------------------ */
static inline float std_min(float a, float b) { return (b < a) ? b : a; }   /* std::min */
static inline float std_max(float a, float b) { return (a < b) ? b : a; }   /* std::max */

typedef struct { float mn[4], mx[4]; } bbox;       /* AABB with .w */

typedef struct {
    bbox aabb;
    float centroid[4];
} tri_cpu;

/* TriangleCPU::computeCentroid / computeAABB (TriangleCPU.cpp:41-72) */
static void tri_prepare(const yn_triangle* t, tri_cpu* o)
{
    const float* v1 = &t->v1.x; const float* v2 = &t->v2.x; const float* v3 = &t->v3.x;
    for (int i = 0; i < 3; ++i) o->centroid[i] = ((v1[i] + v2[i]) + v3[i]) / 3.0f;
    o->centroid[3] = 1.0f;
    float pmn[4] = {0, 0, 0, 1}, pmx[4] = {0, 0, 0, 1};
    for (int i = 0; i < 3; ++i) {
        pmn[i] = std_min(v1[i], std_min(v2[i], v3[i]));
        pmx[i] = std_max(v1[i], std_max(v2[i], v3[i]));
    }
    for (int i = 0; i < 3; ++i)
        if (pmx[i] - pmn[i] == 0.0f) pmx[i] += 0.2f;
    memcpy(o->aabb.mn, pmn, sizeof pmn);
    memcpy(o->aabb.mx, pmx, sizeof pmx);
}

/* BVH::getExtent (BVH.cpp:268-278): copy bb1 then min/max (bb2, bb1). */
static bbox get_extent(const bbox* b1, const bbox* b2)
{
    bbox p = *b1;
    for (int i = 0; i < 3; ++i) {
        p.mn[i] = std_min(b2->mn[i], b1->mn[i]);
        p.mx[i] = std_max(b2->mx[i], b1->mx[i]);
    }
    return p;
}

/* BVH::getSurfaceArea (BVH.cpp:316-320) */
static float surface_area(const bbox* b)
{
    float d0 = b->mx[0] - b->mn[0], d1 = b->mx[1] - b->mn[1], d2 = b->mx[2] - b->mn[2];
    return 2.0f * (d0 * d1 + d0 * d2 + d1 * d2);
}

static bbox empty_box(void)
{
    bbox p;
    p.mn[0] = p.mn[1] = p.mn[2] = FLT_MAX; p.mn[3] = 1.0f;
    p.mx[0] = p.mx[1] = p.mx[2] = -FLT_MAX; p.mx[3] = 1.0f;
    return p;
}

typedef struct { int* v; int n, cap; } ivec;
static void iv_push(ivec* a, int x)
{
    if (a->n == a->cap) { a->cap = a->cap ? a->cap * 2 : 16; a->v = (int*)realloc(a->v, sizeof(int) * a->cap); }
    a->v[a->n++] = x;
}

typedef struct {
    bbox aabb;
    int32_t vert_list[10];
    int32_t child_idx, vert_len;
    ivec prims;
} cnode;

/* BVHNodeCPU(AABB) (BVHNodeCPU.cpp:42-51). The default constructor leaves
 * vert_list[4..9] and the box uninitialised; they are set to -1 / the box
 * assigned later, and only [0, vert_len) is ever compared. */
static void cnode_init(cnode* c, const bbox* b)
{
    memset(c, 0, sizeof *c);
    if (b) c->aabb = *b;
    for (int i = 0; i < 10; ++i) c->vert_list[i] = -1;
    c->vert_len = -1;
    c->child_idx = -2;
}

/* BVH::findSplitAxis (BVH.cpp:280-314) */
static int find_split_axis(const tri_cpu* tc, const ivec* prims)
{
    bbox p = empty_box();
    for (int j = 0; j < prims->n; ++j) p = get_extent(&p, &tc[prims->v[j]].aabb);
    float diag[3] = {p.mx[0] - p.mn[0], p.mx[1] - p.mn[1], p.mx[2] - p.mn[2]};
    float diff = -FLT_MAX;
    int axis = 0;
    for (int i = 0; i < 3; ++i)
        if (diag[i] > diff) { diff = diag[i]; axis = i; }
    return axis;
}

/* BVH::populateChildNodes (BVH.cpp:175-216) */
static void populate(const ivec* parent, cnode* c1, cnode* c2, const tri_cpu* tc)
{
    int child_empty = -1;
    if (surface_area(&c1->aabb) == 0.0f) child_empty = 0;
    else if (surface_area(&c2->aabb) == 0.0f) child_empty = 1;
    if (child_empty == 0) {
        for (int i = 0; i < parent->n; ++i) iv_push(&c2->prims, parent->v[i]);
    } else if (child_empty == 1) {
        for (int i = 0; i < parent->n; ++i) iv_push(&c1->prims, parent->v[i]);
    } else {
        for (int i = 0; i < parent->n; ++i) {
            const float* c = tc[parent->v[i]].centroid;
            int in_c1 = 1;
            for (int k = 0; k < 3; ++k)
                if (c[k] < c1->aabb.mn[k] || c[k] > c1->aabb.mx[k]) { in_c1 = 0; break; }
            iv_push(in_c1 ? &c1->prims : &c2->prims, parent->v[i]);
        }
    }
}

int or_bvh_build(const yn_triangle* tris, int n_tri, int bins, yn_bvh_node* out, int cap)
{
    if (n_tri < 0 || (n_tri > 0 && !tris)) return -1;
    const int leaf_primitives = 10;            /* BVH.cpp:37 */
    const float cost_isect = 1.0f;             /* BVH.cpp:38 */
    const float cost_trav = 1 / 8.0f;          /* BVH.cpp:39 */

    tri_cpu* tc = (tri_cpu*)malloc(sizeof(tri_cpu) * (n_tri > 0 ? n_tri : 1));
    bbox root = empty_box();                   /* Scene.cpp:243-244 */
    for (int i = 0; i < n_tri; ++i) {
        tri_prepare(&tris[i], &tc[i]);
        for (int k = 0; k < 3; ++k) {          /* Scene.cpp:352-356 */
            root.mn[k] = std_min(root.mn[k], tc[i].aabb.mn[k]);
            root.mx[k] = std_max(root.mx[k], tc[i].aabb.mx[k]);
        }
    }

    int n_nodes = 0, node_cap = 1024;
    cnode* nodes = (cnode*)malloc(sizeof(cnode) * node_cap);
    cnode_init(&nodes[0], &root);
    for (int i = 0; i < n_tri; ++i) iv_push(&nodes[0].prims, i);
    n_nodes = 1;
    const long node_limit = 64L * n_tri + 1024;   /* non-terminating inputs */

    for (int i = 0; i < n_nodes; ++i) {           /* BVH.cpp:65-165 */
        int switch_to_median = 0;
        ivec prims = nodes[i].prims;              /* parent_node (read-only copy) */
        bbox parent_aabb = nodes[i].aabb;
        if (prims.n == 0) continue;
        if (prims.n <= leaf_primitives) {
            nodes[i].vert_len = prims.n;
            nodes[i].child_idx = -1;
            for (int j = 0; j < prims.n; ++j) nodes[i].vert_list[j] = prims.v[j];
            continue;
        }
        int axis = find_split_axis(tc, &prims);
        cnode c1, c2;
        cnode_init(&c1, NULL); cnode_init(&c2, NULL);
        if (bins > 2 && prims.n > 20 && !switch_to_median) {
            float cost = surface_area(&parent_aabb) * cost_isect * (float)prims.n;
            float initial_cost = cost;
            float increment = parent_aabb.mx[axis] - parent_aabb.mn[axis];
            increment /= (float)bins;
            for (int b = 0; b < bins - 1; ++b) {
                bbox a1 = parent_aabb, a2 = parent_aabb;
                a1.mx[axis] = parent_aabb.mn[axis] + (float)(b + 1) * increment;
                a2.mn[axis] = a1.mx[axis];
                cnode t1, t2;
                cnode_init(&t1, &a1); cnode_init(&t2, &a2);
                populate(&prims, &t1, &t2, tc);
                float sp = surface_area(&parent_aabb);
                float new_cost = cost_trav
                               + (surface_area(&a1) / sp) * (cost_isect * (float)t1.prims.n)
                               + surface_area(&a2) / sp * (cost_isect * (float)t2.prims.n);
                if (new_cost < cost) {
                    c1.aabb = a1; c2.aabb = a2;
                    free(c1.prims.v); free(c2.prims.v);
                    c1.prims = t1.prims; c2.prims = t2.prims;
                    cost = new_cost;
                } else {
                    free(t1.prims.v); free(t2.prims.v);
                }
            }
            if (cost == initial_cost && prims.n > leaf_primitives) switch_to_median = 1;
        } else {
            switch_to_median = 1;
        }
        if (switch_to_median) {                   /* BVH.cpp:145-160 */
            float center[3];
            for (int k = 0; k < 3; ++k) center[k] = (parent_aabb.mn[k] + parent_aabb.mx[k]) / 2.0f;
            c1.aabb = parent_aabb; c1.aabb.mx[axis] = center[axis];
            c2.aabb = parent_aabb; c2.aabb.mn[axis] = center[axis];
            populate(&prims, &c1, &c2, tc);
        }
        if (n_nodes + 2 > node_limit) {
            for (int k = 0; k < n_nodes; ++k) free(nodes[k].prims.v);
            free(c1.prims.v); free(c2.prims.v); free(nodes); free(tc);
            return -1;
        }
        nodes[i].child_idx = n_nodes;
        if (n_nodes + 2 > node_cap) {
            node_cap *= 2;
            nodes = (cnode*)realloc(nodes, sizeof(cnode) * node_cap);
        }
        nodes[n_nodes++] = c1;
        nodes[n_nodes++] = c2;
    }

    /* BVH::resizeBvh (BVH.cpp:218-247) */
    for (int j = n_nodes - 1; j >= 0; --j) {
        if (nodes[j].vert_len > 0) {
            bbox p = empty_box();
            for (int k = 0; k < nodes[j].prims.n; ++k) p = get_extent(&p, &tc[nodes[j].prims.v[k]].aabb);
            memcpy(nodes[j].aabb.mn, p.mn, sizeof p.mn);
            memcpy(nodes[j].aabb.mx, p.mx, sizeof p.mx);
        } else if (nodes[j].child_idx > 0) {
            int c = nodes[j].child_idx;
            nodes[j].aabb = get_extent(&nodes[c].aabb, &nodes[c + 1].aabb);
        }
    }

    for (int i = 0; i < n_nodes && i < cap; ++i) {
        yn_bvh_node* g = &out[i];
        memcpy(&g->aabb.p_min, nodes[i].aabb.mn, 16);
        memcpy(&g->aabb.p_max, nodes[i].aabb.mx, 16);
        memcpy(g->vert_list, nodes[i].vert_list, sizeof g->vert_list);
        g->child_idx = nodes[i].child_idx;
        g->vert_len = nodes[i].vert_len;
    }
    for (int i = 0; i < n_nodes; ++i) free(nodes[i].prims.v);
    free(nodes);
    free(tc);
    return n_nodes;
}

/* ------------------------------------------------------------------ */
/* Ray tracing (udpt.cl:240-431)                                        */
/* ------------------------------------------------------------------ */
typedef struct { v4 o, d; float len; int shadow; } ray_t;
typedef struct { int tri, light; v4 p, n; } hit_t;

static inline hit_t hit_init(void) { hit_t h = {-1, -1, {0, 0, 0, 1}, {0, 0, 0, 0}}; return h; }

typedef struct {
    const or_scene* s;
    int* cand;          /* BFS candidate list, heap+2 ints */
    int heap;
    uint64_t rays;      /* BVH queries issued */
    int last_trunc;
    uint64_t truncs;    /* queries whose BFS list reached HEAP_SIZE */
    uint64_t conn_traced, conn_occluded;   /* BDPT connection rays */
} tctx;

/* rayAabbIntersection (udpt.cl:392-431); min/max are fminf/fmaxf. */
static int ray_box(const ray_t* r, const yn_float4* bmn, const yn_float4* bmx)
{
    const float mn[3] = {bmn->x, bmn->y, bmn->z}, mx[3] = {bmx->x, bmx->y, bmx->z};
    float t_max = INFINITY, t_min = -INFINITY;
    float ix = 1 / r->d.x, iy = 1 / r->d.y, iz = 1 / r->d.z;
    float mnx = (mn[0] - r->o.x) * ix, mny = (mn[1] - r->o.y) * iy, mnz = (mn[2] - r->o.z) * iz;
    float mxx = (mx[0] - r->o.x) * ix, mxy = (mx[1] - r->o.y) * iy, mxz = (mx[2] - r->o.z) * iz;
    if (!isnan(mnx)) { t_min = fmaxf(fminf(mnx, mxx), t_min); t_max = fminf(fmaxf(mnx, mxx), t_max); }
    if (!isnan(mny)) { t_min = fmaxf(fminf(mny, mxy), t_min); t_max = fminf(fmaxf(mny, mxy), t_max); }
    if (t_max < t_min) return 0;
    if (!isnan(mnz)) { t_min = fmaxf(fminf(mnz, mxz), t_min); t_max = fminf(fmaxf(mnz, mxz), t_max); }
    return t_max > fmaxf(t_min, 0.0f);
}

/* rayTriangleIntersection (udpt.cl:326-390) */
static int ray_tri(const or_scene* s, ray_t* r, hit_t* h, int idx)
{
    const yn_triangle* T = &s->tris[idx];
    v4 v1 = F4(T->v1);
    v4 v1v2 = sub(F4(T->v2), v1);
    v4 v1v3 = sub(F4(T->v3), v1);
    v4 pvec = cross(r->d, v1v3);
    float det = dot(v1v2, pvec);
    float inv_det = 1.0f / det;
    v4 dist = sub(r->o, v1);
    float u = dot(pvec, dist) * inv_det;
    if (u < 0.0f || u > 1.0f) return 0;
    v4 qvec = cross(dist, v1v2);
    float v = dot(qvec, r->d) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return 0;
    float t = dot(v1v3, qvec) * inv_det;
    if (t > 0 && t < r->len) {
        r->len = t;
        v4 N1 = normalize(F4(T->vn1)), N2 = normalize(F4(T->vn2)), N3 = normalize(F4(T->vn3));
        float w = 1 - u - v;
        h->p = add(r->o, muls(r->d, t));
        h->n = normalize(add(add(muls(N1, w), muls(N2, u)), muls(N3, v)));
        h->tri = idx;
        h->light = -1;
        return 1;
    }
    return 0;
}

/* analytic quad light test, first part of traceRay (udpt.cl:244-276) */
static int ray_quad(const yn_light_quad* L, ray_t* r, hit_t* h)
{
    v4 n = F4(L->normal), pos = F4(L->pos), el = F4(L->edge_l), ew = F4(L->edge_w);
    float DdotN = dot(r->d, n);
    if ((double)fabsf(DdotN) > 0.0001) {
        float t = dot(n, sub(pos, r->o)) / DdotN;
        if (t > 0.0f && t < r->len) {
            v4 temp = add(r->o, muls(r->d, t));
            temp = sub(temp, pos);
            float proj1 = dot(temp, el), proj2 = dot(temp, ew);
            float la = length(el), lb = length(ew);
            proj1 /= la;
            proj2 /= lb;
            if ((proj1 >= 0.0f && proj2 >= 0.0f) && (proj1 <= la && proj2 <= lb)) {
                r->len = t;
                h->p = add(r->o, muls(r->d, t));
                h->light = 0;
                h->tri = -1;
                return 1;
            }
        }
    }
    return 0;
}

/* traverseBVH (udpt.cl:288-324): BFS candidate list, never popped,
 * capped at HEAP_SIZE; shadow rays return on the first hit. */
static int traverse(tctx* c, ray_t* r, hit_t* h)
{
    const yn_bvh_node* bvh = c->s->nodes;
    int* cand = c->cand;
    int heap = c->heap;
    cand[0] = 0;
    int len = 1;
    int intersect = 0;
    c->last_trunc = 0;
    if (!ray_box(r, &bvh[0].aabb.p_min, &bvh[0].aabb.p_max)) return intersect;
    for (int i = 0; i < len && len < heap; i++) {
        const yn_bvh_node* nd = &bvh[cand[i]];
        float c_idx = (float)nd->child_idx;
        if (c_idx == -1 && nd->vert_len > 0) {
            for (int j = 0; j < nd->vert_len; j++) {
                intersect |= ray_tri(c->s, r, h, nd->vert_list[j]);
                if (r->shadow && intersect) return 1;
            }
            continue;
        }
        for (int j = (int)c_idx; j < c_idx + 2; j++) {
            const yn_bvh_node* ch = &bvh[j];
            if ((ch->vert_len > 0 || ch->child_idx > 0) &&
                ray_box(r, &ch->aabb.p_min, &ch->aabb.p_max)) {
                cand[len] = j;   /* index `heap` at most: the array has heap+2 slots */
                len++;
            }
        }
    }
    c->last_trunc = len >= heap;
    c->truncs += (uint64_t)c->last_trunc;
    return intersect;
}

/* traceRay (udpt.cl:240-286) */
static int trace(tctx* c, ray_t* r, hit_t* h)
{
    c->rays++;
    int flag = ray_quad(&c->s->light, r, h);
    c->last_trunc = 0;
    if (c->s->n_node > 0) flag |= traverse(c, r, h);
    else for (int i = 0; i < c->s->n_tri; i++) flag |= ray_tri(c->s, r, h, i);
    return flag;
}

static void tctx_init(tctx* c, const or_scene* s)
{
    c->s = s;
    c->heap = s->heap_size > 0 ? s->heap_size : (s->n_node + 2);
    c->cand = (int*)malloc(sizeof(int) * (c->heap + 2));
    c->rays = 0;
    c->last_trunc = 0;
    c->truncs = c->conn_traced = c->conn_occluded = 0;
}
static void tctx_free(tctx* c) { free(c->cand); }

void or_trace(const or_scene* s, const float* rays, int n,
              int32_t* tri, int32_t* light, float* t, int32_t* flags)
{
    tctx c;
    tctx_init(&c, s);
    for (int i = 0; i < n; ++i) {
        const float* q = rays + 8 * (size_t)i;
        ray_t r;
        r.o = V(q[0], q[1], q[2], 1.0f);
        r.len = q[3];
        r.d = V(q[4], q[5], q[6], 0.0f);
        r.shadow = q[7] != 0.0f;
        hit_t h = hit_init();
        trace(&c, &r, &h);
        tri[i] = h.tri; light[i] = h.light; t[i] = r.len;
        flags[i] = c.last_trunc;
    }
    tctx_free(&c);
}

/* ------------------------------------------------------------------ */
/* Materials and sampling (udpt.cl:611-1151)                           */
/* ------------------------------------------------------------------ */
static inline const yn_material* mat_of(const or_scene* s, int tri) { return &s->mats[s->tris[tri].matID]; }

static inline float luminance(v4 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }  /* :1121 */

/* reflect (udpt.cl:949-961): flips the normal towards w_i. */
static v4 reflect_pt(v4 w_i, v4 normal)
{
    if (dot(w_i, normal) < 0) normal = muls(normal, -1.0f);
    v4 refl = sub(muls(normal, 2 * (dot(w_i, normal))), w_i);
    return normalize(refl);
}

/* bdpt.cl keeps the normal as is (bdpt.cl:718-722, phongSampleHemisphere). */
static v4 reflect_raw(v4 w_i, v4 normal)
{
    return sub(muls(normal, 2 * dot(w_i, normal)), w_i);
}

/* toShadingSpace / sinTheta / cosPhi / sinPhi (udpt-primitives.cl:1147-1210) */
static void frame_of(v4 Nz, v4* Nx, v4* Ny)
{
    if (fabsf(Nz.y) > fabsf(Nz.z)) *Nx = V(Nz.y, -Nz.x, 0, 0.f);
    else *Nx = V(Nz.z, 0, -Nz.x, 0.f);
    *Nx = normalize(*Nx);
    *Ny = normalize(cross(Nz, *Nx));
}
static v4 to_shading(v4 w, v4 p, v4 n)
{
    v4 Nx, Ny, Nz = n;
    frame_of(Nz, &Nx, &Ny);
    v4 r1 = V(Nx.x, Nx.y, Nx.z, dot(p, Nx));
    v4 r2 = V(Ny.x, Ny.y, Ny.z, dot(p, Ny));
    v4 r3 = V(Nz.x, Nz.y, Nz.z, dot(p, Nz));
    v4 r4 = V(0.f, 0.f, 0.f, 1.0f);
    return V(dot(r1, w), dot(r2, w), dot(r3, w), dot(r4, w));
}
static inline float sin_theta(v4 w) { return sqrtf(1 - w.z * w.z); }
static inline float clampf(float x, float a, float b) { return fminf(fmaxf(x, a), b); }
static float cos_phi(v4 w)
{
    float st = sin_theta(w);
    if (st <= EPSILON && st >= -EPSILON) return 0;
    return clampf(w.x / st, -1.0f, 1.0f);
}
static float sin_phi(v4 w)
{
    float st = sin_theta(w);
    if (st <= EPSILON && st >= -EPSILON) return 0;
    return clampf(w.y / st, -1.0f, 1.0f);
}

/* OrenNayarBRDF (udpt-primitives.cl:696-725); sigma^2 = alpha_x. */
static v4 oren_nayar(const yn_material* m, v4 w_i, v4 w_o, v4 p, v4 n)
{
    w_i = to_shading(w_i, p, n);
    w_o = to_shading(w_o, p, n);
    float costerm = 0.0f;
    if (sin_theta(w_i) >= EPSILON && sin_theta(w_o) >= EPSILON)
        costerm = fmaxf(0.0f, ((cos_phi(w_i) * cos_phi(w_o)) + (sin_phi(w_i) * sin_phi(w_o))));
    float sin_alpha, tan_beta;
    if (fabsf(w_i.z) < fabsf(w_o.z)) {
        sin_alpha = sin_theta(w_i);
        tan_beta = sin_theta(w_o) / fabsf(w_o.z);
    } else {
        sin_alpha = sin_theta(w_o);
        tan_beta = sin_theta(w_i) / fabsf(w_i.z);
    }
    float sigma_sq = m->alpha_x;
    float A = (float)(1 - (sigma_sq / (2 * (sigma_sq + 0.33))));
    float B = (float)(0.45 * sigma_sq / (sigma_sq + 0.09));
    return muls(muls(F4(m->kd), INV_PI), (A + B * (costerm * sin_alpha * tan_beta)));
}

/* evaluateBRDF (udpt.cl:611-630; ON selection udpt-primitives.cl:673-694;
 * bdpt.cl:718-737 when `bdpt`). */
static v4 eval_brdf(const yn_material* m, v4 w_i, v4 w_o, v4 p, v4 n, int glossy,
                    float rr_prob, int on, int bdpt)
{
    v4 refl = bdpt ? reflect_raw(w_i, n) : reflect_pt(w_i, n);
    refl = normalize(refl);
    if (!glossy) {
        if (on && rr_prob == 1.0f) return oren_nayar(m, w_i, w_o, p, n);
        return divs(muls(F4(m->kd), INV_PI), rr_prob);
    }
    float cos_alpha = ynf_powf(fmaxf(dot(w_o, refl), 0.0f), m->px + m->py);
    int phong_exp = (int)(m->px + m->py);
    return divs(muls(muls(muls(muls(F4(m->ks), cos_alpha), (float)(phong_exp + (OR_PLANT == 11 ? 1 : 2))), INV_PI),
                     0.5f),
                rr_prob);
}

/* sampleGlossyPdf (udpt.cl:1062-1119; bdpt.cl:1048-1105 when `bdpt`) */
static int sample_glossy_pdf(const yn_material* m, uint32_t* seed, float* prob, int bdpt)
{
    float r = next_rand(seed);
    v4 ks = F4(m->ks), kd = F4(m->kd);
    if (length3(ks) == 0.0f) { *prob = 1.0f; return 0; }
    else if (length3(kd) == 0.0f || kd.x + kd.y + kd.z == 0.0f) { *prob = 1.0f; return 1; }
    v4 sum = add(ks, kd);
    float max_val = fmaxf(sum.x, fmaxf(sum.y, sum.z));
    float pd, ps;
    if (max_val == sum.x) { pd = kd.x; ps = ks.x; }
    else if (max_val == sum.y) { pd = kd.y; ps = ks.y; }
    else { pd = kd.z; ps = ks.z; }
    if (bdpt) {
        if (max_val < 1.0f) { pd += (1 - max_val) / 2.0f; ps += (1 - max_val) / 2.0f; }
        if (r < pd) { *prob = pd; return 0; }
        *prob = ps; return 1;
    }
    if (r < pd) { *prob = pd; return 0; }
    else if (r < pd + ps && r >= pd) { *prob = ps; return 1; }
    *prob = 0; return 0;
}

/* Hemisphere rows: normal_to_world (udpt.cl:718-721, 858-861) */
static v4 to_world(v4 Nx, v4 Ny, v4 Nz, v4 p, v4 local)
{
    v4 r1 = V(Nx.x, Ny.x, Nz.x, p.x), r2 = V(Nx.y, Ny.y, Nz.y, p.y), r3 = V(Nx.z, Ny.z, Nz.z, p.z);
    v4 d = V(dot(r1, local), dot(r2, local), dot(r3, local), 0);
    return normalize(d);
}

/* phongSampleHemisphere (udpt.cl:700-771; bdpt.cl variant reflects
 * without flipping the normal and normalises once). */
static void phong_sample(const yn_material* m, ray_t* ray, float* pdf, v4 w_i, v4 p, v4 n,
                         uint32_t* seed, int bdpt)
{
    v4 Nz = bdpt ? reflect_raw(w_i, n) : reflect_pt(w_i, n);
    Nz = normalize(Nz);
    v4 Nx, Ny;
    frame_of(Nz, &Nx, &Ny);
    float r1 = next_rand(seed);
    float r2 = next_rand(seed);
    int phong_exponent = (int)(m->px + m->py);
    float phi = 2 * PI * r2;
    float costheta = ynf_powf(r1, 1.0f / (float)(phong_exponent + 1));
    float sintheta = 1 - ynf_powf(r1, 2.0f / (float)(phong_exponent + 1));
    sintheta = sqrtf(sintheta);
    float x = sintheta * ynf_cosf(phi), y = sintheta * ynf_sinf(phi), z = costheta;
    ray->d = to_world(Nx, Ny, Nz, p, V(x, y, z, 0));
    ray->o = add(p, muls(ray->d, EPSILON));
    ray->shadow = 0;
    ray->len = INFINITY;
    if (dot(ray->d, n) < 0) *pdf = 0;
    else *pdf = (float)((double)(phong_exponent + 1) * 0.5 * (double)INV_PI *
                        (double)ynf_powf(costheta, (float)phong_exponent));
}

/* cosineWeightedHemisphere (udpt.cl:843-910) */
static void cosine_sample(ray_t* ray, float* pdf, v4 p, v4 n, uint32_t* seed)
{
    v4 Nx, Ny, Nz = n;
    frame_of(Nz, &Nx, &Ny);
    float r1 = next_rand(seed);
    float r2 = next_rand(seed);
    float phi = 2 * PI * r2;
    float sinTheta = sqrtf(r1);
    float x = sinTheta * ynf_cosf(phi), y = sinTheta * ynf_sinf(phi), z = sqrtf(1 - r1);
    ray->d = to_world(Nx, Ny, Nz, p, V(x, y, z, 0));
    ray->o = add(p, muls(ray->d, EPSILON));
    ray->shadow = 0;
    ray->len = INFINITY;
    *pdf = z * INV_PI;
}

/* refract (udpt.cl:963-990) */
static v4 refract_dir(const yn_material* m, v4 w_i, v4 normal)
{
    float n1, n2;
    if (dot(w_i, normal) < 0) { n1 = m->n; n2 = 1; normal = muls(normal, -1.0f); }
    else { n1 = 1; n2 = m->n; }
    v4 wt_perp = muls(sub(muls(normal, dot(w_i, normal)), w_i), n1 / n2);
    float lp = length(wt_perp);
    v4 wt_parallel = muls(neg(normal), sqrtf(1 - lp * lp));
    return normalize(add(wt_perp, wt_parallel));
}

/* evalFresnelReflectance (udpt.cl:992-1031) */
static float fresnel(const yn_material* m, v4 w_i, v4 normal, float* ior_factor)
{
    float n1, n2;
    if (dot(w_i, normal) < 0) { n1 = m->n; n2 = 1; normal = muls(normal, -1.0f); }
    else { n1 = 1; n2 = m->n; }
    float cosThetaI = dot(w_i, normal);
    float sinThetaI = sqrtf(1 - cosThetaI * cosThetaI);
    float sinThetaT = n1 * sinThetaI / n2;
    float cosThetaT = sqrtf(1 - sinThetaT * sinThetaT);
    if (sinThetaT >= 1.0f && n1 > n2) return 1.0f;
    float r0 = m->ks.x + m->ks.y + m->ks.z;
    r0 /= 3.0f;
    *ior_factor = (n2 * n2) / (n1 * n1);
    if (n1 > n2) return (r0 + (1 - r0) * (1 - ynf_powf(cosThetaI, 5.0f)));
    return (r0 + (1 - r0) * (1 - ynf_powf(cosThetaT, 5.0f)));
}

/* sampleFresnelIncidence (udpt.cl:912-947) */
static void fresnel_sample(const yn_material* m, ray_t* ray, v4 p, v4 n, v4 w_i,
                           float* ior_factor, uint32_t* seed)
{
    if (m->is_transmissive) {
        float pdf = fresnel(m, w_i, n, ior_factor);
        if (pdf == 1.0f) { ray->d = reflect_pt(w_i, n); *ior_factor = 1.0f; }
        else {
            float r = next_rand(seed);
            if (r < pdf) { ray->d = reflect_pt(w_i, n); *ior_factor = 1.0f; }
            else ray->d = refract_dir(m, w_i, n);
        }
    } else {
        *ior_factor = 1.0f;
        ray->d = reflect_pt(w_i, n);
    }
    ray->len = INFINITY;
    ray->shadow = 0;
    ray->o = add(p, muls(ray->d, EPSILON));
}

/* calcPhongPDF (udpt.cl:1033-1045), including its cos(dot(...)). */
static float phong_pdf(const yn_material* m, v4 w_i, v4 w_o, v4 n)
{
    v4 refl = sub(muls(n, 2 * (dot(w_o, n))), w_o);
    refl = normalize(refl);
    float costheta = fmaxf(0.0f, ynf_cosf(dot(refl, w_i)));
    float pe = m->px + m->py;
    return (float)((double)(pe + 1) * 0.5 * (double)INV_PI * (double)ynf_powf(costheta, pe));
}

static inline float cos_pdf(v4 w_i, v4 n) { return fmaxf(dot(w_i, n), 0.0f) * INV_PI; }   /* :1047 */

static inline float power_heuristic(float w, float p1, float p2)   /* :1149-1151, beta 2 */
{
    if (OR_PLANT == 10) return w / (p1 + p2);
    return (w * w) / (p1 * p1 + p2 * p2);
}

/* sampleLights, LIGHT_SIZE == 1 branch (udpt.cl:632-667) */
static int sample_lights(const yn_light_quad* L, v4 p, v4 n, uint32_t* seed, float* light_pdf, v4* w_i)
{
    v4 lpos = F4(L->pos), el = F4(L->edge_l), ew = F4(L->edge_w), ln = F4(L->normal);
    float r1 = next_rand(seed);
    float r2 = next_rand(seed);
    v4 temp_wi = sub(add(add(lpos, muls(el, r1)), muls(ew, r2)), p);
    float distance = dot(temp_wi, temp_wi);
    v4 wis = temp_wi;
    temp_wi = normalize(temp_wi);
    float cosine_falloff = fmaxf(dot(temp_wi, n), 0.0f) * fmaxf(dot(neg(temp_wi), ln), 0.0f);
    if (cosine_falloff <= 0.0f) return -1;   /* weights=0, sum==0 -> -1 (:654-675) */
    float area = length(el) * length(ew);
    *w_i = wis;
    *light_pdf = 1 / area;
    *light_pdf *= (OR_PLANT == 9 ? 1.0f : distance) / fmaxf(dot(neg(temp_wi), ln), 0.0f);
    return 0;
}

/* ------------------------------------------------------------------ */
/* evaluateDirectLighting (udpt.cl:535-609, bdpt.cl same body)         */
/* ------------------------------------------------------------------ */
typedef struct {
    const yn_render_params* p;
    uint32_t* cap_shadow_valid;  /* set when the first shadow ray is recorded */
    float* cap_shadow_ray;
    int32_t* cap_shadow_occ;
    int32_t* cap_trunc;
} edl_cap;

static v4 edl(tctx* c, const yn_render_params* P, v4 w_o, const hit_t* hit, uint32_t* seed,
              int bdpt, edl_cap* cap)
{
    const or_scene* s = c->s;
    const yn_material* m = mat_of(s, hit->tri);
    v4 emission = F4(m->ke);
    v4 light_sample = V(0, 0, 0, 0);
    v4 w_i;
    float light_pdf, brdf_prob = 0.0f;
    int sample_glossy = 0;
    int j = sample_lights(&s->light, hit->p, hit->n, seed, &light_pdf, &w_i);
    if (j == -1 || light_pdf <= 0.0f) return emission;
    float len = length(w_i);
    len -= EPSILON * 1.5f;
    w_i = normalize(w_i);
    ray_t shadow_ray = {add(hit->p, muls(w_i, EPSILON)), w_i, INFINITY, 1};
    hit_t sh = hit_init();
    shadow_ray.len = len;
    if (P->fix_mis) {
        /* fixed mode: pick the lobe before the shadow test, so the MIS
         * BRDF sample below always has its selection probability. */
        sample_glossy = sample_glossy_pdf(m, seed, &brdf_prob, bdpt);
        if (brdf_prob == 0.0f) return emission;
    }
    int occluded = trace(c, &shadow_ray, &sh);
    if (cap && !*cap->cap_shadow_valid) {
        *cap->cap_shadow_valid = 1;
        float* q = cap->cap_shadow_ray;
        q[0] = hit->p.x + w_i.x * EPSILON; q[1] = hit->p.y + w_i.y * EPSILON;
        q[2] = hit->p.z + w_i.z * EPSILON; q[3] = len;
        q[4] = w_i.x; q[5] = w_i.y; q[6] = w_i.z; q[7] = 1.0f;
        *cap->cap_shadow_occ = occluded;
        *cap->cap_trunc |= c->last_trunc ? 2 : 0;
    }
    if (!occluded) {
        if (!P->fix_mis) sample_glossy = sample_glossy_pdf(m, seed, &brdf_prob, bdpt);
        if (brdf_prob == 0.0f) return emission;
        light_sample = muls(mul(eval_brdf(m, w_i, w_o, hit->p, hit->n, sample_glossy, brdf_prob,
                                          P->oren_nayar && !bdpt, bdpt),
                                F4(s->light.ke)),
                            fmaxf(dot(w_i, hit->n), 0.0f));
        light_sample = muls(light_sample, 1 / light_pdf);
    }
    if (!P->mis) return add(light_sample, emission);

    v4 brdf_sample = V(0, 0, 0, 0);
    float brdf_pdf, mis_weight;
    if (sample_glossy) brdf_pdf = phong_pdf(m, w_i, w_o, hit->n);
    else brdf_pdf = cos_pdf(w_i, hit->n);
    mis_weight = light_pdf;
    mis_weight = power_heuristic(mis_weight, light_pdf, brdf_pdf);
    light_sample = muls(light_sample, mis_weight);
    ray_t bray;
    if (sample_glossy) phong_sample(m, &bray, &brdf_pdf, w_o, hit->p, hit->n, seed, bdpt);
    else cosine_sample(&bray, &brdf_pdf, hit->p, hit->n, seed);
    if (brdf_pdf <= 0.0f) return add(light_sample, emission);
    w_i = bray.d;
    /* traceRay can only return light_ID == j if the ray meets the quad, so
     * a ray that misses the quad needs no BVH query (same outcome). */
    {
        ray_t q = bray;
        hit_t qh = hit_init();
        if (!ray_quad(&s->light, &q, &qh)) return add(light_sample, emission);
    }
    hit_t nh = hit_init();
    int got = trace(c, &bray, &nh);
    if (!got || nh.light != j) return add(light_sample, emission);
    mis_weight = brdf_pdf;
    mis_weight = power_heuristic(mis_weight, light_pdf, brdf_pdf);
    brdf_sample = muls(mul(eval_brdf(m, w_i, w_o, hit->p, hit->n, sample_glossy, brdf_prob,
                                     P->oren_nayar && !bdpt, bdpt),
                           F4(s->light.ke)),
                       fmaxf(dot(w_i, hit->n), 0.0f));
    brdf_sample = muls(brdf_sample, mis_weight / brdf_pdf);
    return add(add(light_sample, brdf_sample), emission);
}

/* ------------------------------------------------------------------ */
/* Unidirectional shading (udpt.cl:433-533)                             */
/* ------------------------------------------------------------------ */
typedef struct {
    int32_t* prim_tri; int32_t* prim_light; float* prim_t; int32_t* trunc;
    edl_cap e;
} pix_cap;

static v4 shade_pt(tctx* c, const yn_render_params* P, ray_t ray, uint32_t* seed, pix_cap* pc)
{
    const or_scene* s = c->s;
    hit_t hit = hit_init();
    int got = trace(c, &ray, &hit);
    if (pc) {
        *pc->prim_tri = hit.tri; *pc->prim_light = hit.light; *pc->prim_t = ray.len;
        *pc->trunc |= c->last_trunc ? 1 : 0;
    }
    edl_cap* ec = pc ? &pc->e : NULL;
    if (!got) return V(P->background[0], P->background[1], P->background[2], P->background[3]);
    if (hit.light >= 0) {
        if (dot(ray.d, F4(s->light.normal)) < 0) return V(1, 1, 1, 1);
        return V(0.1f, .1f, .1f, 1);
    }
    v4 direct = V(0, 0, 0, 0), indirect = V(0, 0, 0, 0), throughput = V(1, 1, 1, 1);
    int matID = s->tris[hit.tri].matID;
    /* the capture records the NEE shadow ray of the primary vertex only */
    if (!s->mats[matID].is_specular) direct = edl(c, P, neg(ray.d), &hit, seed, 0, ec);
    if (P->gi) {
        for (int i = 0; i < P->max_bounces; i++) {
            hit_t nh = hit_init();
            ray_t nr;
            float pdf = 1, brdf_prob = 0.0f, ior_factor = 1.0f;
            int is_glossy = 0;
            const yn_material* m = &s->mats[matID];
            if (m->is_specular) fresnel_sample(m, &nr, hit.p, hit.n, neg(ray.d), &ior_factor, seed);
            else {
                is_glossy = sample_glossy_pdf(m, seed, &brdf_prob, 0);
                if (brdf_prob == 0.0f) break;
                if (is_glossy) phong_sample(m, &nr, &pdf, neg(ray.d), hit.p, hit.n, seed, 0);
                else cosine_sample(&nr, &pdf, hit.p, hit.n, seed);
            }
            if (pdf <= 0.0f || !trace(c, &nr, &nh) || nh.light >= 0) {
                if (m->is_specular && nh.light >= 0)
                    indirect = add(indirect, mul(throughput, F4(s->light.ke)));
                break;
            }
            indirect = add(indirect, mul(throughput, F4(mat_of(s, nh.tri)->ke)));
            if (m->is_specular) throughput = muls(throughput, ior_factor);
            else
                throughput = divs(muls(mul(throughput, eval_brdf(m, nr.d, neg(ray.d), hit.p, hit.n,
                                                                 is_glossy, brdf_prob, P->oren_nayar, 0)),
                                       fmaxf(dot(nr.d, hit.n), 0.0f)),
                                  pdf);
            if (i > P->rr_threshold) {
                float p = fminf(luminance(throughput), 0.95f);
                float r = next_rand(seed);
                if (r >= p) break;
                throughput = muls(throughput, 1 / p);
            }
            matID = s->tris[nh.tri].matID;
            if (!s->mats[matID].is_specular)
                indirect = add(indirect, mul(throughput, edl(c, P, neg(nr.d), &nh, seed, 0, NULL)));
            hit = nh;
            ray = nr;
        }
    }
    return add(direct, indirect);
}

/* ------------------------------------------------------------------ */
/* BDPT (bdpt.cl:432-640)                                               */
/* ------------------------------------------------------------------ */
typedef struct { hit_t hit; v4 dir, contrib; float fwd_pdf, rev_pdf; } pathinfo;

/* createLightPath (bdpt.cl:432-511).  Vertex 0 is a uniform point on the
 * quad: r1 by an FP64 division (:439), r2 on edge_l and r1 on edge_w (:443-
 * 444, swapped against sampleLights); contrib ke / (1/A) = ke * A.  Vertex 1
 * comes from a cosine-weighted ray: contrib max(0, n0 . d) / pdf * contrib0
 * (pi * ke * A up to rounding).  Later vertices: BRDF * cos / pdf *
 * previous contrib, with Russian roulette after RR_THRESHOLD. */
static void create_light_path(tctx* c, const yn_render_params* P, pathinfo* lp, int* len, uint32_t* seed)
{
    const yn_light_quad* L = &c->s->light;
    ray_t lr;
    hit_t hit = {-1, -1, {0, 0, 0, 1}, F4(L->normal)};
    float pdf = 1;
    *seed = or_xor_shift(*seed);
    float r1 = (float)((double)*seed / (double)4294967295u);   /* bdpt.cl:439, FP64 */
    *seed = or_xor_shift(*seed);
    float r2 = or_u2f(*seed);
    v4 A = muls(F4(L->edge_l), OR_PLANT == 4 ? r1 : r2);
    v4 B = muls(F4(L->edge_w), OR_PLANT == 4 ? r2 : r1);
    hit.p = add(add(A, B), F4(L->pos));
    hit.n = F4(L->normal);
    lp[0].hit = hit;
    lp[0].hit.light = 0;
    float area = length(F4(L->edge_l)) * length(F4(L->edge_w));
    lp[0].fwd_pdf = OR_PLANT == 5 ? 1.0f : 1.0f / area;
    lp[0].rev_pdf = 1.0f;
    lp[0].contrib = divs(F4(L->ke), lp[0].fwd_pdf);
    lp[0].dir = V(0, 0, 0, 0);
    cosine_sample(&lr, &pdf, hit.p, hit.n, seed);
    if (pdf <= 0.0f || !trace(c, &lr, &hit) || hit.light >= 0) return;
    lp[1].hit = hit;
    lp[1].dir = lr.d;
    lp[1].fwd_pdf = pdf;
    lp[1].rev_pdf = 1.0f;
    lp[1].contrib = bcast(OR_PLANT == 6 ? 1.0f : fmaxf(0.0f, dot(lp[0].hit.n, lp[1].dir)) / pdf);
    lp[1].contrib = mul(lp[1].contrib, lp[0].contrib);
    float brdf_prob = 0.0f;
    (*len)++;
    for (int i = 2; i < P->bdpt_bounces; i++) {
        const yn_material* m = mat_of(c->s, hit.tri);
        int g = sample_glossy_pdf(m, seed, &brdf_prob, 1);
        if (brdf_prob == 0.0f) break;
        if (g) phong_sample(m, &lr, &pdf, neg(lp[i - 1].dir), hit.p, hit.n, seed, 1);
        else cosine_sample(&lr, &pdf, hit.p, hit.n, seed);
        if (!trace(c, &lr, &hit) || hit.light >= 0 || pdf <= 0.0f) break;
        (*len)++;
        lp[i].hit = hit;
        lp[i].dir = lr.d;
        lp[i].fwd_pdf = pdf;
        lp[i].rev_pdf = 1.0;
        const hit_t* ph = &lp[i - 1].hit;
        lp[i].contrib = muls(eval_brdf(mat_of(c->s, ph->tri), neg(lp[i - 1].dir), lp[i].dir, ph->p, ph->n,
                                       g, brdf_prob, 0, 1),
                             fmaxf(0.0f, dot(lp[i].dir, ph->n)));
        lp[i].contrib = divs(lp[i].contrib, pdf);
        lp[i].contrib = mul(lp[i].contrib, lp[i - 1].contrib);
        if (i > P->rr_threshold) {
            float r = next_rand(seed);
            float p = fminf(luminance(lp[i].contrib), 0.95f);
            if (r >= p) break;
            lp[i].contrib = muls(lp[i].contrib, 1.0f / p);
        }
    }
}

/* createEyePath (bdpt.cl:513-560).  A primary hit on the light ends the eye
 * path here; the reference goes on to read scene_data[-1] (out of bounds)
 * but its pixel is then the constant light colour either way. */
static void create_eye_path(tctx* c, const yn_render_params* P, pathinfo* ep, int* len, ray_t eye_ray,
                            uint32_t* seed, pix_cap* pc)
{
    hit_t hit = hit_init();
    ep[0].hit = hit;
    int got = trace(c, &eye_ray, &ep[0].hit);
    if (pc) {
        *pc->prim_tri = ep[0].hit.tri; *pc->prim_light = ep[0].hit.light; *pc->prim_t = eye_ray.len;
        *pc->trunc |= c->last_trunc ? 1 : 0;
    }
    if (!got) return;
    ep[0].dir = eye_ray.d;
    ep[0].fwd_pdf = 1.0f;
    ep[0].contrib = bcast(1.0f);
    if (ep[0].hit.light >= 0) return;
    float pdf = 1, brdf_prob = 0.0f;
    for (int i = 1; i < P->bdpt_bounces; i++) {
        const hit_t* ph = &ep[i - 1].hit;
        const yn_material* m = mat_of(c->s, ph->tri);
        int g = sample_glossy_pdf(m, seed, &brdf_prob, 1);
        if (brdf_prob == 0.0f) break;
        if (g) phong_sample(m, &eye_ray, &pdf, neg(ep[i - 1].dir), ph->p, ph->n, seed, 1);
        else cosine_sample(&eye_ray, &pdf, ph->p, ph->n, seed);
        if (pdf <= 0.0f || !trace(c, &eye_ray, &hit) || hit.light >= 0) break;
        (*len)++;
        ep[i].hit = hit;
        ep[i].dir = eye_ray.d;
        ep[i].fwd_pdf = pdf;
        ep[i].contrib = muls(eval_brdf(m, ep[i].dir, neg(ep[i - 1].dir), ph->p, ph->n, g, brdf_prob, 0, 1),
                             fmaxf(0.0f, dot(ep[i].dir, ph->n)));
        ep[i].contrib = divs(ep[i].contrib, pdf);
        ep[i].contrib = mul(ep[i].contrib, ep[i - 1].contrib);
        if (i > P->rr_threshold) {
            float r = next_rand(seed);
            float p = fminf(luminance(ep[i].contrib), 0.95f);
            if (r >= p) break;
            ep[i].contrib = muls(ep[i].contrib, 1.0f / p);
        }
    }
}

#define OR_MAX_BDPT 64

/* The eye-vertex loop of shading (bdpt.cl:574-640) over given subpaths:
 * at eye vertex i, T_i * (ke + evaluateDirectLighting) * w_i, plus the
 * connections to light vertices j = lp_len-1 .. 1 that face each other and
 * are unoccluded, G * f_eye * f_light * contrib_j * T_i summed and weighted
 * w_i * (1 - ks_i); w_0 = 1, w_{i+1} = w_i * ks_i with ks_i the largest
 * specular channel floored at 0.1.  A connection ray runs to the light
 * vertex's exact distance, so the surface under the light vertex can
 * occlude it by rounding; tctx counts connection rays and occlusions. */
static v4 bdpt_combine(tctx* c, const yn_render_params* P, const pathinfo* lp, int lp_len, const pathinfo* ep,
                       int ep_len, uint32_t* seed, edl_cap* ec)
{
    const or_scene* s = c->s;
    v4 throughput, color = V(0, 0, 0, 1), subpaths;
    float eye_path_weight = 1.0f, ks = 0.0f;
    for (int i = 0; i < ep_len; i++) {
        if (eye_path_weight == 0) break;
        subpaths = V(0, 0, 0, 1);
        const yn_material* m = mat_of(s, ep[i].hit.tri);
        v4 spec = F4(m->ks), emission = F4(m->ke);
        ks = fmaxf(fmaxf(spec.x, fmaxf(spec.y, spec.z)), OR_PLANT == 3 ? 0.0f : 0.1f);
        throughput = ep[i].contrib;
        color = add(color, muls(mul(throughput, add(emission, edl(c, P, neg(ep[i].dir), &ep[i].hit, seed, 1,
                                                                 i == 0 ? ec : NULL))),
                                eye_path_weight));
        for (int j = lp_len - 1; j > 0; j--) {
            ray_t dr;
            hit_t dh = hit_init();
            v4 tlp = bcast(1.0f);
            dr.d = normalize(sub(lp[j].hit.p, ep[i].hit.p));
            dr.o = add(ep[i].hit.p, muls(dr.d, EPSILON));
            float dist = length(sub(lp[j].hit.p, ep[i].hit.p));
            dist *= dist;
            if (OR_PLANT == 1) dist = 1.0f;
            dr.len = length(sub(lp[j].hit.p, dr.o));
            dr.shadow = 1;
            if (dot(dr.d, ep[i].hit.n) <= 0 || dot(neg(dr.d), lp[j].hit.n) <= 0) continue;
            c->conn_traced++;
            if (!trace(c, &dr, &dh)) {
                tlp = lp[j].contrib;
                v4 w_i = dr.d, w_o = neg(ep[i].dir);
                const hit_t* hh = &ep[i].hit;
                float gf = fmaxf(dot(w_i, hh->n), 0.0f) *
                           (OR_PLANT == 7 ? 1.0f : fmaxf(dot(neg(w_i), lp[j].hit.n), 0.0f)) / dist;
                float prob = 0.0f;
                int g = sample_glossy_pdf(m, seed, &prob, 1);
                v4 e2l = eval_brdf(m, w_i, w_o, hh->p, hh->n, g, prob, 0, 1);
                const yn_material* lm = mat_of(s, lp[j].hit.tri);
                g = sample_glossy_pdf(lm, seed, &prob, 1);
                v4 l2e = eval_brdf(lm, neg(lp[j].dir), neg(w_i), lp[j].hit.p, lp[j].hit.n, g, prob, 0, 1);
                if (OR_PLANT == 8) l2e = bcast(1.0f);
                tlp = mul(tlp, mul(muls(e2l, gf), l2e));
                tlp = mul(tlp, throughput);
                subpaths = add(subpaths, tlp);
            } else {
                c->conn_occluded++;
            }
        }
        color = add(color, muls(subpaths, OR_PLANT == 2 ? eye_path_weight : (eye_path_weight * (1 - ks))));
        eye_path_weight = eye_path_weight * ks;
    }
    return color;
}

/* shading (bdpt.cl:562-640) */
static v4 shade_bdpt(tctx* c, const yn_render_params* P, ray_t ray, uint32_t* seed, pix_cap* pc)
{
    const or_scene* s = c->s;
    pathinfo lp[OR_MAX_BDPT], ep[OR_MAX_BDPT];
    int lp_len = 1, ep_len = 1;
    create_light_path(c, P, lp, &lp_len, seed);
    create_eye_path(c, P, ep, &ep_len, ray, seed, pc);
    edl_cap* ec = pc ? &pc->e : NULL;
    if (ep[0].hit.light >= 0) {
        if (dot(ray.d, F4(s->light.normal)) < 0) return V(1, 1, 1, 1);
        return V(0.1f, 0.1f, 0.1f, 1);
    } else if (ep[0].hit.tri < 0)
        return V(P->background[0], P->background[1], P->background[2], P->background[3]);
    return bdpt_combine(c, P, lp, lp_len, ep, ep_len, seed, ec);
}

/* ------------------------------------------------------------------ */
/* Kernel entry: one sample per pixel (udpt.cl:158-211, bdpt.cl:158-211) */
/* ------------------------------------------------------------------ */
/* createRay (udpt.cl:213-238): the 1.0 / 2.0 literals make the image-plane
 * coordinates double precision. */
static ray_t create_ray(float px, float py, int W, int H, const yn_cam* cam)
{
    ray_t r;
    float aspect = (float)((W * 1.0) / H);
    v4 dir;
    dir.x = (float)((double)aspect * ((2.0 * px / W) - 1));
    dir.y = (float)((2.0 * py / H) - 1);
    dir.z = -cam->view_plane_dist;
    dir.w = 0;
    r.d = V(dot(F4(cam->r1), dir), dot(F4(cam->r2), dir), dot(F4(cam->r3), dir), dot(F4(cam->r4), dir));
    r.d = normalize(r.d);
    r.o = V(cam->r1.w, cam->r2.w, cam->r3.w, cam->r4.w);
    r.shadow = 0;
    r.len = INFINITY;
    return r;
}

static v4 sample_pixel(tctx* c, const yn_cam* cam, int W, int H, const yn_render_params* P,
                       uint32_t rand, int x, int y, pix_cap* pc)
{
    uint32_t seed = or_pixel_seed(rand, x, y, W);
    float r1 = next_rand(&seed);
    float r2 = next_rand(&seed);
    ray_t eye = create_ray((float)x + r1, (float)y + r2, W, H, cam);
    if (P->mode == YN_MODE_BDPT) return shade_bdpt(c, P, eye, &seed, pc);
    return shade_pt(c, P, eye, &seed, pc);
}

void or_sample(const or_scene* s, const yn_cam* cam, int W, int H, const yn_render_params* p,
               uint32_t rand, int x, int y, float out[4])
{
    tctx c;
    tctx_init(&c, s);
    v4 col = sample_pixel(&c, cam, W, H, p, rand, x, y, NULL);
    out[0] = col.x; out[1] = col.y; out[2] = col.z; out[3] = col.w;
    tctx_free(&c);
}

typedef struct {
    const or_scene* s; const yn_cam* cam; int W, H; const yn_render_params* p;
    const uint32_t* seeds; int n_frames, reset_first, x0, y0, x1, y1;
    float* accum; or_capture* cap;
    int next_row; pthread_mutex_t mu;
    uint64_t rays, samples, truncs, conn_traced, conn_occluded;
} job_t;

static void* worker(void* arg)
{
    job_t* J = (job_t*)arg;
    tctx c;
    tctx_init(&c, J->s);
    uint64_t samples = 0;
    const v4 PINK = V(0.988f, 0.0588f, 0.7529f, 1.0f);    /* udpt.cl:110 */
    for (;;) {
        pthread_mutex_lock(&J->mu);
        int y = J->next_row++;
        pthread_mutex_unlock(&J->mu);
        if (y >= J->y1) break;
        for (int x = J->x0; x < J->x1; ++x) {
            size_t pix = (size_t)y * J->W + x;
            float* acc = J->accum + 4 * pix;
            for (int f = 0; f < J->n_frames; ++f) {
                pix_cap pc, *ppc = NULL;
                uint32_t valid = 0;
                if (J->cap && f == 0) {
                    pc.prim_tri = &J->cap->prim_tri[pix]; pc.prim_light = &J->cap->prim_light[pix];
                    pc.prim_t = &J->cap->prim_t[pix]; pc.trunc = &J->cap->trunc[pix];
                    *pc.trunc = 0;
                    pc.e.p = J->p; pc.e.cap_shadow_valid = &valid;
                    pc.e.cap_shadow_ray = &J->cap->shadow_ray[8 * pix];
                    pc.e.cap_shadow_occ = &J->cap->shadow_occ[pix];
                    pc.e.cap_trunc = &J->cap->trunc[pix];
                    for (int k = 0; k < 8; ++k) pc.e.cap_shadow_ray[k] = 0.0f;
                    *pc.e.cap_shadow_occ = -1;
                    ppc = &pc;
                }
                v4 color = sample_pixel(&c, J->cam, J->W, J->H, J->p, J->seeds[f], x, y, ppc);
                samples++;
                if (any_nan(color)) color = PINK;
                if (f == 0 && J->reset_first) {
                    color.w = 1;
                } else {
                    v4 prev = V(acc[0], acc[1], acc[2], acc[3]);
                    int num_passes = (int)prev.w;
                    color = add(color, muls(prev, (float)num_passes));
                    color = divs(color, (float)(num_passes + 1));
                    color.w = (float)(num_passes + 1);
                }
                acc[0] = color.x; acc[1] = color.y; acc[2] = color.z; acc[3] = color.w;
            }
        }
    }
    pthread_mutex_lock(&J->mu);
    J->rays += c.rays;
    J->samples += samples;
    J->truncs += c.truncs;
    J->conn_traced += c.conn_traced;
    J->conn_occluded += c.conn_occluded;
    pthread_mutex_unlock(&J->mu);
    tctx_free(&c);
    return NULL;
}

int or_render(const or_scene* s, const yn_cam* cam, int W, int H, const yn_render_params* p,
              const uint32_t* frame_seeds, int n_frames, int reset_first, int x0, int y0,
              int x1, int y1, int n_threads, float* accum, or_capture* cap, uint64_t* rays_out)
{
    if (W <= 0 || H <= 0 || !accum || !frame_seeds || n_frames < 0) return -1;
    if (p->mode == YN_MODE_BDPT && p->bdpt_bounces > OR_MAX_BDPT) return -1;
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 > W) x1 = W;
    if (y1 > H) y1 = H;
    job_t J;
    memset(&J, 0, sizeof J);
    J.s = s; J.cam = cam; J.W = W; J.H = H; J.p = p; J.seeds = frame_seeds; J.n_frames = n_frames;
    J.reset_first = reset_first; J.x0 = x0; J.y0 = y0; J.x1 = x1; J.y1 = y1;
    J.accum = accum; J.cap = cap; J.next_row = y0;
    pthread_mutex_init(&J.mu, NULL);
    if (n_threads <= 1) {
        worker(&J);
    } else {
        pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * n_threads);
        for (int i = 0; i < n_threads; ++i) pthread_create(&th[i], NULL, worker, &J);
        for (int i = 0; i < n_threads; ++i) pthread_join(th[i], NULL);
        free(th);
    }
    pthread_mutex_destroy(&J.mu);
    if (rays_out) {
        rays_out[0] = J.rays; rays_out[1] = J.samples; rays_out[2] = J.truncs;
        rays_out[3] = J.conn_traced; rays_out[4] = J.conn_occluded;
    }
    return 0;
}

/* tonemap.cl:14-47 */
void or_tonemap(const float* in, int n, float* out)
{
    const float EPS_T = 0.001f;
    for (int i = 0; i < n; ++i) {
        v4 col = V(in[4 * i], in[4 * i + 1], in[4 * i + 2], in[4 * i + 3]);
        float lum_world = 0.212671f * col.x + 0.715160f * col.y + 0.072169f * col.z + EPS_T;
        float lum_white = 1.0f;
        float lum_display = lum_world * (1 + lum_world / (lum_white * lum_white)) / (1 + lum_world);
        v4 sc = divs(col, lum_world);
        v4 ldr = muls(V(ynf_powf(sc.x, 1.0f), ynf_powf(sc.y, 1.0f), ynf_powf(sc.z, 1.0f), ynf_powf(sc.w, 1.0f)), lum_display);
        const float g = 1 / 2.2f;
        out[4 * i] = ynf_powf(ldr.x, g); out[4 * i + 1] = ynf_powf(ldr.y, g);
        out[4 * i + 2] = ynf_powf(ldr.z, g); out[4 * i + 3] = ynf_powf(ldr.w, g);
    }
}

/* tonemap.cl:49-58 (tonemapJohnHable, the variant tonemap.cl:29 leaves
 * commented out), then the same gamma as tonemap.cl:33. */
static float hable(float x)
{
    const float A = 0.15f, B = 0.50f, Cc = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    return ((x * (A * x + Cc * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}

void or_tonemap_hable(const float* in, int n, float* out)
{
    const float g = 1 / 2.2f;
    for (int i = 0; i < 4 * n; ++i) out[i] = ynf_powf(hable(in[i]), g);
}

/* ------------------------------------------------------------------ */
/* exported pieces for closed-form tests                                */
/* ------------------------------------------------------------------ */
static v4 A4(const float a[4]) { return V(a[0], a[1], a[2], a[3]); }
static void S4(v4 a, float o[4]) { o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w; }

void or_cosine_sample(const float p[4], const float n[4], uint32_t* seed, float dir[4], float* pdf)
{
    ray_t r;
    cosine_sample(&r, pdf, A4(p), A4(n), seed);
    S4(r.d, dir);
}

void or_phong_sample(const yn_material* m, const float p[4], const float n[4], const float w_i[4],
                     uint32_t* seed, float dir[4], float* pdf)
{
    ray_t r;
    phong_sample(m, &r, pdf, A4(w_i), A4(p), A4(n), seed, 0);
    S4(r.d, dir);
}

void or_eval_brdf(const yn_material* m, const float n[4], const float w_i[4], const float w_o[4],
                  int glossy, float rr_prob, int on, float out[4])
{
    v4 p = V(0, 0, 0, 1);
    S4(eval_brdf(m, A4(w_i), A4(w_o), p, A4(n), glossy, rr_prob, on, 0), out);
}

float or_fresnel(const yn_material* m, const float n[4], const float w_i[4], float* ior_factor)
{
    return fresnel(m, A4(w_i), A4(n), ior_factor);
}

void or_refract(const yn_material* m, const float n[4], const float w_i[4], float out[4])
{
    S4(refract_dir(m, A4(w_i), A4(n)), out);
}

int or_sample_lights(const yn_light_quad* L, const float p[4], const float n[4], uint32_t* seed,
                     float* light_pdf, float w_i[4])
{
    v4 wi = V(0, 0, 0, 0);
    int j = sample_lights(L, A4(p), A4(n), seed, light_pdf, &wi);
    S4(wi, w_i);
    return j;
}

int or_ray_box(const float o[4], const float d[4], const yn_aabb* bb)
{
    ray_t r = {A4(o), A4(d), INFINITY, 0};
    return ray_box(&r, &bb->p_min, &bb->p_max);
}

int or_ray_tri(const yn_triangle* tri, const float o[4], const float d[4], float* len)
{
    or_scene s;
    memset(&s, 0, sizeof s);
    s.tris = tri; s.n_tri = 1;
    ray_t r = {A4(o), A4(d), *len, 0};
    hit_t h = hit_init();
    int got = ray_tri(&s, &r, &h, 0);
    *len = r.len;
    return got;
}

/* createRay of a pixel's first sample (udpt.cl:175-189, 213-238) */
void or_primary_ray(const yn_cam* cam, int W, int H, uint32_t rand, int x, int y, float o[4], float d[4])
{
    uint32_t seed = or_pixel_seed(rand, x, y, W);
    float r1 = next_rand(&seed);
    float r2 = next_rand(&seed);
    ray_t r = create_ray((float)x + r1, (float)y + r2, W, H, cam);
    S4(r.o, o);
    S4(r.d, d);
}

float or_phong_pdf(const yn_material* m, const float n[4], const float w_i[4], const float w_o[4])
{
    return phong_pdf(m, A4(w_i), A4(w_o), A4(n));
}

/* the shared sin / cos / pow, exported for tests/test_fmath.py */
float or_fm_sinf(float x) { return ynf_sinf(x); }
float or_fm_cosf(float x) { return ynf_cosf(x); }
float or_fm_powf(float x, float y) { return ynf_powf(x, y); }

/* ---- BDPT pieces (bdpt.cl:432-640) on explicit vertices ---- */
static void pv_out(const pathinfo* v, or_pvertex* o)
{
    S4(v->hit.p, o->p); S4(v->hit.n, o->n); S4(v->dir, o->dir); S4(v->contrib, o->contrib);
    o->tri = v->hit.tri; o->light = v->hit.light; o->pad[0] = o->pad[1] = 0;
}
static void pv_in(const or_pvertex* o, pathinfo* v)
{
    memset(v, 0, sizeof *v);
    v->hit.p = A4(o->p); v->hit.n = A4(o->n); v->dir = A4(o->dir); v->contrib = A4(o->contrib);
    v->hit.tri = o->tri; v->hit.light = o->light;
    v->fwd_pdf = 1.0f; v->rev_pdf = 1.0f;
}

int or_bdpt_light_path(const or_scene* s, const yn_render_params* p, uint32_t* seed, or_pvertex* out)
{
    if (p->bdpt_bounces > OR_MAX_BDPT) return -1;
    tctx c;
    tctx_init(&c, s);
    pathinfo lp[OR_MAX_BDPT];
    int len = 1;
    create_light_path(&c, p, lp, &len, seed);
    for (int i = 0; i < len; ++i) pv_out(&lp[i], &out[i]);
    tctx_free(&c);
    return len;
}

int or_bdpt_eye_path(const or_scene* s, const yn_render_params* p, const float o[4], const float d[4],
                     uint32_t* seed, or_pvertex* out)
{
    if (p->bdpt_bounces > OR_MAX_BDPT) return -1;
    tctx c;
    tctx_init(&c, s);
    pathinfo ep[OR_MAX_BDPT];
    int len = 1;
    ray_t r = {A4(o), A4(d), INFINITY, 0};
    create_eye_path(&c, p, ep, &len, r, seed, NULL);
    if (ep[0].hit.tri < 0) len = 0;     /* the camera ray found no triangle */
    for (int i = 0; i < len; ++i) pv_out(&ep[i], &out[i]);
    tctx_free(&c);
    return len;
}

void or_bdpt_combine(const or_scene* s, const yn_render_params* p, const or_pvertex* lpv, int lp_len,
                     const or_pvertex* epv, int ep_len, uint32_t* seed, float out[4], uint64_t counts[2])
{
    tctx c;
    tctx_init(&c, s);
    pathinfo lp[OR_MAX_BDPT], ep[OR_MAX_BDPT];
    if (lp_len > OR_MAX_BDPT) lp_len = OR_MAX_BDPT;
    if (ep_len > OR_MAX_BDPT) ep_len = OR_MAX_BDPT;
    for (int i = 0; i < lp_len; ++i) pv_in(&lpv[i], &lp[i]);
    for (int i = 0; i < ep_len; ++i) pv_in(&epv[i], &ep[i]);
    S4(bdpt_combine(&c, p, lp, lp_len, ep, ep_len, seed, NULL), out);
    if (counts) { counts[0] = c.conn_traced; counts[1] = c.conn_occluded; }
    tctx_free(&c);
}
