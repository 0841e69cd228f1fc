/*
 * yune_types.h — the host<->device data contract of the path-tracing hot path.
 *
 * These are the structs a caller of the reference fills and hands to the
 * OpenCL kernel; the MI355X library accepts exactly the same bytes.
 *
 *   yn_float4    cl_float4                        include/CL_headers.h (OpenCL)
 *   yn_cam       struct Cam         (80 B)        include/CL_headers.h:57-64, udpt.cl:87-91
 *   yn_triangle  struct TriangleGPU (112 B)       include/CL_headers.h:66-76, udpt.cl:47-57
 *   yn_aabb      struct AABB        (32 B)        include/CL_headers.h:78-83, udpt.cl:75-78
 *   yn_bvh_node  struct BVHNodeGPU  (80 B)        include/CL_headers.h:85-91, udpt.cl:80-85
 *   yn_material  struct Material    (80 B)        include/CL_headers.h:93-107, udpt.cl:60-73
 *   yn_light_quad  struct Quad minus unused kd/ks/exponent (80 B)
 *                                                 kernels/legacy/udpt.cl:34-45, 97-106
 *
 * The light quad is a compile-time constant (`light_sources[0]`) in the
 * reference kernels; here it is a per-scene value so one library serves
 * scenes with different lights.  Field meanings are unchanged.
 *
 * Plain C, usable from C, C++ and HIP device code.  No torch types.
 */
#ifndef YUNE_TYPES_H
#define YUNE_TYPES_H

#include <stdint.h>

/* gcc / clang / hipcc: the same attribute works in C, C++ and HIP. */
#define YN_ALIGN16 __attribute__((aligned(16)))

#ifdef __cplusplus
extern "C" {
#endif

typedef struct YN_ALIGN16 yn_float4 {
    float x, y, z, w;
} yn_float4;

/* Camera: rows of the view->world matrix, eye position in the .w column. */
typedef struct YN_ALIGN16 yn_cam {
    yn_float4 r1, r2, r3, r4;
    float view_plane_dist;          /* 1/tan(yFOV*3.14/360), Camera.cpp:49 */
    float pad[3];
} yn_cam;

typedef struct YN_ALIGN16 yn_triangle {
    yn_float4 v1, v2, v3;           /* positions, .w = 1 */
    yn_float4 vn1, vn2, vn3;        /* vertex normals (not normalised), .w = 0 */
    int32_t matID;
    float pad[3];
} yn_triangle;

typedef struct YN_ALIGN16 yn_aabb {
    yn_float4 p_min, p_max;
} yn_aabb;

/* Level-order SAH-BVH node exactly as BVH::createBVH emits it.
 *  leaf:     child_idx == -1, vert_len = number of triangles (1..10)
 *  interior: child_idx  >  0, children at child_idx and child_idx+1
 *  empty:    child_idx == -2, vert_len == -1 (kept in the array) */
typedef struct YN_ALIGN16 yn_bvh_node {
    yn_aabb aabb;
    int32_t vert_list[10];
    int32_t child_idx;
    int32_t vert_len;
} yn_bvh_node;

typedef struct YN_ALIGN16 yn_material {
    yn_float4 ke, kd, ks;
    float n, k, px, py, alpha_x, alpha_y;
    int32_t is_specular, is_transmissive;
} yn_material;

typedef struct YN_ALIGN16 yn_light_quad {
    yn_float4 pos;                  /* corner, .w = 1 */
    yn_float4 normal;               /* emitting side, .w = 0 */
    yn_float4 ke;                   /* radiance, .w = 0 */
    yn_float4 edge_l, edge_w;       /* the two edges from `pos`, .w = 0 */
} yn_light_quad;

/* Render semantics.  Defaults (yn_params_reference) reproduce
 * kernels/legacy/udpt.cl / bdpt.cl as shipped. */
enum yn_mode { YN_MODE_PT = 0, YN_MODE_BDPT = 1 };

typedef struct yn_render_params {
    int32_t mode;           /* YN_MODE_PT (udpt.cl) or YN_MODE_BDPT (bdpt.cl)              */
    int32_t gi;             /* GI_CHECK, kernel arg 8 (udpt.cl:458)                        */
    int32_t mis;            /* `#define MIS` (udpt.cl:9, off as shipped)                    */
    int32_t max_bounces;    /* bounce-loop bound, udpt.cl:463 (100000 as shipped)           */
    int32_t rr_threshold;   /* RR when bounce index i > rr_threshold (udpt.cl:514, 6;
                               bdpt.cl RR_THRESHOLD 4)                                    */
    int32_t bdpt_bounces;   /* BDPT_BOUNCES, bdpt.cl:7 (20 as shipped)                      */
    int32_t oren_nayar;     /* diffuse lobe at rr_prob==1 uses Oren-Nayar with sigma^2 =
                               alpha_x (udpt-primitives.cl:673-725); 0 = udpt.cl Lambert     */
    int32_t fix_mis;        /* 0: reference MIS (lobe pick only after a visible shadow ray,
                               udpt.cl:557-605); 1: lobe picked before the shadow test      */
    float   background[4];  /* BACKGROUND_COLOR, udpt.cl:109                                */
} yn_render_params;

#ifdef __cplusplus
}  /* extern "C" */
#endif

#ifdef __cplusplus
static_assert(sizeof(yn_float4) == 16, "cl_float4 is 16 B");
static_assert(sizeof(yn_cam) == 80, "Cam is 80 B (CL_headers.h:57-64)");
static_assert(sizeof(yn_triangle) == 112, "TriangleGPU is 112 B (CL_headers.h:66-76)");
static_assert(sizeof(yn_aabb) == 32, "AABB is 32 B (CL_headers.h:78-83)");
static_assert(sizeof(yn_bvh_node) == 80, "BVHNodeGPU is 80 B (CL_headers.h:85-91)");
static_assert(sizeof(yn_material) == 80, "Material is 80 B (CL_headers.h:93-107)");
static_assert(sizeof(yn_light_quad) == 80, "light quad is 80 B");
static_assert(__builtin_offsetof(yn_triangle, matID) == 96, "matID at 96");
static_assert(__builtin_offsetof(yn_bvh_node, vert_list) == 32, "vert_list at 32");
static_assert(__builtin_offsetof(yn_bvh_node, child_idx) == 72, "child_idx at 72");
static_assert(__builtin_offsetof(yn_bvh_node, vert_len) == 76, "vert_len at 76");
static_assert(__builtin_offsetof(yn_material, n) == 48, "n at 48");
static_assert(__builtin_offsetof(yn_material, is_specular) == 72, "is_specular at 72");
static_assert(__builtin_offsetof(yn_cam, view_plane_dist) == 64, "vpd at 64");
#endif

#endif /* YUNE_TYPES_H */
