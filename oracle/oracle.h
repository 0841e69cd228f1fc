/*
 * oracle.h — CPU restatement of the reference path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the tests, the
 * smoke test and the CPU-baseline leg of bench.py compare the HIP path
 * against.  Nothing in the product (yune_amd/, include/) links, loads or
 * calls it, and it is never the thing being measured on the GPU.
 *
 * Every function restates one piece of the reference and cites it
 * (file:line, relative to the reference repository):
 *   kernels/legacy/udpt.cl            unidirectional path tracer
 *   kernels/legacy/bdpt.cl            naive Lafortune BDPT
 *   kernels/legacy/udpt-primitives.cl Oren-Nayar diffuse lobe
 *   kernels/post-proc/tonemap.cl      tonemap
 *   src/BVH.cpp, src/TriangleCPU.cpp, src/BVHNodeCPU.cpp   SAH-BVH build
 *   src/Camera.cpp                    camera packing
 *   src/RendererCore.cpp              frame protocol (seed, reset, accumulate)
 *
 * Floating point: built with -ffp-contract=off on x86-64 SSE2, so every
 * float operation rounds once, in the order the reference source writes
 * it.  OpenCL built-ins the reference relies on are given one fixed
 * definition (dot = ((x+y)+z)+w products, normalize = v * (1/sqrt(dot)),
 * min/max = fminf/fmaxf, pow/sin/cos = libm); the HIP kernels use the same
 * definitions, which is what makes primary and shadow rays bit-identical.
 *
 * Pinning (see tests/test_oracle_*.py): RNG and mt19937 by integer known
 * answers; BVH by bit-exact comparison with the reference BVH.cpp compiled
 * from its own sources (oracle/_ref) on the reference's Cornell fixtures
 * and procedural scenes; intersections by hand-built cases and brute force;
 * BSDF sampling by closed forms, pdf normalisation and histograms; NEE by
 * the analytic irradiance of a rectangular emitter; full images only
 * statistically (convergence).  The exact per-pixel radiance of the OpenCL
 * kernels under a real OpenCL runtime is parity unpinned (no CPU OpenCL
 * device exists to run them).
 */
#ifndef YUNE_ORACLE_H
#define YUNE_ORACLE_H

#include <stdint.h>
#include "../include/yune_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG (udpt.cl:1131-1147; RendererCore.cpp:61-63, 268) ---- */
uint32_t or_wang_hash(uint32_t seed);
uint32_t or_xor_shift(uint32_t seed);
float    or_u2f(uint32_t seed);                 /* seed / (float)UINT_MAX */
void     or_mt19937(uint32_t seed, int n, uint32_t* out);
uint32_t or_pixel_seed(uint32_t rand, int x, int y, int width);   /* udpt.cl:175-180 */

/* ---- camera (Camera.cpp:66-82, 105-117, 49) ---- */
void or_camera(const float side[3], const float up[3], const float look_at[3],
               const float eye[3], float yfov_deg, yn_cam* out);

/* ---- SAH BVH (BVH.cpp:56-247, TriangleCPU.cpp:41-72, Scene.cpp:243-356) ----
 * Returns the node count; writes at most `cap` nodes.  -1 on bad input. */
int or_bvh_build(const yn_triangle* tris, int n_tri, int bins,
                 yn_bvh_node* out, int cap);

/* ---- scene for tracing/rendering ---- */
typedef struct or_scene {
    const yn_triangle* tris; int n_tri;
    const yn_material* mats; int n_mat;
    const yn_bvh_node* nodes; int n_node;   /* n_node == 0: brute force (udpt.cl:280-284) */
    yn_light_quad light;
    int heap_size;                          /* BFS cap, HEAP_SIZE=1500 (udpt.cl:8);
                                               0 = no cap (exact closest hit) */
} or_scene;

/* rays: n x 8 floats (ox,oy,oz,len, dx,dy,dz, is_shadow).  Per ray output:
 * tri id, light id, t (final ray length), flags bit0 = BFS truncated. */
void or_trace(const or_scene* s, const float* rays, int n,
              int32_t* tri, int32_t* light, float* t, int32_t* flags);

/* Per-pixel capture of the first frame rendered, for hit-ID parity. */
typedef struct or_capture {
    int32_t* prim_tri;      /* W*H: primary hit triangle (-1 none)          */
    int32_t* prim_light;    /* W*H: primary hit light (-1 none)             */
    float*   prim_t;        /* W*H: primary hit distance                    */
    float*   shadow_ray;    /* W*H*8: the primary vertex's NEE shadow ray:
                               o.xyz, len, d.xyz, 1                          */
    int32_t* shadow_occ;    /* W*H: its result (1 occluded, 0 visible, -1 none) */
    int32_t* trunc;         /* W*H: bit0 primary truncated, bit1 shadow truncated */
} or_capture;

/* Render n_frames progressive frames of the pixels in [x0,x1)x[y0,y1),
 * following RendererCore::enqueueKernels/updateRenderKernelArgs: frame f
 * uses rand = frame_seeds[f]; reset on frame 0 when `reset_first`.
 * accum: W*H*4 floats (the RGBA32F image, .w = sample count).
 * rays_out (optional, 5 entries): [0] BVH queries issued, [1] samples,
 * [2] queries whose BFS list reached heap_size (HEAP_SIZE truncation),
 * [3] BDPT connection rays traced, [4] of those occluded.
 * n_threads <= 0 selects 1.  cap may be NULL. */
int or_render(const or_scene* s, const yn_cam* cam, int W, int H,
              const yn_render_params* p, const uint32_t* frame_seeds, int n_frames,
              int reset_first, int x0, int y0, int x1, int y1, int n_threads,
              float* accum, or_capture* cap, uint64_t* rays_out);

/* One sample's radiance (before NaN->PINK), for unit tests. */
void or_sample(const or_scene* s, const yn_cam* cam, int W, int H,
               const yn_render_params* p, uint32_t rand, int x, int y, float out[4]);

/* tonemap.cl:14-47 (Reinhard, lum_white = 1, then gamma 1/2.2). */
void or_tonemap(const float* in_rgba, int n_pixels, float* out_rgba);
/* tonemap.cl:49-58 (John Hable's curve), then gamma 1/2.2. */
void or_tonemap_hable(const float* in_rgba, int n_pixels, float* out_rgba);

/* ---- BSDF / sampling pieces exported for closed-form tests ---- */
/* hit: p[4], n[4], material; returns direction in dir[4] and pdf */
void or_cosine_sample(const float p[4], const float n[4], uint32_t* seed,
                      float dir[4], float* pdf);
void or_phong_sample(const yn_material* m, const float p[4], const float n[4],
                     const float w_i[4], uint32_t* seed, float dir[4], float* pdf);
void or_eval_brdf(const yn_material* m, const float n[4], const float w_i[4],
                  const float w_o[4], int glossy, float rr_prob, int oren_nayar,
                  float out[4]);
float or_fresnel(const yn_material* m, const float n[4], const float w_i[4],
                 float* ior_factor);
void or_refract(const yn_material* m, const float n[4], const float w_i[4], float out[4]);
int  or_sample_lights(const yn_light_quad* L, const float p[4], const float n[4],
                      uint32_t* seed, float* light_pdf, float w_i[4]);
int  or_ray_box(const float o[4], const float d[4], const yn_aabb* bb);
void or_primary_ray(const yn_cam* cam, int W, int H, uint32_t rand, int x, int y, float o[4], float d[4]);
float or_phong_pdf(const yn_material* m, const float n[4], const float w_i[4], const float w_o[4]);
int  or_ray_tri(const yn_triangle* tri, const float o[4], const float d[4], float* len);
float or_fm_sinf(float x);
float or_fm_cosf(float x);
float or_fm_powf(float x, float y);

/* ---- BDPT pieces (bdpt.cl:432-640), for closed-form tests ----
 * A subpath vertex (PathInfo, bdpt.cl:117-125): hit point, shading normal,
 * direction of the ray that found it, contribution, triangle id (-1 on the
 * light) and light id (0 for light vertex 0, else -1). */
typedef struct or_pvertex {
    float p[4], n[4], dir[4], contrib[4];
    int32_t tri, light, pad[2];
} or_pvertex;
/* createLightPath (bdpt.cl:432-511) from *seed; returns the subpath length
 * (vertex 0 on the light), -1 if bdpt_bounces exceeds 64.  out: 64 slots. */
int  or_bdpt_light_path(const or_scene* s, const yn_render_params* p, uint32_t* seed, or_pvertex* out);
/* createEyePath (bdpt.cl:513-560) from the camera ray (o, d); returns the
 * length, 0 when the camera ray finds no triangle. */
int  or_bdpt_eye_path(const or_scene* s, const yn_render_params* p, const float o[4], const float d[4],
                      uint32_t* seed, or_pvertex* out);
/* The eye-vertex loop of shading (bdpt.cl:574-640) over given subpaths:
 * emission + direct light and all light-vertex connections, with the eye
 * weights.  counts (optional): connection rays traced, occluded. */
void or_bdpt_combine(const or_scene* s, const yn_render_params* p, const or_pvertex* lp, int lp_len,
                     const or_pvertex* ep, int ep_len, uint32_t* seed, float out[4], uint64_t counts[2]);

#ifdef __cplusplus
}
#endif
#endif
