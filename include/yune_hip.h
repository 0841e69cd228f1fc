/*
 * yune_hip.h — C ABI of libyune_hip, the MI355X path-tracing library.
 *
 * It replaces the OpenCL dispatch of the reference (CLManager + RendererCore
 * feeding kernels/legacy/udpt.cl or bdpt.cl) and its CL-GL interop with
 * device buffers and headless HDR/PNG output.  Plain C: pointers, sizes and
 * the structs of yune_types.h; no exceptions cross it.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative yn_status; the
 *     message is then available from yn_last_error(ctx) (or yn_last_error(NULL)
 *     for calls without a context).  This stands in for
 *     CLManager::checkError's exceptions (src/CLManager.cpp:755-855) and the
 *     bool + message of the setup* methods (src/CLManager.cpp:420-484).
 *   - Host arrays passed in are copied; the caller keeps ownership
 *     (CL_MEM_COPY_HOST_PTR semantics, src/CLManager.cpp:431, 455, 475).
 *   - One context per host thread, or external serialisation.
 *   - Images are RGBA32F, row 0 = bottom (the GL image the reference
 *     renders into); .w holds the sample count (udpt.cl:196-210).
 */
#ifndef YUNE_HIP_H
#define YUNE_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "yune_types.h"

#ifdef __cplusplus
extern "C" {
#endif

enum yn_status {
    YN_OK = 0,
    YN_ERR_HIP = -1,          /* HIP runtime error (message names the call)        */
    YN_ERR_INVALID = -2,      /* bad argument / malformed scene or BVH             */
    YN_ERR_NOT_FOUND = -3,    /* unknown scene / job id, missing file              */
    YN_ERR_NO_MEMORY = -4,    /* device allocation or memory budget exceeded       */
    YN_ERR_IO = -5,           /* file read / write failure, bad OBJ/MTL            */
    YN_ERR_NO_DEVICE = -6,    /* no gfx950 device                                  */
};

typedef struct yn_ctx yn_ctx;

typedef struct yn_config {
    int32_t device;           /* HIP device ordinal                                 */
    int32_t max_paths;        /* paths in flight per batch (0: default, 64M)        */
    size_t  mem_budget;       /* bytes for job buffers; 0: 64 GiB                   */
    int32_t bvh_width;        /* GPU node width 2 or 4 (0: 4)                       */
    int32_t trace_block;      /* trace kernel block size 64/128/256 (0: 256)        */
    int32_t trace_chunk;      /* rays per wave per queue fetch (0: 64)              */
    int32_t trace_grid;       /* trace blocks per launch (0: the blocks that fit)  */
    float   hoist_frac;       /* triangles whose box area exceeds this fraction of the
                                 root box's are tested by every ray instead of living
                                 in the tree (0: 0.01; < 0: off)                    */
    int32_t leaf_max;         /* triangles per leaf at most (1..10; 0: 1 for the
                                 library's own tree, 2 for the caller's)           */
    int32_t lds_stack;        /* traversal stack entries per lane kept in LDS:
                                 8, 12, 16 or 32 (0: per launch, 12 when the largest
                                 scene's nodes exceed 4 MB, else 16; the
                                 no-allocation trace: 16); deeper entries spill to HBM  */
    int32_t shade_waves;      /* occupancy the shade kernel is compiled for: 2, 3, 4
                                 (0: 4)                                             */
    int32_t batch_factor;     /* path-tracing work per batch over paths in flight:
                                 1 (0: default) or 2: the rest enters as paths end,
                                 behind the rays already queued (regeneration)     */
    int32_t trace_overlap;    /* 1: each iteration's closest-hit and any-hit traces
                                 run on two streams and overlap; 0: one stream     */
    int32_t graph;            /* 1: groups of bounce iterations are captured once in
                                 a HIP graph and replayed; 0: launched one by one  */
    int32_t trace_xcd;        /* ray-queue slices of a trace launch, one per XCD with
                                 its own work counter (1..16; 0: 8)                */
    int32_t trace_refill;     /* trace lanes whose ray has ended take new rays once
                                 this many lanes of the wave are idle (1..64);
                                 -1: off, a wave runs its rays to the last one;
                                 0: 48                                             */
    int32_t bvh_build;        /* GPU tree of a scene uploaded with a BVH: 1: the
                                 library's own binned SAH build over the triangles
                                 (3 axes, leaves of leaf_max); 2: the caller's
                                 tree (the reference's), re-laid out; 0: 1         */
    int32_t serve_mode;       /* yn_serve: 0: continuous batching (path-tracing jobs
                                 join one persistent wavefront as paths end); 1:
                                 rounds that each run until every path has ended */
} yn_config;

/* Parameters of udpt.cl / bdpt.cl as shipped: GI on, MIS off, no bounce
 * limit, RR after bounce 6 (BDPT: 4, 20 bounces), Lambert, background 0.4. */
void yn_params_reference(yn_render_params* p, int32_t mode);

/* ---- context: CLManager::setup / ~CLManager (src/CLManager.cpp:83-156) ---- */
int  yn_create(const yn_config* cfg, yn_ctx** out);
void yn_destroy(yn_ctx* ctx);
const char* yn_last_error(yn_ctx* ctx);
int  yn_device_info(yn_ctx* ctx, char* name, int name_len, int32_t* n_cu, size_t* hbm_bytes);

/* ---- host-side scene preparation (L2 of the reference) ---- */
/* BVH::createBVH (src/BVH.cpp:56-173) with its TriangleCPU prep and the
 * root box of Scene::loadModel (src/Scene.cpp:351-356): bit-identical node
 * array.  *nodes is allocated by the library; release with yn_free. */
int  yn_bvh_build(const yn_triangle* tris, int32_t n_tri, int32_t bins,
                  yn_bvh_node** nodes, int32_t* n_nodes);
/* No counterpart in the reference: the tree a context with bvh_build = 1
 * lays out on the GPU (before hoisting), the library's own binned SAH
 * build over all three axes with leaves of at most leaf_max (1..10)
 * triangles, in the reference's node format (BVHNodeGPU,
 * include/CL_headers.h:85-91; children at child_idx and child_idx+1), so
 * that a traversal of the reference's kind can walk it.  Release with
 * yn_free. */
int  yn_bvh_build_gpu(const yn_triangle* tris, int32_t n_tri, int32_t leaf_max,
                      yn_bvh_node** nodes, int32_t* n_nodes);
void yn_free(void* p);

typedef struct yn_host_scene {
    yn_triangle* tris; int32_t n_tri;
    yn_material* mats; int32_t n_mat;
    yn_light_quad light;
    yn_cam cam;
} yn_host_scene;
/* Scene::loadModel OBJ/MTL reader (src/Scene.cpp:133-410).  The camera is
 * the reference's default (Camera.cpp:93-103) and the light the kernels'
 * constant quad (udpt.cl:96-104), as the reference renders an OBJ scene. */
int  yn_load_obj(const char* obj_path, yn_host_scene* out);
/* Procedural scenes: Cornell box (36 triangles) and icosphere rooms. */
int  yn_gen_cornell(yn_host_scene* out);
int  yn_gen_icosphere_room(uint32_t seed, int32_t n_spheres, int32_t subdiv, yn_host_scene* out);
void yn_host_scene_free(yn_host_scene* s);
/* Camera::setViewMatrix + setBuffer (src/Camera.cpp:66-117, 49). */
int  yn_camera(const float side[3], const float up[3], const float look_at[3],
               const float eye[3], float yfov_deg, yn_cam* out);

/* ---- camera controls: Camera (src/Camera.cpp:36-166) without glm.  The
 * state holds the basis (side, up, look_at), the eye and the view-to-world
 * matrix (columns side, up, -look_at, eye).  yn_camctl_orient is
 * Camera::setOrientation: a step of move_speed along +-look_at, +-side or
 * +-up (the first non-zero component of dir, tested z, x, y), then a pitch
 * of pitch * rotation_speed radians about the camera's side axis (skipped
 * when it would turn `up` below the horizon) and a yaw of
 * yaw * rotation_speed about world +Y. ---- */
typedef struct yn_camctl {
    yn_float4 side, up, look_at, eye;
    yn_float4 view2world[4];        /* columns */
    float y_fov, rotation_speed, move_speed, view_plane_dist;
} yn_camctl;
int  yn_camctl_init(float y_fov, float rotation_speed, float move_speed, yn_camctl* out);  /* identity view */
int  yn_camctl_set_view(yn_camctl* c, const float side[3], const float up[3], const float look_at[3],
                        const float eye[3]);
int  yn_camctl_orient(yn_camctl* c, const float dir[3], float pitch, float yaw);
int  yn_camctl_pack(const yn_camctl* c, yn_cam* out);                                     /* Camera::setBuffer */

/* ---- scenes on the device: CLManager::setupVertexBuffer / setupMatBuffer /
 *      setupBVHBuffer (src/CLManager.cpp:420-484), kernel args 3-7 ----
 * n_node == 0 selects brute force over all triangles (arg 6 == 0,
 * udpt.cl:278-284).  The reference's compile-time light_sources[0]
 * (udpt.cl:97-106) is passed as `light`. */
int  yn_scene_create(yn_ctx* ctx, const yn_triangle* tris, int32_t n_tri,
                     const yn_material* mats, int32_t n_mat,
                     const yn_bvh_node* nodes, int32_t n_node,
                     const yn_light_quad* light, int32_t* scene_id);
int  yn_scene_release(yn_ctx* ctx, int32_t scene_id);

typedef struct yn_scene_info {
    int32_t n_tri, n_ref_nodes, n_gpu_nodes, width, max_depth, n_hoist;
    size_t  node_bytes, tri_bytes, shade_bytes;
} yn_scene_info;
int  yn_scene_info_get(yn_ctx* ctx, int32_t scene_id, yn_scene_info* out);

/* ---- the kernel contract: pathtracer(args 0-13) (udpt.cl:158-160,
 * launched by RendererCore::enqueueKernels, src/RendererCore.cpp:248-300) ----
 * One frame over the pixels of tile `block` of a block_x x block_y grid,
 * on caller-owned device images (out_img, in_img: W*H float4).  Scene
 * buffers (args 3-7) are the uploaded scene; GI_CHECK, reset, rand, block,
 * block_x, block_y keep their meaning.  Synchronous. */
int  yn_pathtracer(yn_ctx* ctx, float* out_img, const float* in_img, const yn_cam* cam,
                   int32_t scene_id, int32_t width, int32_t height,
                   const yn_render_params* params, int32_t gi_check, int32_t reset,
                   uint32_t rand, int32_t block, int32_t block_x, int32_t block_y);
/* tonemap(current, prev, out, reset) (kernels/post-proc/tonemap.cl:14-35) on
 * device images. */
int  yn_tonemap(yn_ctx* ctx, const float* in_img, float* out_img, int32_t width, int32_t height);
/* The same with the operator chosen: the Reinhard operator the reference
 * ships (tonemapReinhard, tonemap.cl:42-47, lum_white 1) or John Hable's
 * curve (tonemapJohnHable, tonemap.cl:49-58, commented out at :29). */
enum { YN_TONEMAP_REINHARD = 0, YN_TONEMAP_HABLE = 1 };
int  yn_tonemap_op(yn_ctx* ctx, const float* in_img, float* out_img, int32_t width, int32_t height, int32_t op);

/* ---- progressive jobs: RendererCore's frame loop (src/RendererCore.cpp:
 * 248-469, 507-579) without the GUI.  Frame f uses rand = f-th output of
 * std::mt19937(base_seed) (the reference seeds from the clock); frame 0 of
 * a job resets the image. ---- */
int  yn_job_create(yn_ctx* ctx, int32_t scene_id, const yn_cam* cam, int32_t width, int32_t height,
                   const yn_render_params* params, uint32_t base_seed, int32_t* job_id);
int  yn_job_render(yn_ctx* ctx, int32_t job_id, int32_t n_frames);
int  yn_job_read(yn_ctx* ctx, int32_t job_id, float* rgba);           /* W*H*4 floats  */
int  yn_job_read_tonemapped(yn_ctx* ctx, int32_t job_id, float* rgba);
/* Operator used by yn_job_read_tonemapped and the .png of yn_job_save
 * (default YN_TONEMAP_REINHARD, as the shipped tonemap.cl). */
int  yn_job_set_tonemap(yn_ctx* ctx, int32_t job_id, int32_t op);
/* RendererCore::saveImage (src/RendererCore.cpp:608-646): .hdr writes the
 * linear image, .png the tonemapped one as 8-bit, both flipped vertically. */
int  yn_job_save(yn_ctx* ctx, int32_t job_id, const char* path);
int  yn_job_release(yn_ctx* ctx, int32_t job_id);

typedef struct yn_stats {
    uint64_t frames, samples;
    uint64_t rays_ext, rays_shadow;       /* BVH queries (closest / any hit)      */
    uint64_t nodes_ext, tris_ext;         /* counting mode only: interior nodes    */
    uint64_t nodes_shadow, tris_shadow;   /* visited and triangles tested          */
    uint64_t trace_launches_ext, trace_launches_shadow, iterations;
    double   ms_total;                    /* wall time of yn_job_render calls      */
    double   ms_trace_ext, ms_trace_shadow, ms_shade, ms_other;   /* HIP events */
    /* continuous serving (yn_serve_stats): groups of bounce iterations run,
     * the sum of the paths in flight at each group's start, refills, groups
     * whose planning stopped for want of samples-buffer room or piece
     * entries, and host milliseconds spent between groups */
    uint64_t groups, paths_in_flight, refills, plan_stalls;
    double   ms_host;
} yn_stats;
int  yn_job_stats(yn_ctx* ctx, int32_t job_id, yn_stats* out);
/* 1: next renders also count node visits and triangle tests (slower). */
int  yn_job_set_counting(yn_ctx* ctx, int32_t job_id, int32_t on);
/* Record, for the next frame rendered, each pixel's primary hit and the
 * NEE shadow ray cast from it (for parity tests).  Arrays of W*H (shadow:
 * W*H*8 floats o.xyz, len, d.xyz, recorded-flag; occ -1 when the primary
 * vertex cast no shadow ray). */
int  yn_job_capture(yn_ctx* ctx, int32_t job_id, int32_t* prim_tri, int32_t* prim_light,
                    float* prim_t, float* shadow_ray, int32_t* shadow_occ);

/* ---- serving: continuous batching of many progressive jobs.  No
 * counterpart in the reference (its RendererCore renders one image); the
 * jobs keep the frame protocol of yn_job_render, so a job's image is the
 * same whether it runs alone or among others. ----
 * yn_submit queues a job of n_frames frames (spp); its accumulation image is
 * admitted against cfg.mem_budget, oldest first, as memory frees up.
 * yn_serve runs up to max_rounds scheduler rounds; each packs whole frames
 * of the oldest admitted jobs (any scenes of the context) into one
 * wavefront batch and folds them into the jobs' images.  *pending: served
 * jobs not finished.  A finished job is read and released like any job. */
int  yn_submit(yn_ctx* ctx, int32_t scene_id, const yn_cam* cam, int32_t width, int32_t height,
               const yn_render_params* params, uint32_t base_seed, int32_t n_frames, int32_t* job_id);
int  yn_serve(yn_ctx* ctx, int32_t max_rounds, int32_t* pending);
typedef struct yn_job_status {
    int32_t frames_target, frames_done, admitted, done;
    double  t_submit_ms, t_start_ms, t_done_ms;   /* yn_clock_ms times; -1 = not yet */
} yn_job_status;
int  yn_job_status_get(yn_ctx* ctx, int32_t job_id, yn_job_status* out);
int  yn_serve_stats(yn_ctx* ctx, yn_stats* out);     /* rays and kernel times of all rounds */
double yn_clock_ms(yn_ctx* ctx);                     /* ms since yn_create */

/* ---- autotuning: no counterpart in the reference (its launch shape is
 * fixed: one work-item per pixel of a tile).  yn_autotune renders short
 * timed runs of the given view (frames per run) and keeps, knob by knob,
 * the fastest BVH width (the scenes' GPU layout is rebuilt), trace block
 * size, rays per queue fetch, LDS stack depth, shade occupancy, paths
 * per batch, batch_factor, trace_overlap, graph, trace_xcd and trace_refill in
 * the context's configuration.  Not while
 * served jobs are pending.  yn_config_get reads the configuration in effect. ---- */
typedef struct yn_tune_result {
    int32_t bvh_width, trace_block, trace_chunk, lds_stack, shade_waves, n_tried;
    int64_t batch_paths;      /* paths in flight per batch                              */
    int32_t batch_factor;     /* work items per batch over paths in flight; 2: the
                                 rest enters as paths end (regeneration)             */
    int32_t trace_overlap;    /* closest-hit and any-hit traces on two streams        */
    int32_t graph;            /* bounce loop replayed from a HIP graph                */
    int32_t trace_xcd;        /* ray-queue slices per trace launch                    */
    int32_t trace_refill;     /* idle lanes before a trace wave takes new rays (-1 off) */
    int32_t pad;
    double  mrays_per_s;      /* the chosen configuration on the tuning view */
} yn_tune_result;
int  yn_autotune(yn_ctx* ctx, int32_t scene_id, const yn_cam* cam, int32_t width, int32_t height,
                 const yn_render_params* params, int32_t frames, yn_tune_result* out);
int  yn_config_get(yn_ctx* ctx, yn_config* out, int64_t* batch_paths);

/* ---- diagnostics: the kernels' sin (op 0), cos (op 1) and pow (op 2) on
 * the device for n host inputs (x, and y for pow).  OpenCL leaves these
 * built-ins' exact results to the runtime; this library defines each once
 * and evaluates it identically on host and device. ---- */
int  yn_fmath_eval(yn_ctx* ctx, int32_t op, const float* x, const float* y, int32_t n, float* out);

/* ---- traceRay (udpt.cl:240-286) on caller rays: n x 8 floats
 * (o.xyz, length, d.xyz, is_shadow).  Outputs per ray: triangle id (-1),
 * light id (-1 / 0), final length.  Rays with is_shadow != 0 take the
 * any-hit path; their tri is one occluder. ---- */
int  yn_trace(yn_ctx* ctx, int32_t scene_id, const float* rays, int32_t n,
              int32_t* tri, int32_t* light, float* t);

/* ---- No counterpart in the reference: yn_trace for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  rays, tri, light and t are
 * DEVICE pointers in yn_trace's layouts; n <= YN_TRACE_NOALLOC_MAX_RAYS,
 * larger n is refused with YN_ERR_INVALID.  The call allocates nothing:
 * its ray queues, hit records, counters and traversal-stack overflow area
 * live in one static device array per GPU (35,663,104 bytes, held until
 * the process ends), whose address the context takes at creation.  Every
 * call on a GPU shares that array and it keeps the last call's contents,
 * so calls that overlap on one GPU (two contexts, two threads) race.  A
 * traversal deeper than the array's stack (the LDS part plus 32 entries)
 * fails with YN_ERR_INVALID.  Synchronous on the context's stream, like
 * yn_trace, and bit-identical to it. ---- */
#define YN_TRACE_NOALLOC_MAX_RAYS 262144
int  yn_trace_noalloc(yn_ctx* ctx, int32_t scene_id, const float* rays, int32_t n,
                      int32_t* tri, int32_t* light, float* t);

#ifdef __cplusplus
}
#endif
#endif /* YUNE_HIP_H */
