// yn_device.h — device-side data shared by the HIP kernels and the host
// runtime that launches them.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/yune_types.h"

namespace yune {

constexpr float kPI = 3.14159265359f;       // udpt.cl:3
constexpr float kInvPI = 0.31830988618f;    // udpt.cl:4
constexpr float kEps = 0.0001f;             // udpt.cl:5

// One uploaded scene in its GPU layout.
struct DevScene {
    const float4* nodes;        // width 2: 4 float4 per node, width 4: 8 float4 per node
    const float4* tri_isect;    // 3 float4 per slot: v1 (w = tri id bits), e1, e2
    const float4* tri_shade;    // 4 float4 per triangle: normalised vertex normals, {matID bits, 0, 0, 0}
    const int32_t* tri_mat;     // matID per triangle
    const yn_material* mats;
    const float4* hoist_box;    // 2 float4 per hoisted slot: lo, hi
    yn_light_quad light;
    uint32_t root_ref;
    float root_box[6];
    int32_t width;              // 2 or 4
    int32_t brute;              // 1: no BVH, test every slot in order
    int32_t n_slots;
    int32_t n_hoist;            // slots [0, n_hoist) are tested before traversal
    int32_t fma_slab;           // 1: every child box coordinate is finite and within
                                //    2^60 (the traversal's FMA slab form cannot overflow)
    int32_t mat_base;           // first entry of the scene's materials in the context's
                                //    table of every live scene's materials
};

// Per-job constants (a job = one image being accumulated).
struct DevJob {
    yn_cam cam;
    yn_render_params p;
    int32_t W, H;
    int32_t scene;              // index into the scene table
    int32_t pad;
    float4* samples;            // [slots][W*H] one sample colour per frame slot
};

// A queue of rays: rays[i] = {o.xyz, tmax}, {d.xyz, w}.  w: the path index
// (bidirectional paths, caller rays), or the job-table entry of the path
// whose state travels with the ray (unidirectional paths).
struct RayQueue {
    float4* o_t;
    float4* d_p;
    uint32_t* count;            // number of rays in the queue
};

// Unidirectional path state, one record per ray in a pair of queues: the
// record of extension-queue entry i is record i, that of shadow-queue entry
// j is record cap - 1 - j.  A shade step reads the record beside the ray
// whose result came back and writes it, updated, beside the ray it makes,
// so records move with their rays in queue order and every access is
// coalesced.
struct PathRecs {
    // Thirteen 16-byte fields per record in one allocation, in tiles of
    // kTile records: tile t holds field 0 of its records, then field 1, ...
    // (field f of record r at base[((r / kTile) * FIELDS + f) * kTile +
    // r % kTile]).  A wave's 64 consecutive records still read each field
    // as two whole 512-byte runs, and all thirteen fields of a record lie
    // within +-3 KB of its field 6, so one address per record and immediate
    // offsets reach them (a field-major layout needs one 64-bit base per
    // field, which the shade kernels could only keep in spill lanes).
    static constexpr int kTile = 32;
    float4* base;
    int64_t cap;                // records (paths in flight at most)
    enum Field {
        M0,                     // int4 {seed, state | flags, bounce, material of the current vertex}
        M1,                     // int4 {pixel, job-table entry, frame slot, NEE lobe probability bits}
        HP,                     // vertex position
        HN,                     // shading normal
        RD,                     // incoming direction (the camera ray's for a new path)
        TP,                     // throughput
        DIRECT,                 // direct light of the first vertex
        INDIRECT,               // everything else
        ND,                     // pending bounce direction
        BNC,                    // pending bounce: pdf, lobe probability, ior
        WI,                     // NEE: light direction, w = light pdf
        LS,                     // NEE: MIS-weighted light sample
        MIS,                    // NEE: MIS BRDF direction, w = BRDF pdf
        FIELDS
    };
    __host__ __device__ float4* f(int field, int64_t r) const
    {
        return base + ((size_t)(r / kTile) * FIELDS + (size_t)field) * kTile + (size_t)(r % kTile);
    }
    // float4 entries to allocate for n records (whole tiles)
    static constexpr int64_t alloc_entries(int64_t n) { return (n + kTile - 1) / kTile * kTile * FIELDS; }
};

// Path states, structure of arrays.
struct PathSoA {
    int32_t* job;               // bidirectional: job slot per path (the trace kernels' scene lookup)
    float4* hit;                // closest-hit result: t, tri bits, u, v (per path / queue entry)
    int32_t* occ;               // any-hit result (per shadow-queue entry)
    PathRecs R[2];              // unidirectional: records beside queues 0 and 1
};

// Bidirectional paths (bdpt.cl): the per-path state, both subpaths'
// vertices and the bookkeeping of the connection rounds, in one allocation
// of 16-byte fields in tiles of 32 paths, as PathRecs (field f of path p at
// base[((p / 32) * fields + f) * 32 + p % 32]): a path's fields lie within
// a few KB of each other, so the shade keeps few addresses and no base
// pointer per field.
struct BdptRecs {
    static constexpr int kTile = 32;
    enum Field {
        I0,                     // int4 {seed, state | flags, bounce (vertex being built), pixel y*W + x}
        I1,                     // int4 {frame slot, light subpath length, eye subpath length, eye_path_weight bits}
        I2,                     // int4 {connection slots of the round (64 bits), NEE lobe probability bits, -}
        DIRECT,                 // colour accumulated so far
        ND,                     // pending bounce direction
        BNC,                    // pending bounce: pdf, brdf_prob, -, -
        WI,                     // NEE: normalised w_i, w = light pdf
        LS,                     // NEE: light sample (MIS-weighted)
        MIS,                    // NEE: MIS brdf dir xyz, w = brdf pdf
        EYE_O,                  // camera ray, kept until the eye subpath starts
        EYE_D,
        VTX                     // vertex v = side * K + k (side 0: light, 1: eye) at VTX + 4 v:
    };                          //   position (w = its triangle's material id bits, -1: on the light), normal,
                                //   direction of the ray that found it, contribution
    float4* base;
    int32_t fields;             // VTX + 8 K
    __host__ __device__ float4* f(int field, int64_t p) const
    {
        return base + ((size_t)(p / kTile) * (size_t)fields + (size_t)field) * kTile + (size_t)(p % kTile);
    }
    static int64_t alloc_entries(int64_t n, int fields) { return (n + kTile - 1) / kTile * kTile * fields; }
};

struct BdptSoA {
    BdptRecs R;
    int32_t* occ;               // [cap << shift] any-hit results of a round
    int32_t K;                  // vertices kept per subpath
    int32_t shift;              // log2 of shadow-ray slots per path and round
    int32_t cap;                // paths
};

// Counters and stats in device memory.
constexpr int kMaxXcd = 16;     // work counters per trace launch (one per queue slice)

// Device-scope atomics on one cache line serialise, so every counter that
// many waves update sits on a 256-byte line of its own: the queue counts
// (appended to by each block of a shade pass) and the work counters of the
// trace's queue slices.  The work counters come last, after the part the
// host reads back.
constexpr int kWorkStride = 64;

struct DevCounters {
    unsigned long long rays_ext;
    unsigned long long rays_shadow;
    unsigned long long nodes_ext;     // counting mode only: interior nodes visited
    unsigned long long tris_ext;      // counting mode only: triangle tests
    unsigned long long nodes_shadow;
    unsigned long long tris_shadow;
    unsigned long long pad[26];
    uint32_t q[12][kWorkStride];      // queue counts and flags (see runtime), q[k][0] used
    uint32_t work[2][kMaxXcd][kWorkStride];   // chunk counters of the extend / shadow traces, per queue slice
};
constexpr size_t kCtrHostBytes = offsetof(DevCounters, work);   // what the host reads back

}  // namespace yune
