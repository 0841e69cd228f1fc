// yn_internal.h — declarations shared by the host-side C++ of libyune_hip.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/yune_types.h"

namespace yune {

// SAH builder bit-identical to src/BVH.cpp (bvh_build.cpp).  Returns the
// node count, or -1 on bad input / a node count that shows the reference
// algorithm would not terminate.
int build_sah_bvh(const yn_triangle* tris, int n_tri, int bins, int nthreads,
                  std::vector<yn_bvh_node>& out);

// ---- GPU layout (bvh_layout.cpp) --------------------------------------
// Child reference of a wide node:
//   >= 0                interior node index
//   kLeafBit|first<<4|n leaf: n (1..15) triangle slots starting at `first`
//   kEmptyRef           no child (empty reference node, or unused slot)
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmptyRef = 0x80000000u;   // a leaf of 0 triangles
constexpr int kMaxLeafFirst = (1 << 27) - 1;

struct GpuLayout {
    int width = 2;                      // 2 or 4 children per node
    std::vector<float> nodes;           // width 2: 16 floats/node; width 4: 32 floats/node
    int n_nodes = 0;
    uint32_t root_ref = kEmptyRef;
    float root_box[6] = {0, 0, 0, 0, 0, 0};   // tested only when the root is a leaf
    int max_depth = 0;                  // stack entries traversal can need
    int brute = 0;                      // no BVH: every triangle, in index order
    int n_hoist = 0;                    // slots [0, n_hoist): tested by every ray first
    std::vector<float> hoist_box;       // per hoisted slot: lo.xyz, -, hi.xyz, -
    std::vector<float> tri_isect;       // per slot: v1 (w = id bits), e1, e2 -> 12 floats
    std::vector<float> tri_shade;       // per triangle: normalize(vn1..3), matID bits -> 16 floats
    std::vector<int32_t> tri_mat;       // per triangle: matID
    std::vector<int32_t> slot_tri;      // per slot: triangle id
};

// Validates the reference arrays (every index in range, leaves 1..10) and
// builds the GPU layout.  Returns 0 or a negative status with `err` set.
// hoist_frac > 0: triangles whose box area exceeds hoist_frac x the root's
// (at most kMaxHoist, largest first) leave their leaves and are tested by
// every ray before traversal; the tree is refitted without them.
// leaf_max < 10: larger leaves are split into subtrees.
constexpr int kMaxHoist = 64;
struct LayoutOpts {
    int width = 4;
    float hoist_frac = 0.01f;
    int leaf_max = 10;
    bool rebuild = false;       // the library's own SAH tree instead of the caller's
};
// The library's own binned-SAH tree (3 axes, leaves of at most leaf_max
// triangles) in the reference's node format; returns the node count.
int build_gpu_bvh(const yn_triangle* tris, int n_tri, int leaf_max, std::vector<yn_bvh_node>& out);
int make_layout(const yn_triangle* tris, int n_tri, int n_mat, const yn_bvh_node* nodes,
                int n_node, const LayoutOpts& opt, GpuLayout& out, std::string& err);

// ---- scenes (scene_io.cpp) ----------------------------------------------
struct HostScene {
    std::vector<yn_triangle> tris;
    std::vector<yn_material> mats;
    yn_light_quad light{};
    yn_cam cam{};
};

yn_material new_material();     // newMaterial(), include/CL_headers.h:109-124
int load_obj(const std::string& obj_path, HostScene& out, std::string& err);
void gen_cornell(HostScene& out);
void gen_icosphere_room(uint32_t seed, int n_spheres, int subdiv, HostScene& out);
void make_camera(const float side[3], const float up[3], const float look_at[3],
                 const float eye[3], float yfov_deg, yn_cam& out);

// ---- images (image_io.cpp) ----------------------------------------------
int write_hdr(const std::string& path, int w, int h, const float* rgb_top_down);
int write_png(const std::string& path, int w, int h, const uint8_t* rgb_top_down);

// Message for yn_last_error(NULL): errors of calls made without a context.
void set_global_error(const std::string& msg);

}  // namespace yune
