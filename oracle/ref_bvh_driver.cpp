/*
 * ref_bvh_driver.cpp — C entry point around the reference's own SAH-BVH
 * builder, compiled from the reference sources where they lie
 * (src/BVH.cpp, src/BVHNodeCPU.cpp, src/TriangleCPU.cpp) by oracle/Makefile
 * into oracle/_ref/libref_bvh.so.
 *
 * TEST INFRASTRUCTURE ONLY: it pins the oracle's and the product's BVH
 * builders to the reference bit for bit.  Nothing here is shipped.
 *
 * The root box is computed as Scene::loadModel does (Scene.cpp:243-244,
 * 351-356): std::min/std::max over the triangle AABBs in load order.
 */
#include "BVH.h"

#include <algorithm>
#include <cstring>
#include <limits>
#include <vector>

#include "../include/yune_types.h"

static_assert(sizeof(TriangleGPU) == sizeof(yn_triangle), "TriangleGPU layout");
static_assert(sizeof(BVHNodeGPU) == sizeof(yn_bvh_node), "BVHNodeGPU layout");

extern "C" int ref_bvh_build(const yn_triangle* tris, int n_tri, int bins,
                             yn_bvh_node* out, int cap)
{
    std::vector<yune::TriangleCPU> list(n_tri);
    float inf = std::numeric_limits<float>::max();
    AABB root;
    root.p_min = {inf, inf, inf, 1.0};
    root.p_max = {-inf, -inf, -inf, 1.0};
    for (int i = 0; i < n_tri; ++i) {
        std::memcpy(&list[i].props, &tris[i], sizeof(TriangleGPU));
        list[i].computeCentroid();
        for (int k = 0; k < 3; k++) {
            root.p_min.s[k] = std::min(root.p_min.s[k], list[i].aabb.p_min.s[k]);
            root.p_max.s[k] = std::max(root.p_max.s[k], list[i].aabb.p_max.s[k]);
        }
    }
    yune::BVH bvh;
    bvh.createBVH(root, list, bins);
    int n = (int)bvh.gpu_node_list.size();
    for (int i = 0; i < n && i < cap; ++i) {
        std::memcpy(&out[i], &bvh.gpu_node_list[i], sizeof(yn_bvh_node));
        /* vert_list beyond vert_len is uninitialised in the reference
         * (BVHNodeCPU.cpp:30-38); blank it so callers compare only data. */
        int vl = out[i].vert_len > 0 ? out[i].vert_len : 0;
        for (int k = vl; k < 10; ++k) out[i].vert_list[k] = -1;
    }
    return n;
}
