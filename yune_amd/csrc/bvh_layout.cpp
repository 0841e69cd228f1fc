// bvh_layout.cpp — the GPU node layout built from the reference's flattened
// level-order BVH (the BVHNodeGPU array of include/CL_headers.h:85-91 that
// src/BVH.cpp emits and kernels/legacy/udpt.cl:288-324 walks).
//
// The reference node stores its own box plus an index to two children; the
// kernel re-reads each child's 80-byte node just to test its box.  Here each
// node carries its children's boxes (copied bit for bit from the reference
// nodes, so every box test sees the same numbers), laid out per axis so one
// 16-byte load feeds the same slab of every child:
//   width 2 (64 B):  [lo_x0 lo_x1 hi_x0 hi_x1] [.. y ..] [.. z ..] [c0 c1 - -]
//   width 4 (128 B): [lo_x0..3] [hi_x0..3] [lo_y..] [hi_y..] [lo_z..] [hi_z..] [c0..c3] [-]
// Width 4 collapses two reference levels: an interior child with the largest
// surface area is replaced by its own two children until four are held.
// Nodes are numbered depth first, so a ray that descends touches nearby
// memory.  Leaf triangles are copied into slots in the same depth-first
// order, keeping the reference's vert_list order inside a leaf:
//   tri_isect[slot] = v1 (w = triangle id bits), v2 - v1, v3 - v1
// (the reference computes the same two differences in rayTriangleIntersection,
// udpt.cl:328-329).  Shading data stays indexed by triangle id and is read
// once per closest hit: normalize(vn1..3) (udpt.cl:377-379) and matID.
#include "yn_internal.h"

#include <cmath>
#include <cstring>

namespace yune {

namespace {

inline float u2f_bits(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

inline float box_area(const yn_aabb& b)
{
    float dx = b.p_max.x - b.p_min.x, dy = b.p_max.y - b.p_min.y, dz = b.p_max.z - b.p_min.z;
    return dx * dy + dx * dz + dy * dz;
}

}  // namespace

int make_layout(const yn_triangle* tris, int n_tri, int n_mat, const yn_bvh_node* nodes, int n_node,
                int width, GpuLayout& L, std::string& err)
{
    L = GpuLayout();
    if (width != 2 && width != 4) { err = "BVH width must be 2 or 4"; return -2; }
    L.width = width;
    if (n_tri < 0 || n_node < 0 || (n_tri > 0 && !tris) || (n_node > 0 && !nodes)) {
        err = "bad scene arrays"; return -2;
    }
    if (n_tri > kMaxLeafFirst) { err = "too many triangles"; return -2; }
    for (int i = 0; i < n_tri; ++i)
        if (tris[i].matID < 0 || tris[i].matID >= n_mat) {
            err = "triangle " + std::to_string(i) + " has matID out of range"; return -2;
        }

    // Shading arrays, by triangle id.
    L.tri_shade.resize((size_t)n_tri * 12);
    L.tri_mat.resize(n_tri);
    for (int i = 0; i < n_tri; ++i) {
        const yn_float4* vn[3] = {&tris[i].vn1, &tris[i].vn2, &tris[i].vn3};
        for (int k = 0; k < 3; ++k) {
            float x = vn[k]->x, y = vn[k]->y, z = vn[k]->z, w = vn[k]->w;
            float inv = 1.0f / sqrtf(x * x + y * y + z * z + w * w);
            float* o = &L.tri_shade[(size_t)i * 12 + 4 * k];
            o[0] = x * inv; o[1] = y * inv; o[2] = z * inv; o[3] = w * inv;
        }
        L.tri_mat[i] = tris[i].matID;
    }
    auto add_slot = [&](int tri) {
        const yn_triangle& t = tris[tri];
        float v[12] = {t.v1.x, t.v1.y, t.v1.z, u2f_bits((uint32_t)tri),
                       t.v2.x - t.v1.x, t.v2.y - t.v1.y, t.v2.z - t.v1.z, 0.0f,
                       t.v3.x - t.v1.x, t.v3.y - t.v1.y, t.v3.z - t.v1.z, 0.0f};
        L.tri_isect.insert(L.tri_isect.end(), v, v + 12);
        L.slot_tri.push_back(tri);
    };

    if (n_node == 0) {          // brute force over every triangle (udpt.cl:280-284)
        L.brute = 1;
        for (int i = 0; i < n_tri; ++i) add_slot(i);
        L.root_ref = kEmptyRef;
        return 0;
    }

    // Classify reference nodes exactly as traverseBVH reads them.
    enum { EMPTY = 0, LEAF = 1, INNER = 2 };
    std::vector<uint8_t> kind(n_node);
    for (int j = 0; j < n_node; ++j) {
        const yn_bvh_node& nd = nodes[j];
        if (nd.child_idx == -1 && nd.vert_len > 0) {
            if (nd.vert_len > 10) { err = "node " + std::to_string(j) + ": vert_len > 10"; return -2; }
            for (int k = 0; k < nd.vert_len; ++k)
                if (nd.vert_list[k] < 0 || nd.vert_list[k] >= n_tri) {
                    err = "node " + std::to_string(j) + ": triangle index out of range"; return -2;
                }
            kind[j] = LEAF;
        } else if (nd.child_idx > 0) {
            if (nd.child_idx <= j || nd.child_idx + 1 >= n_node) {
                err = "node " + std::to_string(j) + ": child_idx out of range"; return -2;
            }
            kind[j] = INNER;
        } else if (nd.vert_len > 0) {
            err = "node " + std::to_string(j) + ": neither leaf nor interior"; return -2;
        } else {
            kind[j] = EMPTY;
        }
    }
    std::vector<uint8_t> seen(n_node, 0);

    if (kind[0] == LEAF) {
        uint32_t first = (uint32_t)L.slot_tri.size();
        for (int k = 0; k < nodes[0].vert_len; ++k) add_slot(nodes[0].vert_list[k]);
        L.root_ref = kLeafBit | (first << 4) | (uint32_t)nodes[0].vert_len;
        const yn_aabb& b = nodes[0].aabb;
        float rb[6] = {b.p_min.x, b.p_min.y, b.p_min.z, b.p_max.x, b.p_max.y, b.p_max.z};
        std::memcpy(L.root_box, rb, sizeof rb);
        L.max_depth = 0;
        return 0;
    }
    if (kind[0] == EMPTY) { L.root_ref = kEmptyRef; return 0; }

    const int fpn = width == 2 ? 16 : 32;   // floats per node
    // Children of a wide node: reference node ids.
    auto gather = [&](int j, int* ch) -> int {
        int n = 0;
        int c = nodes[j].child_idx;
        if (kind[c] != EMPTY || width == 2) ch[n++] = c;
        if (kind[c + 1] != EMPTY || width == 2) ch[n++] = c + 1;
        while (width == 4 && n < 4) {
            int best = -1;
            float best_a = -1.0f;
            for (int k = 0; k < n; ++k)
                if (kind[ch[k]] == INNER) {
                    float a = box_area(nodes[ch[k]].aabb);
                    if (best < 0 || a > best_a) { best = k; best_a = a; }
                }
            if (best < 0) break;
            int p = ch[best];
            if (seen[p]) return -1;
            seen[p] = 1;
            int cc = nodes[p].child_idx;
            int add[2], na = 0;
            if (kind[cc] != EMPTY) add[na++] = cc;
            if (kind[cc + 1] != EMPTY) add[na++] = cc + 1;
            if (na == 0) { ch[best] = ch[--n]; continue; }
            ch[best] = add[0];
            if (na == 2) ch[n++] = add[1];
        }
        return n;
    };

    // Iterative depth-first emission.
    struct Frame { int ref_node; int gnode; int depth; };
    std::vector<Frame> todo;
    auto new_gnode = [&]() {
        int id = L.n_nodes++;
        L.nodes.resize((size_t)L.n_nodes * fpn, 0.0f);
        return id;
    };
    seen[0] = 1;
    int root = new_gnode();
    L.root_ref = (uint32_t)root;
    todo.push_back({0, root, 1});
    while (!todo.empty()) {
        Frame f = todo.back();
        todo.pop_back();
        int ch[4];
        int n = gather(f.ref_node, ch);
        if (n < 0) { err = "BVH is not a tree"; return -2; }
        if (f.depth * (width - 1) > L.max_depth) L.max_depth = f.depth * (width - 1);
        float* g = &L.nodes[(size_t)f.gnode * fpn];
        uint32_t refs[4] = {kEmptyRef, kEmptyRef, kEmptyRef, kEmptyRef};
        // First pass: leaves take slots now; interior children get node ids
        // in order, and are expanded depth first (first child next).
        std::vector<Frame> kids;
        for (int k = 0; k < n; ++k) {
            int c = ch[k];
            if (seen[c] && kind[c] != EMPTY) { err = "BVH is not a tree"; return -2; }
            seen[c] = 1;
            const yn_aabb& b = nodes[c].aabb;
            float lo[3] = {b.p_min.x, b.p_min.y, b.p_min.z}, hi[3] = {b.p_max.x, b.p_max.y, b.p_max.z};
            if (width == 2) {
                for (int a = 0; a < 3; ++a) { g[4 * a + k] = lo[a]; g[4 * a + 2 + k] = hi[a]; }
            } else {
                for (int a = 0; a < 3; ++a) { g[8 * a + k] = lo[a]; g[8 * a + 4 + k] = hi[a]; }
            }
            if (kind[c] == LEAF) {
                uint32_t first = (uint32_t)L.slot_tri.size();
                if (first > (uint32_t)kMaxLeafFirst) { err = "too many triangle slots"; return -2; }
                for (int q = 0; q < nodes[c].vert_len; ++q) add_slot(nodes[c].vert_list[q]);
                refs[k] = kLeafBit | (first << 4) | (uint32_t)nodes[c].vert_len;
            } else if (kind[c] == INNER) {
                int id = new_gnode();
                g = &L.nodes[(size_t)f.gnode * fpn];   // storage may have moved
                refs[k] = (uint32_t)id;
                kids.push_back({c, id, f.depth + 1});
            } else {
                refs[k] = kEmptyRef;
            }
        }
        int* gi = reinterpret_cast<int*>(g);
        if (width == 2) { gi[12] = (int)refs[0]; gi[13] = (int)refs[1]; gi[14] = 0; gi[15] = 0; }
        else { for (int k = 0; k < 4; ++k) gi[24 + k] = (int)refs[k]; }
        for (size_t k = kids.size(); k-- > 0;) todo.push_back(kids[k]);
    }
    return 0;
}

}  // namespace yune
