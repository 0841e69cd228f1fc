// bvh_build.cpp — level-order binned-SAH BVH, bit-identical to the
// reference builder (src/BVH.cpp:56-247, src/TriangleCPU.cpp:41-72,
// src/BVHNodeCPU.cpp:30-51, root box as src/Scene.cpp:243-244, 351-356).
//
// The reference copies primitive vectors for each of the bins-1 candidate
// planes of every node.  This builder keeps one index array, counts the
// candidates with a per-node histogram (a primitive in child 1 of plane i is
// in child 1 of every later plane, since plane positions only grow), and
// partitions each node stably in place.  Nodes of one level are built in
// parallel; child indices are handed out in level order afterwards, which is
// the order the reference appends them in.  Every float expression that
// reaches the output is evaluated with the reference's operands and order
// (compiled with -ffp-contract=off), and min/max ties resolve as std::min /
// std::max do, so node boxes match bit for bit.
#include "yn_internal.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cstring>
#include <thread>
#include <vector>

namespace yune {

namespace {

inline float std_min(float a, float b) { return (b < a) ? b : a; }
inline float std_max(float a, float b) { return (a < b) ? b : a; }

struct Box {
    float mn[4], mx[4];
};

inline Box empty_box()
{
    Box b;
    b.mn[0] = b.mn[1] = b.mn[2] = FLT_MAX; b.mn[3] = 1.0f;
    b.mx[0] = b.mx[1] = b.mx[2] = -FLT_MAX; b.mx[3] = 1.0f;
    return b;
}

// BVH::getExtent: copy of bb1, then min/max(bb2, bb1) per axis.
inline Box get_extent(const Box& b1, const Box& b2)
{
    Box p = b1;
    for (int i = 0; i < 3; ++i) {
        p.mn[i] = std_min(b2.mn[i], b1.mn[i]);
        p.mx[i] = std_max(b2.mx[i], b1.mx[i]);
    }
    return p;
}

inline float surface_area(const Box& b)
{
    float d0 = b.mx[0] - b.mn[0], d1 = b.mx[1] - b.mn[1], d2 = b.mx[2] - b.mn[2];
    return 2.0f * (d0 * d1 + d0 * d2 + d1 * d2);
}

struct Prim {
    Box aabb;
    float c[3];
};

struct BNode {
    Box aabb;
    int32_t child_idx = -2, vert_len = -1;
    int32_t vert_list[10];
    uint32_t begin = 0, end = 0;        // primitive range in the index array
    int32_t split = 0;                  // 1 if this node splits (set while building its level)
    Box c1, c2;                          // child boxes chosen for a split
    uint32_t n1 = 0;                     // primitives going to child 1
};

// Centroid-in-box test of populateChildNodes (BVH.cpp:196-208), written as
// the reference's negated comparison so NaN centroids behave the same.
inline bool in_box(const float c[3], const Box& b)
{
    for (int k = 0; k < 3; ++k)
        if (c[k] < b.mn[k] || c[k] > b.mx[k]) return false;
    return true;
}

template <class F>
void parallel_for(int nthreads, size_t n, size_t grain, F f)
{
    if (nthreads <= 1 || n <= grain) { f(size_t(0), n, 0); return; }
    size_t chunks = std::min<size_t>((size_t)nthreads, (n + grain - 1) / grain);
    std::vector<std::thread> th;
    for (size_t t = 0; t < chunks; ++t) {
        size_t b = n * t / chunks, e = n * (t + 1) / chunks;
        th.emplace_back(f, b, e, (int)t);
    }
    for (auto& x : th) x.join();
}

struct Builder {
    const yn_triangle* tris;
    int n_tri, bins, nthreads;
    std::vector<Prim> prims;
    std::vector<uint32_t> idx, tmp;
    std::vector<BNode> nodes;

    static constexpr int kLeaf = 10;            // BVH.cpp:37
    static constexpr float kIsect = 1.0f;       // BVH.cpp:38
    static constexpr float kTrav = 1 / 8.0f;    // BVH.cpp:39

    void prepare()
    {
        prims.resize(n_tri);
        parallel_for(nthreads, (size_t)n_tri, 1 << 14, [&](size_t b, size_t e, int) {
            for (size_t i = b; i < e; ++i) {
                const float* v1 = &tris[i].v1.x; const float* v2 = &tris[i].v2.x; const float* v3 = &tris[i].v3.x;
                Prim& p = prims[i];
                for (int k = 0; k < 3; ++k) p.c[k] = ((v1[k] + v2[k]) + v3[k]) / 3.0f;
                float mn[4] = {0, 0, 0, 1}, mx[4] = {0, 0, 0, 1};
                for (int k = 0; k < 3; ++k) {
                    mn[k] = std_min(v1[k], std_min(v2[k], v3[k]));
                    mx[k] = std_max(v1[k], std_max(v2[k], v3[k]));
                }
                for (int k = 0; k < 3; ++k)
                    if (mx[k] - mn[k] == 0.0f) mx[k] += 0.2f;
                std::memcpy(p.aabb.mn, mn, sizeof mn);
                std::memcpy(p.aabb.mx, mx, sizeof mx);
            }
        });
    }

    // Decide leaf / split for one node; partition its primitives stably.
    // `par` lets the biggest nodes use several threads for their loops.
    void process(BNode& nd, int par)
    {
        uint32_t n = nd.end - nd.begin;
        if (n == 0) return;
        if (n <= (uint32_t)kLeaf) {
            nd.vert_len = (int32_t)n;
            nd.child_idx = -1;
            for (uint32_t j = 0; j < n; ++j) nd.vert_list[j] = (int32_t)idx[nd.begin + j];
            return;
        }
        const uint32_t* P = idx.data() + nd.begin;
        // findSplitAxis (BVH.cpp:280-314): longest extent of the union of
        // primitive boxes; min/max values do not depend on order.
        float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        {
            int nt = par > 1 && n > (1u << 15) ? par : 1;
            std::vector<float> part(6 * nt);
            parallel_for(nt, n, 1 << 14, [&](size_t b, size_t e, int t) {
                float a[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, z[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
                for (size_t j = b; j < e; ++j) {
                    const Box& bb = prims[P[j]].aabb;
                    for (int k = 0; k < 3; ++k) {
                        a[k] = std_min(bb.mn[k], a[k]);
                        z[k] = std_max(bb.mx[k], z[k]);
                    }
                }
                for (int k = 0; k < 3; ++k) { part[6 * t + k] = a[k]; part[6 * t + 3 + k] = z[k]; }
            });
            int used = nt > 1 ? (int)std::min<size_t>((size_t)nt, (n + (1 << 14) - 1) / (1 << 14)) : 1;
            for (int t = 0; t < used; ++t)
                for (int k = 0; k < 3; ++k) {
                    mn[k] = std_min(part[6 * t + k], mn[k]);
                    mx[k] = std_max(part[6 * t + 3 + k], mx[k]);
                }
        }
        int axis = 0;
        {
            float diff = -FLT_MAX;
            for (int i = 0; i < 3; ++i) {
                float d = mx[i] - mn[i];
                if (d > diff) { diff = d; axis = i; }
            }
        }
        const Box& pb = nd.aabb;
        bool median = true;
        Box c1 = pb, c2 = pb;
        if (bins > 2 && n > 20) {
            // SAH over bins-1 planes (BVH.cpp:89-143).
            float cost = surface_area(pb) * kIsect * (float)n;
            float initial_cost = cost;
            float increment = pb.mx[axis] - pb.mn[axis];
            increment /= (float)bins;
            int np = bins - 1;
            std::vector<float> plane(np);
            for (int i = 0; i < np; ++i) plane[i] = pb.mn[axis] + (float)(i + 1) * increment;
            // hist[i]: primitives whose first containing plane is i.
            std::vector<uint32_t> hist(np + 1, 0);
            {
                int nt = par > 1 && n > (1u << 15) ? par : 1;
                std::vector<std::vector<uint32_t>> h(nt, std::vector<uint32_t>(np + 1, 0));
                parallel_for(nt, n, 1 << 14, [&](size_t b, size_t e, int t) {
                    uint32_t* hh = h[t].data();
                    for (size_t j = b; j < e; ++j) {
                        const float* c = prims[P[j]].c;
                        bool ok = true;
                        for (int k = 0; k < 3; ++k) {
                            if (k == axis) continue;
                            if (c[k] < pb.mn[k] || c[k] > pb.mx[k]) { ok = false; break; }
                        }
                        if (!ok || c[axis] < pb.mn[axis]) { hh[np]++; continue; }
                        // first plane with !(c > plane)
                        int lo = 0, hi = np;
                        while (lo < hi) {
                            int mid = (lo + hi) >> 1;
                            if (!(c[axis] > plane[mid])) hi = mid; else lo = mid + 1;
                        }
                        hh[lo]++;
                    }
                });
                for (int t = 0; t < nt; ++t)
                    for (int i = 0; i <= np; ++i) hist[i] += h[t][i];
            }
            float sp = surface_area(pb);
            uint32_t acc = 0;
            int best = -1;
            for (int i = 0; i < np; ++i) {
                acc += hist[i];
                Box a1 = pb, a2 = pb;
                a1.mx[axis] = plane[i];
                a2.mn[axis] = a1.mx[axis];
                uint32_t n1;
                // populateChildNodes' empty-child rule (BVH.cpp:177-193)
                if (surface_area(a1) == 0.0f) n1 = 0;
                else if (surface_area(a2) == 0.0f) n1 = n;
                else n1 = acc;
                uint32_t n2 = n - n1;
                float new_cost = kTrav + (surface_area(a1) / sp) * (kIsect * (float)n1)
                               + surface_area(a2) / sp * (kIsect * (float)n2);
                if (new_cost < cost) { best = i; cost = new_cost; c1 = a1; c2 = a2; }
            }
            median = (cost == initial_cost);
            (void)best;
        }
        if (median) {   // BVH.cpp:145-160
            float center[3];
            for (int k = 0; k < 3; ++k) center[k] = (pb.mn[k] + pb.mx[k]) / 2.0f;
            c1 = pb; c1.mx[axis] = center[axis];
            c2 = pb; c2.mn[axis] = center[axis];
        }
        // populateChildNodes with the chosen boxes: stable partition.
        int empty = -1;
        if (surface_area(c1) == 0.0f) empty = 0;
        else if (surface_area(c2) == 0.0f) empty = 1;
        uint32_t n1 = 0;
        if (empty == 0) n1 = 0;
        else if (empty == 1) n1 = n;
        else {
            int nt = par > 1 && n > (1u << 15) ? par : 1;
            size_t chunks = nt > 1 ? std::min<size_t>((size_t)nt, (n + (1 << 14) - 1) / (1 << 14)) : 1;
            std::vector<uint32_t> cnt(chunks + 1, 0);
            std::vector<uint8_t> flag(n);
            parallel_for((int)chunks, n, 1 << 14, [&](size_t b, size_t e, int t) {
                uint32_t k = 0;
                for (size_t j = b; j < e; ++j) { flag[j] = in_box(prims[P[j]].c, c1); k += flag[j]; }
                cnt[t] = k;
            });
            std::vector<uint32_t> off1(chunks), off2(chunks);
            uint32_t s = 0;
            for (size_t t = 0; t < chunks; ++t) { off1[t] = s; s += cnt[t]; }
            n1 = s;
            uint32_t s2 = n1;
            for (size_t t = 0; t < chunks; ++t) {
                off2[t] = s2;
                size_t b = n * t / chunks, e = n * (t + 1) / chunks;
                s2 += (uint32_t)(e - b) - cnt[t];
            }
            uint32_t* T = tmp.data() + nd.begin;
            parallel_for((int)chunks, n, 1 << 14, [&](size_t b, size_t e, int t) {
                uint32_t a = off1[t], z = off2[t];
                for (size_t j = b; j < e; ++j) {
                    if (flag[j]) T[a++] = P[j]; else T[z++] = P[j];
                }
            });
            std::memcpy(idx.data() + nd.begin, T, sizeof(uint32_t) * n);
        }
        nd.split = 1;
        nd.c1 = c1;
        nd.c2 = c2;
        nd.n1 = n1;
    }

    int build(std::vector<yn_bvh_node>& out)
    {
        prepare();
        Box root = empty_box();
        for (int i = 0; i < n_tri; ++i)
            for (int k = 0; k < 3; ++k) {
                root.mn[k] = std_min(root.mn[k], prims[i].aabb.mn[k]);
                root.mx[k] = std_max(root.mx[k], prims[i].aabb.mx[k]);
            }
        idx.resize(n_tri);
        tmp.resize(n_tri);
        for (int i = 0; i < n_tri; ++i) idx[i] = (uint32_t)i;
        nodes.clear();
        nodes.reserve((size_t)n_tri / 2 + 16);
        BNode r;
        r.aabb = root;
        for (int k = 0; k < 10; ++k) r.vert_list[k] = -1;
        r.begin = 0; r.end = (uint32_t)n_tri;
        nodes.push_back(r);
        const size_t node_limit = 64ull * (size_t)n_tri + 1024;
        size_t lb = 0, le = 1;
        while (lb < le) {
            // Big nodes one at a time with threaded loops, the rest spread
            // over threads one node each.
            std::vector<size_t> big, small;
            for (size_t i = lb; i < le; ++i)
                ((nodes[i].end - nodes[i].begin) > (1u << 17) ? big : small).push_back(i);
            for (size_t i : big) process(nodes[i], nthreads);
            if (!small.empty()) {
                std::atomic<size_t> next{0};
                int nt = std::min<int>(nthreads, (int)((small.size() + 63) / 64));
                if (nt < 1) nt = 1;
                auto work = [&]() {
                    for (;;) {
                        size_t k = next.fetch_add(64);
                        if (k >= small.size()) break;
                        size_t e = std::min(small.size(), k + 64);
                        for (size_t j = k; j < e; ++j) process(nodes[small[j]], 1);
                    }
                };
                if (nt == 1) work();
                else {
                    std::vector<std::thread> th;
                    for (int t = 0; t < nt; ++t) th.emplace_back(work);
                    for (auto& x : th) x.join();
                }
            }
            // Append children in node order (BVH.cpp:161-163).
            size_t next_level = nodes.size();
            for (size_t i = lb; i < le; ++i) {
                if (!nodes[i].split) continue;
                if (nodes.size() + 2 > node_limit) return -1;
                BNode a, b;
                for (int k = 0; k < 10; ++k) a.vert_list[k] = b.vert_list[k] = -1;
                a.aabb = nodes[i].c1;
                b.aabb = nodes[i].c2;
                a.begin = nodes[i].begin; a.end = nodes[i].begin + nodes[i].n1;
                b.begin = a.end; b.end = nodes[i].end;
                nodes[i].child_idx = (int32_t)nodes.size();
                nodes.push_back(a);
                nodes.push_back(b);
            }
            lb = next_level;
            le = nodes.size();
        }
        // resizeBvh (BVH.cpp:218-247): bottom-up refit.
        for (size_t j = nodes.size(); j-- > 0;) {
            BNode& nd = nodes[j];
            if (nd.vert_len > 0) {
                Box p = empty_box();
                for (int k = 0; k < nd.vert_len; ++k) p = get_extent(p, prims[nd.vert_list[k]].aabb);
                nd.aabb = p;
            } else if (nd.child_idx > 0) {
                nd.aabb = get_extent(nodes[nd.child_idx].aabb, nodes[nd.child_idx + 1].aabb);
            }
        }
        out.resize(nodes.size());
        for (size_t i = 0; i < nodes.size(); ++i) {
            yn_bvh_node& g = out[i];
            std::memcpy(&g.aabb.p_min, nodes[i].aabb.mn, 16);
            std::memcpy(&g.aabb.p_max, nodes[i].aabb.mx, 16);
            std::memcpy(g.vert_list, nodes[i].vert_list, sizeof g.vert_list);
            g.child_idx = nodes[i].child_idx;
            g.vert_len = nodes[i].vert_len;
        }
        return (int)nodes.size();
    }
};

}  // namespace

int build_sah_bvh(const yn_triangle* tris, int n_tri, int bins, int nthreads,
                  std::vector<yn_bvh_node>& out)
{
    if (n_tri < 0 || (n_tri > 0 && !tris)) return -1;
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    Builder b{tris, n_tri, bins, nthreads, {}, {}, {}, {}};
    return b.build(out);
}

}  // namespace yune
