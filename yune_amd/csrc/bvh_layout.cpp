// bvh_layout.cpp — the GPU node layout built from the reference's flattened
// level-order BVH (the BVHNodeGPU array of include/CL_headers.h:85-91 that
// src/BVH.cpp emits and kernels/legacy/udpt.cl:288-324 walks).
//
// The reference node stores its own box plus an index to two children; the
// kernel re-reads each child's 80-byte node just to test its box.  Here each
// node carries its children's boxes, laid out per axis so one 16-byte load
// feeds the same slab of every child:
//   width 2 (64 B):  [lo_x0 lo_x1 hi_x0 hi_x1] [.. y ..] [.. z ..] [c0 c1 - -]
//   width 4 (128 B): [lo_x0..3] [hi_x0..3] [lo_y..] [hi_y..] [lo_z..] [hi_z..] [c0..c3] [-]
// Width 4 collapses two binary levels: an interior child with the largest
// surface area is replaced by its own two children until four are held.
// Nodes are numbered depth first; leaf triangles are copied into slots in
// the same order, keeping vert_list order inside a leaf:
//   tri_isect[slot] = v1 (w = triangle id bits), v2 - v1, v3 - v1
// (the two differences rayTriangleIntersection computes, udpt.cl:328-329).
// Shading data stays indexed by triangle id and is read once per closest
// hit: normalize(vn1..3) (udpt.cl:377-379) and matID.
//
// Two optional transformations of the tree, neither of which changes which
// triangle is closest along a ray (every triangle stays reachable through
// boxes that contain it):
//  * hoisting: triangles whose box covers a large part of the scene (walls
//    of a room) inflate every ancestor box of their leaf, so nearly every
//    ray walks those chains.  They are moved to a short list, each with its
//    own box, tested by every ray before the traversal, and the tree is
//    refitted without them;
//  * leaf splitting: leaves of more than `leaf_max` triangles (the
//    reference allows 10) become small subtrees split at the median
//    centroid along their longest centroid extent.
// With both off the node boxes are the reference's boxes bit for bit.
// Alternatively (`rebuild`) the tree is the library's own SAH build over the
// triangles (sah_tree below), hoisting applied the same way.
#include "yn_internal.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <system_error>
#include <thread>

namespace yune {

namespace {

inline float u2f_bits(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

inline float box_area(const yn_aabb& b)
{
    float dx = b.p_max.x - b.p_min.x, dy = b.p_max.y - b.p_min.y, dz = b.p_max.z - b.p_min.z;
    return dx * dy + dx * dz + dy * dz;
}

// Padded triangle box of TriangleCPU::computeAABB (TriangleCPU.cpp:51-72).
void tri_box(const yn_triangle& t, float mn[3], float mx[3])
{
    const float* v1 = &t.v1.x; const float* v2 = &t.v2.x; const float* v3 = &t.v3.x;
    for (int k = 0; k < 3; ++k) {
        float a = v2[k] < v1[k] ? v2[k] : v1[k];
        mn[k] = v3[k] < a ? v3[k] : a;
        float b = v1[k] < v2[k] ? v2[k] : v1[k];
        mx[k] = b < v3[k] ? v3[k] : b;
        if (mx[k] - mn[k] == 0.0f) mx[k] += 0.2f;
    }
}

float area3(const float mn[3], const float mx[3])
{
    float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    return dx * dy + dx * dz + dy * dz;
}

inline bool live(const yn_bvh_node& n) { return n.vert_len > 0 || n.child_idx > 0; }

void set_empty(yn_bvh_node& n)
{
    n.vert_len = -1;
    n.child_idx = -2;
    for (int q = 0; q < 10; ++q) n.vert_list[q] = -1;
}

// Copy of the reference tree with hoisted triangles removed and large
// leaves split, refitted bottom up.  Children always have larger indices
// than their parent, in the input and in the appended nodes.
void derive_tree(const yn_triangle* tris, const yn_bvh_node* in, int n_node, const std::vector<uint8_t>& hoisted,
                 int leaf_max, std::vector<yn_bvh_node>& out)
{
    out.assign(in, in + n_node);
    for (int j = 0; j < n_node; ++j) {
        yn_bvh_node& nd = out[j];
        if (!(nd.child_idx == -1 && nd.vert_len > 0)) continue;
        int k = 0;
        for (int q = 0; q < nd.vert_len; ++q)
            if (!hoisted[nd.vert_list[q]]) nd.vert_list[k++] = nd.vert_list[q];
        for (int q = k; q < 10; ++q) nd.vert_list[q] = -1;
        nd.vert_len = k;
        if (k == 0) set_empty(nd);
    }
    // split leaves (worklist; appended nodes may split again)
    for (size_t j = 0; j < out.size(); ++j) {
        if (!(out[j].child_idx == -1 && out[j].vert_len > leaf_max)) continue;
        yn_bvh_node nd = out[j];
        int n = nd.vert_len;
        float cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
        float cen[10][3];
        for (int q = 0; q < n; ++q) {
            const yn_triangle& t = tris[nd.vert_list[q]];
            const float* a = &t.v1.x; const float* b = &t.v2.x; const float* c = &t.v3.x;
            for (int k = 0; k < 3; ++k) {
                cen[q][k] = ((a[k] + b[k]) + c[k]) / 3.0f;
                cmn[k] = std::fmin(cmn[k], cen[q][k]);
                cmx[k] = std::fmax(cmx[k], cen[q][k]);
            }
        }
        int axis = 0;
        for (int k = 1; k < 3; ++k)
            if (cmx[k] - cmn[k] > cmx[axis] - cmn[axis]) axis = k;
        int ord[10];
        for (int q = 0; q < n; ++q) ord[q] = q;
        std::stable_sort(ord, ord + n, [&](int a, int b) { return cen[a][axis] < cen[b][axis]; });
        yn_bvh_node a = nd, b = nd;
        int na = n / 2;
        for (int q = 0; q < 10; ++q) a.vert_list[q] = b.vert_list[q] = -1;
        for (int q = 0; q < na; ++q) a.vert_list[q] = nd.vert_list[ord[q]];
        for (int q = na; q < n; ++q) b.vert_list[q - na] = nd.vert_list[ord[q]];
        a.vert_len = na;
        b.vert_len = n - na;
        a.child_idx = b.child_idx = -1;
        int c = (int)out.size();
        out[j].child_idx = c;
        out[j].vert_len = -1;
        for (int q = 0; q < 10; ++q) out[j].vert_list[q] = -1;
        out.push_back(a);
        out.push_back(b);
    }
    // refit
    for (size_t jj = out.size(); jj-- > 0;) {
        yn_bvh_node& nd = out[jj];
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        bool any = false;
        if (nd.child_idx == -1 && nd.vert_len > 0) {
            for (int q = 0; q < nd.vert_len; ++q) {
                float a[3], b[3];
                tri_box(tris[nd.vert_list[q]], a, b);
                for (int c = 0; c < 3; ++c) { mn[c] = std::fmin(mn[c], a[c]); mx[c] = std::fmax(mx[c], b[c]); }
            }
            any = true;
        } else if (nd.child_idx > 0) {
            for (int c = nd.child_idx; c <= nd.child_idx + 1; ++c) {
                const yn_bvh_node& ch = out[c];
                if (!live(ch)) continue;
                any = true;
                const float* a = &ch.aabb.p_min.x; const float* b = &ch.aabb.p_max.x;
                for (int k = 0; k < 3; ++k) { mn[k] = std::fmin(mn[k], a[k]); mx[k] = std::fmax(mx[k], b[k]); }
            }
            if (!any) set_empty(nd);
        }
        if (any) {
            nd.aabb.p_min.x = mn[0]; nd.aabb.p_min.y = mn[1]; nd.aabb.p_min.z = mn[2];
            nd.aabb.p_max.x = mx[0]; nd.aabb.p_max.y = mx[1]; nd.aabb.p_max.z = mx[2];
        }
    }
}

// The library's own binary tree over the triangles not hoisted, in the
// node format of the reference's tree (children at child_idx and
// child_idx+1, leaves of at most leaf_max triangles), for the layout below
// to collapse like the reference's.  Binned SAH (Wald, "On fast Construction
// of SAH-based Bounding Volume Hierarchies", 2007): 32 bins of the centroid
// bounds (n bins for a node of n < 32 triangles) on each of the three axes,
// filled in one pass over the node's triangles, split at the plane of least
//   area(left) * n(left) + area(right) * n(right)
// over the padded triangle boxes, down to leaf_max triangles.  The
// reference's builder (src/BVH.cpp:56-173) bins only the longest axis and
// keeps leaves of up to 10 triangles; on the 983,054-triangle room of
// icospheres its tree has 1.6 times the summed interior-box area of this one
// (the SAH estimate of node visits per ray).  Closest hits do not depend on
// the tree (every triangle stays reachable through boxes that contain it);
// only the order among exactly tied triangles may differ.  Deterministic:
// one thread, fixed bin arithmetic; ~1.4 s for a million triangles, then
// three sweeps of treelet restructuring (below: summed interior area 63.0
// -> 58.9 root areas on that scene, 17.05 -> 16.88 BVH4 node visits per
// ray; ~1.2 s more).
void treelet_optimize(std::vector<yn_bvh_node>& out, int passes);
constexpr int kTreeletPasses = 3;

void sah_tree(const yn_triangle* tris, int n_tri, const std::vector<uint8_t>& hoisted, int leaf_max,
              std::vector<yn_bvh_node>& out)
{
    constexpr int NB = 32;
    struct Box3 {       // plain data: cleared explicitly where a box starts empty
        float lo[3], hi[3];
        void clear()
        {
            for (int k = 0; k < 3; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
        }
        void grow(const float* a, const float* b)     // NaN coordinates leave the box as it was
        {
            for (int k = 0; k < 3; ++k) {
                lo[k] = a[k] < lo[k] ? a[k] : lo[k];
                hi[k] = b[k] > hi[k] ? b[k] : hi[k];
            }
        }
        void grow(const Box3& o) { grow(o.lo, o.hi); }
        double area() const
        {
            if (!(hi[0] >= lo[0])) return 0.0;
            double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
            return dx * dy + dx * dz + dy * dz;
        }
    };
    std::vector<int> ids;
    std::vector<Box3> tb(n_tri);
    std::vector<float> cen((size_t)n_tri * 3);
    for (int t = 0; t < n_tri; ++t) {
        tri_box(tris[t], tb[t].lo, tb[t].hi);
        const float* a = &tris[t].v1.x; const float* b = &tris[t].v2.x; const float* c = &tris[t].v3.x;
        for (int k = 0; k < 3; ++k) cen[3 * (size_t)t + k] = ((a[k] + b[k]) + c[k]) / 3.0f;
        if (!hoisted[t]) ids.push_back(t);
    }
    out.clear();
    yn_bvh_node root{};
    set_empty(root);
    out.push_back(root);
    if (ids.empty()) return;
    struct Item { int node, lo, hi; };
    std::vector<Item> todo{{0, 0, (int)ids.size()}};
    Box3 bins[3][NB];
    int cnts[3][NB];
    while (!todo.empty()) {
        Item it = todo.back();
        todo.pop_back();
        Box3 b, cb;
        b.clear();
        cb.clear();
        for (int i = it.lo; i < it.hi; ++i) {
            b.grow(tb[ids[i]]);
            const float* c = &cen[3 * (size_t)ids[i]];
            cb.grow(c, c);
        }
        yn_bvh_node& nd = out[it.node];
        nd.aabb.p_min.x = b.lo[0]; nd.aabb.p_min.y = b.lo[1]; nd.aabb.p_min.z = b.lo[2]; nd.aabb.p_min.w = 0.0f;
        nd.aabb.p_max.x = b.hi[0]; nd.aabb.p_max.y = b.hi[1]; nd.aabb.p_max.z = b.hi[2]; nd.aabb.p_max.w = 0.0f;
        const int n = it.hi - it.lo;
        if (n <= leaf_max) {
            nd.child_idx = -1;
            nd.vert_len = n;
            for (int q = 0; q < 10; ++q) nd.vert_list[q] = q < n ? ids[it.lo + q] : -1;
            continue;
        }
        // the three axes binned in one pass over the node's triangles; a
        // node of n < 32 triangles uses n bins
        const int nb = n < NB ? n : NB;
        auto bin_of = [&](int t, int ax, float scale) {
            const float f = (cen[3 * (size_t)t + ax] - cb.lo[ax]) * scale;   // NaN goes to bin 0
            return f > 0.0f ? (f < (float)nb ? (int)f : nb - 1) : 0;
        };
        float scales[3];
        for (int ax = 0; ax < 3; ++ax) {
            const float ext = cb.hi[ax] - cb.lo[ax];
            scales[ax] = ext > 0.0f ? (float)nb / ext : 0.0f;
            for (int k = 0; k < nb; ++k) { bins[ax][k].clear(); cnts[ax][k] = 0; }
        }
        for (int i = it.lo; i < it.hi; ++i) {
            const int t = ids[i];
            for (int ax = 0; ax < 3; ++ax) {
                const int k = bin_of(t, ax, scales[ax]);
                cnts[ax][k]++;
                bins[ax][k].grow(tb[t]);
            }
        }
        double best = INFINITY;
        int bax = -1, bsplit = 0;
        for (int ax = 0; ax < 3; ++ax) {
            if (!(scales[ax] > 0.0f)) continue;
            const Box3* bb = bins[ax];
            const int* cnt = cnts[ax];
            double la[NB];
            int lc[NB];
            Box3 acc;
            acc.clear();
            int c = 0;
            for (int k = 0; k < nb; ++k) { acc.grow(bb[k]); c += cnt[k]; la[k] = acc.area(); lc[k] = c; }
            Box3 r;
            r.clear();
            int rc = 0;
            for (int k = nb - 1; k > 0; --k) {
                r.grow(bb[k]);
                rc += cnt[k];
                if (lc[k - 1] == 0 || rc == 0) continue;
                double cost = la[k - 1] * lc[k - 1] + r.area() * rc;
                if (cost < best) { best = cost; bax = ax; bsplit = k; }
            }
        }
        int mid = it.lo + n / 2;
        if (bax >= 0) {
            const float scale = scales[bax];
            auto m = std::partition(ids.begin() + it.lo, ids.begin() + it.hi,
                                    [&](int t) { return bin_of(t, bax, scale) < bsplit; });
            mid = (int)(m - ids.begin());
        }
        if (mid <= it.lo || mid >= it.hi) mid = it.lo + n / 2;   // every centroid in one bin
        const int c = (int)out.size();
        nd.child_idx = c;
        nd.vert_len = -1;
        for (int q = 0; q < 10; ++q) nd.vert_list[q] = -1;
        out.push_back(root);
        out.push_back(root);
        todo.push_back({c + 1, mid, it.hi});
        todo.push_back({c, it.lo, mid});
    }
    treelet_optimize(out, kTreeletPasses);
}

// Treelet restructuring of a binary tree in the reference's node format
// (Karras and Aila, "Fast Parallel Construction of High-Quality Bounding
// Volume Hierarchies", HPG 2013).  For each interior node, bottom up, the
// treelet below it is grown to 7 subtrees by opening the interior subtree
// root of largest area, and rebuilt with the topology of least summed
// interior-box area: a dynamic program over the 127 subsets of its
// subtrees (each split taken over the sub-masks holding the subset's
// lowest member).  Leaves and the subtrees below the treelet are kept, so
// the SAH cost (interior areas + the fixed leaf areas) falls with the
// interior areas alone.  `passes` sweeps over the tree; disjoint subtrees
// below the top 256 are swept on several threads (deterministic: a sweep
// of a subtree touches nothing outside it), the top nodes after them.
// The tree is re-emitted depth first, children in pairs, as sah_tree does.
void treelet_optimize(std::vector<yn_bvh_node>& out, int passes)
{
    struct TN { float lo[3], hi[3]; int l, r; };
    const int n = (int)out.size();
    if (n < 5 || out[0].vert_len > 0 || out[0].child_idx <= 0) return;
    std::vector<TN> T(n);
    for (int i = 0; i < n; ++i) {
        const yn_bvh_node& s = out[i];
        TN& t = T[i];
        t.lo[0] = s.aabb.p_min.x; t.lo[1] = s.aabb.p_min.y; t.lo[2] = s.aabb.p_min.z;
        t.hi[0] = s.aabb.p_max.x; t.hi[1] = s.aabb.p_max.y; t.hi[2] = s.aabb.p_max.z;
        const bool inner = s.vert_len <= 0 && s.child_idx > 0;
        t.l = inner ? s.child_idx : -1;
        t.r = inner ? s.child_idx + 1 : -1;
    }
    auto area = [](const float* lo, const float* hi) {
        const double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        return dx * dy + dx * dz + dy * dz;
    };
    auto optimize = [&](int root) {
        int lv[7], in[6];
        int nl = 2, ni = 1;
        in[0] = root;
        lv[0] = T[root].l;
        lv[1] = T[root].r;
        while (nl < 7) {
            int bk = -1;
            double ba = -1.0;
            for (int k = 0; k < nl; ++k)
                if (T[lv[k]].l >= 0) {
                    const double a = area(T[lv[k]].lo, T[lv[k]].hi);
                    if (a > ba) { ba = a; bk = k; }
                }
            if (bk < 0) break;
            const int x = lv[bk];
            in[ni++] = x;
            lv[bk] = T[x].l;
            lv[nl++] = T[x].r;
        }
        if (nl < 3) return;
        double cur = 0.0;
        for (int k = 0; k < ni; ++k) cur += area(T[in[k]].lo, T[in[k]].hi);
        const int full = (1 << nl) - 1;
        float blo[128][3], bhi[128][3];
        double c[128];
        uint8_t part[128];
        for (int m = 1; m <= full; ++m) {
            const int low = m & -m;
            if (m == low) {
                const TN& t = T[lv[__builtin_ctz(m)]];
                for (int k = 0; k < 3; ++k) { blo[m][k] = t.lo[k]; bhi[m][k] = t.hi[k]; }
                c[m] = 0.0;
                continue;
            }
            const int rest = m ^ low;
            for (int k = 0; k < 3; ++k) {
                blo[m][k] = std::min(blo[rest][k], blo[low][k]);
                bhi[m][k] = std::max(bhi[rest][k], bhi[low][k]);
            }
            double best = INFINITY;
            int bp = low;
            for (int q = rest;; q = (q - 1) & rest) {
                const int pm = q | low;
                if (pm != m) {
                    const double v = c[pm] + c[m ^ pm];
                    if (v < best) { best = v; bp = pm; }
                }
                if (q == 0) break;
            }
            c[m] = area(blo[m], bhi[m]) + best;
            part[m] = (uint8_t)bp;
        }
        if (!(c[full] < cur * (1.0 - 1e-9))) return;
        int next = 1;
        // rebuild: subset m becomes node `id` (the treelet's interior ids reused)
        auto emit = [&](auto&& self, int m, int id) -> int {
            if ((m & (m - 1)) == 0) return lv[__builtin_ctz(m)];
            if (id < 0) id = in[next++];
            const int pm = part[m];
            const int l = self(self, pm, -1), r = self(self, m ^ pm, -1);
            TN& t = T[id];
            t.l = l;
            t.r = r;
            for (int k = 0; k < 3; ++k) { t.lo[k] = blo[m][k]; t.hi[k] = bhi[m][k]; }
            return id;
        };
        emit(emit, full, root);
    };
    // children before parents over the subtree at `top`
    auto sweep = [&](int top) {
        std::vector<std::pair<int, bool>> st{{top, false}};
        while (!st.empty()) {
            auto [x, done] = st.back();
            st.pop_back();
            if (T[x].l < 0) continue;
            if (done) { optimize(x); continue; }
            st.push_back({x, true});
            st.push_back({T[x].r, false});
            st.push_back({T[x].l, false});
        }
    };
    for (int pass = 0; pass < passes; ++pass) {
        // the top nodes (breadth first) until 256 subtrees hang below them
        std::vector<int> top{0}, front;
        size_t h = 0;
        std::vector<int> q{0};
        top.clear();
        while (h < q.size() && q.size() - h < 256) {
            const int x = q[h++];
            if (T[x].l < 0) { front.push_back(x); continue; }
            top.push_back(x);
            q.push_back(T[x].l);
            q.push_back(T[x].r);
        }
        for (; h < q.size(); ++h) front.push_back(q[h]);
        const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::atomic<int> nx{0};
        auto work = [&]() {
            for (int k; (k = nx.fetch_add(1)) < (int)front.size();) sweep(front[k]);
        };
        std::vector<std::thread> th;
        try {
            for (int t = 1; t < nth; ++t) th.emplace_back(work);
        } catch (const std::system_error&) {
            // no more threads: the ones started and this one share the rest
        }
        work();
        for (auto& t : th) t.join();
        for (int k = (int)top.size() - 1; k >= 0; --k) optimize(top[k]);
    }
    // re-emit depth first, children in pairs
    std::vector<yn_bvh_node> res;
    res.reserve(n);
    yn_bvh_node inner = out[0];
    for (int q2 = 0; q2 < 10; ++q2) inner.vert_list[q2] = -1;
    inner.vert_len = -1;
    res.push_back(inner);
    std::vector<std::pair<int, int>> st{{0, 0}};    // (tree node, output index)
    while (!st.empty()) {
        auto [x, o] = st.back();
        st.pop_back();
        const TN& t = T[x];
        yn_bvh_node nd = t.l < 0 ? out[x] : inner;
        nd.aabb.p_min.x = t.lo[0]; nd.aabb.p_min.y = t.lo[1]; nd.aabb.p_min.z = t.lo[2]; nd.aabb.p_min.w = 0.0f;
        nd.aabb.p_max.x = t.hi[0]; nd.aabb.p_max.y = t.hi[1]; nd.aabb.p_max.z = t.hi[2]; nd.aabb.p_max.w = 0.0f;
        if (t.l >= 0) {
            const int cidx = (int)res.size();
            nd.child_idx = cidx;
            res.push_back(inner);
            res.push_back(inner);
            st.push_back({t.r, cidx + 1});
            st.push_back({t.l, cidx});
        }
        res[o] = nd;
    }
    out.swap(res);
}

}  // namespace

int build_gpu_bvh(const yn_triangle* tris, int n_tri, int leaf_max, std::vector<yn_bvh_node>& out)
{
    if (leaf_max < 1 || leaf_max > 10) return -2;
    std::vector<uint8_t> none(n_tri, 0);
    sah_tree(tris, n_tri, none, leaf_max, out);
    return (int)out.size();
}

int make_layout(const yn_triangle* tris, int n_tri, int n_mat, const yn_bvh_node* nodes_in, int n_node_in,
                const LayoutOpts& opt, GpuLayout& L, std::string& err)
{
    L = GpuLayout();
    const int width = opt.width;
    if (width != 2 && width != 4) { err = "BVH width must be 2 or 4"; return -2; }
    if (opt.leaf_max < 1) { err = "leaf_max must be >= 1"; return -2; }
    L.width = width;
    if (n_tri < 0 || n_node_in < 0 || (n_tri > 0 && !tris) || (n_node_in > 0 && !nodes_in)) {
        err = "bad scene arrays"; return -2;
    }
    if (n_tri > kMaxLeafFirst) { err = "too many triangles"; return -2; }
    for (int i = 0; i < n_tri; ++i)
        if (tris[i].matID < 0 || tris[i].matID >= n_mat) {
            err = "triangle " + std::to_string(i) + " has matID out of range"; return -2;
        }

    // Shading arrays, by triangle id: one 64-byte record per triangle (the
    // three normals and the material id, read together at a closest hit
    // in one memory sector), and matID alone.
    L.tri_shade.assign((size_t)n_tri * 16, 0.0f);
    L.tri_mat.resize(n_tri);
    for (int i = 0; i < n_tri; ++i) {
        const yn_float4* vn[3] = {&tris[i].vn1, &tris[i].vn2, &tris[i].vn3};
        for (int k = 0; k < 3; ++k) {
            float x = vn[k]->x, y = vn[k]->y, z = vn[k]->z, w = vn[k]->w;
            float inv = 1.0f / sqrtf(x * x + y * y + z * z + w * w);
            float* o = &L.tri_shade[(size_t)i * 16 + 4 * k];
            o[0] = x * inv; o[1] = y * inv; o[2] = z * inv; o[3] = w * inv;
        }
        L.tri_shade[(size_t)i * 16 + 12] = u2f_bits((uint32_t)tris[i].matID);
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

    if (n_node_in == 0) {       // brute force over every triangle (udpt.cl:280-284)
        L.brute = 1;
        for (int i = 0; i < n_tri; ++i) add_slot(i);
        L.root_ref = kEmptyRef;
        return 0;
    }

    // Classify nodes exactly as traverseBVH reads them; reject anything the
    // reference would index out of bounds with.
    enum { EMPTY = 0, LEAF = 1, INNER = 2 };
    std::vector<uint8_t> kind;
    auto classify = [&](const yn_bvh_node* nds, int n) -> int {
        kind.assign(n, EMPTY);
        for (int j = 0; j < n; ++j) {
            const yn_bvh_node& nd = nds[j];
            if (nd.child_idx == -1 && nd.vert_len > 0) {
                if (nd.vert_len > 10) { err = "node " + std::to_string(j) + ": vert_len > 10"; return -2; }
                for (int k = 0; k < nd.vert_len; ++k)
                    if (nd.vert_list[k] < 0 || nd.vert_list[k] >= n_tri) {
                        err = "node " + std::to_string(j) + ": triangle index out of range"; return -2;
                    }
                kind[j] = LEAF;
            } else if (nd.child_idx > 0) {
                if (nd.child_idx <= j || nd.child_idx + 1 >= n) {
                    err = "node " + std::to_string(j) + ": child_idx out of range"; return -2;
                }
                kind[j] = INNER;
            } else if (nd.vert_len > 0) {
                err = "node " + std::to_string(j) + ": neither leaf nor interior"; return -2;
            }
        }
        return 0;
    };
    if (int r = classify(nodes_in, n_node_in)) return r;

    const yn_bvh_node* nodes = nodes_in;
    int n_node = n_node_in;
    std::vector<yn_bvh_node> derived;
    std::vector<uint8_t> hoisted(n_tri, 0);
    if (opt.hoist_frac > 0.0f && n_tri > 0 && kind[0] == INNER) {
        const yn_aabb& rb = nodes_in[0].aabb;
        float rmn[3] = {rb.p_min.x, rb.p_min.y, rb.p_min.z}, rmx[3] = {rb.p_max.x, rb.p_max.y, rb.p_max.z};
        float root_a = area3(rmn, rmx);
        std::vector<std::pair<float, int>> big;
        for (int t = 0; t < n_tri; ++t) {
            float a[3], b[3];
            tri_box(tris[t], a, b);
            float ar = area3(a, b);
            if (ar > opt.hoist_frac * root_a) big.push_back({-ar, t});
        }
        std::sort(big.begin(), big.end());
        // the largest, at most kMaxHoist and an eighth of the scene: a scene
        // made mostly of big triangles (a Cornell box) is better served by
        // a tree over them than by testing each of them before every walk
        // (all 36 of the Cornell box hoisted: 8% slower than 4 or none)
        const size_t cap = std::min<size_t>(kMaxHoist, std::max<size_t>(1, (size_t)n_tri / 8));
        if (big.size() > cap) big.resize(cap);
        std::vector<int> ids;
        for (auto& b : big) ids.push_back(b.second);
        std::sort(ids.begin(), ids.end());
        for (int t : ids) {
            hoisted[t] = 1;
            add_slot(t);
            float a[3], b[3];
            tri_box(tris[t], a, b);
            float hb[8] = {a[0], a[1], a[2], 0.0f, b[0], b[1], b[2], 0.0f};
            L.hoist_box.insert(L.hoist_box.end(), hb, hb + 8);
        }
        L.n_hoist = (int)ids.size();
    }
    bool any_split = false;
    for (int j = 0; j < n_node_in && !any_split; ++j) any_split = kind[j] == LEAF && nodes_in[j].vert_len > opt.leaf_max;
    if (opt.rebuild && kind[0] != EMPTY) {
        sah_tree(tris, n_tri, hoisted, std::min(opt.leaf_max, 10), derived);
        nodes = derived.data();
        n_node = (int)derived.size();
        if (int r = classify(nodes, n_node)) return r;
    } else if (L.n_hoist > 0 || any_split) {
        derive_tree(tris, nodes_in, n_node_in, hoisted, opt.leaf_max, derived);
        nodes = derived.data();
        n_node = (int)derived.size();
        if (int r = classify(nodes, n_node)) return r;
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
    // The library's own tree is collapsed optimally: T[j][k] is the least
    // cost of binary node j's subtree spread over at most k child slots of
    // a 4-wide node, where each 4-wide node costs its box area (the chance
    // that a ray visits it) and a leaf costs nothing (it fills one slot
    // wherever it goes).  split[j][k]: slots given to j's first child when j
    // is opened, 0 when j stays one slot; open4[j]: the same for the four
    // slots of j as a wide node.  On the 983,054-triangle scene this lowers
    // the summed wide-node area by 2.3% against the largest-area-first rule
    // below, which the reference's tree keeps (it may hold empty nodes).
    const bool dp = width == 4 && opt.rebuild;
    std::vector<std::array<double, 5>> T;
    std::vector<std::array<uint8_t, 5>> split;
    std::vector<uint8_t> open4;
    if (dp) {
        T.assign(n_node, std::array<double, 5>{});
        split.assign(n_node, std::array<uint8_t, 5>{});
        open4.assign(n_node, 1);
        for (int j = n_node - 1; j >= 0; --j) {   // children have larger indices
            if (kind[j] != INNER) continue;        // a leaf: one slot at no cost
            const int l = nodes[j].child_idx, r = l + 1;
            double best = INFINITY;
            for (int i = 1; i <= 3; ++i) {
                const double v = T[l][i] + T[r][4 - i];
                if (v < best) { best = v; open4[j] = (uint8_t)i; }
            }
            T[j][1] = (double)box_area(nodes[j].aabb) + best;
            for (int k = 2; k <= 4; ++k) {
                T[j][k] = T[j][1];
                for (int i = 1; i < k; ++i) {
                    const double v = T[l][i] + T[r][k - i];
                    if (v < T[j][k]) { T[j][k] = v; split[j][k] = (uint8_t)i; }
                }
            }
        }
    }
    auto gather_dp = [&](int j, int* ch) -> int {
        int n = 0;
        struct Part { int node, slots; };
        Part st[8];
        int sp = 0;
        const int c = nodes[j].child_idx;
        st[sp++] = {c + 1, 4 - open4[j]};
        st[sp++] = {c, open4[j]};
        while (sp > 0) {
            const Part q = st[--sp];
            const int i = kind[q.node] == INNER && q.slots > 1 ? split[q.node][q.slots] : 0;
            if (i == 0) { ch[n++] = q.node; continue; }
            if (seen[q.node]) return -1;
            seen[q.node] = 1;
            const int cc = nodes[q.node].child_idx;
            st[sp++] = {cc + 1, q.slots - i};
            st[sp++] = {cc, i};
        }
        return n;
    };
    // Children of a wide node: node ids.
    auto gather = [&](int j, int* ch) -> int {
        if (dp) return gather_dp(j, ch);
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
    std::vector<Frame> kids;
    while (!todo.empty()) {
        Frame f = todo.back();
        todo.pop_back();
        int ch[4];
        int n = gather(f.ref_node, ch);
        if (n < 0) { err = "BVH is not a tree"; return -2; }
        if (f.depth * (width - 1) > L.max_depth) L.max_depth = f.depth * (width - 1);
        uint32_t refs[4] = {kEmptyRef, kEmptyRef, kEmptyRef, kEmptyRef};
        kids.clear();
        for (int k = 0; k < n; ++k) {
            int c = ch[k];
            if (seen[c] && kind[c] != EMPTY) { err = "BVH is not a tree"; return -2; }
            seen[c] = 1;
            const yn_aabb& b = nodes[c].aabb;
            float lo[3] = {b.p_min.x, b.p_min.y, b.p_min.z}, hi[3] = {b.p_max.x, b.p_max.y, b.p_max.z};
            float* g = &L.nodes[(size_t)f.gnode * fpn];
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
                refs[k] = (uint32_t)id;
                kids.push_back({c, id, f.depth + 1});
            }
        }
        int* gi = reinterpret_cast<int*>(&L.nodes[(size_t)f.gnode * fpn]);
        if (width == 2) { gi[12] = (int)refs[0]; gi[13] = (int)refs[1]; gi[14] = 0; gi[15] = 0; }
        else { for (int k = 0; k < 4; ++k) gi[24 + k] = (int)refs[k]; }
        for (size_t k = kids.size(); k-- > 0;) todo.push_back(kids[k]);
    }
    return 0;
}

}  // namespace yune
