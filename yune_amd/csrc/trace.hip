// trace.hip — extend (closest hit) and shadow (any hit) kernels.
//
// Replaces traceRay/traverseBVH/rayAabbIntersection/rayTriangleIntersection
// of kernels/legacy/udpt.cl:240-431 for one queue of rays.  The light quad
// is tested by the shading kernel when it creates a ray (udpt.cl:244-276),
// which then passes the quad distance as the ray's tmax; so "no triangle
// nearer than tmax" reads back as the quad hit, exactly as in traceRay.
//
// Result semantics follow the reference:
//  * closest hit: the triangle with the smallest t in (0, tmax) over every
//    leaf whose box the ray meets.  The reference's breadth-first list never
//    shrinks and ignores t; this traversal visits children near-first and
//    skips a box whose entry distance is beyond the best hit so far (with a
//    1e-4 relative margin, so rounding in the box entry can never hide a
//    triangle the reference would find).  Moller-Trumbore uses the
//    reference's operands and order, so t, u, v and the accept/reject of
//    each triangle are bit-identical to the CPU restatement.  Boxes are
//    tested with the reference's slab arithmetic or with a widened FMA form
//    that enters every box the reference's arithmetic enters (yn_slab.h).
//    Only equal-t ties may resolve to a different triangle (test order).
//  * any hit (shadow rays, udpt.cl:307-308): true if some triangle has t in
//    (0, tmax).
//
// Each lane owns one ray.  Its traversal stack keeps the top LST entries in
// LDS, interleaved by lane so a wave's 64 lanes hit 64 banks, and spills
// deeper entries to a per-thread global area (rarely touched: stack depth
// stays far below the tree depth bound).  An entry is a child reference and
// a 16-bit lower bound of its entry distance, so a deferred child that now
// lies beyond the best hit is dropped on pop without fetching it.  Waves pull
// chunks of rays from the queue with one atomic per chunk.
#include "yn_device.h"
#include "yn_internal_dev.h"
#include "yn_slab.h"

namespace yune {

namespace {

struct RayState {
    float ox, oy, oz, dx, dy, dz, ix, iy, iz;
    float lx, ly, lz;       // FMA slab form: o * dir_inv moved by 2^-21 |o * dir_inv|,
    float hx, hy, hz;       //   one way for the box's lo planes, the other for its hi planes
    float best;             // current tmax
    float cull;             // nodes entered beyond it are skipped: best * (1 + 1e-4) * kWiden
    int tri;                // best hit: triangle id (-1: none), u, v; for any-hit rays
    float u, v;             //   the slot, with id, u and v recomputed at the end if wanted
};

// A child box of an interior node: the reference's slab arithmetic (EXACT)
// or the widened FMA form.
template <bool EXACT>
__device__ __forceinline__ bool child_hit(float lox, float loy, float loz, float hix, float hiy, float hiz,
                                          const RayState& r, float& enter)
{
    return EXACT ? box_hit(lox, loy, loz, hix, hiy, hiz, r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, enter)
                 : box_hit_fma(lox, loy, loz, hix, hiy, hiz, r.ix, r.iy, r.iz, r.lx, r.ly, r.lz, r.hx, r.hy, r.hz,
                               enter);
}

// A ray's traversal state from its queue entry.  Returns whether the FMA
// slab form may be used (see box_hit_fma).
__device__ __forceinline__ bool start_ray(RayState& r, const float4& ot, const float4& dp, const DevScene& S)
{
    r.ox = ot.x; r.oy = ot.y; r.oz = ot.z;
    r.dx = dp.x; r.dy = dp.y; r.dz = dp.z;
    r.ix = 1 / r.dx; r.iy = 1 / r.dy; r.iz = 1 / r.dz;   // dir_inv (udpt.cl:395)
    fma_offsets(r.ox, r.ix, r.lx, r.hx);
    fma_offsets(r.oy, r.iy, r.ly, r.hy);
    fma_offsets(r.oz, r.iz, r.lz, r.hz);
    r.best = ot.w;
    r.cull = cull_of(ot.w);
    r.tri = -1; r.u = 0.0f; r.v = 0.0f;
    return S.fma_slab && fma_form_ok(r.ox, r.oy, r.oz, r.ix, r.iy, r.iz);
}

// A load from a scene array as a global-memory load.  Through a generic
// pointer (the scene table holds them) the compiler emits FLAT loads, which
// also count against the LDS wait counter: every access to the LDS stack
// would then wait for the node fetches still in flight.
// (Through a native vector type: a cast of float4* itself still copies
// through a generic reference.)
typedef float yn_v4f __attribute__((ext_vector_type(4)));
typedef int yn_v4i __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 gload(const float4* p)
{
    yn_v4f v = *(const __attribute__((address_space(1))) yn_v4f*)(const void*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ int4 gload(const int4* p)
{
    yn_v4i v = *(const __attribute__((address_space(1))) yn_v4i*)(const void*)p;
    return make_int4(v.x, v.y, v.z, v.w);
}

// rayTriangleIntersection (udpt.cl:326-390) on a stored slot; xyz only
// (the w terms of the reference's dots are exact zeros).  Returns whether
// (u, v) lies in the triangle, with t; id: the triangle id.
__device__ __forceinline__ bool tri_uvt(const float4* __restrict__ tri, int s, const RayState& r, float& t, float& u,
                                        float& v, int& id)
{
    float4 A = gload(tri + 3 * s), B = gload(tri + 3 * s + 1), C = gload(tri + 3 * s + 2);
    id = __float_as_int(A.w);
    // pvec = cross(dir, v1v3)
    float px = r.dy * C.z - r.dz * C.y;
    float py = r.dz * C.x - r.dx * C.z;
    float pz = r.dx * C.y - r.dy * C.x;
    float det = B.x * px + B.y * py + B.z * pz;
    float inv_det = 1.0f / det;
    float sx = r.ox - A.x, sy = r.oy - A.y, sz = r.oz - A.z;
    u = (px * sx + py * sy + pz * sz) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    // qvec = cross(dist, v1v2)
    float qx = sy * B.z - sz * B.y;
    float qy = sz * B.x - sx * B.z;
    float qz = sx * B.y - sy * B.x;
    v = (qx * r.dx + qy * r.dy + qz * r.dz) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = (C.x * qx + C.y * qy + C.z * qz) * inv_det;
    return true;
}

template <bool ANY>
__device__ __forceinline__ bool tri_test(const float4* __restrict__ tri, int s, RayState& r)
{
    float t, u, v;
    int id;
    if (tri_uvt(tri, s, r, t, u, v, id) && t > 0.0f && t < r.best) {
        r.best = t;
        r.cull = cull_of(t);
        if (ANY) {
            r.tri = s;
        } else {
            r.tri = id;
            r.u = u;
            r.v = v;
        }
        return true;
    }
    return false;
}

// The result record {t, triangle id bits, u, v} of a finished traversal.
// Any-hit rays kept the slot: its id, u and v come from the same test on
// the same inputs as when it was found.
template <bool ANY>
__device__ __forceinline__ float4 hit_record(const float4* __restrict__ tri, const RayState& r)
{
    if (!ANY || r.tri < 0) return make_float4(r.best, __int_as_float(r.tri), r.u, r.v);
    float t, u, v;
    int id;
    tri_uvt(tri, r.tri, r, t, u, v, id);
    return make_float4(r.best, __int_as_float(id), u, v);
}

struct Counts { unsigned nodes, tris; };

template <bool ANY, bool COUNT>
__device__ __forceinline__ bool leaf_test(const float4* __restrict__ tri, uint32_t ref, RayState& r, Counts& c)
{
    int first = (int)((ref >> 4) & 0x7FFFFFFu);
    int cnt = (int)(ref & 15u);
    bool got = false;
    for (int k = 0; k < cnt; ++k) {
        if (COUNT) c.tris++;
        got |= tri_test<ANY>(tri, first + k, r);
        if (ANY && got) return true;
    }
    return got;
}

// The stack's two parts in explicit address spaces: through plain (generic)
// pointers the compiler merges a pop's LDS and global branches into one
// FLAT load, which goes through the vector-memory pipeline the node fetches
// already saturate (and waits on both counters).
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) uint16_t lds_u16;
typedef __attribute__((address_space(1))) uint32_t gbl_u32;
typedef __attribute__((address_space(1))) uint16_t gbl_u16;

// Per-lane traversal stack: LDS for the top LST entries, global beyond.
template <int LST, int BLOCK>
struct Stack {
    lds_u32* lref;          // LDS, the block's arrays: entry e of column c at e * BLOCK + c
    lds_u16* lt;
    uint32_t col;           // this lane's column (one register for both arrays)
    gbl_u32* gref;          // global overflow area: entry k of grid thread g at
    gbl_u16* gt;            //   k * gstride + g, g = gbase + column()
    uint32_t gbase;         // the block's first grid thread
    uint32_t gstride;       // threads in the grid
    int gcap;               // overflow entries available
    uint32_t* full;         // set to 1 by a traversal that would overflow the stack (it ends there)
    int sp;

    // the column, opaque to the compiler at each use: otherwise it keeps
    // one precomputed address per array live through the whole loop, and
    // at the kernel's register bound one of them spills
    __device__ __forceinline__ uint32_t column() const
    {
        uint32_t c = col;
        __asm__ volatile("" : "+v"(c));
        return c;
    }

    __device__ __forceinline__ bool push(uint32_t ref, float t)
    {
        uint16_t tb = (uint16_t)(__float_as_uint(t) >> 16);   // truncation: a lower bound of t >= 0
        if (sp < LST) {
            const uint32_t i = sp * BLOCK + column();
            lref[i] = ref;
            lt[i] = tb;
        } else {
            int k = sp - LST;
            if (k >= gcap) return false;
            const size_t i = (size_t)k * gstride + gbase + column();
            gref[i] = ref;
            gt[i] = tb;
        }
        ++sp;
        return true;
    }
    __device__ __forceinline__ uint32_t pop(float& t)
    {
        --sp;
        uint32_t ref;
        uint16_t tb;
        if (sp < LST) {
            const uint32_t i = sp * BLOCK + column();
            ref = lref[i];
            tb = lt[i];
        } else {
            const size_t i = (size_t)(sp - LST) * gstride + gbase + column();
            ref = gref[i];
            tb = gt[i];
        }
        t = __uint_as_float((uint32_t)tb << 16);
        return ref;
    }
};

// Traverse one ray; returns true if a hit was found.
template <int W, bool ANY, int LST, int BLOCK, bool COUNT, bool EXACT>
__device__ bool traverse(const DevScene& S, RayState& r, Stack<LST, BLOCK>& st, Counts& c)
{
    const float4* __restrict__ nodes = S.nodes;
    const float4* __restrict__ tri = S.tri_isect;
    bool got = false;
    if (S.brute) {
        for (int s = 0; s < S.n_slots; ++s) {
            if (COUNT) c.tris++;
            got |= tri_test<ANY>(tri, s, r);
            if (ANY && got) return true;
        }
        return got;
    }
    // oversized triangles hoisted out of the tree, each behind its own box
    for (int s = 0; s < S.n_hoist; ++s) {
        float4 lo = gload(S.hoist_box + 2 * s), hi = gload(S.hoist_box + 2 * s + 1);
        float e;
        if (!box_hit(lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, e) || e > r.cull)
            continue;
        if (COUNT) c.tris++;
        got |= tri_test<ANY>(tri, s, r);
        if (ANY && got) return true;
    }
    uint32_t node = S.root_ref;
    if (node == kEmptyRef) return got;
    if (node & kLeafBit) {
        float e;
        if (!box_hit(S.root_box[0], S.root_box[1], S.root_box[2], S.root_box[3], S.root_box[4], S.root_box[5],
                     r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, e))
            return got;
        return leaf_test<ANY, COUNT>(tri, node, r, c) || got;
    }
    for (;;) {
        if (!(node & kLeafBit)) {
            if (COUNT) c.nodes++;
            if (W == 2) {
                const float4* n = nodes + 4 * (size_t)node;
                float4 X = gload(n), Y = gload(n + 1), Z = gload(n + 2);
                int4 C = gload(reinterpret_cast<const int4*>(n + 3));
                float e0, e1;
                bool h0 = child_hit<EXACT>(X.x, Y.x, Z.x, X.z, Y.z, Z.z, r, e0);
                bool h1 = child_hit<EXACT>(X.y, Y.y, Z.y, X.w, Y.w, Z.w, r, e1);
                h0 = h0 && e0 <= r.cull && (uint32_t)C.x != kEmptyRef;
                h1 = h1 && e1 <= r.cull && (uint32_t)C.y != kEmptyRef;
                if (h0 && h1) {
                    uint32_t nearc = (uint32_t)C.x, farc = (uint32_t)C.y;
                    float fe = e1;
                    if (e1 < e0) { nearc = (uint32_t)C.y; farc = (uint32_t)C.x; fe = e0; }
                    if (!st.push(farc, fe)) { atomicOr(st.full, 1u); return got; }   // cannot happen: depth bound
                    node = nearc;
                    continue;
                } else if (h0) { node = (uint32_t)C.x; continue; }
                else if (h1) { node = (uint32_t)C.y; continue; }
            } else {
                const float4* n = nodes + 8 * (size_t)node;
                float4 LX = gload(n), HX = gload(n + 1), LY = gload(n + 2), HY = gload(n + 3), LZ = gload(n + 4), HZ = gload(n + 5);
                int4 C = gload(reinterpret_cast<const int4*>(n + 6));
                float k0, k1, k2, k3;
                bool h0 = child_hit<EXACT>(LX.x, LY.x, LZ.x, HX.x, HY.x, HZ.x, r, k0);
                bool h1 = child_hit<EXACT>(LX.y, LY.y, LZ.y, HX.y, HY.y, HZ.y, r, k1);
                bool h2 = child_hit<EXACT>(LX.z, LY.z, LZ.z, HX.z, HY.z, HZ.z, r, k2);
                bool h3 = child_hit<EXACT>(LX.w, LY.w, LZ.w, HX.w, HY.w, HZ.w, r, k3);
                uint32_t c0 = (uint32_t)C.x, c1 = (uint32_t)C.y, c2 = (uint32_t)C.z, c3 = (uint32_t)C.w;
                // missed, culled or absent children sort last (key +inf)
                k0 = (h0 && k0 <= r.cull && c0 != kEmptyRef) ? k0 : INFINITY;
                k1 = (h1 && k1 <= r.cull && c1 != kEmptyRef) ? k1 : INFINITY;
                k2 = (h2 && k2 <= r.cull && c2 != kEmptyRef) ? k2 : INFINITY;
                k3 = (h3 && k3 <= r.cull && c3 != kEmptyRef) ? k3 : INFINITY;
                // 4-input sorting network, ascending entry distance
#define YN_CS(ka, ca, kb, cb)                                              \
    {                                                                      \
        bool sw = kb < ka;                                                 \
        float tk = sw ? kb : ka; kb = sw ? ka : kb; ka = tk;               \
        uint32_t tc = sw ? cb : ca; cb = sw ? ca : cb; ca = tc;            \
    }
                YN_CS(k0, c0, k1, c1);
                YN_CS(k2, c2, k3, c3);
                YN_CS(k0, c0, k2, c2);
                YN_CS(k1, c1, k3, c3);
                YN_CS(k1, c1, k2, c2);
#undef YN_CS
                if (k0 != INFINITY) {
                    bool ok = true;
                    if (k3 != INFINITY) ok &= st.push(c3, k3);
                    if (k2 != INFINITY) ok &= st.push(c2, k2);
                    if (k1 != INFINITY) ok &= st.push(c1, k1);
                    if (!ok) { atomicOr(st.full, 1u); return got; }   // cannot happen: depth bound
                    node = c0;
                    continue;
                }
            }
        } else {
            got |= leaf_test<ANY, COUNT>(tri, node, r, c);
            if (ANY && got) return true;
        }
        // pop; a deferred child whose entry now lies beyond the best hit is dropped
        for (;;) {
            if (st.sp == 0) return got;
            float et;
            uint32_t ref = st.pop(et);
            if (et > r.cull) continue;
            node = ref;
            break;
        }
    }
}


// One step of a ray's traversal: an interior node, descending to its
// nearest child; when that child is a leaf (or the step began at one), the
// leaf's triangles are tested in the same step and the next node is popped.
// The nodes and triangles are visited in traverse()'s order, so results and
// counts are the same; only the steps are fewer (a wave that runs interior
// and leaf lanes side by side no longer spends a step on each).  Returns true
// while the ray has nodes left; `node`, the stack and the best hit carry over
// between steps.
template <int W, bool ANY, int LST, int BLOCK, bool COUNT, bool EXACT>
__device__ __forceinline__ bool trav_step(const float4* __restrict__ nodes, const float4* __restrict__ tri,
                                          uint32_t& node, RayState& r, Stack<LST, BLOCK>& st, bool& got, Counts& c)
{
    if (!(node & kLeafBit)) {
        if (COUNT) c.nodes++;
        if (W == 2) {
            const float4* n = nodes + 4 * (size_t)node;
            float4 X = gload(n), Y = gload(n + 1), Z = gload(n + 2);
            int4 C = gload(reinterpret_cast<const int4*>(n + 3));
            float e0, e1;
            bool h0 = child_hit<EXACT>(X.x, Y.x, Z.x, X.z, Y.z, Z.z, r, e0);
            bool h1 = child_hit<EXACT>(X.y, Y.y, Z.y, X.w, Y.w, Z.w, r, e1);
            h0 = h0 && e0 <= r.cull && (uint32_t)C.x != kEmptyRef;
            h1 = h1 && e1 <= r.cull && (uint32_t)C.y != kEmptyRef;
            if (h0 && h1) {
                uint32_t nearc = (uint32_t)C.x, farc = (uint32_t)C.y;
                float fe = e1;
                if (e1 < e0) { nearc = (uint32_t)C.y; farc = (uint32_t)C.x; fe = e0; }
                if (!st.push(farc, fe)) { atomicOr(st.full, 1u); return false; }   // cannot happen: depth bound
                node = nearc;
                if (!(node & kLeafBit)) return true;
            } else if (h0) { node = (uint32_t)C.x; if (!(node & kLeafBit)) return true; }
            else if (h1) { node = (uint32_t)C.y; if (!(node & kLeafBit)) return true; }
            else goto pop;
        } else {
            const float4* n = nodes + 8 * (size_t)node;
            float4 LX = gload(n), HX = gload(n + 1), LY = gload(n + 2), HY = gload(n + 3), LZ = gload(n + 4), HZ = gload(n + 5);
            int4 C = gload(reinterpret_cast<const int4*>(n + 6));
            float k0, k1, k2, k3;
            bool h0 = child_hit<EXACT>(LX.x, LY.x, LZ.x, HX.x, HY.x, HZ.x, r, k0);
            bool h1 = child_hit<EXACT>(LX.y, LY.y, LZ.y, HX.y, HY.y, HZ.y, r, k1);
            bool h2 = child_hit<EXACT>(LX.z, LY.z, LZ.z, HX.z, HY.z, HZ.z, r, k2);
            bool h3 = child_hit<EXACT>(LX.w, LY.w, LZ.w, HX.w, HY.w, HZ.w, r, k3);
            uint32_t c0 = (uint32_t)C.x, c1 = (uint32_t)C.y, c2 = (uint32_t)C.z, c3 = (uint32_t)C.w;
            // missed, culled or absent children sort last (key +inf)
            k0 = (h0 && k0 <= r.cull && c0 != kEmptyRef) ? k0 : INFINITY;
            k1 = (h1 && k1 <= r.cull && c1 != kEmptyRef) ? k1 : INFINITY;
            k2 = (h2 && k2 <= r.cull && c2 != kEmptyRef) ? k2 : INFINITY;
            k3 = (h3 && k3 <= r.cull && c3 != kEmptyRef) ? k3 : INFINITY;
            // 4-input sorting network, ascending entry distance
#define YN_CS(ka, ca, kb, cb)                                              \
{                                                                      \
    bool sw = kb < ka;                                                 \
    float tk = sw ? kb : ka; kb = sw ? ka : kb; ka = tk;               \
    uint32_t tc = sw ? cb : ca; cb = sw ? ca : cb; ca = tc;            \
}
            YN_CS(k0, c0, k1, c1);
            YN_CS(k2, c2, k3, c3);
            YN_CS(k0, c0, k2, c2);
            YN_CS(k1, c1, k3, c3);
            YN_CS(k1, c1, k2, c2);
#undef YN_CS
            if (k0 != INFINITY) {
                bool ok = true;
                if (k3 != INFINITY) ok &= st.push(c3, k3);
                if (k2 != INFINITY) ok &= st.push(c2, k2);
                if (k1 != INFINITY) ok &= st.push(c1, k1);
                if (!ok) { atomicOr(st.full, 1u); return false; }   // cannot happen: depth bound
                node = c0;
                if (!(node & kLeafBit)) return true;
            } else goto pop;
        }
    }
    // a leaf: this iteration's node, or the nearest child just reached
    got |= leaf_test<ANY, COUNT>(tri, node, r, c);
    if (ANY && got) return false;
pop:
    // pop; a deferred child whose entry now lies beyond the best hit is dropped
    for (;;) {
        if (st.sp == 0) return false;
        float et;
        uint32_t ref = st.pop(et);
        if (et > r.cull) continue;
        node = ref;
        return true;
    }
}

}  // namespace

// Occupancy the kernel is compiled for, per SIMD: the LDS stack (6 B per
// entry and lane, 160 KB per CU) allows 13 / 8 / 6 / 3 waves per SIMD for
// 8 / 12 / 16 / 32 entries, and registers are held to what 7 / 7 / 6 / 3
// waves leave (72 / 72 / 80 / 168).  More waves in flight hide more of the
// node-fetch latency; past 7 the register spills cost more than they hide.
template <int LST>
constexpr int trace_waves() { return LST <= 12 ? 7 : (LST == 16 ? 6 : 3); }

template <int W, bool ANY, int LST, int BLOCK, bool COUNT>
__global__ __launch_bounds__(BLOCK, trace_waves<LST>()) void k_trace(TraceArgs a)
{
    __shared__ uint32_t stk_all[LST * BLOCK];
    __shared__ uint16_t stkt_all[LST * BLOCK];
    Stack<LST, BLOCK> st;
    st.lref = (lds_u32*)stk_all;
    st.lt = (lds_u16*)stkt_all;
    st.col = threadIdx.x;
    st.gstride = gridDim.x * BLOCK;
    st.gref = (gbl_u32*)a.ovf_ref;
    st.gt = (gbl_u16*)a.ovf_t;
    st.gbase = blockIdx.x * BLOCK;
    st.gcap = a.ovf_depth;
    st.full = &a.ctr->q[7][0];
    const int lane = (int)__lane_id();
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.zero0) { *a.zero0 = 0; if (a.zero1) *a.zero1 = 0; }
    const uint32_t n = *a.q.count;
    Counts c{0, 0};
    unsigned nrays = 0;
    // The queue is cut into one slice of chunks per XCD: workgroups go to
    // the XCDs round-robin, so block b runs on XCD b % n_slices and takes
    // its chunks from slice b % n_slices, where neighbouring rays (neighbouring
    // pixels) share that XCD's L2.  Each wave first takes its static chunk of
    // its slice without an atomic, then chunks past the slice's static ones
    // from the slice's counter; when the slice is used up it moves on to the
    // next slices in turn (never back), so every chunk is taken exactly once
    // and every wave ends after at most n_slices empty requests.  When the
    // static chunks cover the whole queue, a wave leaves without touching a
    // counter (same-address atomics from every wave would otherwise set a
    // floor on the launch time of short queues).
    const uint32_t ns = (uint32_t)a.n_slices;
    const uint32_t G = gridDim.x, wpb = BLOCK >> 6;
    const uint32_t xs = blockIdx.x % ns;
    const uint32_t wx = (blockIdx.x / ns) * wpb + (threadIdx.x >> 6);     // wave within its slice's XCD
    const uint32_t chunk = (uint32_t)a.chunk;
    const uint32_t n_chunks = (n + chunk - 1) / chunk;
    auto waves_of = [&](uint32_t y) { return ((G - y + ns - 1) / ns) * wpb; };
    auto slice_lo = [&](uint32_t y) { return (uint32_t)(((uint64_t)y * n_chunks) / ns); };
    const bool all_static = n_chunks <= ns * ((G / ns) * wpb);
    uint32_t y = xs, moved = 0;
    for (bool first = true;; first = false) {
        uint32_t ci;
        for (;;) {
            const uint32_t c0 = slice_lo(y), c1 = slice_lo(y + 1);
            if (first) {
                ci = c0 + wx;
            } else if (all_static) {
                ci = n_chunks;          // nothing past the static chunks
                break;
            } else {
                uint32_t k = 0;
                if (lane == 0) k = atomicAdd(a.work + y * kWorkStride, 1u);
                ci = c0 + waves_of(y) + __shfl(k, 0);
            }
            if (ci < c1) break;
            if (all_static) { ci = n_chunks; break; }
            first = false;
            if (++moved >= ns) { ci = n_chunks; break; }
            y = (y + 1) % ns;
        }
        if (ci >= n_chunks) break;
        const uint32_t base = ci * chunk;
        for (uint32_t off = (uint32_t)lane; off < (uint32_t)a.chunk; off += 64) {
            uint32_t i = base + off;
            if (i >= n) break;
            float4 ot = a.q.o_t[i];
            float4 dp = a.q.d_p[i];
            int pidx = __float_as_int(dp.w);
            int sc = a.single_scene >= 0 ? a.single_scene
                                         : a.jobs[a.by_queue ? pidx : a.path_job[pidx >> a.id_shift]].scene;
            if (a.by_queue) pidx = (int)i;      // the result goes beside the ray
            const DevScene& S = a.scenes[sc];
            RayState r;
            const bool fma_ok = start_ray(r, ot, dp, S);
            st.sp = 0;
            // A ray with a NaN in its origin, direction or length hits nothing:
            // u, v or t of every triangle test is NaN, and each acceptance
            // comparison is false.  Its traversal would still visit every
            // node (the NaN guard keeps each slab), one lane for the whole
            // tree, so it is answered without one.
            bool nan_ray = !(ot.x == ot.x && ot.y == ot.y && ot.z == ot.z && ot.w == ot.w && dp.x == dp.x &&
                             dp.y == dp.y && dp.z == dp.z);
            bool got = false;
            if (!nan_ray)
                got = __ballot(!fma_ok) ? traverse<W, ANY, LST, BLOCK, COUNT, true>(S, r, st, c)
                                        : traverse<W, ANY, LST, BLOCK, COUNT, false>(S, r, st, c);
            ++nrays;
            if (ANY) {
                a.occ_out[pidx] = got ? 1 : 0;
                if (a.hit_out) a.hit_out[pidx] = hit_record<ANY>(S.tri_isect, r);
            } else {
                a.hit_out[pidx] = hit_record<ANY>(S.tri_isect, r);
            }
        }
    }
    // one atomic per wave for the statistics
    unsigned long long tot = wave_sum_u64(nrays);
    if (lane == 0 && tot) atomicAdd(ANY ? &a.ctr->rays_shadow : &a.ctr->rays_ext, tot);
    if (COUNT) {
        unsigned long long tn = wave_sum_u64(c.nodes), tt = wave_sum_u64(c.tris);
        if (lane == 0) {
            atomicAdd(ANY ? &a.ctr->nodes_shadow : &a.ctr->nodes_ext, tn);
            atomicAdd(ANY ? &a.ctr->tris_shadow : &a.ctr->tris_ext, tt);
        }
    }
}

// Result of one ray into the path's slot.
template <bool ANY>
__device__ __forceinline__ void store_hit(const TraceArgs& a, int pidx, const float4* __restrict__ tri,
                                          const RayState& r, bool got)
{
    if (ANY) {
        a.occ_out[pidx] = got ? 1 : 0;
        if (a.hit_out) a.hit_out[pidx] = hit_record<ANY>(tri, r);
    } else {
        a.hit_out[pidx] = hit_record<ANY>(tri, r);
    }
}

// Traversal with lanes refilled.  k_trace runs each lane's ray to its end
// while the wave waits for its slowest ray, so about a quarter of the vector
// lanes do work.  Here the loop takes one traversal step per lane per
// iteration (traverse()'s loop body), and once `a.refill` lanes of the wave
// have finished, those lanes take new rays from the queue, one atomic for
// all of them.  Each ray's own steps, tests and results are those of
// k_trace.  The first fetch of every wave is static (rays wave*64 ..),
// later ones come after the static range from the work counter.
// SS: every ray renders the one scene a.single_scene, whose node and
// triangle arrays are then wave-uniform (scalar registers; per-lane
// pointers would take four vector registers from a kernel held to 72).
template <int W, bool ANY, int LST, int BLOCK, bool COUNT, bool SS>
__global__ __launch_bounds__(BLOCK, trace_waves<LST>()) void k_trace_rf(TraceArgs a)
{
    __shared__ uint32_t stk_all[LST * BLOCK];
    __shared__ uint16_t stkt_all[LST * BLOCK];
    Stack<LST, BLOCK> st;
    st.lref = (lds_u32*)stk_all;
    st.lt = (lds_u16*)stkt_all;
    st.col = threadIdx.x;
    st.gstride = gridDim.x * BLOCK;
    st.gref = (gbl_u32*)a.ovf_ref;
    st.gt = (gbl_u16*)a.ovf_t;
    st.gbase = blockIdx.x * BLOCK;
    st.gcap = a.ovf_depth;
    st.full = &a.ctr->q[7][0];
    st.sp = 0;
    const int lane = (int)__lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.zero0) { *a.zero0 = 0; if (a.zero1) *a.zero1 = 0; }
    const uint32_t n = *a.q.count;
    // a wave with every lane idle always fetches, so the queue drains
    const int refill = a.refill < 1 ? 1 : (a.refill > 64 ? 64 : a.refill);
    // Queue slices as in k_trace, in rays: the wave's first fetch is its
    // static 64 rays of its XCD's slice, later ones as many rays as idle
    // lanes from the slice's counter, past the slice's static rays; a used-up
    // slice sends the wave on to the next ones (never back).
    const uint32_t ns = (uint32_t)a.n_slices;
    const uint32_t G = gridDim.x, wpb = BLOCK >> 6;
    const uint32_t wx = (blockIdx.x / ns) * wpb + (threadIdx.x >> 6);
    auto waves_of = [&](uint32_t y) { return ((G - y + ns - 1) / ns) * wpb; };
    auto slice_lo = [&](uint32_t y) { return (uint32_t)(((uint64_t)y * n) / ns); };
    const bool all_static = (uint64_t)n <= (uint64_t)ns * ((G / ns) * wpb) * 64u;
    uint32_t y = blockIdx.x % ns, moved = 0;
    // the current slice's bounds and wave count, recomputed only when the
    // wave moves on (slice_lo divides 64-bit)
    auto sreg = [](uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); };   // wave-uniform
    uint32_t y_lo = sreg(slice_lo(y)), y_hi = sreg(slice_lo(y + 1)), y_waves = sreg(waves_of(y));
    Counts c{0, 0};
    unsigned nrays = 0;
    RayState r{};
    uint32_t node = 0;
    int pidx = 0;
    bool live = false, got = false, exact = false;
    const int ss = SS ? __builtin_amdgcn_readfirstlane(a.single_scene) : 0;
    const float4* nodes = SS ? a.scenes[ss].nodes : nullptr;
    const float4* tri = SS ? a.scenes[ss].tri_isect : nullptr;
    bool more = true, first = true;
    for (;;) {
        const unsigned long long idle = __ballot(!live);
        const int n_idle = __popcll(idle);
        if (more && n_idle >= refill) {
            uint32_t base = 0, lim = 0;
            for (;;) {
                if (first) {
                    base = y_lo + wx * 64u;          // every lane is idle on the first fetch
                    first = false;
                } else {
                    uint32_t b = 0;
                    if (lane == 0) b = atomicAdd(a.work + y * kWorkStride, (uint32_t)n_idle);
                    base = y_lo + y_waves * 64u + __shfl(b, 0);
                }
                if (base < y_hi) { lim = y_hi; break; }
                if (all_static || ++moved >= ns) { more = false; break; }
                y = (y + 1) % ns;
                y_lo = sreg(slice_lo(y));
                y_hi = sreg(slice_lo(y + 1));
                y_waves = sreg(waves_of(y));
            }
            if (all_static) more = false;            // nothing past the static rays
            const uint32_t i = base + (uint32_t)__popcll(idle & below);
            if (!live && i < lim) {
                float4 ot = a.q.o_t[i];
                float4 dp = a.q.d_p[i];
                pidx = __float_as_int(dp.w);
                int sc = SS ? ss
                            : a.single_scene >= 0 ? a.single_scene
                                                  : a.jobs[a.by_queue ? pidx : a.path_job[pidx >> a.id_shift]].scene;
                if (a.by_queue) pidx = (int)i;      // the result goes beside the ray
                const DevScene& S = a.scenes[sc];
                exact = !start_ray(r, ot, dp, S);
                st.sp = 0;
                got = false;
                ++nrays;
                // a ray with a NaN component hits nothing (see k_trace)
                live = ot.x == ot.x && ot.y == ot.y && ot.z == ot.z && ot.w == ot.w && dp.x == dp.x &&
                       dp.y == dp.y && dp.z == dp.z;
                if (!SS) {
                    nodes = S.nodes;
                    tri = S.tri_isect;
                }
                // the tests ahead of the tree, as in traverse(): every slot
                // without a BVH; the hoisted slots; a root that is one leaf
                if (live && S.brute) {
                    for (int s = 0; s < S.n_slots; ++s) {
                        if (COUNT) c.tris++;
                        got |= tri_test<ANY>(tri, s, r);
                        if (ANY && got) break;
                    }
                    live = false;
                }
                if (live) {
                    for (int s = 0; s < S.n_hoist; ++s) {
                        float4 lo = gload(S.hoist_box + 2 * s), hi = gload(S.hoist_box + 2 * s + 1);
                        float e;
                        if (!box_hit(lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, e) ||
                            e > r.cull)
                            continue;
                        if (COUNT) c.tris++;
                        got |= tri_test<ANY>(tri, s, r);
                        if (ANY && got) break;
                    }
                    node = S.root_ref;
                    if ((ANY && got) || node == kEmptyRef) {
                        live = false;
                    } else if (node & kLeafBit) {
                        float e;
                        if (box_hit(S.root_box[0], S.root_box[1], S.root_box[2], S.root_box[3], S.root_box[4],
                                    S.root_box[5], r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, e))
                            got |= leaf_test<ANY, COUNT>(tri, node, r, c);
                        live = false;
                    }
                }
                if (!live) store_hit<ANY>(a, pidx, tri, r, got);
            }
        }
        if (__ballot(live) == 0) {
            if (!more) break;
            continue;
        }
        const bool wexact = __ballot(live && exact) != 0;
        if (live) {
            live = wexact ? trav_step<W, ANY, LST, BLOCK, COUNT, true>(nodes, tri, node, r, st, got, c)
                          : trav_step<W, ANY, LST, BLOCK, COUNT, false>(nodes, tri, node, r, st, got, c);
            if (!live) store_hit<ANY>(a, pidx, tri, r, got);
        }
    }
    unsigned long long tot = wave_sum_u64(nrays);
    if (lane == 0 && tot) atomicAdd(ANY ? &a.ctr->rays_shadow : &a.ctr->rays_ext, tot);
    if (COUNT) {
        unsigned long long tn = wave_sum_u64(c.nodes), tt = wave_sum_u64(c.tris);
        if (lane == 0) {
            atomicAdd(ANY ? &a.ctr->nodes_shadow : &a.ctr->nodes_ext, tn);
            atomicAdd(ANY ? &a.ctr->tris_shadow : &a.ctr->tris_ext, tt);
        }
    }
}

namespace {
template <int W, bool ANY, int LST, int BL>
void launch_b(const TraceArgs& a, int grid, bool count, hipStream_t s)
{
    if (a.refill > 0) {
        bool done = false;
        if constexpr (W == 4) {
            if (a.single_scene >= 0) {
                if (count) hipLaunchKernelGGL((k_trace_rf<W, ANY, LST, BL, true, true>), dim3(grid), dim3(BL), 0, s, a);
                else hipLaunchKernelGGL((k_trace_rf<W, ANY, LST, BL, false, true>), dim3(grid), dim3(BL), 0, s, a);
                done = true;
            }
        }
        if (!done) {
            if (count) hipLaunchKernelGGL((k_trace_rf<W, ANY, LST, BL, true, false>), dim3(grid), dim3(BL), 0, s, a);
            else hipLaunchKernelGGL((k_trace_rf<W, ANY, LST, BL, false, false>), dim3(grid), dim3(BL), 0, s, a);
        }
    } else {
        if (count) hipLaunchKernelGGL((k_trace<W, ANY, LST, BL, true>), dim3(grid), dim3(BL), 0, s, a);
        else hipLaunchKernelGGL((k_trace<W, ANY, LST, BL, false>), dim3(grid), dim3(BL), 0, s, a);
    }
}

template <bool ANY, int LST>
void launch_w(const TraceArgs& a, int width, int block, int grid, bool count, hipStream_t s)
{
    if (width == 2) {
        if (block == 64) launch_b<2, ANY, LST, 64>(a, grid, count, s);
        else if (block == 128) launch_b<2, ANY, LST, 128>(a, grid, count, s);
        else launch_b<2, ANY, LST, 256>(a, grid, count, s);
    } else {
        if (block == 64) launch_b<4, ANY, LST, 64>(a, grid, count, s);
        else if (block == 128) launch_b<4, ANY, LST, 128>(a, grid, count, s);
        else launch_b<4, ANY, LST, 256>(a, grid, count, s);
    }
}

template <bool ANY>
void launch_l(const TraceArgs& a, int width, int block, int grid, bool count, int lds, hipStream_t s)
{
    if (lds == 8) launch_w<ANY, 8>(a, width, block, grid, count, s);
    else if (lds == 12) launch_w<ANY, 12>(a, width, block, grid, count, s);
    else if (lds == 32) launch_w<ANY, 32>(a, width, block, grid, count, s);
    else launch_w<ANY, 16>(a, width, block, grid, count, s);
}
}  // namespace

// LDS part of the traversal stack: 8, 12, 16 or 32 entries per lane.
int trace_lds_stack(int lds) { return lds == 8 || lds == 12 || lds == 32 ? lds : 16; }

// Waves per SIMD a trace launch runs at for a configured LDS stack.
int trace_resident_waves(int lds)
{
    lds = trace_lds_stack(lds);
    return lds == 8 ? trace_waves<8>() : lds == 12 ? trace_waves<12>() : (lds == 32 ? trace_waves<32>() : trace_waves<16>());
}

// `stack`: traversal stack entries the scene can need; the launch is refused
// when LDS plus the overflow area cannot hold them.
int launch_trace(const TraceArgs& a, bool any_hit, int width, int stack, int block, int grid, bool count, int lds,
                 hipStream_t stream)
{
    lds = trace_lds_stack(lds);
    if (stack > lds + a.ovf_depth) return -1;
    if (any_hit) launch_l<true>(a, width, block, grid, count, lds, stream);
    else launch_l<false>(a, width, block, grid, count, lds, stream);
    return 0;
}

}  // namespace yune
