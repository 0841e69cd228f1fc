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
//    triangle the reference would find).  Boxes, slab test and
//    Moller-Trumbore use the reference's operands and order, so t, u, v and
//    the accept/reject decisions are bit-identical to the CPU restatement.
//    Only equal-t ties may resolve to a different triangle (test order).
//  * any hit (shadow rays, udpt.cl:307-308): true if some triangle has t in
//    (0, tmax).
//
// Layout: per node the children's boxes are stored per axis (64 B for two
// children, 128 B for four), child references carry leaf ranges inline.
// Each lane owns one ray; its traversal stack lives in LDS, interleaved by
// lane so the 64 lanes of a wave hit 64 different banks.  Waves pull chunks
// of 64 rays from the queue with one atomic per chunk.
#include "yn_device.h"
#include "yn_internal_dev.h"

namespace yune {

namespace {

__device__ __forceinline__ void slab(float lo, float hi, float o, float inv, float& tmin, float& tmax)
{
    float a = (lo - o) * inv;           // min_diff (udpt.cl:397)
    float b = (hi - o) * inv;           // max_diff (udpt.cl:398)
    if (!(a != a)) {                    // !isnan(min_diff)
        tmin = fmaxf(fminf(a, b), tmin);
        tmax = fminf(fmaxf(a, b), tmax);
    }
}

// rayAabbIntersection (udpt.cl:392-431) + entry distance for culling.
__device__ __forceinline__ bool box_hit(float lox, float loy, float loz, float hix, float hiy, float hiz,
                                        float ox, float oy, float oz, float ix, float iy, float iz,
                                        float& enter)
{
    float tmin = -INFINITY, tmax = INFINITY;
    slab(lox, hix, ox, ix, tmin, tmax);
    slab(loy, hiy, oy, iy, tmin, tmax);
    slab(loz, hiz, oz, iz, tmin, tmax);
    enter = fmaxf(tmin, 0.0f);
    return tmax > enter;
}

struct RayState {
    float ox, oy, oz, dx, dy, dz, ix, iy, iz;
    float best;             // current tmax
    float cull;             // best * (1 + 1e-4)
    int slot;               // slot of the best hit, -1 none
    int tri;
    float u, v;
    int overflow;           // traversal stack would have overflowed
};

// rayTriangleIntersection (udpt.cl:326-390) on a stored slot; xyz only
// (the w terms of the reference's dots are exact zeros).
template <bool ANY>
__device__ __forceinline__ bool tri_test(const float4* __restrict__ tri, int s, RayState& r)
{
    float4 A = tri[3 * s], B = tri[3 * s + 1], C = tri[3 * s + 2];
    // pvec = cross(dir, v1v3)
    float px = r.dy * C.z - r.dz * C.y;
    float py = r.dz * C.x - r.dx * C.z;
    float pz = r.dx * C.y - r.dy * C.x;
    float det = B.x * px + B.y * py + B.z * pz;
    float inv_det = 1.0f / det;
    float sx = r.ox - A.x, sy = r.oy - A.y, sz = r.oz - A.z;
    float u = (px * sx + py * sy + pz * sz) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    // qvec = cross(dist, v1v2)
    float qx = sy * B.z - sz * B.y;
    float qy = sz * B.x - sx * B.z;
    float qz = sx * B.y - sy * B.x;
    float v = (qx * r.dx + qy * r.dy + qz * r.dz) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = (C.x * qx + C.y * qy + C.z * qz) * inv_det;
    if (t > 0.0f && t < r.best) {
        r.best = t;
        r.cull = t * 1.0001f;
        r.slot = s;
        r.tri = __float_as_int(A.w);
        r.u = u;
        r.v = v;
        return true;
    }
    return false;
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

// Traverse one ray; returns true if a hit was found.
template <int W, bool ANY, int STACK, int BLOCK, bool COUNT>
__device__ bool traverse(const DevScene& S, RayState& r, uint32_t* stk, Counts& c)
{
    const float4* __restrict__ nodes = S.nodes;
    const float4* __restrict__ tri = S.tri_isect;
    if (S.brute) {
        bool got = false;
        for (int s = 0; s < S.n_slots; ++s) {
            if (COUNT) c.tris++;
            got |= tri_test<ANY>(tri, s, r);
            if (ANY && got) return true;
        }
        return got;
    }
    uint32_t node = S.root_ref;
    if (node == kEmptyRef) return false;
    if (node & kLeafBit) {
        float e;
        if (!box_hit(S.root_box[0], S.root_box[1], S.root_box[2], S.root_box[3], S.root_box[4], S.root_box[5],
                     r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, e))
            return false;
        return leaf_test<ANY, COUNT>(tri, node, r, c);
    }
    int sp = 0;
    bool got = false;
    for (;;) {
        if (!(node & kLeafBit)) {
            if (COUNT) c.nodes++;
            if (W == 2) {
                const float4* n = nodes + 4 * (size_t)node;
                float4 X = n[0], Y = n[1], Z = n[2];
                int4 C = *reinterpret_cast<const int4*>(n + 3);
                float e0, e1;
                bool h0 = box_hit(X.x, Y.x, Z.x, X.z, Y.z, Z.z, r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, e0);
                bool h1 = box_hit(X.y, Y.y, Z.y, X.w, Y.w, Z.w, r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, e1);
                h0 = h0 && e0 <= r.cull && (uint32_t)C.x != kEmptyRef;
                h1 = h1 && e1 <= r.cull && (uint32_t)C.y != kEmptyRef;
                if (h0 && h1) {
                    uint32_t nearc = (uint32_t)C.x, farc = (uint32_t)C.y;
                    if (e1 < e0) { nearc = (uint32_t)C.y; farc = (uint32_t)C.x; }
                    if (sp >= STACK) { r.overflow = 1; return got; }   // cannot happen: depth <= STACK
                    stk[sp * BLOCK] = farc;
                    ++sp;
                    node = nearc;
                    continue;
                } else if (h0) { node = (uint32_t)C.x; continue; }
                else if (h1) { node = (uint32_t)C.y; continue; }
            } else {
                const float4* n = nodes + 8 * (size_t)node;
                float4 LX = n[0], HX = n[1], LY = n[2], HY = n[3], LZ = n[4], HZ = n[5];
                int4 C = *reinterpret_cast<const int4*>(n + 6);
                float e[4];
                bool h[4];
                h[0] = box_hit(LX.x, LY.x, LZ.x, HX.x, HY.x, HZ.x, r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, e[0]);
                h[1] = box_hit(LX.y, LY.y, LZ.y, HX.y, HY.y, HZ.y, r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, e[1]);
                h[2] = box_hit(LX.z, LY.z, LZ.z, HX.z, HY.z, HZ.z, r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, e[2]);
                h[3] = box_hit(LX.w, LY.w, LZ.w, HX.w, HY.w, HZ.w, r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, e[3]);
                uint32_t cr[4] = {(uint32_t)C.x, (uint32_t)C.y, (uint32_t)C.z, (uint32_t)C.w};
                // order hit children by entry distance, far ones pushed first
                float key[4];
                uint32_t ch[4];
                int nh = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    bool ok = h[k] && e[k] <= r.cull && cr[k] != kEmptyRef;
                    if (ok) { key[nh] = e[k]; ch[nh] = cr[k]; ++nh; }
                }
                if (nh > 0) {
                    // insertion sort, descending entry distance
                    for (int a = 1; a < nh; ++a) {
                        float kk = key[a];
                        uint32_t cc = ch[a];
                        int b = a - 1;
                        while (b >= 0 && key[b] < kk) { key[b + 1] = key[b]; ch[b + 1] = ch[b]; --b; }
                        key[b + 1] = kk;
                        ch[b + 1] = cc;
                    }
                    if (sp + nh - 1 > STACK) { r.overflow = 1; return got; }   // cannot happen: depth <= STACK
                    for (int a = 0; a < nh - 1; ++a) { stk[sp * BLOCK] = ch[a]; ++sp; }
                    node = ch[nh - 1];
                    continue;
                }
            }
        } else {
            bool g = leaf_test<ANY, COUNT>(tri, node, r, c);
            got |= g;
            if (ANY && got) return true;
        }
        // pop
        for (;;) {
            if (sp == 0) return got;
            --sp;
            node = stk[sp * BLOCK];
            // a deferred child may now lie beyond the best hit: skip it
            break;
        }
    }
}

}  // namespace

template <int W, bool ANY, int STACK, int BLOCK, bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_trace(TraceArgs a)
{
    __shared__ uint32_t stk_all[STACK * BLOCK];
    uint32_t* stk = stk_all + threadIdx.x;
    const int lane = (int)__lane_id();
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.zero0) { *a.zero0 = 0; if (a.zero1) *a.zero1 = 0; }
    const uint32_t n = *a.q.count;
    Counts c{0, 0};
    unsigned nrays = 0;
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(a.work, (uint32_t)a.chunk);
        base = __shfl(base, 0);
        if (base >= n) break;
        for (uint32_t off = (uint32_t)lane; off < (uint32_t)a.chunk; off += 64) {
            uint32_t i = base + off;
            if (i >= n) break;
            float4 ot = a.q.o_t[i];
            float4 dp = a.q.d_p[i];
            int pidx = __float_as_int(dp.w);
            int sc = a.single_scene >= 0 ? a.single_scene : a.jobs[a.path_job[pidx]].scene;
            const DevScene& S = a.scenes[sc];
            RayState r;
            r.ox = ot.x; r.oy = ot.y; r.oz = ot.z;
            r.dx = dp.x; r.dy = dp.y; r.dz = dp.z;
            r.ix = 1 / r.dx; r.iy = 1 / r.dy; r.iz = 1 / r.dz;   // dir_inv (udpt.cl:395)
            r.best = ot.w;
            r.cull = ot.w * 1.0001f;
            r.slot = -1; r.tri = -1; r.u = 0.0f; r.v = 0.0f; r.overflow = 0;
            bool got = traverse<W, ANY, STACK, BLOCK, COUNT>(S, r, stk, c);
            ++nrays;
            if (r.overflow) atomicOr(&a.ctr->q[7], 1u);
            if (ANY) {
                a.occ_out[pidx] = got ? 1 : 0;
                if (a.hit_out) a.hit_out[pidx] = make_float4(r.best, __int_as_float(got ? r.tri : -1), r.u, r.v);
            } else {
                a.hit_out[pidx] = make_float4(r.best, __int_as_float(r.tri), r.u, r.v);
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

// Host-side launcher: picks the instantiation for width / stack depth.
#define YN_TRACE_CASE(W, ANY, ST, BL, CNT)                                                        \
    hipLaunchKernelGGL((k_trace<W, ANY, ST, BL, CNT>), dim3(grid), dim3(BL), 0, stream, a)

template <int W, bool ANY, int BL>
static int launch_stack(const TraceArgs& a, int stack, int grid, bool count, hipStream_t stream)
{
    if (stack <= 32) {
        if (count) YN_TRACE_CASE(W, ANY, 32, BL, true); else YN_TRACE_CASE(W, ANY, 32, BL, false);
    } else if (stack <= 64) {
        if (count) YN_TRACE_CASE(W, ANY, 64, BL, true); else YN_TRACE_CASE(W, ANY, 64, BL, false);
    } else if (stack <= 96) {
        if (count) YN_TRACE_CASE(W, ANY, 96, BL, true); else YN_TRACE_CASE(W, ANY, 96, BL, false);
    } else {
        return -1;
    }
    return 0;
}

template <bool ANY>
static int launch_w(const TraceArgs& a, int width, int stack, int block, int grid, bool count, hipStream_t s)
{
    if (width == 2) {
        if (block == 64) return launch_stack<2, ANY, 64>(a, stack, grid, count, s);
        if (block == 128) return launch_stack<2, ANY, 128>(a, stack, grid, count, s);
        return launch_stack<2, ANY, 256>(a, stack, grid, count, s);
    }
    if (block == 64) return launch_stack<4, ANY, 64>(a, stack, grid, count, s);
    if (block == 128) return launch_stack<4, ANY, 128>(a, stack, grid, count, s);
    return launch_stack<4, ANY, 256>(a, stack, grid, count, s);
}

int launch_trace(const TraceArgs& a, bool any_hit, int width, int stack, int block, int grid, bool count,
                 hipStream_t stream)
{
    return any_hit ? launch_w<true>(a, width, stack, block, grid, count, stream)
                   : launch_w<false>(a, width, stack, block, grid, count, stream);
}

}  // namespace yune
