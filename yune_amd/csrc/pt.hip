// pt.hip — generate, shade, resolve and tonemap kernels of the wavefront
// unidirectional path tracer (kernels/legacy/udpt.cl semantics).
//
// The reference runs one path per work-item from camera ray to the end
// (pathtracer -> shading -> evaluateDirectLighting, udpt.cl:158-609).  Here
// each path is a small state machine stored in HBM; a shade pass consumes
// the result of the path's one outstanding ray and runs until the path
// needs its next ray (or ends):
//
//   PRIMARY  --hit--> [NEE start] --shadow ray--> NEE_SH --> [lobe pick,
//   light sample, MIS direction] --MIS ray--> NEE_MIS --> [BOUNCE: sample
//   direction] --extension ray--> BOUNCE --hit--> [emission, throughput,
//   RR] --> NEE start at the new vertex ...
//
// The per-path xorshift state travels with the path and is consumed in the
// reference's order (including the lobe pick that only happens after an
// unoccluded shadow ray, udpt.cl:557-559), so primary and first shadow rays
// are bit-identical to the CPU restatement and later decisions agree up to
// libm-vs-device ulps in sin/cos/pow.
//
// Rays never recorded by the reference as separate entities:
//  * the quad light is tested here when a ray is made (udpt.cl:244-276);
//    a hit shortens the ray to the quad and marks it, so an extension ray
//    that finds no nearer triangle is a light hit;
//  * an NEE shadow ray the quad already blocks is occluded without a BVH
//    query (traceRay would return true either way);
//  * the MIS BRDF ray only matters if it reaches the light first
//    (udpt.cl:599), so it is an any-hit query up to the quad, skipped when
//    the quad is missed.
#include "yn_internal_dev.h"
#include "yn_shade.h"

namespace yune {

namespace {

enum : int {
    ST_DONE = 0,
    ST_PRIMARY = 1,
    ST_BOUNCE = 2,
    ST_NEE_SH = 3,
    ST_NEE_MIS = 4,
    ST_MASK = 0xff,
    F_DIRECT = 1 << 8,      // NEE result is the first vertex's direct light
    F_LCAND = 1 << 9,       // pending extension ray ends on the quad
    F_NGLOSSY = 1 << 10,    // NEE lobe: glossy
    F_BGLOSSY = 1 << 11,    // bounce lobe: glossy
    F_CAPSH = 1 << 12,      // first shadow ray already captured
    SCENE_SHIFT = 16,       // bits 16..23: the path's scene (its job's), read without the job
};

constexpr float kPinkR = 0.988f, kPinkG = 0.0588f, kPinkB = 0.7529f;   // udpt.cl:110

struct Out {
    int kind = 0;           // 0 none, 1 extension, 2 shadow
    bool done = false;      // the path ended and wrote its sample
    int job = -1;           // the path's job-table entry
    float4 o_t;
    float4 d_p;
};

// A path's state between two shade steps (the fields of PathRecs).
struct Rec {
    int4 m0, m1;
    f4 hp, hn, rd, tp, direct, indirect, nd, bnc, wi, ls, mis;
};

__device__ __forceinline__ void write_sample(const DevJob& J, int slot, int pix, f4 c)
{
    if (any_nan(c)) c = mk(kPinkR, kPinkG, kPinkB, 1.0f);     // udpt.cl:193-194
    J.samples[(size_t)slot * (size_t)(J.W * J.H) + (size_t)pix] = st4(c);
}

// The ray's w carries the path's job-table entry: the trace kernels look
// the scene up through it.
__device__ __forceinline__ Out make_ray(int kind, f4 o, f4 d, float tmax, int job)
{
    Out r;
    r.kind = kind;
    r.o_t = make_float4(o.x, o.y, o.z, tmax);
    r.d_p = make_float4(d.x, d.y, d.z, __int_as_float(job));
    return r;
}

// The record a step reads.  The fields most states use are loaded without
// waiting for the state (a new path's record holds only m0, m1 and the
// camera ray's direction; the rest is not used then); only the MIS ray's
// fields, used by one state, wait for it.  The shade pass is bound by the
// latency of such chains of loads.  Which queue the result came from is
// known before the state: a bounce result (extension queue) reads the
// pending bounce and never the NEE light direction, a shadow result the
// light direction and never the pending bounce (each is rewritten before
// the step stores it), so each loads only its own.
__device__ __forceinline__ void load_rec(const PathRecs& I, uint32_t ri, Rec& R, bool ext, bool primary)
{
    using F = PathRecs;
    R.m0 = *reinterpret_cast<const int4*>(I.f(F::M0, ri));
    R.m1 = *reinterpret_cast<const int4*>(I.f(F::M1, ri));
    R.rd = ld(*I.f(F::RD, ri));
    if (primary) {          // a camera ray's record: the rest is set by its step
        R.hp = R.hn = R.tp = R.direct = R.indirect = R.nd = R.bnc = R.wi = mk(0, 0, 0, 0);
        return;
    }
    R.hp = ld(*I.f(F::HP, ri));
    R.hn = ld(*I.f(F::HN, ri));
    R.tp = ld(*I.f(F::TP, ri));
    R.direct = ld(*I.f(F::DIRECT, ri));
    R.indirect = ld(*I.f(F::INDIRECT, ri));
    R.nd = R.bnc = R.wi = mk(0, 0, 0, 0);
    if (ext) {
        R.nd = ld(*I.f(F::ND, ri));
        R.bnc = ld(*I.f(F::BNC, ri));
    } else {
        R.wi = ld(*I.f(F::WI, ri));
    }
    if ((R.m0.y & ST_MASK) == ST_NEE_MIS) {
        R.ls = ld(*I.f(F::LS, ri));
        R.mis = ld(*I.f(F::MIS, ri));
    }
}

// The record beside a ray just made: what the next step's state uses.
__device__ __forceinline__ void store_rec(const PathRecs& O, uint32_t ri, const Rec& R)
{
    using F = PathRecs;
    *reinterpret_cast<int4*>(O.f(F::M0, ri)) = R.m0;
    *reinterpret_cast<int4*>(O.f(F::M1, ri)) = R.m1;
    *O.f(F::HP, ri) = st4(R.hp);
    *O.f(F::HN, ri) = st4(R.hn);
    *O.f(F::RD, ri) = st4(R.rd);
    *O.f(F::TP, ri) = st4(R.tp);
    *O.f(F::DIRECT, ri) = st4(R.direct);
    *O.f(F::INDIRECT, ri) = st4(R.indirect);
    const int code = R.m0.y & ST_MASK;
    if (code == ST_BOUNCE) {
        *O.f(F::ND, ri) = st4(R.nd);
        *O.f(F::BNC, ri) = st4(R.bnc);
    } else {
        *O.f(F::WI, ri) = st4(R.wi);
        if (code == ST_NEE_MIS) {
            *O.f(F::LS, ri) = st4(R.ls);
            *O.f(F::MIS, ri) = st4(R.mis);
        }
    }
}

// One shade step of the path whose ray sits at entry qi of the extension
// queue (ext) or of the shadow queue; R holds its record, updated in place
// for the ray the step makes.  The ray whose result arrived is rebuilt from
// the path state, bit for bit as it was made: a camera ray starts at the
// camera's eye (createRay, udpt.cl:236) along the stored direction rd; a
// bounce ray at hp + nd * EPSILON along the pending bounce direction nd
// (udpt.cl:766, 906, 945).
template <bool CAP, int LM>
__device__ Out advance(const ShadeArgs& a, const DevCapture& cap, bool ext, uint32_t qi, Rec& R, const LdsScene& lds)
{
    load_rec(a.in, ext ? qi : (uint32_t)(a.in.cap - 1 - (int64_t)qi), R, ext, ext && a.ext_primary);
    const int4 m0 = R.m0, m1 = R.m1;
    const int jid = m1.y;
    const DevJob& J = a.jobs[jid];
    const yn_render_params& RP = J.p;
    const int sid = (m0.y >> SCENE_SHIFT) & 0xff;
    const DevScene& S = a.scenes[sid];                                 // loaded beside J, not after it
    // LM: materials and light in LDS, the launch's one scene's (1) or every scene's (2)
    auto mat_of = [&](int id) {
        return LM == 1 ? load_mat(lds.mats, id) : LM == 2 ? load_mat(lds.mats, S.mat_base + id) : load_mat(S.mats, id);
    };
    auto& L = light_of<LM>(S.light, lds, sid);
    const bool on = RP.oren_nayar != 0;
    const bool fix = RP.fix_mis != 0;

    int st = m0.y;
    int code = st & ST_MASK;
    uint32_t seed = (uint32_t)m0.x;
    const int pix = m1.x;
    const int slot = m1.z;
    const bool cap_on = CAP && cap.prim_tri != nullptr && slot == 0;

    int vmat = m0.w;                    // material of the path's current vertex
    f4& hp = R.hp;
    f4& hn = R.hn;
    f4& rd = R.rd;
    f4& tp = R.tp;
    f4& direct = R.direct;
    f4& indirect = R.indirect;
    float nee_prob = __int_as_float(m1.w);
    bool v_dirty = false;
    int bounce = m0.z;

    // NEE scratch
    f4 wi = mk(0, 0, 0, 0);
    float lpdf = 0.0f, nprob = 0.0f;
    f4 ls = mk(0, 0, 0, 0);
    f4 em = mk(0, 0, 0, 0);
    f4 nee_res = mk(0, 0, 0, 0);
    int occ = 0;

    enum Step { S_NEE_START, S_NEE_SHADOW, S_NEE_MIS, S_NEE_DONE, S_BOUNCE_SAMPLE, S_FINISH, S_FINISH_CONST };
    Step step;
    f4 const_color = mk(0, 0, 0, 0);

    if (code == ST_PRIMARY) {
        float4 h = a.hit[qi];
        int htri = __float_as_int(h.y);
        if (cap_on) {
            cap.prim_tri[pix] = htri;
            cap.prim_light[pix] = (htri < 0 && (st & F_LCAND)) ? 0 : -1;
            cap.prim_t[pix] = h.x;
        }
        f4 d = rd;
        if (htri < 0) {
            if (st & F_LCAND) {     // udpt.cl:440-446
                const_color = dot(d, ld(L.normal)) < 0 ? mk(1, 1, 1, 1) : mk(0.1f, .1f, .1f, 1);
            } else {
                const_color = mk(RP.background[0], RP.background[1], RP.background[2], RP.background[3]);
            }
            step = S_FINISH_CONST;
        } else {
            f4 o = mk(J.cam.r1.w, J.cam.r2.w, J.cam.r3.w, 1.0f);
            vmat = tri_material(S.tri_shade, htri);
            hp = o + d * h.x;
            hn = shade_normal(S.tri_shade, htri, h.z, h.w);
            rd = d;
            tp = mk(1, 1, 1, 1);
            direct = mk(0, 0, 0, 0);
            indirect = mk(0, 0, 0, 0);
            v_dirty = true;
            bounce = 0;
            Mat m = mat_of(vmat);
            if (!m.spec) { st |= F_DIRECT; step = S_NEE_START; }
            else { step = S_BOUNCE_SAMPLE; }
        }
    } else {
        if (code == ST_NEE_SH) {
            occ = a.occ[qi];
            step = S_NEE_SHADOW;
        } else if (code == ST_NEE_MIS) {
            occ = a.occ[qi];
            step = S_NEE_MIS;
        } else {   // ST_BOUNCE: extension result (udpt.cl:488-529)
            float4 h = a.hit[qi];
            int htri = __float_as_int(h.y);
            Mat m = mat_of(vmat);
            f4 nd = R.nd;
            float4 bnc = st4(R.bnc);
            if (htri < 0) {
                if ((st & F_LCAND) && m.spec) indirect = indirect + tp * ld(L.ke);
                step = S_FINISH;
            } else {
                f4 o = hp + nd * kEps;
                f4 nhp = o + nd * h.x;
                f4 nhn = shade_normal(S.tri_shade, htri, h.z, h.w);
                const int nmat = tri_material(S.tri_shade, htri);
                Mat nm = mat_of(nmat);
                indirect = indirect + tp * nm.ke;
                if (m.spec) tp = tp * bnc.z;
                else
                    tp = ((tp * eval_brdf(m, nd, -rd, hp, hn, (st & F_BGLOSSY) != 0, bnc.y, on, false)) *
                          fmaxf(dot(nd, hn), 0.0f)) / bnc.x;
                step = S_NEE_START;
                if (bounce > RP.rr_threshold) {    // udpt.cl:514-523
                    float pr = fminf(luminance(tp), 0.95f);
                    float r = next_rand(seed);
                    if (r >= pr) step = S_FINISH;
                    else tp = tp * (1 / pr);
                }
                if (step != S_FINISH) {
                    v_dirty = true;
                    vmat = nmat;
                    hp = nhp;
                    hn = nhn;
                    rd = nd;
                    bounce++;
                    st &= ~F_DIRECT;
                    if (nm.spec) step = S_BOUNCE_SAMPLE;
                }
            }
        }
    }

    Out out;
    bool ended = false;
    // The steps only ever move forward (NEE start, shadow, MIS, done, bounce
    // sample, finish), so they run as one forward sequence: a step that
    // hands over sets `step` to a later one and leaves its block, a step that
    // makes a ray jumps past the rest.  Without a loop back-edge, values of
    // one step are not kept live through the others.
    if (step == S_NEE_START) do {   // evaluateDirectLighting (udpt.cl:535-555)
        Mat m = mat_of(vmat);
        em = m.ke;
        f4 wraw;
        if (!sample_lights(L, hp, hn, seed, lpdf, wraw) || lpdf <= 0.0f) { nee_res = em; step = S_NEE_DONE; break; }
        float len = length(wraw);
        len -= kEps * 1.5f;
        wi = normalize(wraw);
        bool glossy = false;
        nprob = 0.0f;
        if (fix) {
            glossy = sample_glossy_pdf(m, seed, nprob, false);
            if (nprob == 0.0f) { nee_res = em; step = S_NEE_DONE; break; }
            st = glossy ? (st | F_NGLOSSY) : (st & ~F_NGLOSSY);
        }
        f4 so = hp + wi * kEps;
        float tq;
        if (cap_on && (st & F_DIRECT) && !(st & F_CAPSH)) {
            cap.shadow_o[pix] = make_float4(so.x, so.y, so.z, len);
            cap.shadow_d[pix] = make_float4(wi.x, wi.y, wi.z, 1.0f);
        }
        if (quad_hit(L, so, wi, len, tq)) { occ = 1; step = S_NEE_SHADOW; break; }
        R.wi = mk(wi.x, wi.y, wi.z, lpdf);
        nee_prob = nprob;
        st = (st & ~ST_MASK) | ST_NEE_SH;
        out = make_ray(2, so, wi, len, jid);
        goto steps_done;
    } while (0);
    if (step == S_NEE_SHADOW) do {  // udpt.cl:557-565 and the MIS light weight
        if (code == ST_NEE_SH) {
            wi = mk(R.wi.x, R.wi.y, R.wi.z, 0.0f);
            lpdf = R.wi.w;
            nprob = nee_prob;
            code = ST_DONE;     // consumed
        }
        if (cap_on && (st & F_DIRECT) && !(st & F_CAPSH)) { cap.shadow_occ[pix] = occ; st |= F_CAPSH; }
        Mat m = mat_of(vmat);
        em = m.ke;
        f4 w_o = -rd;
        bool glossy = fix ? ((st & F_NGLOSSY) != 0) : false;
        ls = mk(0, 0, 0, 0);
        if (!occ) {
            if (!fix) {
                glossy = sample_glossy_pdf(m, seed, nprob, false);
                if (nprob == 0.0f) { nee_res = em; step = S_NEE_DONE; break; }
            }
            ls = (eval_brdf(m, wi, w_o, hp, hn, glossy, nprob, on, false) * ld(L.ke)) * fmaxf(dot(wi, hn), 0.0f);
            ls = ls * (1 / lpdf);
        }
        if (!RP.mis) { nee_res = ls + em; step = S_NEE_DONE; break; }
        float bpdf = glossy ? phong_pdf(m, wi, w_o, hn) : cos_pdf(wi, hn);
        float mw = power_heuristic(lpdf, lpdf, bpdf);
        ls = ls * mw;
        f4 bd = glossy ? phong_sample(m, bpdf, w_o, hp, hn, seed, false) : cosine_sample(bpdf, hp, hn, seed);
        if (bpdf <= 0.0f) { nee_res = ls + em; step = S_NEE_DONE; break; }
        f4 bo = hp + bd * kEps;
        float tq;
        if (!quad_hit(L, bo, bd, INFINITY, tq)) { nee_res = ls + em; step = S_NEE_DONE; break; }
        st = glossy ? (st | F_NGLOSSY) : (st & ~F_NGLOSSY);
        R.ls = ls;
        R.mis = mk(bd.x, bd.y, bd.z, bpdf);
        R.wi = mk(wi.x, wi.y, wi.z, lpdf);
        nee_prob = nprob;
        st = (st & ~ST_MASK) | ST_NEE_MIS;
        out = make_ray(2, bo, bd, tq, jid);
        goto steps_done;
    } while (0);
    if (step == S_NEE_MIS) do {     // udpt.cl:599-607
        Mat m = mat_of(vmat);
        em = m.ke;
        ls = R.ls;
        if (occ) { nee_res = ls + em; step = S_NEE_DONE; break; }
        f4 bd = mk(R.mis.x, R.mis.y, R.mis.z, 0.0f);
        float bpdf = R.mis.w;
        lpdf = R.wi.w;
        nprob = nee_prob;
        bool glossy = (st & F_NGLOSSY) != 0;
        float mw = power_heuristic(bpdf, lpdf, bpdf);
        f4 bs = (eval_brdf(m, bd, -rd, hp, hn, glossy, nprob, on, false) * ld(L.ke)) * fmaxf(dot(bd, hn), 0.0f);
        bs = bs * (mw / bpdf);
        nee_res = (ls + bs) + em;
        step = S_NEE_DONE;
        break;
    } while (0);
    if (step == S_NEE_DONE) do {
        if (st & F_DIRECT) direct = nee_res;
        else indirect = indirect + tp * nee_res;
        step = S_BOUNCE_SAMPLE;
        break;
    } while (0);
    if (step == S_BOUNCE_SAMPLE) do {   // udpt.cl:458-487
        if (!RP.gi || bounce >= RP.max_bounces) { step = S_FINISH; break; }
        Mat m = mat_of(vmat);
        float pdf = 1.0f, bprob = 0.0f, ior = 1.0f;
        bool glossy = false;
        f4 nd;
        if (m.spec) nd = fresnel_sample(m, hn, -rd, ior, seed);
        else {
            glossy = sample_glossy_pdf(m, seed, bprob, false);
            if (bprob == 0.0f) { step = S_FINISH; break; }
            nd = glossy ? phong_sample(m, pdf, -rd, hp, hn, seed, false) : cosine_sample(pdf, hp, hn, seed);
        }
        if (pdf <= 0.0f) { step = S_FINISH; break; }
        f4 no = hp + nd * kEps;
        float tq, tmax = INFINITY;
        st &= ~F_LCAND;
        if (quad_hit(L, no, nd, INFINITY, tq)) { tmax = tq; st |= F_LCAND; }
        st = glossy ? (st | F_BGLOSSY) : (st & ~F_BGLOSSY);
        R.nd = nd;
        R.bnc = mk(pdf, bprob, ior, 0.0f);
        st = (st & ~ST_MASK) | ST_BOUNCE;
        out = make_ray(1, no, nd, tmax, jid);
        goto steps_done;
    } while (0);
    if (step == S_FINISH) do {
        write_sample(J, slot, pix, direct + indirect);
        st = ST_DONE;
        ended = true;
        goto steps_done;
    } while (0);
    if (step == S_FINISH_CONST) { write_sample(J, slot, pix, const_color); st = ST_DONE; ended = true; }
steps_done:
    if (out.kind == 0 && !ended) {   // unreachable by construction; never leave a path hanging
        write_sample(J, slot, pix, mk(NAN, NAN, NAN, NAN));
        st = ST_DONE;
    }
    out.done = out.kind == 0;
    out.job = jid;
    R.m0 = make_int4((int)seed, st, v_dirty ? bounce : m0.z, v_dirty ? vmat : m0.w);
    R.m1 = make_int4(m1.x, m1.y, m1.z, __float_as_int(nee_prob));
    return out;
}

}  // namespace

// WAVES: the occupancy the register allocation is held to, per SIMD.
// Blocks of shade_block_pt(WAVES) threads walk the queues together, one atomic per
// block and output queue.  A path's record is read beside the ray whose
// result it consumes and written, once the block knows where the path's
// next ray goes, beside that ray.
// CAP: the variant that records the first frame's primary hits and shadow
// rays for the parity tests (without it, the kernel holds fewer values).
template <int WAVES, bool CAP, int LM>
__global__ __launch_bounds__(shade_block_pt(WAVES), WAVES) void k_shade_pt(ShadeArgs a, DevCapture cap)
{
    constexpr int BL = shade_block_pt(WAVES);
    __shared__ BlockReserveLds<2, BL> rsv;
    // LM: the materials of a.lmats (a.n_lmats <= kLdsMats) and the light quads
    __shared__ float4 smat4[LM ? kLdsMats * 5 : 1];
    __shared__ float4 slight4[LM ? kLdsLights * 5 : 1];
    if (LM) {
        for (int k = (int)threadIdx.x; k < a.n_lmats * 5; k += BL) smat4[k] = a.lmats[k];
        for (int k = (int)threadIdx.x; k < a.n_lights * 5; k += BL)
            slight4[k] = reinterpret_cast<const float4*>(&a.scenes[LM == 1 ? a.lscene : k / 5].light)[k % 5];
        __syncthreads();
    }
    const LdsScene lds{(lds_mat*)(void*)smat4, (lds_light*)(void*)slight4};
    // continuous serving: ended paths per job-table entry, gathered per block
    // in a small table and added to the device counters once at the end
    // (device-scope atomics on one cache line, one per wave, serialise)
    constexpr int kDoneSlots = 16;
    __shared__ int32_t dkey[kDoneSlots];
    __shared__ uint32_t dval[kDoneSlots];
    const uint32_t n_ext = *a.ext_in.count, n_sh = *a.sh_in.count;
    const uint32_t total = n_ext + n_sh;
    if (blockIdx.x == 0 && (int)threadIdx.x < a.n_zero) a.zero[threadIdx.x * kWorkStride] = 0;
    if (a.done_ctr) {
        if (threadIdx.x < (unsigned)kDoneSlots) { dkey[threadIdx.x] = -1; dval[threadIdx.x] = 0; }
        __syncthreads();
    }
    auto done_add = [&](int e, uint32_t n) {
        int h = e & (kDoneSlots - 1);
        for (int k = 0; k < kDoneSlots; ++k, h = (h + 1) & (kDoneSlots - 1)) {
            const int old = atomicCAS(&dkey[h], -1, e);
            if (old == -1 || old == e) { atomicAdd(&dval[h], n); return; }
        }
        atomicAdd(a.done_ctr + (e - a.done_base), n);     // table full: straight to the counter
    };
    uint32_t* const ctr[2] = {a.ext_out.count, a.sh_out.count};
    int par = 0;
    // the block walks the queues together: its appends are one atomic per queue
    constexpr uint32_t B = BL;
    for (uint32_t base = blockIdx.x * B; base < total; base += gridDim.x * B, par ^= 1) {
        uint32_t i = base + threadIdx.x;
        Out o;
        Rec R;
        if (i < total) {
            const bool ext = i < n_ext;
            o = advance<CAP, LM>(a, cap, ext, ext ? i : i - n_ext, R, lds);
        }
        const uint32_t n[2] = {o.kind == 1 ? 1u : 0u, o.kind == 2 ? 1u : 0u};
        uint32_t at[2];
        block_reserve<2, BL>(rsv, par, ctr, n, at);
        if (o.kind == 1) {
            a.ext_out.o_t[at[0]] = o.o_t;
            a.ext_out.d_p[at[0]] = o.d_p;
            store_rec(a.out, at[0], R);
        }
        if (o.kind == 2) {
            a.sh_out.o_t[at[1]] = o.o_t;
            a.sh_out.d_p[at[1]] = o.d_p;
            store_rec(a.out, (uint32_t)(a.out.cap - 1 - (int64_t)at[1]), R);
        }
        if (a.done_ctr) {
            // ended paths per job-table entry, one table update per entry
            // present in the wave (mostly one: a queue's paths come in runs of one job)
            const bool dn = i < total && o.done;
            unsigned long long m = __ballot(dn);
            while (m) {
                const int leader = __ffsll((long long)m) - 1;
                const int e = __shfl(o.job, leader);
                const unsigned long long g = __ballot(dn && o.job == e);
                if ((int)__lane_id() == leader && e >= a.done_base) done_add(e, (uint32_t)__popcll(g));
                m &= ~g;
            }
        }
    }
    if (a.done_ctr) {       // the loop is block-uniform: every thread gets here
        __syncthreads();
        if (threadIdx.x < (unsigned)kDoneSlots && dkey[threadIdx.x] >= 0 && dval[threadIdx.x])
            atomicAdd(a.done_ctr + (dkey[threadIdx.x] - a.done_base), dval[threadIdx.x]);
    }
}

// Primary rays (pathtracer entry, udpt.cl:162-189) for a list of segments:
// work item -> (segment, pixel), a ray at the item's queue position and the
// new path's record beside it.
__global__ __launch_bounds__(256) void k_generate(GenArgs a, DevCapture cap)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_paths) return;
    const int item = a.item0 + i;
    // behind the rays already queued: k_queue_reserve took the count before
    // this launch and added n_paths to it
    const uint32_t qpos = a.queue_base ? *a.queue_base + (uint32_t)i : (uint32_t)i;
    int lo = 0, hi = a.n_segs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a.segs[mid].path_base <= item) lo = mid; else hi = mid - 1;
    }
    const GenSeg& g = a.segs[lo];
    int x, y;
    seg_pixel(item - g.path_base + g.first, g.w, g.h, x, y);
    x += g.x0;
    y += g.y0;
    const DevJob& J = a.jobs[g.job];
    const DevScene& S = a.scenes[J.scene];
    uint32_t seed = pixel_seed(g.rand, x, y, J.W);
    float r1 = next_rand(seed);
    float r2 = next_rand(seed);
    f4 o, d;
    create_ray((float)x + r1, (float)y + r2, J.W, J.H, J.cam, o, d);
    float tq, tmax = INFINITY;
    int st = ST_PRIMARY | (J.scene << SCENE_SHIFT);
    if (quad_hit(S.light, o, d, INFINITY, tq)) { tmax = tq; st |= F_LCAND; }
    *reinterpret_cast<int4*>(a.rec.f(PathRecs::M0, qpos)) = make_int4((int)seed, st, 0, -1);
    *reinterpret_cast<int4*>(a.rec.f(PathRecs::M1, qpos)) = make_int4(y * J.W + x, g.job, g.slot, 0);
    *a.rec.f(PathRecs::RD, qpos) = st4(d);
    a.ext_out.o_t[qpos] = make_float4(o.x, o.y, o.z, tmax);
    a.ext_out.d_p[qpos] = make_float4(d.x, d.y, d.z, __int_as_float(g.job));
    if (!a.queue_base && i == 0) *a.ext_out.count = (uint32_t)a.n_paths;
    if (cap.prim_tri && g.slot == 0) {
        cap.shadow_occ[y * J.W + x] = -1;
        cap.shadow_d[y * J.W + x] = make_float4(0, 0, 0, 0);
    }
}

// Running average of the batch's frames, oldest first (udpt.cl:196-210).
__global__ __launch_bounds__(256) void k_resolve(ResolveArgs a)
{
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.w * a.h) return;
    const DevJob& J = a.jobs[a.job];
    int x = a.x0 + q % a.w, y = a.y0 + q / a.w;
    size_t pix = (size_t)y * J.W + x;
    size_t npix = (size_t)J.W * J.H;
    f4 acc = ld(a.accum[pix]);
    for (int k = 0; k < a.n_slots; ++k) {
        f4 c = ld(J.samples[(size_t)k * npix + pix]);
        if (k == 0 && a.reset_first) {
            c.w = 1;
        } else {
            int num_passes = (int)acc.w;
            c = c + acc * (float)num_passes;
            c = c / (float)(num_passes + 1);
            c.w = (float)(num_passes + 1);
        }
        acc = c;
    }
    a.accum[pix] = st4(acc);
}

// John Hable's filmic curve (tonemap.cl:49-58) on one channel: shoulder
// 0.15, linear 0.50 at angle 0.10, toe 0.20 with 0.02 / 0.30; no white-point
// division, as in the reference.
__device__ __forceinline__ float hable_curve(float x)
{
    const float sA = 0.15f, sB = 0.50f, sC = 0.10f, sD = 0.20f, sE = 0.02f, sF = 0.30f;
    return (x * (sA * x + sC * sB) + sD * sE) / (x * (sA * x + sB) + sD * sF) - sE / sF;
}

// tonemap.cl:14-47, op 0: Reinhard with lum_white = 1 (identity up to
// rounding); op 1: the Hable curve the reference ships commented out
// (tonemap.cl:29).  The clamp is discarded as in the reference, then gamma
// 1/2.2.
__global__ __launch_bounds__(256) void k_tonemap(const float4* in, float4* out, int n, int op)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f4 col = ld(in[i]);
    f4 ldr;
    if (op == 1) {
        ldr = mk(hable_curve(col.x), hable_curve(col.y), hable_curve(col.z), hable_curve(col.w));
    } else {
        float lum_world = 0.212671f * col.x + 0.715160f * col.y + 0.072169f * col.z + 0.001f;
        float lum_white = 1.0f;
        float lum_display = lum_world * (1 + lum_world / (lum_white * lum_white)) / (1 + lum_world);
        f4 sc = col / lum_world;
        ldr = mk(ynf_powf(sc.x, 1.0f), ynf_powf(sc.y, 1.0f), ynf_powf(sc.z, 1.0f), ynf_powf(sc.w, 1.0f)) * lum_display;
    }
    const float g = 1 / 2.2f;
    out[i] = make_float4(ynf_powf(ldr.x, g), ynf_powf(ldr.y, g), ynf_powf(ldr.z, g), ynf_powf(ldr.w, g));
}

// Rays given by a caller (n x 8 floats: o, len, d, is_shadow): quad test
// (traceRay, udpt.cl:244-276) and queue fill.  hit/occ slots are indexed by
// ray number.
__global__ __launch_bounds__(256) void k_prepare_rays(const DevScene* scene, const float* rays, int n, RayQueue q,
                                                      int32_t* light_out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = rays + 8 * (size_t)i;
    f4 o = mk(r[0], r[1], r[2], 1.0f), d = mk(r[4], r[5], r[6], 0.0f);
    float len = r[3], t;
    int light = -1;
    if (quad_hit(scene->light, o, d, len, t)) { len = t; light = 0; }
    light_out[i] = light;
    q.o_t[i] = make_float4(o.x, o.y, o.z, len);
    q.d_p[i] = make_float4(d.x, d.y, d.z, __int_as_float(i));
    if (i == 0) *q.count = (uint32_t)n;
}

// The kernels' sin / cos / pow evaluated on the device (diagnostics).
__global__ __launch_bounds__(256) void k_fmath(int op, const float* x, const float* y, int n, float* out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = op == 0 ? ynf_sinf(x[i]) : op == 1 ? ynf_cosf(x[i]) : ynf_powf(x[i], y[i]);
}

void launch_fmath(int op, const float* x, const float* y, int n, float* out, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fmath, dim3((n + 255) / 256), dim3(256), 0, s, op, x, y, n, out);
}

// Reserves n entries at the end of a queue: base = count, count += n.
__global__ void k_queue_reserve(uint32_t* count, uint32_t n, uint32_t* base)
{
    if (threadIdx.x == 0) {
        const uint32_t c = *count;
        *base = c;
        *count = c + n;
    }
}

void launch_generate(const GenArgs& a, const DevCapture& cap, hipStream_t s)
{
    int grid = (a.n_paths + 255) / 256;
    if (grid <= 0) return;
    if (a.queue_base)
        hipLaunchKernelGGL(k_queue_reserve, dim3(1), dim3(64), 0, s, a.ext_out.count, (uint32_t)a.n_paths,
                           a.queue_base);
    hipLaunchKernelGGL(k_generate, dim3(grid), dim3(256), 0, s, a, cap);
}
void launch_shade_pt(const ShadeArgs& a, const DevCapture& cap, int grid, int waves, hipStream_t s)
{
    if (cap.prim_tri) {
        hipLaunchKernelGGL((k_shade_pt<4, true, 0>), dim3(grid), dim3(shade_block_pt(4)), 0, s, a, cap);
        return;
    }
    const int lm = a.lmats != nullptr && a.n_lmats <= kLdsMats ? a.lmode : 0;
    if (waves == 4) {
        if (lm == 1) hipLaunchKernelGGL((k_shade_pt<4, false, 1>), dim3(grid), dim3(shade_block_pt(4)), 0, s, a, cap);
        else if (lm == 2) hipLaunchKernelGGL((k_shade_pt<4, false, 2>), dim3(grid), dim3(shade_block_pt(4)), 0, s, a, cap);
        else hipLaunchKernelGGL((k_shade_pt<4, false, 0>), dim3(grid), dim3(shade_block_pt(4)), 0, s, a, cap);
    } else if (waves == 3) {
        hipLaunchKernelGGL((k_shade_pt<3, false, 0>), dim3(grid), dim3(shade_block_pt(3)), 0, s, a, cap);
    } else {
        hipLaunchKernelGGL((k_shade_pt<2, false, 0>), dim3(grid), dim3(shade_block_pt(2)), 0, s, a, cap);
    }
}
void launch_resolve(const ResolveArgs& a, hipStream_t s)
{
    int n = a.w * a.h;
    if (n > 0) hipLaunchKernelGGL(k_resolve, dim3((n + 255) / 256), dim3(256), 0, s, a);
}
void launch_tonemap(const float4* in, float4* out, int n, int op, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_tonemap, dim3((n + 255) / 256), dim3(256), 0, s, in, out, n, op);
}
void launch_prepare(const DevScene* scene, const float* rays, int n, RayQueue q, int32_t* light_out, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_prepare_rays, dim3((n + 255) / 256), dim3(256), 0, s, scene, rays, n, q, light_out);
}

}  // namespace yune
