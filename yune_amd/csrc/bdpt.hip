// bdpt.hip — generate and shade kernels of the wavefront bidirectional path
// tracer (kernels/legacy/bdpt.cl semantics: createLightPath :432-511,
// createEyePath :513-560, shading :562-640).
//
// The reference builds, per work-item, a light subpath and an eye subpath
// (each up to BDPT_BOUNCES vertices, one traceRay per vertex), then walks
// the eye vertices: evaluateDirectLighting at each, and a shadow-ray
// connection to every light vertex j >= 1.  Here each path is a state
// machine over HBM-resident vertex arrays:
//
//   generate: camera ray kept, light vertex 0 on the quad, first light ray
//   LIGHT   --ext--> vertex k of the light subpath, next light ray ...
//   EYE     --ext--> vertex k of the eye subpath (k = 0: the camera ray) ...
//   ROUND i --shadow--> the NEE shadow ray of eye vertex i (slot 0) and the
//                      connection rays to light vertices j (slot j), all in
//                      one batch; [MIS --shadow--> the MIS ray (slot 0)];
//                      then NEE, connections and the eye weight, in the
//                      reference's order, and on to eye vertex i + 1.
//
// The per-path xorshift state is consumed in the reference's order: a
// connection's visibility does not depend on the random stream, so the
// connection rays of a vertex are traced together with its NEE ray and
// their results used afterwards, in the reference's j order, each drawing
// its two lobe picks only when unoccluded (bdpt.cl:611-620).
#include "yn_internal_dev.h"
#include "yn_shade.h"

namespace yune {

namespace {

enum : int {
    B_DONE = 0,
    B_LIGHT = 1,            // waiting for the ray that makes light vertex `bounce`
    B_EYE = 2,              // waiting for the ray that makes eye vertex `bounce`
    B_ROUND = 3,            // waiting for the NEE + connection rays of eye vertex `bounce`
    B_MIS = 4,              // waiting for the MIS ray of eye vertex `bounce`
    B_MASK = 0xff,
    BF_LCAND = 1 << 8,      // pending extension ray ends on the quad
    BF_BGLOSSY = 1 << 9,    // lobe of the pending extension ray
    BF_NGLOSSY = 1 << 10,   // NEE lobe
    BF_NEE_Q = 1 << 11,     // NEE shadow ray blocked by the quad itself (occluded, no query)
    BF_NEE_DONE = 1 << 12,  // NEE finished before its shadow ray; result in P.ls
    BF_CAPSH = 1 << 13,     // first shadow ray already captured
};

constexpr float kPinkR = 0.988f, kPinkG = 0.0588f, kPinkB = 0.7529f;   // bdpt.cl:110

struct Vtx {
    f4 p, n, d, c;
    int mat;                // material of the vertex's triangle (-1: light vertex 0)
};

using BF = BdptRecs;

// the position field of vertex k of a side; normal, direction and
// contribution follow it
__device__ __forceinline__ float4* vfield(const BdptSoA& B, int side, int k, int p)
{
    return B.R.f(BF::VTX + 4 * (side * B.K + k), p);
}

__device__ __forceinline__ Vtx load_v(const BdptSoA& B, int side, int k, int p)
{
    const float4* f = vfield(B, side, k, p);
    float4 q = f[0];
    Vtx v;
    v.p = mk(q.x, q.y, q.z, 1.0f);
    v.mat = __float_as_int(q.w);
    v.n = ld(f[BF::kTile]);
    v.d = ld(f[2 * BF::kTile]);
    v.c = ld(f[3 * BF::kTile]);
    return v;
}

__device__ __forceinline__ void store_v(const BdptSoA& B, int side, int k, int p, const Vtx& v)
{
    float4* f = vfield(B, side, k, p);
    f[0] = make_float4(v.p.x, v.p.y, v.p.z, __int_as_float(v.mat));
    f[BF::kTile] = st4(v.n);
    f[2 * BF::kTile] = st4(v.d);
    f[3 * BF::kTile] = st4(v.c);
}

__device__ __forceinline__ int4 ld_i4(const BdptSoA& B, int field, int p)
{
    return *reinterpret_cast<const int4*>(B.R.f(field, p));
}

__device__ __forceinline__ void write_sample(const DevJob& J, int slot, int pix, f4 c)
{
    if (any_nan(c)) c = mk(kPinkR, kPinkG, kPinkB, 1.0f);     // bdpt.cl:193-194
    J.samples[(size_t)slot * (size_t)(J.W * J.H) + (size_t)pix] = st4(c);
}

// Connection ray from eye vertex e to light vertex l (bdpt.cl:597-606).
// Returns false when either end faces away (the reference's `continue`).
__device__ __forceinline__ bool conn_ray(const Vtx& e, const Vtx& l, f4& o, f4& d, float& len)
{
    d = normalize(l.p - e.p);
    o = e.p + d * kEps;
    len = length(l.p - o);
    return !(dot(d, e.n) <= 0 || dot(-d, l.n) <= 0);
}

struct BOut {
    bool ext = false;       // one extension ray (o_t, d_p)
    uint64_t emit = 0;      // shadow-ray slots to emit this round (bit j: slot j, j < 64)
    float4 o_t, d_p;        // the extension ray, or the slot-0 shadow ray
    int eye_i = 0;          // eye vertex of the round (slots >= 1 are connections)
};

__device__ __forceinline__ void set_ray(BOut& r, f4 o, f4 d, float tmax, int id)
{
    r.o_t = make_float4(o.x, o.y, o.z, tmax);
    r.d_p = make_float4(d.x, d.y, d.z, __int_as_float(id));
}

template <bool CAP, int LM>
__device__ BOut advance_bdpt(const BdptShadeArgs& a, const DevCapture& cap, int p, const LdsScene& lds)
{
    const PathSoA& P = a.P;
    const BdptSoA& B = a.B;
    const DevJob& J = a.jobs[P.job[p]];
    const yn_render_params& RP = J.p;
    const DevScene& S = a.scenes[J.scene];
    // LM: materials and light in LDS, the launch's one scene's (1) or every scene's (2)
    auto mat_of = [&](int id) {
        return LM == 1 ? load_mat(lds.mats, id) : LM == 2 ? load_mat(lds.mats, S.mat_base + id) : load_mat(S.mats, id);
    };
    auto& L = light_of<LM>(S.light, lds, J.scene);
    const bool fix = RP.fix_mis != 0;
    const int KB = RP.bdpt_bounces;

    const int4 i0 = ld_i4(B, BF::I0, p), i1 = ld_i4(B, BF::I1, p), i2 = ld_i4(B, BF::I2, p);
    int st = i0.y;
    const int code = st & B_MASK;
    uint32_t seed = (uint32_t)i0.x;
    const int pix = i0.w;
    const int slot = i1.x;
    const bool cap_on = CAP && cap.prim_tri != nullptr && slot == 0;
    int k = i0.z;
    int lp_len = i1.y, ep_len = i1.z;
    float w_eye = __int_as_float(i1.w);
    f4 color = ld(*B.R.f(BF::DIRECT, p));
    uint64_t conn = (uint64_t)(uint32_t)i2.x | ((uint64_t)(uint32_t)i2.y << 32);
    float nee_prob = __int_as_float(i2.z);     // NEE lobe probability, carried between rounds

    enum Step { T_LIGHT_NEXT, T_EYE_START, T_EYE_NEXT, T_ROUND_START, T_NEE_SHADOW, T_NEE_MIS, T_CONNECT,
                T_FINISH, T_FINISH_CONST };
    Step step = T_FINISH;
    f4 const_color = mk(0, 0, 0, 0);
    Vtx cur;                // the vertex the next step starts from
    f4 nee_res = mk(0, 0, 0, 0);
    BOut out;

    if (code == B_LIGHT || code == B_EYE) {
        const int side = code == B_LIGHT ? 0 : 1;
        float4 h = P.hit[p];
        int htri = __float_as_int(h.y);
        if (side == 1 && k == 0) {          // camera ray (createEyePath :517-528, shading :566-571)
            if (cap_on) {
                cap.prim_tri[pix] = htri;
                cap.prim_light[pix] = (htri < 0 && (st & BF_LCAND)) ? 0 : -1;
                cap.prim_t[pix] = h.x;
            }
            f4 d = ld(*B.R.f(BF::EYE_D, p));
            if (htri < 0) {
                if (st & BF_LCAND)
                    const_color = dot(d, ld(L.normal)) < 0 ? mk(1, 1, 1, 1) : mk(0.1f, 0.1f, 0.1f, 1);
                else
                    const_color = mk(RP.background[0], RP.background[1], RP.background[2], RP.background[3]);
                step = T_FINISH_CONST;
            } else {
                f4 o = ld(*B.R.f(BF::EYE_O, p));
                cur.p = o + d * h.x;
                cur.n = shade_normal(S.tri_shade, htri, h.z, h.w);
                cur.d = d;
                cur.c = mk(1.0f, 1.0f, 1.0f, 1.0f);
                cur.mat = S.tri_mat[htri];
                store_v(B, 1, 0, p, cur);
                ep_len = 1;
                step = 1 < KB ? T_EYE_NEXT : T_ROUND_START;
            }
        } else if (htri < 0) {              // miss or the quad: the subpath ends
            step = side == 0 ? T_EYE_START : T_ROUND_START;
        } else {
            Vtx pv = load_v(B, side, k - 1, p);
            f4 d = ld(*B.R.f(BF::ND, p));
            float4 bnc = *B.R.f(BF::BNC, p);
            float pdf = bnc.x, prob = bnc.y;
            bool g = (st & BF_BGLOSSY) != 0;
            f4 o = pv.p + d * kEps;
            cur.p = o + d * h.x;
            cur.n = shade_normal(S.tri_shade, htri, h.z, h.w);
            cur.d = d;
            cur.mat = S.tri_mat[htri];
            if (side == 0 && k == 1) {      // :453-458
                cur.c = bc(fmaxf(0.0f, dot(pv.n, d)) / pdf);
                cur.c = cur.c * pv.c;
            } else {
                Mat pm = mat_of(pv.mat);
                f4 f = side == 0 ? eval_brdf(pm, -pv.d, d, pv.p, pv.n, g, prob, false, true)      // :486-490
                                 : eval_brdf(pm, d, -pv.d, pv.p, pv.n, g, prob, false, true);     // :546-550
                cur.c = f * fmaxf(0.0f, dot(d, pv.n));
                cur.c = cur.c / pdf;
                cur.c = cur.c * pv.c;
            }
            if (side == 0) lp_len = k + 1; else ep_len = k + 1;
            bool ended = false;
            if ((side == 1 || k >= 2) && k > RP.rr_threshold) {     // :494-502, :552-558
                float r = next_rand(seed);
                float pr = fminf(luminance(cur.c), 0.95f);
                if (r >= pr) ended = true;
                else cur.c = cur.c * (1.0f / pr);
            }
            store_v(B, side, k, p, cur);
            if (!ended && k + 1 < KB) step = side == 0 ? T_LIGHT_NEXT : T_EYE_NEXT;
            else step = side == 0 ? T_EYE_START : T_ROUND_START;
        }
        if (step == T_ROUND_START) { k = 0; w_eye = 1.0f; color = mk(0, 0, 0, 1); }
    } else if (code == B_ROUND) {
        step = T_NEE_SHADOW;
    } else if (code == B_MIS) {
        step = T_NEE_MIS;
    }

    bool ended = false;
    // Subpath steps and the camera ray either make a ray or hand over to a
    // later step; only the connection step goes back (to the next eye
    // vertex's round), so only the round steps sit in a loop and the
    // subpath vertex is not kept live through them.
    if (step == T_LIGHT_NEXT || step == T_EYE_NEXT) do {   // :474-484, :532-544
        const bool light = step == T_LIGHT_NEXT;
        Mat m = mat_of(cur.mat);
        float prob = 0.0f, pdf = 1.0f;
        bool g = sample_glossy_pdf(m, seed, prob, true);
        Step stop = light ? T_EYE_START : T_ROUND_START;
        if (prob == 0.0f) { step = stop; if (!light) { k = 0; w_eye = 1.0f; color = mk(0, 0, 0, 1); } break; }
        f4 d = g ? phong_sample(m, pdf, -cur.d, cur.p, cur.n, seed, true) : cosine_sample(pdf, cur.p, cur.n, seed);
        if (pdf <= 0.0f) { step = stop; if (!light) { k = 0; w_eye = 1.0f; color = mk(0, 0, 0, 1); } break; }
        f4 o = cur.p + d * kEps;
        float tq, tmax = INFINITY;
        st &= ~(BF_LCAND | BF_BGLOSSY);
        if (quad_hit(L, o, d, INFINITY, tq)) { tmax = tq; st |= BF_LCAND; }
        if (g) st |= BF_BGLOSSY;
        *B.R.f(BF::ND, p) = st4(d);
        *B.R.f(BF::BNC, p) = make_float4(pdf, prob, 0.0f, 0.0f);
        st = (st & ~B_MASK) | (light ? B_LIGHT : B_EYE);
        k = k + 1;
        out.ext = true;
        set_ray(out, o, d, tmax, p);
        goto emitted;
    } while (0);
    if (step == T_EYE_START) do {          // the camera ray (createEyePath :517)
        f4 o = ld(*B.R.f(BF::EYE_O, p)), d = ld(*B.R.f(BF::EYE_D, p));
        float tq, tmax = INFINITY;
        st &= ~(BF_LCAND | BF_BGLOSSY);
        if (quad_hit(L, o, d, INFINITY, tq)) { tmax = tq; st |= BF_LCAND; }
        st = (st & ~B_MASK) | B_EYE;
        k = 0;
        out.ext = true;
        set_ray(out, o, d, tmax, p);
        goto emitted;
    } while (0);
    // a no-ray round is 3 steps per eye vertex (<= 64): the bound is never reached
    for (int guard = 0; guard < 1024; ++guard) {
        if (step == T_ROUND_START) do {        // eye vertex k: NEE start + connection rays
            if (k >= ep_len || w_eye == 0) { step = T_FINISH; break; }
            Vtx e = load_v(B, 1, k, p);
            Mat m = mat_of(e.mat);
            conn = 0;
            st &= ~(BF_NEE_Q | BF_NEE_DONE | BF_NGLOSSY);
            // evaluateDirectLighting start (bdpt.cl's copy of udpt.cl:535-555)
            f4 wraw;
            float lpdf = 0.0f, nprob = 0.0f;
            bool nee_done = false;
            if (!sample_lights(L, e.p, e.n, seed, lpdf, wraw) || lpdf <= 0.0f) nee_done = true;
            f4 wi = mk(0, 0, 0, 0);
            float len = 0.0f;
            if (!nee_done) {
                len = length(wraw);
                len -= kEps * 1.5f;
                wi = normalize(wraw);
                if (fix) {
                    bool g = sample_glossy_pdf(m, seed, nprob, true);
                    if (nprob == 0.0f) nee_done = true;
                    else if (g) st |= BF_NGLOSSY;
                }
            }
            if (nee_done) {
                *B.R.f(BF::LS, p) = st4(m.ke);
                st |= BF_NEE_DONE;
            } else {
                f4 so = e.p + wi * kEps;
                if (cap_on && k == 0 && !(st & BF_CAPSH)) {
                    cap.shadow_o[pix] = make_float4(so.x, so.y, so.z, len);
                    cap.shadow_d[pix] = make_float4(wi.x, wi.y, wi.z, 1.0f);
                }
                *B.R.f(BF::WI, p) = make_float4(wi.x, wi.y, wi.z, lpdf);
                nee_prob = nprob;
                float tq;
                if (quad_hit(L, so, wi, len, tq)) st |= BF_NEE_Q;
                else { conn |= 1ull; set_ray(out, so, wi, len, p << B.shift); }
            }
            // connection rays (shading :597-609)
            for (int j = lp_len - 1; j > 0; --j) {
                Vtx l = load_v(B, 0, j, p);
                f4 o, d;
                float cl;
                float tq;
                if (conn_ray(e, l, o, d, cl) && !quad_hit(L, o, d, cl, tq)) conn |= 1ull << j;
            }
            if (conn == 0) { step = T_NEE_SHADOW; break; }
            st = (st & ~B_MASK) | B_ROUND;
            out.emit = conn;
            out.eye_i = k;
            goto emitted;
        } while (0);
        if (step == T_NEE_SHADOW) do {         // udpt.cl:557-597 (bdpt variant of the BRDF pieces)
            Vtx e = load_v(B, 1, k, p);
            Mat m = mat_of(e.mat);
            if (st & BF_NEE_DONE) { nee_res = ld(*B.R.f(BF::LS, p)); step = T_CONNECT; break; }
            int occ = (st & BF_NEE_Q) ? 1 : B.occ[((size_t)p << B.shift)];
            if (cap_on && k == 0 && !(st & BF_CAPSH)) { cap.shadow_occ[pix] = occ; st |= BF_CAPSH; }
            float4 w4 = *B.R.f(BF::WI, p);
            f4 wi = mk(w4.x, w4.y, w4.z, 0.0f);
            float lpdf = w4.w;
            float nprob = nee_prob;
            f4 em = m.ke;
            f4 w_o = -e.d;
            bool glossy = fix ? ((st & BF_NGLOSSY) != 0) : false;
            f4 ls = mk(0, 0, 0, 0);
            if (!occ) {
                if (!fix) {
                    glossy = sample_glossy_pdf(m, seed, nprob, true);
                    if (nprob == 0.0f) { nee_res = em; step = T_CONNECT; break; }
                }
                ls = (eval_brdf(m, wi, w_o, e.p, e.n, glossy, nprob, false, true) * ld(L.ke)) * fmaxf(dot(wi, e.n), 0.0f);
                ls = ls * (1 / lpdf);
            }
            if (!RP.mis) { nee_res = ls + em; step = T_CONNECT; break; }
            float bpdf = glossy ? phong_pdf(m, wi, w_o, e.n) : cos_pdf(wi, e.n);
            float mw = power_heuristic(lpdf, lpdf, bpdf);
            ls = ls * mw;
            f4 bd = glossy ? phong_sample(m, bpdf, w_o, e.p, e.n, seed, true) : cosine_sample(bpdf, e.p, e.n, seed);
            if (bpdf <= 0.0f) { nee_res = ls + em; step = T_CONNECT; break; }
            f4 bo = e.p + bd * kEps;
            float tq;
            if (!quad_hit(L, bo, bd, INFINITY, tq)) { nee_res = ls + em; step = T_CONNECT; break; }
            st = glossy ? (st | BF_NGLOSSY) : (st & ~BF_NGLOSSY);
            *B.R.f(BF::LS, p) = st4(ls);
            *B.R.f(BF::MIS, p) = make_float4(bd.x, bd.y, bd.z, bpdf);
            nee_prob = nprob;
            st = (st & ~B_MASK) | B_MIS;
            conn |= 1ull;
            out.emit = 1ull;
            out.eye_i = k;
            set_ray(out, bo, bd, tq, p << B.shift);
            goto emitted;
        } while (0);
        if (step == T_NEE_MIS) do {            // udpt.cl:599-607
            Vtx e = load_v(B, 1, k, p);
            Mat m = mat_of(e.mat);
            f4 em = m.ke;
            f4 ls = ld(*B.R.f(BF::LS, p));
            int occ = B.occ[((size_t)p << B.shift)];
            if (occ) { nee_res = ls + em; step = T_CONNECT; break; }
            float4 mb = *B.R.f(BF::MIS, p);
            f4 bd = mk(mb.x, mb.y, mb.z, 0.0f);
            float bpdf = mb.w;
            float lpdf = B.R.f(BF::WI, p)->w;
            float nprob = nee_prob;
            bool glossy = (st & BF_NGLOSSY) != 0;
            float mw = power_heuristic(bpdf, lpdf, bpdf);
            f4 bs = (eval_brdf(m, bd, -e.d, e.p, e.n, glossy, nprob, false, true) * ld(L.ke)) * fmaxf(dot(bd, e.n), 0.0f);
            bs = bs * (mw / bpdf);
            nee_res = (ls + bs) + em;
            step = T_CONNECT;
            break;
        } while (0);
        if (step != T_CONNECT) break;
        {            // shading :578-637 for eye vertex k
            Vtx e = load_v(B, 1, k, p);
            Mat m = mat_of(e.mat);
            f4 spec = m.ks, em = m.ke;
            float ks = fmaxf(fmaxf(spec.x, fmaxf(spec.y, spec.z)), 0.1f);
            f4 tp = e.c;
            color = color + (tp * (em + nee_res)) * w_eye;
            f4 sub = mk(0, 0, 0, 1);
            for (int j = lp_len - 1; j > 0; --j) {
                if (!((conn >> j) & 1ull)) continue;
                if (B.occ[((size_t)p << B.shift) + (size_t)j]) continue;
                Vtx l = load_v(B, 0, j, p);
                f4 d = normalize(l.p - e.p);
                float dist = length(l.p - e.p);
                dist *= dist;
                f4 tlp = l.c;
                f4 w_i = d, w_o = -e.d;
                float gf = fmaxf(dot(w_i, e.n), 0.0f) * fmaxf(dot(-w_i, l.n), 0.0f) / dist;
                float prob = 0.0f;
                bool g = sample_glossy_pdf(m, seed, prob, true);
                f4 e2l = eval_brdf(m, w_i, w_o, e.p, e.n, g, prob, false, true);
                Mat lm = mat_of(l.mat);
                g = sample_glossy_pdf(lm, seed, prob, true);
                f4 l2e = eval_brdf(lm, -l.d, -w_i, l.p, l.n, g, prob, false, true);
                tlp = tlp * ((e2l * gf) * l2e);
                tlp = tlp * tp;
                sub = sub + tlp;
            }
            color = color + sub * (w_eye * (1 - ks));
            w_eye = w_eye * ks;
            k = k + 1;
            step = T_ROUND_START;
            continue;
        }
    }
    if (step == T_FINISH) { write_sample(J, slot, pix, color); ended = true; }
    else if (step == T_FINISH_CONST) { write_sample(J, slot, pix, const_color); ended = true; }
emitted:
    if (!out.ext && out.emit == 0 && !ended) {   // unreachable by construction; never leave a path hanging
        write_sample(J, slot, pix, mk(NAN, NAN, NAN, NAN));
        ended = true;
    }
    *reinterpret_cast<int4*>(B.R.f(BF::I0, p)) = make_int4((int)seed, ended ? B_DONE : st, k, pix);
    *reinterpret_cast<int4*>(B.R.f(BF::I1, p)) = make_int4(slot, lp_len, ep_len, __float_as_int(w_eye));
    *reinterpret_cast<int4*>(B.R.f(BF::I2, p)) =
        make_int4((int)(uint32_t)conn, (int)(uint32_t)(conn >> 32), __float_as_int(nee_prob), 0);
    *B.R.f(BF::DIRECT, p) = st4(color);
    return out;
}

}  // namespace

// Compiled for 3 waves per SIMD (168 VGPRs, a few spilled) and launched as
// exactly the blocks that fit: unconstrained it takes 204 VGPRs and runs 2
// waves per SIMD, 9% slower on the Cornell box.
// CAP: the variant that records the first frame's camera hits and first
// shadow rays for the parity tests.
template <bool CAP, int LM>
__global__ __launch_bounds__(256, 3) void k_shade_bdpt(BdptShadeArgs a, DevCapture cap)
{
    __shared__ BlockReserveLds<3, 256> rsv;
    // LM: the materials of a.lmats (a.n_lmats <= kLdsMats) and the light quads
    __shared__ float4 smat4[LM ? kLdsMats * 5 : 1];
    __shared__ float4 slight4[LM ? kLdsLights * 5 : 1];
    if (LM) {
        for (int k = (int)threadIdx.x; k < a.n_lmats * 5; k += 256) smat4[k] = a.lmats[k];
        for (int k = (int)threadIdx.x; k < a.n_lights * 5; k += 256)
            slight4[k] = reinterpret_cast<const float4*>(&a.scenes[LM == 1 ? a.lscene : k / 5].light)[k % 5];
        __syncthreads();
    }
    const LdsScene lds{(lds_mat*)(void*)smat4, (lds_light*)(void*)slight4};
    const uint32_t n = *a.list_in_count;
    if (blockIdx.x == 0 && (int)threadIdx.x < a.n_zero) a.zero[threadIdx.x * kWorkStride] = 0;
    const int shift = a.B.shift;
    uint32_t* const ctr[3] = {a.ext_out.count, a.sh_out.count, a.list_out_count};
    int par = 0;
    // the block walks the list together: its appends are one atomic per queue
    for (uint32_t base = blockIdx.x * 256u; base < n; base += gridDim.x * 256u, par ^= 1) {
        uint32_t i = base + threadIdx.x;
        BOut o;
        int p = -1;
        if (i < n) {
            p = (int)a.list_in[i];
            o = advance_bdpt<CAP, LM>(a, cap, p, lds);
        }
        const bool more = o.ext || o.emit != 0;
        const uint32_t cnt[3] = {o.ext ? 1u : 0u, (uint32_t)__popcll(o.emit), more ? 1u : 0u};
        uint32_t at[3];
        block_reserve<3, 256>(rsv, par, ctr, cnt, at);
        if (o.ext) { a.ext_out.o_t[at[0]] = o.o_t; a.ext_out.d_p[at[0]] = o.d_p; }
        uint32_t s = at[1];
        if (o.emit) {
            uint64_t m = o.emit;
            if (m & 1ull) {
                a.sh_out.o_t[s] = o.o_t;
                a.sh_out.d_p[s] = o.d_p;
                ++s;
                m &= ~1ull;
            }
            if (m) {
                // connection rays, rebuilt exactly as when they were counted
                const BdptSoA& B = a.B;
                Vtx ev = load_v(B, 1, o.eye_i, p);
                while (m) {
                    int j = __ffsll((long long)m) - 1;
                    m &= m - 1;
                    Vtx l = load_v(B, 0, j, p);
                    f4 ro, rd;
                    float len;
                    conn_ray(ev, l, ro, rd, len);
                    a.sh_out.o_t[s] = make_float4(ro.x, ro.y, ro.z, len);
                    a.sh_out.d_p[s] = make_float4(rd.x, rd.y, rd.z, __int_as_float((p << shift) | j));
                    ++s;
                }
            }
        }
        if (more) a.list_out[at[2]] = (uint32_t)p;
    }
}

// Camera ray, light vertex 0 and the first light ray (bdpt.cl:162-189,
// createLightPath :432-462) for a list of segments.
__global__ __launch_bounds__(256) void k_generate_bdpt(GenArgs a, DevCapture cap)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n_paths) return;
    int lo = 0, hi = a.n_segs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a.segs[mid].path_base <= p) lo = mid; else hi = mid - 1;
    }
    const GenSeg& g = a.segs[lo];
    int x, y;
    seg_pixel(p - g.path_base + g.first, g.w, g.h, x, y);
    x += g.x0;
    y += g.y0;
    const DevJob& J = a.jobs[g.job];
    const DevScene& S = a.scenes[J.scene];
    const yn_light_quad& L = S.light;
    const BdptSoA& B = a.B;
    uint32_t seed = pixel_seed(g.rand, x, y, J.W);
    float r1 = next_rand(seed);
    float r2 = next_rand(seed);
    f4 eo, ed;
    create_ray((float)x + r1, (float)y + r2, J.W, J.H, J.cam, eo, ed);
    *B.R.f(BF::EYE_O, p) = st4(eo);
    *B.R.f(BF::EYE_D, p) = st4(ed);
    // light vertex 0: r1 by a double division (bdpt.cl:439), r1/r2 swapped on the edges
    seed = xor_shift(seed);
    float lr1 = (float)((double)seed / (double)4294967295u);
    seed = xor_shift(seed);
    float lr2 = u2f(seed);
    f4 A = ld(L.edge_l) * lr2;
    f4 Bv = ld(L.edge_w) * lr1;
    Vtx v0;
    v0.p = (A + Bv) + ld(L.pos);
    v0.n = ld(L.normal);
    v0.d = mk(0, 0, 0, 0);
    float area = length(ld(L.edge_l)) * length(ld(L.edge_w));
    float fwd = 1.0f / area;
    v0.c = ld(L.ke) / fwd;
    v0.mat = -1;
    store_v(B, 0, 0, p, v0);
    float pdf = 1.0f;
    f4 d = cosine_sample(pdf, v0.p, v0.n, seed);
    int st;
    f4 o;
    if (pdf > 0.0f) {
        o = v0.p + d * kEps;
        st = B_LIGHT;
        *B.R.f(BF::ND, p) = st4(d);
        *B.R.f(BF::BNC, p) = make_float4(pdf, 0.0f, 0.0f, 0.0f);
    } else {                                // no light subpath beyond vertex 0: the camera ray
        o = eo;
        d = ed;
        st = B_EYE;
    }
    float tq, tmax = INFINITY;
    if (quad_hit(L, o, d, INFINITY, tq)) { tmax = tq; st |= BF_LCAND; }
    a.P.job[p] = g.job;
    // bounce: 1 when the first light ray leaves vertex 0, else 0 (the camera ray)
    *reinterpret_cast<int4*>(B.R.f(BF::I0, p)) = make_int4((int)seed, st, pdf > 0.0f ? 1 : 0, y * J.W + x);
    *reinterpret_cast<int4*>(B.R.f(BF::I1, p)) = make_int4(g.slot, 1, 0, __float_as_int(1.0f));
    *reinterpret_cast<int4*>(B.R.f(BF::I2, p)) = make_int4(0, 0, 0, 0);
    *B.R.f(BF::DIRECT, p) = make_float4(0, 0, 0, 1);
    a.ext_out.o_t[p] = make_float4(o.x, o.y, o.z, tmax);
    a.ext_out.d_p[p] = make_float4(d.x, d.y, d.z, __int_as_float(p));
    a.list_out[p] = (uint32_t)p;
    if (p == 0) { *a.ext_out.count = (uint32_t)a.n_paths; *a.list_count = (uint32_t)a.n_paths; }
    if (cap.prim_tri && g.slot == 0) {
        cap.shadow_occ[y * J.W + x] = -1;
        cap.shadow_d[y * J.W + x] = make_float4(0, 0, 0, 0);
    }
}

void launch_generate_bdpt(const GenArgs& a, const DevCapture& cap, hipStream_t s)
{
    int grid = (a.n_paths + 255) / 256;
    if (grid > 0) hipLaunchKernelGGL(k_generate_bdpt, dim3(grid), dim3(256), 0, s, a, cap);
}

void launch_shade_bdpt(const BdptShadeArgs& a, const DevCapture& cap, int grid, hipStream_t s)
{
    const int lm = a.lmats != nullptr && a.n_lmats <= kLdsMats ? a.lmode : 0;
    if (cap.prim_tri) hipLaunchKernelGGL((k_shade_bdpt<true, 0>), dim3(grid), dim3(256), 0, s, a, cap);
    else if (lm == 1) hipLaunchKernelGGL((k_shade_bdpt<false, 1>), dim3(grid), dim3(256), 0, s, a, cap);
    else if (lm == 2) hipLaunchKernelGGL((k_shade_bdpt<false, 2>), dim3(grid), dim3(256), 0, s, a, cap);
    else hipLaunchKernelGGL((k_shade_bdpt<false, 0>), dim3(grid), dim3(256), 0, s, a, cap);
}

}  // namespace yune
