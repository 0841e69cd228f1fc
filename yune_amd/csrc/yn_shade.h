// yn_shade.h — materials, sampling and light sampling on the device.
//
// Each function restates the reference kernel function it names, with the
// operand order of the OpenCL source (see yn_math.h for the built-ins).
// `bdpt` selects the kernels/legacy/bdpt.cl variant where the two files
// differ.
#pragma once

#include "yn_device.h"
#include "yn_math.h"
#include "yn_fmath.h"

namespace yune {

struct Mat {
    f4 ke, kd, ks;
    float n, px, py, alpha_x;
    int spec, trans;
};

__device__ __forceinline__ f4 ld(const float4& v) { return mk(v.x, v.y, v.z, v.w); }
__device__ __forceinline__ f4 ld(const yn_float4& v) { return mk(v.x, v.y, v.z, v.w); }
__device__ __forceinline__ float4 st4(f4 v) { return make_float4(v.x, v.y, v.z, v.w); }

__device__ __forceinline__ Mat load_mat(const yn_material* __restrict__ mats, int id)
{
    const yn_material& m = mats[id];
    Mat r;
    r.ke = ld(m.ke); r.kd = ld(m.kd); r.ks = ld(m.ks);
    r.n = m.n; r.px = m.px; r.py = m.py; r.alpha_x = m.alpha_x;
    r.spec = m.is_specular; r.trans = m.is_transmissive;
    return r;
}

// Materials and light quads staged in LDS by the shade kernels (typed, so
// that their reads are LDS reads, not generic loads through the vector
// memory pipeline).
typedef __attribute__((address_space(3))) const yn_material lds_mat;
typedef __attribute__((address_space(3))) const yn_light_quad lds_light;
__device__ __forceinline__ f4 ld(__attribute__((address_space(3))) const yn_float4& v)
{
    return mk(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ Mat load_mat(lds_mat* mats, int id)
{
    lds_mat& m = mats[id];
    Mat r;
    r.ke = ld(m.ke); r.kd = ld(m.kd); r.ks = ld(m.ks);
    r.n = m.n; r.px = m.px; r.py = m.py; r.alpha_x = m.alpha_x;
    r.spec = m.is_specular; r.trans = m.is_transmissive;
    return r;
}

// What a shade kernel staged in LDS (see lds_mode in its arguments).
struct LdsScene {
    lds_mat* mats;
    lds_light* lights;
};

// The light quad of scene `sid`: from the scene table (LM 0), the one
// staged scene's (1), or the staged table of every scene slot's (2).
template <int LM>
__device__ __forceinline__ auto& light_of(const yn_light_quad& scene_light, const LdsScene& ls, int sid)
{
    if constexpr (LM == 0) return scene_light;
    else return ls.lights[LM == 1 ? 0 : sid];
}

// reflect (udpt.cl:949-961)
__device__ __forceinline__ f4 reflect_pt(f4 w_i, f4 normal)
{
    if (dot(w_i, normal) < 0) normal = normal * -1.0f;
    f4 refl = normal * (2 * (dot(w_i, normal))) - w_i;
    return normalize(refl);
}
// bdpt.cl:722 / 832: reflection about the unflipped normal, not normalised
__device__ __forceinline__ f4 reflect_raw(f4 w_i, f4 normal) { return normal * (2 * dot(w_i, normal)) - w_i; }

// orthonormal frame around Nz (udpt.cl:850-856)
__device__ __forceinline__ void frame_of(f4 Nz, f4& Nx, f4& Ny)
{
    if (fabsf(Nz.y) > fabsf(Nz.z)) Nx = mk(Nz.y, -Nz.x, 0, 0.f);
    else Nx = mk(Nz.z, 0, -Nz.x, 0.f);
    Nx = normalize(Nx);
    Ny = normalize(cross(Nz, Nx));
}

// normal_to_world * local, then normalize (udpt.cl:858-904)
__device__ __forceinline__ f4 to_world(f4 Nx, f4 Ny, f4 Nz, f4 p, f4 local)
{
    f4 r1 = mk(Nx.x, Ny.x, Nz.x, p.x), r2 = mk(Nx.y, Ny.y, Nz.y, p.y), r3 = mk(Nx.z, Ny.z, Nz.z, p.z);
    return normalize(mk(dot(r1, local), dot(r2, local), dot(r3, local), 0));
}

// toShadingSpace + Oren-Nayar (udpt-primitives.cl:696-725, 1147-1210)
__device__ __forceinline__ f4 to_shading(f4 w, f4 p, f4 n)
{
    f4 Nx, Ny;
    frame_of(n, Nx, Ny);
    f4 r1 = mk(Nx.x, Nx.y, Nx.z, dot(p, Nx));
    f4 r2 = mk(Ny.x, Ny.y, Ny.z, dot(p, Ny));
    f4 r3 = mk(n.x, n.y, n.z, dot(p, n));
    f4 r4 = mk(0.f, 0.f, 0.f, 1.0f);
    return mk(dot(r1, w), dot(r2, w), dot(r3, w), dot(r4, w));
}
__device__ __forceinline__ float sin_theta(f4 w) { return sqrtf(1 - w.z * w.z); }
__device__ __forceinline__ float cos_phi(f4 w)
{
    float s = sin_theta(w);
    if (s <= kEps && s >= -kEps) return 0;
    return fminf(fmaxf(w.x / s, -1.0f), 1.0f);
}
__device__ __forceinline__ float sin_phi(f4 w)
{
    float s = sin_theta(w);
    if (s <= kEps && s >= -kEps) return 0;
    return fminf(fmaxf(w.y / s, -1.0f), 1.0f);
}
__device__ __forceinline__ f4 oren_nayar(const Mat& m, f4 w_i, f4 w_o, f4 p, f4 n)
{
    w_i = to_shading(w_i, p, n);
    w_o = to_shading(w_o, p, n);
    float costerm = 0.0f;
    if (sin_theta(w_i) >= kEps && sin_theta(w_o) >= kEps)
        costerm = fmaxf(0.0f, ((cos_phi(w_i) * cos_phi(w_o)) + (sin_phi(w_i) * sin_phi(w_o))));
    float sin_alpha, tan_beta;
    if (fabsf(w_i.z) < fabsf(w_o.z)) { sin_alpha = sin_theta(w_i); tan_beta = sin_theta(w_o) / fabsf(w_o.z); }
    else { sin_alpha = sin_theta(w_o); tan_beta = sin_theta(w_i) / fabsf(w_i.z); }
    float s2 = m.alpha_x;
    float A = (float)(1 - (s2 / (2 * (s2 + 0.33))));
    float B = (float)(0.45 * s2 / (s2 + 0.09));
    return (m.kd * kInvPI) * (A + B * (costerm * sin_alpha * tan_beta));
}

// evaluateBRDF (udpt.cl:611-630; bdpt.cl:718-737; ON per udpt-primitives.cl:681-687)
__device__ __forceinline__ f4 eval_brdf(const Mat& m, f4 w_i, f4 w_o, f4 p, f4 n, bool glossy, float rr_prob,
                                        bool on, bool bdpt)
{
    f4 refl = bdpt ? reflect_raw(w_i, n) : reflect_pt(w_i, n);
    refl = normalize(refl);
    if (!glossy) {
        if (on && rr_prob == 1.0f) return oren_nayar(m, w_i, w_o, p, n);
        return (m.kd * kInvPI) / rr_prob;
    }
    float cos_alpha = ynf_powf(fmaxf(dot(w_o, refl), 0.0f), m.px + m.py);
    int phong_exp = (int)(m.px + m.py);
    return ((((m.ks * cos_alpha) * (float)(phong_exp + 2)) * kInvPI) * 0.5f) / rr_prob;
}

// sampleGlossyPdf (udpt.cl:1062-1119; bdpt.cl:1048-1105)
__device__ __forceinline__ bool sample_glossy_pdf(const Mat& m, uint32_t& seed, float& prob, bool bdpt)
{
    float r = next_rand(seed);
    if (length3(m.ks) == 0.0f) { prob = 1.0f; return false; }
    else if (length3(m.kd) == 0.0f || m.kd.x + m.kd.y + m.kd.z == 0.0f) { prob = 1.0f; return true; }
    f4 sum = m.ks + m.kd;
    float max_val = fmaxf(sum.x, fmaxf(sum.y, sum.z));
    float pd, ps;
    if (max_val == sum.x) { pd = m.kd.x; ps = m.ks.x; }
    else if (max_val == sum.y) { pd = m.kd.y; ps = m.ks.y; }
    else { pd = m.kd.z; ps = m.ks.z; }
    if (bdpt) {
        if (max_val < 1.0f) { pd += (1 - max_val) / 2.0f; ps += (1 - max_val) / 2.0f; }
        if (r < pd) { prob = pd; return false; }
        prob = ps; return true;
    }
    if (r < pd) { prob = pd; return false; }
    else if (r < pd + ps && r >= pd) { prob = ps; return true; }
    prob = 0; return false;
}

// phongSampleHemisphere (udpt.cl:700-771; bdpt.cl variant)
__device__ __forceinline__ f4 phong_sample(const Mat& m, float& pdf, f4 w_i, f4 p, f4 n, uint32_t& seed, bool bdpt)
{
    f4 Nz = bdpt ? reflect_raw(w_i, n) : reflect_pt(w_i, n);
    Nz = normalize(Nz);
    f4 Nx, Ny;
    frame_of(Nz, Nx, Ny);
    float r1 = next_rand(seed);
    float r2 = next_rand(seed);
    int pe = (int)(m.px + m.py);
    float phi = 2 * kPI * r2;
    float costheta = ynf_powf(r1, 1.0f / (float)(pe + 1));
    float sintheta = 1 - ynf_powf(r1, 2.0f / (float)(pe + 1));
    sintheta = sqrtf(sintheta);
    f4 d = to_world(Nx, Ny, Nz, p, mk(sintheta * ynf_cosf(phi), sintheta * ynf_sinf(phi), costheta, 0));
    if (dot(d, n) < 0) pdf = 0;
    else pdf = (float)((double)(pe + 1) * 0.5 * (double)kInvPI * (double)ynf_powf(costheta, (float)pe));
    return d;
}

// cosineWeightedHemisphere (udpt.cl:843-910)
__device__ __forceinline__ f4 cosine_sample(float& pdf, f4 p, f4 n, uint32_t& seed)
{
    f4 Nx, Ny;
    frame_of(n, Nx, Ny);
    float r1 = next_rand(seed);
    float r2 = next_rand(seed);
    float phi = 2 * kPI * r2;
    float sinTheta = sqrtf(r1);
    float z = sqrtf(1 - r1);
    f4 d = to_world(Nx, Ny, n, p, mk(sinTheta * ynf_cosf(phi), sinTheta * ynf_sinf(phi), z, 0));
    pdf = z * kInvPI;
    return d;
}

// refract (udpt.cl:963-990)
__device__ __forceinline__ f4 refract_dir(const Mat& m, f4 w_i, f4 normal)
{
    float n1, n2;
    if (dot(w_i, normal) < 0) { n1 = m.n; n2 = 1; normal = normal * -1.0f; }
    else { n1 = 1; n2 = m.n; }
    f4 wt_perp = (normal * dot(w_i, normal) - w_i) * (n1 / n2);
    float lp = length(wt_perp);
    f4 wt_par = (-normal) * sqrtf(1 - lp * lp);
    return normalize(wt_perp + wt_par);
}

// evalFresnelReflectance (udpt.cl:992-1031)
__device__ __forceinline__ float fresnel(const Mat& m, f4 w_i, f4 normal, float& ior_factor)
{
    float n1, n2;
    if (dot(w_i, normal) < 0) { n1 = m.n; n2 = 1; normal = normal * -1.0f; }
    else { n1 = 1; n2 = m.n; }
    float cI = dot(w_i, normal);
    float sI = sqrtf(1 - cI * cI);
    float sT = n1 * sI / n2;
    float cT = sqrtf(1 - sT * sT);
    if (sT >= 1.0f && n1 > n2) return 1.0f;
    float r0 = m.ks.x + m.ks.y + m.ks.z;
    r0 /= 3.0f;
    ior_factor = (n2 * n2) / (n1 * n1);
    if (n1 > n2) return (r0 + (1 - r0) * (1 - ynf_powf(cI, 5.0f)));
    return (r0 + (1 - r0) * (1 - ynf_powf(cT, 5.0f)));
}

// sampleFresnelIncidence (udpt.cl:912-947): returns the new direction.
__device__ __forceinline__ f4 fresnel_sample(const Mat& m, f4 n, f4 w_i, float& ior_factor, uint32_t& seed)
{
    if (m.trans) {
        float pdf = fresnel(m, w_i, n, ior_factor);
        if (pdf == 1.0f) { ior_factor = 1.0f; return reflect_pt(w_i, n); }
        float r = next_rand(seed);
        if (r < pdf) { ior_factor = 1.0f; return reflect_pt(w_i, n); }
        return refract_dir(m, w_i, n);
    }
    ior_factor = 1.0f;
    return reflect_pt(w_i, n);
}

// calcPhongPDF (udpt.cl:1033-1045), calcCosPDF (:1047-1050)
__device__ __forceinline__ float phong_pdf(const Mat& m, f4 w_i, f4 w_o, f4 n)
{
    f4 refl = normalize(n * (2 * (dot(w_o, n))) - w_o);
    float costheta = fmaxf(0.0f, ynf_cosf(dot(refl, w_i)));
    float pe = m.px + m.py;
    return (float)((double)(pe + 1) * 0.5 * (double)kInvPI * (double)ynf_powf(costheta, pe));
}
__device__ __forceinline__ float cos_pdf(f4 w_i, f4 n) { return fmaxf(dot(w_i, n), 0.0f) * kInvPI; }
__device__ __forceinline__ float power_heuristic(float w, float p1, float p2) { return (w * w) / (p1 * p1 + p2 * p2); }

// sampleLights for LIGHT_SIZE == 1 (udpt.cl:632-667).  Returns false for
// the j == -1 case.  wi_raw: un-normalised point-to-light vector.
template <typename LQ>
__device__ __forceinline__ bool sample_lights(const LQ& L, f4 p, f4 n, uint32_t& seed, float& light_pdf, f4& wi_raw)
{
    f4 lpos = ld(L.pos), el = ld(L.edge_l), ew = ld(L.edge_w), ln = ld(L.normal);
    float r1 = next_rand(seed);
    float r2 = next_rand(seed);
    f4 temp = ((lpos + el * r1) + ew * r2) - p;
    float distance = dot(temp, temp);
    f4 wis = temp;
    temp = normalize(temp);
    float cf = fmaxf(dot(temp, n), 0.0f) * fmaxf(dot(-temp, ln), 0.0f);
    if (cf <= 0.0f) return false;
    float area = length(el) * length(ew);
    wi_raw = wis;
    light_pdf = 1 / area;
    light_pdf *= distance / fmaxf(dot(-temp, ln), 0.0f);
    return true;
}

// Quad light test of traceRay (udpt.cl:244-276).  On a hit returns true and
// sets t (the new ray length).
template <typename LQ>
__device__ __forceinline__ bool quad_hit(const LQ& L, f4 o, f4 d, float len, float& t)
{
    f4 n = ld(L.normal), pos = ld(L.pos), el = ld(L.edge_l), ew = ld(L.edge_w);
    float DdotN = dot(d, n);
    if ((double)fabsf(DdotN) > 0.0001) {
        float tt = dot(n, pos - o) / DdotN;
        if (tt > 0.0f && tt < len) {
            f4 temp = o + d * tt;
            temp = temp - pos;
            float proj1 = dot(temp, el), proj2 = dot(temp, ew);
            float la = length(el), lb = length(ew);
            proj1 /= la;
            proj2 /= lb;
            if ((proj1 >= 0.0f && proj2 >= 0.0f) && (proj1 <= la && proj2 <= lb)) { t = tt; return true; }
        }
    }
    return false;
}

// createRay (udpt.cl:213-238), image plane in double as the literals make it.
__device__ __forceinline__ void create_ray(float px, float py, int W, int H, const yn_cam& cam, f4& o, f4& d)
{
    float aspect = (float)((W * 1.0) / H);
    f4 dir;
    dir.x = (float)((double)aspect * ((2.0 * px / W) - 1));
    dir.y = (float)((2.0 * py / H) - 1);
    dir.z = -cam.view_plane_dist;
    dir.w = 0;
    d = mk(dot(ld(cam.r1), dir), dot(ld(cam.r2), dir), dot(ld(cam.r3), dir), dot(ld(cam.r4), dir));
    d = normalize(d);
    o = mk(cam.r1.w, cam.r2.w, cam.r3.w, cam.r4.w);
}

// Shading normal at (u, v) of triangle `tri` (udpt.cl:377-383).
__device__ __forceinline__ f4 shade_normal(const float4* __restrict__ tri_shade, int tri, float u, float v)
{
    f4 N1 = ld(tri_shade[4 * tri]), N2 = ld(tri_shade[4 * tri + 1]), N3 = ld(tri_shade[4 * tri + 2]);
    float w = 1 - u - v;
    return normalize((N1 * w + N2 * u) + N3 * v);
}

// Material id of triangle `tri`, from the same 64-byte record as its normals.
__device__ __forceinline__ int tri_material(const float4* __restrict__ tri_shade, int tri)
{
    return __float_as_int(tri_shade[4 * tri + 3].x);
}

}  // namespace yune
