// yn_math.h — float4 arithmetic with the OpenCL C semantics the reference
// kernels are written in (all four lanes take part in +,-,*,/ and dot).
//
// One fixed definition per OpenCL built-in, shared by host code and HIP
// device code and compiled with -ffp-contract=off everywhere:
//   dot(a,b)     = ((a.x*b.x + a.y*b.y) + a.z*b.z) + a.w*b.w
//   length(a)    = sqrt(dot(a,a))
//   normalize(a) = a * (1 / sqrt(dot(a,a)))
//   cross(a,b)   = 3-D cross product, w = 0
//   min/max      = fminf/fmaxf
// sqrt and division are IEEE correctly rounded on both sides (hipcc default),
// so every expression built only from + - * / sqrt rounds identically on the
// CPU and on gfx950.  pow/sin/cos come from each side's math library.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define YN_HD __host__ __device__ __forceinline__
#else
#define YN_HD inline
#endif

namespace yune {

struct f4 {
    float x, y, z, w;
};

YN_HD f4 mk(float x, float y, float z, float w) { f4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
YN_HD f4 bc(float s) { return mk(s, s, s, s); }
YN_HD f4 operator+(f4 a, f4 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
YN_HD f4 operator-(f4 a, f4 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
YN_HD f4 operator*(f4 a, f4 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
YN_HD f4 operator*(f4 a, float s) { return mk(a.x * s, a.y * s, a.z * s, a.w * s); }
YN_HD f4 operator/(f4 a, float s) { return mk(a.x / s, a.y / s, a.z / s, a.w / s); }
YN_HD f4 operator-(f4 a) { return mk(-a.x, -a.y, -a.z, -a.w); }
YN_HD float dot(f4 a, f4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
YN_HD float length(f4 a) { return sqrtf(dot(a, a)); }
YN_HD float length3(f4 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
YN_HD f4 normalize(f4 a) { float inv = 1.0f / sqrtf(dot(a, a)); return a * inv; }
YN_HD f4 cross(f4 a, f4 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.0f);
}
YN_HD bool any_nan(f4 a) { return isnan(a.x) || isnan(a.y) || isnan(a.z) || isnan(a.w); }

// udpt.cl:1131-1147
YN_HD uint32_t wang_hash(uint32_t seed)
{
    seed = (seed ^ 61u) ^ (seed >> 16);
    seed *= 9u;
    seed = seed ^ (seed >> 4);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15);
    return seed;
}
YN_HD uint32_t xor_shift(uint32_t seed)
{
    seed ^= seed << 13;
    seed ^= seed >> 17;
    seed ^= seed << 5;
    return seed;
}
// `seed / (float) UINT_MAX`; (float)UINT_MAX rounds to 2^32, an exact scale.
YN_HD float u2f(uint32_t s) { return (float)s * 2.3283064365386963e-10f; }
YN_HD float next_rand(uint32_t& s) { s = xor_shift(s); return u2f(s); }

// udpt.cl:175-180
YN_HD uint32_t pixel_seed(uint32_t rand, int x, int y, int width)
{
    uint32_t seed = (uint32_t)((y + 1) * width + (x + 1));
    seed = rand * seed;
    seed = wang_hash(seed);
    if (seed == 0) seed = wang_hash(seed);
    return seed;
}

YN_HD float luminance(f4 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }

}  // namespace yune
