// yn_slab.h — the ray / box slab tests of the traversal, for the device
// kernels and for host-side checks of the one against the other.
//
// box_hit is rayAabbIntersection (udpt.cl:392-431) in the reference's
// arithmetic.  box_hit_fma is the traversal's faster form; it must enter
// every box box_hit enters, which tests/slab_check.cpp checks on the CPU
// with this same code.
#ifndef YN_SLAB_H
#define YN_SLAB_H

#include <math.h>

#if defined(__HIPCC__)
#define YN_SL __host__ __device__ __forceinline__
#else
#define YN_SL inline
#endif

namespace yune {

// One slab of rayAabbIntersection (udpt.cl:397-418).  min_diff can only be
// NaN when dir_inv is infinite (a zero or denormal direction component) and
// the plane passes through the origin, where the reference skips the slab.
YN_SL void slab(float lo, float hi, float o, float inv, float& tmin, float& tmax)
{
    float a = (lo - o) * inv;           // min_diff (udpt.cl:397)
    float b = (hi - o) * inv;           // max_diff (udpt.cl:398)
    if (!(a != a)) {                    // !isnan(min_diff)
        tmin = fmaxf(fminf(a, b), tmin);
        tmax = fminf(fmaxf(a, b), tmax);
    }
}

// rayAabbIntersection (udpt.cl:392-431) + entry distance for culling.
YN_SL bool box_hit(float lox, float loy, float loz, float hix, float hiy, float hiz, float ox, float oy, float oz,
                   float ix, float iy, float iz, float& enter)
{
    float tmin = -INFINITY, tmax = INFINITY;
    slab(lox, hix, ox, ix, tmin, tmax);
    slab(loy, hiy, oy, iy, tmin, tmax);
    slab(loz, hiz, oz, iz, tmin, tmax);
    enter = fmaxf(tmin, 0.0f);
    return tmax > enter;
}

constexpr float kWiden = 1.0f + 0x1p-19f;   // relative widening of the FMA slab form

// The cull distance for a best hit t: the margin 1e-4 the traversal always
// had, widened as the FMA slab form's exit distance is (below), so that a
// node the reference's arithmetic would enter is never culled by it.
YN_SL float cull_of(float t) { return t * 1.0001f * kWiden; }

// The offsets the FMA form subtracts on one axis: o * dir_inv moved by
// 8u |o * dir_inv| (u = 2^-24; the scaling by 2^-21 is exact), towards
// smaller distances for the plane the ray enters through (lo when
// dir_inv > 0) and larger ones for the plane it leaves through.
YN_SL void fma_offsets(float o, float inv, float& l, float& h)
{
    const float oi = o * inv, e = fabsf(oi) * 0x1p-21f;
    l = inv > 0.0f ? oi + e : oi - e;
    h = inv > 0.0f ? oi - e : oi + e;
}

// Whether the FMA form may stand in for box_hit: dir_inv and o within
// 2^40 (and, checked per scene on the host, box coordinates finite and
// within 2^60), where no product or difference can overflow.
YN_SL bool fma_form_ok(float ox, float oy, float oz, float ix, float iy, float iz)
{
    const float lim = 0x1p40f;
    return fabsf(ix) <= lim && fabsf(iy) <= lim && fabsf(iz) <= lim && fabsf(ox) <= lim && fabsf(oy) <= lim &&
           fabsf(oz) <= lim;
}

// The slab test with the offsets above: each slab distance is one FMA,
// plane * dir_inv - offset, where the reference takes a difference and a
// product.  Rounding o * dir_inv costs up to u|o * dir_inv| in absolute
// terms, which the offsets' move covers together with its own rounding.
// What remains is relative: each entry distance is at most 4u of its own
// size above the reference's, each exit distance at most 4u below; min and
// max are monotone, so the same holds for the box's entry and exit.  The
// exit distance is therefore widened by kWiden = 1 + 32u before the
// comparison, with >= where the reference has >, and cull distances
// likewise (cull_of): every box the reference's arithmetic enters is
// entered here.
YN_SL bool box_hit_fma(float lox, float loy, float loz, float hix, float hiy, float hiz, float ix, float iy,
                       float iz, float lx, float ly, float lz, float hx, float hy, float hz, float& enter)
{
    const float ax = fmaf(lox, ix, -lx), bx = fmaf(hix, ix, -hx);
    const float ay = fmaf(loy, iy, -ly), by = fmaf(hiy, iy, -hy);
    const float az = fmaf(loz, iz, -lz), bz = fmaf(hiz, iz, -hz);
    const float tmin = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
    const float tmax = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
    enter = fmaxf(tmin, 0.0f);
    return tmax * kWiden >= enter;
}

}  // namespace yune

#endif  // YN_SLAB_H
