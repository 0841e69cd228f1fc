// Host check of the traversal's FMA slab form (yune_amd/csrc/yn_slab.h)
// against the reference's arithmetic (box_hit, rayAabbIntersection,
// udpt.cl:392-431), on random and adversarial rays and boxes:
//   1. every box box_hit enters, box_hit_fma enters;
//   2. a node the reference would not cull (entry <= best * 1.0001) is not
//      culled by the FMA form (entry <= cull_of(best));
//   3. the FMA form enters few boxes beyond those: among boxes like a BVH's
//      (no flat axis: a triangle's flat extent is widened by 0.2; each
//      extent at least 2^-10 of the box's distance plus one; a room-sized
//      origin not on one of the box's faces), at most 1 in 1000 of the
//      tested.  Boxes far smaller than their distance, far-out origins and
//      a ray leaving a box right at its origin are entered more often: the
//      widening is relative to the distance and to |o / d|.
// Built with -ffp-contract=off, like the kernels, so fmaf / fminf / fmaxf
// and every product and sum round exactly as on the GPU.
// Usage: slab_check <cases> <seed>; prints the counts, exits 1 on a miss.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "yn_slab.h"

using namespace yune;

namespace {

struct Rng {
    uint64_t s;
    uint64_t next()
    {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return s;
    }
    double uni() { return (double)(next() >> 11) * 0x1p-53; }
    float sym(float a) { return (float)((2.0 * uni() - 1.0) * a); }
    // |x| log-uniform in [2^lo, 2^hi], random sign
    float logu(int lo, int hi)
    {
        double e = lo + (hi - lo) * uni();
        float v = (float)std::exp2(e);
        return (next() & 1) ? -v : v;
    }
};

}  // namespace

int main(int argc, char** argv)
{
    const long n = argc > 1 ? std::atol(argv[1]) : 1000000;
    Rng g{argc > 2 ? (uint64_t)std::atoll(argv[2]) * 0x9E3779B97F4A7C15ull + 1 : 88172645463325252ull};
    long tested = 0, ref_hits = 0, fma_extra = 0, misses = 0, cull_tests = 0, cull_misses = 0;
    long solid = 0, solid_extra = 0;
    for (long c = 0; c < n; ++c) {
        // origin: a room-sized scene, or far out (up to the FMA form's 2^40)
        float o[3], d[3];
        const int okind = (int)(g.next() % 4);
        for (int k = 0; k < 3; ++k)
            o[k] = okind == 0 ? g.logu(-20, 40) : g.sym(okind == 1 ? 10.0f : 1000.0f);
        // direction: random, with components often made tiny (grazing rays)
        double dd[3], len = 0;
        for (int k = 0; k < 3; ++k) { dd[k] = 2 * g.uni() - 1; len += dd[k] * dd[k]; }
        len = std::sqrt(len);
        for (int k = 0; k < 3; ++k) {
            d[k] = (float)(dd[k] / len);
            if (g.next() % 4 == 0) d[k] *= (float)std::exp2(-40.0 * g.uni());
        }
        const float ix = 1 / d[0], iy = 1 / d[1], iz = 1 / d[2];
        if (!fma_form_ok(o[0], o[1], o[2], ix, iy, iz)) continue;   // those rays use box_hit itself
        // box: around a point along the ray (or off it), extents from tiny
        // to large, sometimes flat, sometimes with a face through the origin
        const float t0 = g.logu(-10, 12) * (g.next() % 8 == 0 ? -1.0f : 1.0f);
        float lo[3], hi[3];
        bool on_face = false;
        for (int k = 0; k < 3; ++k) {
            float p = o[k] + d[k] * std::fabs(t0);
            if (g.next() % 3 == 0) p += g.logu(-20, 10);
            float ext = g.next() % 6 == 0 ? 0.0f : std::fabs(g.logu(-24, 12));
            float a = p - ext * (float)g.uni(), b = p + ext * (float)g.uni();
            if (g.next() % 16 == 0) { a = o[k]; on_face = true; }
            if (g.next() % 16 == 0) { b = o[k]; on_face = true; }
            lo[k] = std::fmin(a, b);
            hi[k] = std::fmax(a, b);
        }
        float lx, hx, ly, hy, lz, hz;
        fma_offsets(o[0], ix, lx, hx);
        fma_offsets(o[1], iy, ly, hy);
        fma_offsets(o[2], iz, lz, hz);
        float e_ref, e_fma;
        const bool ref = box_hit(lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], o[0], o[1], o[2], ix, iy, iz, e_ref);
        const bool fma = box_hit_fma(lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], ix, iy, iz, lx, ly, lz, hx, hy, hz,
                                     e_fma);
        ++tested;
        const float small = 0x1p-10f * (std::fabs(t0) + 1.0f);
        const bool bvh_like = okind != 0 && !on_face && hi[0] - lo[0] >= small && hi[1] - lo[1] >= small &&
                              hi[2] - lo[2] >= small;
        if (bvh_like) ++solid;
        if (ref) ++ref_hits;
        if (fma && !ref) {
            ++fma_extra;
            if (bvh_like) ++solid_extra;
        }
        if (ref && !fma) {
            if (misses < 10)
                std::printf("MISS o=(%a %a %a) d=(%a %a %a) lo=(%a %a %a) hi=(%a %a %a)\n", o[0], o[1], o[2],
                            d[0], d[1], d[2], lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
            ++misses;
            continue;
        }
        if (!ref) continue;
        // culling against best hits just around the box's entry
        for (int q = 0; q < 4; ++q) {
            const float best = e_ref * (1.0f + g.sym(4e-4f)) + (q == 0 ? 0.0f : g.sym(1e-30f));
            if (!(best > 0.0f) || !std::isfinite(best)) continue;
            ++cull_tests;
            if (e_ref <= best * 1.0001f && !(e_fma <= cull_of(best))) {
                if (cull_misses < 10)
                    std::printf("CULL best=%a e_ref=%a e_fma=%a\n", best, e_ref, e_fma);
                ++cull_misses;
            }
        }
    }
    std::printf("tested %ld boxes, %ld entered by the reference, %ld more by the FMA form (%ld of %ld like a "
                "BVH's), %ld missed; %ld cull tests, %ld missed\n",
                tested, ref_hits, fma_extra, solid_extra, solid, misses, cull_tests, cull_misses);
    return (misses || cull_misses || solid_extra * 1000 > solid) ? 1 : 0;
}
