/*
 * yn_fmath.h — sin, cos and pow for the path-tracing kernels.
 *
 * OpenCL C leaves the exact results of its sin, cos and pow built-ins to
 * the runtime (within a few ulp), so the reference kernels have no single
 * bit-exact answer for them.  This library fixes one definition of each:
 * the argument is widened to IEEE double, the function is evaluated with a
 * fixed sequence of double operations (build with -ffp-contract=off and
 * without fast-math) and the result is rounded once to float.  The same
 * code runs on the GPU and on the host, so results are identical bit for
 * bit wherever they are computed; the error is below 1 ulp of the float
 * result.
 *
 * Methods (textbook): sin/cos by Cody-Waite reduction modulo pi/2 with a
 * two-part pi/2 and Taylor polynomials on [-pi/4, pi/4] (truncation below
 * 1e-16); log by m = x * 2^-e in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1) and
 * the series log m = 2 atanh s; exp by reduction modulo ln 2 (two-part ln 2)
 * and a Taylor polynomial on [-ln2/2, ln2/2]; pow(x, y) = exp(y log x) with
 * the C99 special cases.
 *
 * Plain C, usable from C, C++ and HIP device code.
 */
#ifndef YN_FMATH_H
#define YN_FMATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define YN_FM __host__ __device__ static inline
#else
#define YN_FM static inline
#endif

#define YNF_PIO2_HI 1.5707963267341256           /* pi/2 to 33 bits: k * HI is exact for |k| < 2^20 */
#define YNF_PIO2_LO 6.077100506506192e-11        /* pi/2 - HI */
#define YNF_TWO_OVER_PI 0.6366197723675814
#define YNF_LN2_HI 0.6931471806019545            /* ln 2 to 32 bits */
#define YNF_LN2_LO -4.2009150726810846e-11       /* ln 2 - HI */
#define YNF_INV_LN2 1.4426950408889634
#define YNF_SQRT2 1.4142135623730951

/* Polynomial and reduction constants.  In device code each is made where
 * it is used, by two scalar moves the compiler may not hoist: otherwise it
 * lifts every coefficient to the kernel's entry, and in the shade kernels,
 * which hold many live values, parks them in spill lanes and reads them
 * back around each use.  The bits are those of the literal either way. */
#if defined(__HIP_DEVICE_COMPILE__)
template <unsigned long long B>
__device__ static inline double ynf_kk()
{
    unsigned lo, hi;
    __asm__ volatile("s_mov_b32 %0, %1" : "=s"(lo) : "i"((unsigned)(B & 0xffffffffull)));
    __asm__ volatile("s_mov_b32 %0, %1" : "=s"(hi) : "i"((unsigned)(B >> 32)));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
#define YNF_K(d) (ynf_kk<__builtin_bit_cast(unsigned long long, (double)(d))>())
#else
#define YNF_K(d) (d)
#endif

YN_FM uint64_t ynf_bits(double x)
{
    uint64_t u;
    __builtin_memcpy(&u, &x, sizeof u);
    return u;
}

YN_FM double ynf_from_bits(uint64_t u)
{
    double x;
    __builtin_memcpy(&x, &u, sizeof x);
    return x;
}

/* round to nearest, ties to even, for |x| < 2^51 (adding and removing
 * 1.5 * 2^52 leaves no fraction bits) */
YN_FM double ynf_rint(double x)
{
    const double big = YNF_K(6755399441055744.0);
    double t = x + big;
    return t - big;
}

/* sin r and cos r for |r| <= pi/4 + tiny */
YN_FM double ynf_sin_k(double r)
{
    double z = r * r;
    double p = YNF_K(-7.647163731819816e-13);            /* -1/15! */
    p = p * z + YNF_K(1.6059043836821613e-10);           /*  1/13! */
    p = p * z + YNF_K(-2.505210838544172e-08);           /* -1/11! */
    p = p * z + YNF_K(2.7557319223985893e-06);           /*  1/9!  */
    p = p * z + YNF_K(-0.0001984126984126984);           /* -1/7!  */
    p = p * z + YNF_K(0.008333333333333333);             /*  1/5!  */
    p = p * z + YNF_K(-0.16666666666666666);             /* -1/3!  */
    return r + (r * z) * p;
}

YN_FM double ynf_cos_k(double r)
{
    double z = r * r;
    double p = YNF_K(4.779477332387385e-14);             /*  1/16! */
    p = p * z + YNF_K(-1.1470745597729725e-11);          /* -1/14! */
    p = p * z + YNF_K(2.08767569878681e-09);             /*  1/12! */
    p = p * z + YNF_K(-2.755731922398589e-07);           /* -1/10! */
    p = p * z + YNF_K(2.48015873015873e-05);             /*  1/8!  */
    p = p * z + YNF_K(-0.001388888888888889);            /* -1/6!  */
    p = p * z + YNF_K(0.041666666666666664);             /*  1/4!  */
    p = p * z + -0.5;                             /* -1/2!  */
    return 1.0 + z * p;
}

/* x = k pi/2 + r; returns r, quadrant k mod 4 in *q */
YN_FM double ynf_reduce(double x, int* q)
{
    double k = ynf_rint(x * YNF_K(YNF_TWO_OVER_PI));
    double r = (x - k * YNF_K(YNF_PIO2_HI)) - k * YNF_K(YNF_PIO2_LO);
    *q = (int)((long long)k & 3);
    return r;
}

YN_FM float ynf_sinf(float xf)
{
    if (xf != xf || xf - xf != 0.0f) return __builtin_nanf("");   /* NaN, +-inf */
    if (xf == 0.0f) return xf;                    /* sin(+-0) = +-0 */
    int q;
    double r = ynf_reduce((double)xf, &q);
    double v = (q & 1) ? ynf_cos_k(r) : ynf_sin_k(r);
    return (float)((q & 2) ? -v : v);
}

YN_FM float ynf_cosf(float xf)
{
    if (xf != xf || xf - xf != 0.0f) return __builtin_nanf("");
    int q;
    double r = ynf_reduce((double)xf, &q);
    double v = (q & 1) ? ynf_sin_k(r) : ynf_cos_k(r);
    return (float)((q == 1 || q == 2) ? -v : v);
}

/* natural log of a positive, finite, normal double */
YN_FM double ynf_log(double x)
{
    uint64_t b = ynf_bits(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    double m = ynf_from_bits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);   /* [1, 2) */
    if (m > YNF_K(YNF_SQRT2)) {
        m = m * 0.5;
        e += 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
    double p = YNF_K(0.047619047619047616);              /* 1/21 */
    p = p * z + YNF_K(0.05263157894736842);              /* 1/19 */
    p = p * z + YNF_K(0.058823529411764705);             /* 1/17 */
    p = p * z + YNF_K(0.06666666666666667);              /* 1/15 */
    p = p * z + YNF_K(0.07692307692307693);              /* 1/13 */
    p = p * z + YNF_K(0.09090909090909091);              /* 1/11 */
    p = p * z + YNF_K(0.1111111111111111);               /* 1/9  */
    p = p * z + YNF_K(0.14285714285714285);              /* 1/7  */
    p = p * z + YNF_K(0.2);                              /* 1/5  */
    p = p * z + YNF_K(0.3333333333333333);               /* 1/3  */
    p = p * z + 1.0;
    double lm = 2.0 * (s * p);
    double de = (double)e;
    return de * YNF_K(YNF_LN2_HI) + (lm + de * YNF_K(YNF_LN2_LO));
}

/* e^z for |z| <= 110 */
YN_FM double ynf_exp(double z)
{
    double k = ynf_rint(z * YNF_K(YNF_INV_LN2));
    double r = (z - k * YNF_K(YNF_LN2_HI)) - k * YNF_K(YNF_LN2_LO);
    double p = YNF_K(1.6059043836821613e-10);            /* 1/13! */
    p = p * r + YNF_K(2.08767569878681e-09);             /* 1/12! */
    p = p * r + YNF_K(2.505210838544172e-08);            /* 1/11! */
    p = p * r + YNF_K(2.755731922398589e-07);            /* 1/10! */
    p = p * r + YNF_K(2.7557319223985893e-06);           /* 1/9!  */
    p = p * r + YNF_K(2.48015873015873e-05);             /* 1/8!  */
    p = p * r + YNF_K(0.0001984126984126984);            /* 1/7!  */
    p = p * r + YNF_K(0.001388888888888889);             /* 1/6!  */
    p = p * r + YNF_K(0.008333333333333333);             /* 1/5!  */
    p = p * r + YNF_K(0.041666666666666664);             /* 1/4!  */
    p = p * r + YNF_K(0.16666666666666666);              /* 1/3!  */
    p = p * r + 0.5;                              /* 1/2!  */
    p = p * r + 1.0;
    p = p * r + 1.0;
    /* scale by 2^k: p in [0.7, 1.5] and |k| <= 160, so the result stays a
     * normal double and the exponent add is exact */
    return ynf_from_bits(ynf_bits(p) + ((uint64_t)(int64_t)(long long)k << 52));
}

/* powf with the C99 special cases (Annex F.9.4.4) */
YN_FM float ynf_powf(float x, float y)
{
    if (y == 0.0f) return 1.0f;
    if (x == 1.0f) return 1.0f;
    if (x != x || y != y) return __builtin_nanf("");
    const float ay = y < 0.0f ? -y : y;
    const int y_int = ay >= 16777216.0f || (float)(long long)ay == ay;
    const int y_odd = ay < 16777216.0f && y_int && ((long long)ay & 1);
    const float inf = __builtin_inff();
    if (y - y != 0.0f) {                          /* y = +-inf */
        float ax = x < 0.0f ? -x : x;
        if (ax == 1.0f) return 1.0f;
        return ((ax > 1.0f) == (y > 0.0f)) ? inf : 0.0f;
    }
    float sign = 1.0f;
    if (x < 0.0f) {
        if (!y_int) return __builtin_nanf("");
        if (y_odd) sign = -1.0f;
        x = -x;
    }
    if (x == 0.0f) {                              /* +-0 (sign of x kept for odd y) */
        float r = y > 0.0f ? 0.0f : inf;
        return (y_odd && (sign < 0.0f || 1.0f / x < 0.0f)) ? -r : r;
    }
    if (x - x != 0.0f) return sign * (y > 0.0f ? inf : 0.0f);     /* x = inf */
    double t = (double)y * ynf_log((double)x);
    if (t > 89.0) return sign * inf;              /* beyond FLT_MAX */
    if (t < -104.0) return sign * 0.0f;           /* below half the smallest denormal */
    return sign * (float)ynf_exp(t);
}

#endif /* YN_FMATH_H */
