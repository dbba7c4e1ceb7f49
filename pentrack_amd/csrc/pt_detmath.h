/*
 * pt_detmath.h -- deterministic FP64 elementary functions and the per-particle
 * random stream used on the tracking path.
 *
 * Every function here is built only from IEEE-754 operations that are
 * correctly rounded on both the host CPU and gfx950 (+, -, *, /, sqrt,
 * comparisons, integer and bit operations), with FP contraction disabled by
 * the build.  Host code and device code therefore produce the same bits for
 * the same inputs, which is what lets an FP64 GPU trajectory be compared
 * bit-for-bit with a CPU trajectory.
 *
 * The reference (PENTrack) calls glibc libm for these functions
 * (pow in Boost.Odeint's step adjuster, log in std::exponential_distribution,
 * sin/cos/acos in the Lambert samplers of src/neutron.cpp:151-184 and
 * include/mc.h:162-175, csqrt/complex division in src/neutron.cpp:240-242).
 * The implementations below are accurate to a few ulp (checked against libm by
 * the test suite); they are not bit-identical to glibc.
 *
 * The file is plain C99 so it can be included by C, C++ and HIP translation
 * units.
 */
#ifndef PT_DETMATH_H_
#define PT_DETMATH_H_

#include <stdint.h>

#if defined(__HIPCC__)
#define PTM_FN static inline __host__ __device__
#else
#define PTM_FN static inline
#endif

/* ---- bit helpers ------------------------------------------------------- */

typedef union { double d; uint64_t u; } ptm_bits;

PTM_FN double ptm_from_bits(uint64_t u) { ptm_bits b; b.u = u; return b.d; }
PTM_FN uint64_t ptm_to_bits(double d) { ptm_bits b; b.d = d; return b.u; }

/* 2^k for integer k in the normal range [-1022, 1023] */
PTM_FN double ptm_pow2i(int k) { return ptm_from_bits((uint64_t)(k + 1023) << 52); }

PTM_FN double ptm_fabs(double x) { return ptm_from_bits(ptm_to_bits(x) & 0x7fffffffffffffffULL); }

PTM_FN double ptm_copysign(double x, double s) {
    return ptm_from_bits((ptm_to_bits(x) & 0x7fffffffffffffffULL) | (ptm_to_bits(s) & 0x8000000000000000ULL));
}

/* round-to-nearest-even integer value of x (|x| < 2^51) */
PTM_FN double ptm_rint(double x) {
    const double big = 6755399441055744.0; /* 1.5 * 2^52 */
    double y = (x + big) - big;
    return y;
}

/* ---- natural logarithm -------------------------------------------------- */
/*
 * x = 2^k * m with m in [sqrt(1/2), sqrt(2)); log(m) = 2 atanh(s),
 * s = (m - 1)/(m + 1), |s| <= 0.1716; atanh series to s^25.
 */
PTM_FN double ptm_log(double x) {
    if (!(x > 0.0)) {
        if (x == 0.0) return -ptm_from_bits(0x7ff0000000000000ULL);
        return ptm_from_bits(0x7ff8000000000000ULL);
    }
    if (x == ptm_from_bits(0x7ff0000000000000ULL)) return x;
    uint64_t u = ptm_to_bits(x);
    int k = 0;
    if ((u >> 52) == 0) { /* subnormal: scale up by 2^54 */
        x = x * 18014398509481984.0;
        u = ptm_to_bits(x);
        k = -54;
    }
    k += (int)(u >> 52) - 1023;
    double m = ptm_from_bits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL); /* [1,2) */
    if (m > 1.4142135623730951) { m = m * 0.5; k += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double p = 1.0 / 25.0;
    p = p * z + 1.0 / 23.0;
    p = p * z + 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    /* log(m) = 2s + 2s*z*p ; written as f - s*(f - 2*z*p) to keep the leading term exact-ish */
    double hf = 2.0 * s * z * p;        /* 2 s^3 (1/3 + ...) */
    double logm = 2.0 * s + hf;
    const double ln2_hi = 6.93147180369123816490e-01; /* 0x3fe62e42fee00000: low bits zero */
    const double ln2_lo = 1.90821492927058770002e-10;
    double dk = (double)k;
    return (dk * ln2_hi) + ((dk * ln2_lo) + logm);
}

/* ---- exponential -------------------------------------------------------- */
PTM_FN double ptm_exp(double x) {
    if (x != x) return x;
    if (x > 709.78) return ptm_from_bits(0x7ff0000000000000ULL);
    if (x < -745.2) return 0.0;
    const double inv_ln2 = 1.44269504088896338700e+00;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double kd = ptm_rint(x * inv_ln2);
    int k = (int)kd;
    double r = (x - kd * ln2_hi) - kd * ln2_lo; /* |r| <= 0.3466 */
    /* Taylor series to r^14 */
    double p = 1.0 / 87178291200.0;                 /* 1/14! */
    p = p * r + 1.0 / 6227020800.0;                 /* 1/13! */
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    double er = p * r + 1.0;
    if (k < -1021) { /* result may be subnormal: scale in two steps */
        return (er * ptm_pow2i(k + 600)) * ptm_pow2i(-600);
    }
    if (k > 1023) return (er * ptm_pow2i(k - 1)) * 2.0;
    return er * ptm_pow2i(k);
}

/* x^y for x > 0 (the only use on the path: step-size factors err^(-1/3), err^(-1/5)) */
PTM_FN double ptm_pow(double x, double y) {
    if (x == 1.0 || y == 0.0) return 1.0;
    return ptm_exp(y * ptm_log(x));
}

/* ---- sine / cosine ------------------------------------------------------ */
/* kernels on |r| <= pi/4 */
PTM_FN double ptm_sin_kernel(double r) {
    double z = r * r;
    double p = -1.0 / 1307674368000.0;     /* -1/15! */
    p = p * z + 1.0 / 6227020800.0;        /* 1/13! */
    p = p * z - 1.0 / 39916800.0;          /* -1/11! */
    p = p * z + 1.0 / 362880.0;            /* 1/9! */
    p = p * z - 1.0 / 5040.0;              /* -1/7! */
    p = p * z + 1.0 / 120.0;               /* 1/5! */
    p = p * z - 1.0 / 6.0;
    return r + (r * z) * p;
}
PTM_FN double ptm_cos_kernel(double r) {
    double z = r * r;
    double p = -1.0 / 6402373705728000.0;  /* -1/18! */
    p = p * z + 1.0 / 20922789888000.0;    /* 1/16! */
    p = p * z - 1.0 / 87178291200.0;       /* -1/14! */
    p = p * z + 1.0 / 479001600.0;         /* 1/12! */
    p = p * z - 1.0 / 3628800.0;           /* -1/10! */
    p = p * z + 1.0 / 40320.0;             /* 1/8! */
    p = p * z - 1.0 / 720.0;               /* -1/6! */
    p = p * z + 1.0 / 24.0;                /* 1/4! */
    double hz = 0.5 * z;
    return (1.0 - hz) + (z * z) * p;
}
/* reduce x (|x| < 2^20) by pi/2 with a three-part constant (Cody-Waite) */
PTM_FN double ptm_reduce_pio2(double x, int *quadrant) {
    const double two_over_pi = 0x1.45f306dc9c883p-1;
    const double pio2_1 = 0x1.921fb54400000p+0;   /* pi/2 rounded to 33 bits */
    const double pio2_2 = 0x1.0b4611a600000p-34;  /* next 33 bits */
    const double pio2_3 = 0x1.3198a2e037073p-69;  /* remainder to 53 bits */
    double n = ptm_rint(x * two_over_pi);
    *quadrant = ((int)n) & 3;
    return ((x - n * pio2_1) - n * pio2_2) - n * pio2_3;
}
PTM_FN double ptm_sin(double x) {
    int q;
    double r = ptm_reduce_pio2(x, &q);
    switch (q) {
        case 0: return ptm_sin_kernel(r);
        case 1: return ptm_cos_kernel(r);
        case 2: return -ptm_sin_kernel(r);
        default: return -ptm_cos_kernel(r);
    }
}
PTM_FN double ptm_cos(double x) {
    int q;
    double r = ptm_reduce_pio2(x, &q);
    switch (q) {
        case 0: return ptm_cos_kernel(r);
        case 1: return -ptm_sin_kernel(r);
        case 2: return -ptm_cos_kernel(r);
        default: return ptm_sin_kernel(r);
    }
}

/* ---- arc cosine --------------------------------------------------------- */
/* asin on |z| <= 0.5 by its Maclaurin series (terms to z^61) */
PTM_FN double ptm_asin_small(double z) {
    double t = z * z;
    /* c_n = (2n)! / (4^n (n!)^2 (2n+1)), n = 30 .. 1, Horner in t = z^2 */
    double p = 0x1.b8d2e5667ce6cp-10;
    p = p * t + 0x1.cf7dea5b6e830p-10; /* n=29 */
    p = p * t + 0x1.e82be60d9127ep-10; /* n=28 */
    p = p * t + 0x1.018f963c229bfp-9; /* n=27 */
    p = p * t + 0x1.1052bc5fa960ap-9; /* n=26 */
    p = p * t + 0x1.208d3570ae5a6p-9; /* n=25 */
    p = p * t + 0x1.3275586c5f2f0p-9; /* n=24 */
    p = p * t + 0x1.464c0950f7d47p-9; /* n=23 */
    p = p * t + 0x1.5c5f56efaaaabp-9; /* n=22 */
    p = p * t + 0x1.750de64d7d05fp-9; /* n=21 */
    p = p * t + 0x1.90cb77f60c7cep-9; /* n=20 */
    p = p * t + 0x1.b026f57b13b14p-9; /* n=19 */
    p = p * t + 0x1.d3d2a8e0dd67dp-9; /* n=18 */
    p = p * t + 0x1.fcaf8fb6db6dbp-9; /* n=17 */
    p = p * t + 0x1.15ee9d45d1746p-8; /* n=16 */
    p = p * t + 0x1.31683bdef7bdfp-8; /* n=15 */
    p = p * t + 0x1.51ba308d3dcb1p-8; /* n=14 */
    p = p * t + 0x1.782dda12f684cp-8; /* n=13 */
    p = p * t + 0x1.a6863d70a3d71p-8; /* n=12 */
    p = p * t + 0x1.df3bd37a6f4dfp-8; /* n=11 */
    p = p * t + 0x1.12ef3cf3cf3cfp-7; /* n=10 */
    p = p * t + 0x1.3fde50d79435ep-7; /* n=9 */
    p = p * t + 0x1.7a87878787878p-7; /* n=8 */
    p = p * t + 0x1.c99999999999ap-7; /* n=7 */
    p = p * t + 0x1.1c4ec4ec4ec4fp-6; /* n=6 */
    p = p * t + 0x1.6e8ba2e8ba2e9p-6; /* n=5 */
    p = p * t + 0x1.f1c71c71c71c7p-6; /* n=4 */
    p = p * t + 0x1.6db6db6db6db7p-5; /* n=3 */
    p = p * t + 0x1.3333333333333p-4; /* n=2 */
    p = p * t + 0x1.5555555555555p-3; /* n=1 */
    return z + (z * t) * p;
}
PTM_FN double ptm_acos(double x) {
    const double pio2_hi = 1.57079632679489655800e+00;
    const double pio2_lo = 6.12323399573676603587e-17;
    if (x != x || x > 1.0 || x < -1.0) return ptm_from_bits(0x7ff8000000000000ULL);
    if (x <= 0.5 && x >= -0.5) {
        return pio2_hi - (ptm_asin_small(x) - pio2_lo);
    }
    if (x > 0.5) {
        double z = (1.0 - x) * 0.5;
        double s = __builtin_sqrt(z);
        return 2.0 * ptm_asin_small(s);
    }
    {
        double z = (1.0 + x) * 0.5;
        double s = __builtin_sqrt(z);
        return 2.0 * (pio2_hi - (ptm_asin_small(s) - pio2_lo));
    }
}

/* ---- arc tangent --------------------------------------------------------- */
/* atan on [0, inf): 1/x above 1, pi/4 - atan((1-x)/(1+x)) on (1/sqrt(3), 1], then
   atan(u) = asin(u / sqrt(1 + u^2)) with |u| <= 1/sqrt(3), i.e. asin argument <= 1/2 */
PTM_FN double ptm_atan_pos(double x) {
    const double pio2_hi = 1.57079632679489655800e+00, pio2_lo = 6.12323399573676603587e-17;
    const double pio4_hi = 7.85398163397448278999e-01, pio4_lo = 3.06161699786838301793e-17;
    int inv = 0, fold = 0;
    if (x > 1.0) { x = 1.0 / x; inv = 1; }
    if (x > 0.57735026918962576451) { x = (1.0 - x) / (1.0 + x); fold = 1; }
    double r = ptm_asin_small(x / __builtin_sqrt(1.0 + x * x));
    if (fold) r = pio4_hi - (r - pio4_lo);
    if (inv) r = pio2_hi - (r - pio2_lo);
    return r;
}
PTM_FN double ptm_atan(double x) {
    if (x != x) return x;
    if (x < 0.0) return -ptm_atan_pos(-x);
    if (x == 0.0) return x;
    if (x > 1e300) return 1.57079632679489655800e+00;
    return ptm_atan_pos(x);
}
/* atan2(y, x) with the C quadrant and signed-zero conventions */
PTM_FN double ptm_atan2(double y, double x) {
    const double pi_hi = 3.14159265358979311600e+00, pi_lo = 1.22464679914735317723e-16;
    const double pio2 = 1.57079632679489655800e+00;
    if (x != x || y != y) return x + y;
    const int yneg = (ptm_to_bits(y) >> 63) != 0;
    if (x == 0.0) {
        if (y == 0.0) return (ptm_to_bits(x) >> 63) ? (yneg ? -pi_hi : pi_hi) : y;
        return yneg ? -pio2 : pio2;
    }
    double r = ptm_atan(y / x);
    if (x > 0.0) return r;
    return yneg ? -pi_hi + (r - pi_lo) : pi_hi + (r + pi_lo);
}

/* ---- complex helpers (Fermi-potential reflection probability) ----------- */
PTM_FN double ptm_hypot(double x, double y) {
    double a = ptm_fabs(x), b = ptm_fabs(y);
    if (a < b) { double t = a; a = b; b = t; }
    if (a == 0.0) return 0.0;
    double r = b / a;
    return a * __builtin_sqrt(1.0 + r * r);
}
/* principal square root of (re + i im); sign of a zero imaginary part selects the branch */
PTM_FN void ptm_csqrt(double re, double im, double *ore, double *oim) {
    if (re == 0.0 && im == 0.0) { *ore = 0.0; *oim = im; return; }
    double t = __builtin_sqrt((ptm_fabs(re) + ptm_hypot(re, im)) * 0.5);
    if (re >= 0.0) {
        *ore = t;
        *oim = im / (2.0 * t);
    } else {
        *ore = ptm_fabs(im) / (2.0 * t);
        *oim = ptm_copysign(t, im);
    }
}
/* (a + ib) / (c + id), Smith's algorithm */
PTM_FN void ptm_cdiv(double a, double b, double c, double d, double *ore, double *oim) {
    if (ptm_fabs(c) >= ptm_fabs(d)) {
        double r = d / c;
        double den = c + d * r;
        *ore = (a + b * r) / den;
        *oim = (b - a * r) / den;
    } else {
        double r = c / d;
        double den = c * r + d;
        *ore = (a * r + b) / den;
        *oim = (b * r - a) / den;
    }
}

/* ---- per-particle random stream ---------------------------------------- */
/*
 * The reference draws every random number of a run from one global
 * std::mt19937_64 (include/mc.h:15, src/main.cpp:127-134), consumed in particle
 * order.  A batch of independent particles needs one stream per particle:
 * draw n of particle id is mix64(key(seed, id) + n * golden), the SplitMix64
 * output function (Steele, Lea, Flood 2014) applied to a Weyl sequence.  The
 * draws are consumed at the same points and through the same libstdc++
 * transforms as the reference's distributions.
 */
PTM_FN uint64_t ptm_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
PTM_FN uint64_t ptm_stream_key(uint64_t seed, uint64_t id) {
    return ptm_mix64(ptm_mix64(seed + 0x632be59bd9b4e019ULL) ^ (id * 0xd1b54a32d192ed03ULL + 0x9e3779b97f4a7c15ULL));
}
PTM_FN uint64_t ptm_draw_u64(uint64_t key, uint64_t *counter) {
    *counter += 1;
    return ptm_mix64(key + (*counter) * 0x9e3779b97f4a7c15ULL);
}
/* std::generate_canonical<double, 53 or 64> over a 64-bit engine: r / 2^64, kept below 1 */
PTM_FN double ptm_canonical(uint64_t r) {
    double u = (double)r * 5.42101086242752217004e-20; /* 2^-64 */
    if (u >= 1.0) u = 0.99999999999999988898; /* nextafter(1, 0) */
    return u;
}
PTM_FN double ptm_uniform(uint64_t key, uint64_t *counter, double a, double b) {
    return ptm_canonical(ptm_draw_u64(key, counter)) * (b - a) + a;
}
PTM_FN double ptm_exponential(uint64_t key, uint64_t *counter, double lambda) {
    return -ptm_log(1.0 - ptm_canonical(ptm_draw_u64(key, counter))) / lambda;
}
/* std::uniform_int_distribution<int>(0, 1) over a 64-bit engine (libstdc++ downscaling) */
PTM_FN int ptm_uniform_int01(uint64_t key, uint64_t *counter) {
    const uint64_t scaling = 0x7fffffffffffffffULL;  /* (2^64-1) / 2 */
    const uint64_t past = 0xfffffffffffffffeULL;     /* 2 * scaling */
    uint64_t r;
    do { r = ptm_draw_u64(key, counter); } while (r >= past);
    return (int)(r / scaling);
}

#endif /* PT_DETMATH_H_ */
