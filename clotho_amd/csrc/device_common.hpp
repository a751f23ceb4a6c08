// device_common.hpp — stream draws and deterministic arithmetic for the gfx950
// kernels of the offspring pipeline.
//
// Every random quantity of a generation is a pure function of
// (seed, generation, stream, replicate, index, sub) through Philox4x32-10
// (Salmon et al., SC'11), so the draws do not depend on launch shape or
// thread count.  Transcendentals that feed bit-exact outputs (normal effects,
// fitness) are evaluated by fixed sequences of IEEE +, -, * so that the host
// restatement reproduces them to the last bit; this translation unit must be
// compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clotho {

enum : uint32_t { S_PARENTS = 0, S_GAMETE = 1, S_MUTCOUNT = 2, S_MUTATTR = 3, S_ENV = 4 };
// stream 7 draws synthetic genotypes; stream 8 draws statistics row samples
constexpr uint32_t kStreamSample = 8;

struct Draw4 {
    uint32_t x[4];
};

__host__ __device__ inline Draw4 philox_draw(uint64_t seed, uint32_t gen, uint32_t stream,
                                             uint32_t rep, uint32_t index, uint32_t sub) {
    uint32_t c0 = index, c1 = sub, c2 = gen, c3 = stream | (rep << 8);
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
    Draw4 d;
    d.x[0] = c0;
    d.x[1] = c1;
    d.x[2] = c2;
    d.x[3] = c3;
    return d;
}

__host__ __device__ inline uint64_t k53(uint32_t lo, uint32_t hi) {
    return (((uint64_t)hi << 32) | lo) >> 11;
}
__host__ __device__ inline double u53(uint32_t lo, uint32_t hi) {
    return (double)k53(lo, hi) * 0x1p-53;
}
__host__ __device__ inline double uopen(uint32_t lo, uint32_t hi) {
    const uint64_t k = (((uint64_t)hi << 32) | lo) >> 12;
    return (double)(2 * k + 1) * 0x1p-53;
}

__host__ __device__ inline uint64_t dbits(double x) {
    union {
        double d;
        uint64_t u;
    } v;
    v.d = x;
    return v.u;
}
__host__ __device__ inline double bitsd(uint64_t b) {
    union {
        double d;
        uint64_t u;
    } v;
    v.u = b;
    return v.d;
}

// 1/x: exponent-flip seed, six Newton steps.
__host__ __device__ inline double det_recip(double x) {
    const uint64_t b = dbits(x);
    const uint64_t sign = b & 0x8000000000000000ull;
    const double a = bitsd(b & 0x7FFFFFFFFFFFFFFFull);
    double r = bitsd(0x7FDE623822FC16E6ull - dbits(a));
    for (int i = 0; i < 6; ++i) {
        const double t = a * r;
        r = r * (2.0 - t);
    }
    return bitsd(dbits(r) | sign);
}

__host__ __device__ inline double det_sqrt(double x) {
    if (!(x > 0.0)) return 0.0;
    double y = bitsd(0x5FE6EB50C7B537A9ull - (dbits(x) >> 1));
    const double h = 0.5 * x;
    for (int i = 0; i < 6; ++i) {
        const double yy = y * y;
        y = y * (1.5 - h * yy);
    }
    return x * y;
}

// log via m 2^e, m in [sqrt(1/2), sqrt(2)), 2 atanh((m-1)/(m+1)) series to s^23.
__host__ __device__ inline double det_log(double x) {
    const uint64_t b = dbits(x);
    int e = (int)((b >> 52) & 0x7FF) - 1023;
    double m = bitsd((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    const double f = m - 1.0;
    const double s = f * det_recip(2.0 + f);
    const double z = s * s;
    double p = 0.043478260869565216;
    p = p * z + 0.047619047619047616;
    p = p * z + 0.052631578947368418;
    p = p * z + 0.058823529411764705;
    p = p * z + 0.066666666666666666;
    p = p * z + 0.076923076923076927;
    p = p * z + 0.090909090909090912;
    p = p * z + 0.1111111111111111;
    p = p * z + 0.14285714285714285;
    p = p * z + 0.20000000000000001;
    p = p * z + 0.33333333333333331;
    p = p * z + 1.0;
    const double de = (double)e;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    return de * ln2_hi + (de * ln2_lo + 2.0 * s * p);
}

// exp via k = round(x/ln2) (1.5*2^52 shift), Taylor to r^13, 2^k by exponent bits.
__host__ __device__ inline double det_exp(double x) {
    if (x < -745.0) return 0.0;
    if (x > 709.0) return bitsd(0x7FF0000000000000ull);
    const double shifter = 6755399441055744.0;
    const double kd = (x * 1.4426950408889634 + shifter) - shifter;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double r = (x - kd * ln2_hi) - kd * ln2_lo;
    double p = 1.6059043836821613e-10;
    p = p * r + 2.08767569878681e-09;
    p = p * r + 2.505210838544172e-08;
    p = p * r + 2.755731922398589e-07;
    p = p * r + 2.7557319223985893e-06;
    p = p * r + 2.48015873015873e-05;
    p = p * r + 0.0001984126984126984;
    p = p * r + 0.001388888888888889;
    p = p * r + 0.008333333333333333;
    p = p * r + 0.041666666666666664;
    p = p * r + 0.16666666666666666;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    const int k = (int)kd;
    if (k >= -1022) return p * bitsd((uint64_t)(k + 1023) << 52);
    p = p * bitsd((uint64_t)(k + 1000 + 1023) << 52);
    return p * bitsd((uint64_t)(-1000 + 1023) << 52);
}

// Inverse normal CDF, P. J. Acklam's rational approximation (rel. err < 1.15e-9).
__host__ __device__ inline double normal_ppf(double p) {
    const double a1 = -3.969683028665376e+01, a2 = 2.209460984245205e+02,
                 a3 = -2.759285104469687e+02, a4 = 1.383577518672690e+02,
                 a5 = -3.066479806614716e+01, a6 = 2.506628277459239e+00;
    const double b1 = -5.447609879822406e+01, b2 = 1.615858368580409e+02,
                 b3 = -1.556989798598866e+02, b4 = 6.680131188771972e+01,
                 b5 = -1.328068155288572e+01;
    const double c1 = -7.784894002430293e-03, c2 = -3.223964580411365e-01,
                 c3 = -2.400758277161838e+00, c4 = -2.549732539343734e+00,
                 c5 = 4.374664141464968e+00, c6 = 2.938163982698783e+00;
    const double d1 = 7.784695709041462e-03, d2 = 3.224671290700398e-01,
                 d3 = 2.445134137142996e+00, d4 = 3.754408661907416e+00;
    const double plow = 0.02425;
    if (p < plow) {
        const double q = det_sqrt(-2.0 * det_log(p));
        const double num = ((((c1 * q + c2) * q + c3) * q + c4) * q + c5) * q + c6;
        const double den = (((d1 * q + d2) * q + d3) * q + d4) * q + 1.0;
        return num * det_recip(den);
    }
    if (p <= 1.0 - plow) {
        const double q = p - 0.5;
        const double r = q * q;
        const double num = (((((a1 * r + a2) * r + a3) * r + a4) * r + a5) * r + a6) * q;
        const double den = ((((b1 * r + b2) * r + b3) * r + b4) * r + b5) * r + 1.0;
        return num * det_recip(den);
    }
    const double q = det_sqrt(-2.0 * det_log(1.0 - p));
    const double num = ((((c1 * q + c2) * q + c3) * q + c4) * q + c5) * q + c6;
    const double den = (((d1 * q + d2) * q + d3) * q + d4) * q + 1.0;
    return -(num * det_recip(den));
}

// floor(log2 x) for finite x > 0, subnormals included (std::ilogb).
__host__ __device__ inline int det_ilogb(double x) {
    uint64_t b = dbits(x);
    int e = (int)((b >> 52) & 0x7FF);
    if (e == 0) {
        b = dbits(x * 18446744073709551616.0);  // 2^64, exact
        return (int)((b >> 52) & 0x7FF) - 1023 - 64;
    }
    return e - 1023;
}

// x * 2^n, exact whenever the result is a normal number (std::ldexp).
__host__ __device__ inline double det_ldexp(double x, int n) {
    while (n > 1023) {
        x = x * 0x1p1023;
        n -= 1023;
    }
    while (n < -1022) {
        x = x * 0x1p-1022;
        n += 1022;
    }
    return x * bitsd((uint64_t)(n + 1023) << 52);
}

}  // namespace clotho
