// oracle.cpp — TEST INFRASTRUCTURE ONLY (the checker of the HIP path).
//
// A single-threaded-by-default C++17 restatement of the reference's
// per-generation offspring pipeline, Engine<parallel_pipeline>::simulate
// (examples/QTLSim/engine_parallel_pipeline.hpp:172-239), over the reference's
// dense row-major uint64 bitset (DS/population_space/population_space_row_block_alignment.hpp:70-258).
// DS/ = include/clotho/data_spaces/.
//
// The reference draws from boost::random (absent here, and not reproducible
// bit-for-bit without it).  This restatement instead consumes counter-based
// Philox4x32-10 streams keyed by (seed, generation, stream, replicate, index):
// the same streams the GPU library draws, so GPU outputs can be compared with
// these bit for bit.  Each function cites the reference code it restates.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library.

#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Philox4x32-10, from Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
// as easy as 1, 2, 3" (SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
// increments 0x9E3779B9 / 0xBB67AE85, ten rounds.
// ---------------------------------------------------------------------------
inline void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
    }
}

enum : uint32_t { S_PARENTS = 0, S_GAMETE = 1, S_MUTCOUNT = 2, S_MUTATTR = 3, S_ENV = 4 };

struct Draw {
    uint32_t x[4];
};

inline Draw draw(uint64_t seed, uint32_t gen, uint32_t stream, uint32_t rep, uint32_t index,
                 uint32_t sub) {
    Draw d;
    d.x[0] = index;
    d.x[1] = sub;
    d.x[2] = gen;
    d.x[3] = stream | (rep << 8);
    philox10(d.x, (uint32_t)seed, (uint32_t)(seed >> 32));
    return d;
}

// 53-bit integer from two words, its [0,1) dyadic, and an open (0,1) dyadic
// built from 52 bits so that 2k+1 stays exact.
inline uint64_t k53(uint32_t lo, uint32_t hi) { return (((uint64_t)hi << 32) | lo) >> 11; }
inline double u53(uint32_t lo, uint32_t hi) { return (double)k53(lo, hi) * 0x1p-53; }
inline double uopen(uint32_t lo, uint32_t hi) {
    const uint64_t k = (((uint64_t)hi << 32) | lo) >> 12;
    return (double)(2 * k + 1) * 0x1p-53;
}

// ---------------------------------------------------------------------------
// Deterministic elementary functions: fixed sequences of +, -, * (IEEE
// round-to-nearest, no contraction) so that CPU and GPU evaluate them to the
// same bits.  Accuracy is pinned against libm / scipy in tests.
// ---------------------------------------------------------------------------
inline uint64_t dbits(double x) {
    uint64_t b;
    std::memcpy(&b, &x, 8);
    return b;
}
inline double bitsd(uint64_t b) {
    double x;
    std::memcpy(&x, &b, 8);
    return x;
}

// 1/x for finite nonzero x: exponent-flip seed, six Newton steps r(2 - x r).
double det_recip(double x) {
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

// sqrt(x) for x >= 0: reciprocal-square-root seed, six Newton steps, x * y.
double det_sqrt(double x) {
    if (!(x > 0.0)) return 0.0;
    double y = bitsd(0x5FE6EB50C7B537A9ull - (dbits(x) >> 1));
    const double h = 0.5 * x;
    for (int i = 0; i < 6; ++i) {
        const double yy = y * y;
        y = y * (1.5 - h * yy);
    }
    return x * y;
}

// log(x), x > 0 and normal: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s), s = (m-1)/(m+1), series to s^23; ln2 split (fdlibm).
double det_log(double x) {
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
    double p = 0.043478260869565216;  // 1/23
    p = p * z + 0.047619047619047616; // 1/21
    p = p * z + 0.052631578947368418; // 1/19
    p = p * z + 0.058823529411764705; // 1/17
    p = p * z + 0.066666666666666666; // 1/15
    p = p * z + 0.076923076923076927; // 1/13
    p = p * z + 0.090909090909090912; // 1/11
    p = p * z + 0.1111111111111111;   // 1/9
    p = p * z + 0.14285714285714285;  // 1/7
    p = p * z + 0.20000000000000001;  // 1/5
    p = p * z + 0.33333333333333331;  // 1/3
    p = p * z + 1.0;
    const double de = (double)e;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    return de * ln2_hi + (de * ln2_lo + 2.0 * s * p);
}

// exp(x): k = round(x/ln2) by the 1.5*2^52 shift, r = x - k ln2 (split),
// Taylor polynomial to r^13, scale by 2^k through the exponent field.
double det_exp(double x) {
    if (x < -745.0) return 0.0;
    if (x > 709.0) return bitsd(0x7FF0000000000000ull);
    const double shifter = 6755399441055744.0;  // 1.5 * 2^52
    const double kd = (x * 1.4426950408889634 + shifter) - shifter;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double r = (x - kd * ln2_hi) - kd * ln2_lo;
    double p = 1.6059043836821613e-10;    // 1/13!
    p = p * r + 2.08767569878681e-09;      // 1/12!
    p = p * r + 2.505210838544172e-08;     // 1/11!
    p = p * r + 2.755731922398589e-07;     // 1/10!
    p = p * r + 2.7557319223985893e-06;    // 1/9!
    p = p * r + 2.48015873015873e-05;      // 1/8!
    p = p * r + 0.0001984126984126984;     // 1/7!
    p = p * r + 0.001388888888888889;      // 1/6!
    p = p * r + 0.008333333333333333;      // 1/5!
    p = p * r + 0.041666666666666664;      // 1/4!
    p = p * r + 0.16666666666666666;       // 1/3!
    p = p * r + 0.5;                       // 1/2!
    p = p * r + 1.0;
    p = p * r + 1.0;
    int k = (int)kd;
    if (k >= -1022) return p * bitsd((uint64_t)(k + 1023) << 52);
    // subnormal result: two exact-power steps
    p = p * bitsd((uint64_t)(k + 1000 + 1023) << 52);
    return p * bitsd((uint64_t)(-1000 + 1023) << 52);
}

// Inverse standard-normal CDF by P. J. Acklam's rational approximation
// (relative error < 1.15e-9), evaluated with det_log/det_sqrt/det_recip.
double normal_ppf(double p) {
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

// Poisson(lambda) CDF for inversion, last entry forced to 1 (the draw is
// truncated there; the tail mass beyond is below 1e-16 for the table length
// chosen).  Counts are min{k : u < cdf[k]}.
int poisson_table(double lambda, double* cdf, int cap) {
    if (!(lambda >= 0.0) || lambda > 64.0 || cap < 2) return -1;
    double pk = std::exp(-lambda);
    double c = pk;
    int k = 0;
    cdf[0] = c;
    while (k + 1 < cap && c < 1.0 - 1e-17 && !(k > lambda && pk < 1e-18)) {
        ++k;
        pk = pk * lambda / (double)k;
        c = c + pk;
        cdf[k] = c;
    }
    cdf[k] = 1.0;
    return k + 1;
}

inline uint32_t poisson_draw(const std::vector<double>& cdf, double u) {
    uint32_t k = 0;
    while (!(u < cdf[k])) ++k;
    return k;
}

inline int ctz64(uint64_t x) { return __builtin_ctzll(x); }

// Effects are integer multiples of 2^q: round-half-even of (mean + sigma z) 2^-q,
// clamped to |n| <= 2^30 - 1.
inline double quantize_effect(double w, double scale, double quantum) {
    double n = std::nearbyint(w * scale);
    const double lim = 1073741823.0;
    if (n > lim) n = lim;
    if (n < -lim) n = -lim;
    return n * quantum;
}

// ---------------------------------------------------------------------------
// block_crossover over one word (DS/crossover/block_crossover.hpp:46-63): for
// each heterozygous bit (ascending, as debruijn_bit_walker::next_and_shift
// walks it, include/clotho/utility/debruijn_bit_walker.hpp:111-127) the
// classifier is odd(#{events <= position}), region_classifier's upper_bound
// (include/clotho/classifiers/region_classifier.hpp:70-77) under is_odd_tag.
// ---------------------------------------------------------------------------
inline uint64_t crossover_word(uint64_t t, uint64_t b, uint32_t offset, const double* positions,
                               uint32_t alleles, const double* ev, uint32_t ne) {
    uint64_t hets = t ^ b;
    uint64_t mask = 0;
    while (hets) {
        const int j = ctz64(hets);
        hets &= hets - 1;
        const uint32_t slot = offset + (uint32_t)j;
        // PositionClassifier uses positions->at(offset); a het bit beyond the
        // allele table would throw there: it cannot occur for valid populations.
        const double p = slot < alleles ? positions[slot] : 0.0;
        uint32_t cnt = 0;
        for (uint32_t e = 0; e < ne; ++e) cnt += (ev[e] <= p) ? 1u : 0u;
        if (cnt & 1u) mask |= (uint64_t)1 << j;
    }
    return (t & ~mask) | (b & mask);
}

}  // namespace

// ===========================================================================
// exported scalar helpers
// ===========================================================================
extern "C" void or_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox10(c, key[0], key[1]);
    std::memcpy(out, c, 16);
}
extern "C" void or_draw(uint64_t seed, uint32_t generation, uint32_t stream, uint32_t replicate,
                        uint32_t index, uint32_t sub, uint32_t* out) {
    Draw d = draw(seed, generation, stream, replicate, index, sub);
    std::memcpy(out, d.x, 16);
}
extern "C" double or_det_log(double x) { return det_log(x); }
extern "C" double or_det_exp(double x) { return det_exp(x); }
extern "C" double or_det_sqrt(double x) { return det_sqrt(x); }
extern "C" double or_det_recip(double x) { return det_recip(x); }
extern "C" double or_normal_ppf(double p) { return normal_ppf(p); }
extern "C" int or_poisson_table(double lambda, double* cdf, int cap) {
    return poisson_table(lambda, cdf, cap);
}

// Synthetic genotype rows [r0, r1): bit j of row r set when Philox word
// (seed, generation 0, stream 7, replicate, index r, sub j/4)[j%4] < thr[j].
extern "C" void or_synth_rows(uint64_t seed, uint32_t replicate, const uint32_t* thr,
                              uint32_t sites, uint32_t words, uint32_t r0, uint32_t r1,
                              uint64_t* out) {
    for (uint32_t r = r0; r < r1; ++r) {
        uint64_t* row = out + (size_t)(r - r0) * words;
        std::memset(row, 0, 8 * (size_t)words);
        for (uint32_t j0 = 0; j0 < sites; j0 += 4) {
            Draw d = draw(seed, 0, 7, replicate, r, j0 / 4);
            for (uint32_t k = 0; k < 4 && j0 + k < sites; ++k)
                if (d.x[k] < thr[j0 + k]) row[(j0 + k) / 64] |= (uint64_t)1 << ((j0 + k) % 64);
        }
    }
}

// block_crossover_method::operator() (DS/crossover/block_crossover_method.hpp:36-70)
// with common_crossover::run_crossover_task's zero-event copy
// (DS/crossover/common_crossover_row_block_alignment.hpp:71-93).
extern "C" void or_crossover(const uint64_t* top, const uint64_t* bottom, uint32_t wp,
                             uint64_t* child, uint32_t wc, const double* positions,
                             uint32_t alleles, const double* events, uint32_t n_events) {
    for (uint32_t w = 0; w < wc; ++w) {
        const uint64_t t = w < wp ? top[w] : 0;
        const uint64_t b = w < wp ? bottom[w] : 0;
        child[w] = n_events == 0 ? t
                                 : crossover_word(t, b, 64u * w, positions, alleles, events,
                                                  n_events);
    }
}

// free_space_accumulator_mt::{resetBuffers,updateVariableBuffer,updateFixedBuffer,finalize}
// (DS/free_space/free_space_accumulator.hpp:112-192) and
// free_space_details::analyze_free_indices(fixed, variable, W, M)
// (DS/free_space/free_space_details.hpp:359-408): fixed = AND of all rows,
// lost = ~OR, indices < M ascending, fixed first.
extern "C" void or_free_space(const uint64_t* rows, uint32_t h, uint32_t w, uint32_t max_alleles,
                              uint32_t* fixed, uint32_t* n_fixed, uint32_t* lost,
                              uint32_t* n_lost) {
    std::vector<uint64_t> var(w, 0), fix(w, ~(uint64_t)0);
    for (uint32_t r = 0; r < h; ++r) {
        const uint64_t* row = rows + (size_t)r * w;
        for (uint32_t k = 0; k < w; ++k) {
            var[k] |= row[k];
            fix[k] &= row[k];
        }
    }
    uint32_t nf = 0, nl = 0;
    for (uint32_t k = 0; k < w; ++k) {
        uint64_t fx = fix[k];
        while (fx) {
            const uint32_t j = 64u * k + (uint32_t)ctz64(fx);
            fx &= fx - 1;
            if (j < max_alleles) fixed[nf++] = j;
        }
    }
    for (uint32_t k = 0; k < w; ++k) {
        uint64_t ls = ~var[k];
        while (ls) {
            const uint32_t j = 64u * k + (uint32_t)ctz64(ls);
            ls &= ls - 1;
            if (j < max_alleles) lost[nl++] = j;
        }
    }
    *n_fixed = nf;
    *n_lost = nl;
}

// common_phenotype_accumulator<row_block_alignment>::operator()
// (DS/phenotype_evaluator/common_phenotype_accumulator_row_block_alignment.hpp:47-66):
// words ascending, bits ascending, FP64 running sums per trait.
extern "C" void or_phenotype(const uint64_t* rows, uint32_t h, uint32_t w, const double* effects,
                             uint32_t alleles, uint32_t t, double* traits) {
    std::vector<double> buf(t);
    for (uint32_t r = 0; r < h; ++r) {
        std::fill(buf.begin(), buf.end(), 0.0);
        const uint64_t* row = rows + (size_t)r * w;
        for (uint32_t k = 0; k < w; ++k) {
            uint64_t b = row[k];
            while (b) {
                const uint32_t j = 64u * k + (uint32_t)ctz64(b);
                b &= b - 1;
                if (j >= alleles) continue;
                for (uint32_t x = 0; x < t; ++x) buf[x] += effects[(size_t)j * t + x];
            }
        }
        for (uint32_t x = 0; x < t; ++x) traits[(size_t)r * t + x] = buf[x];
    }
}

namespace {

// make_selection_generator (DS/selection/selection_generator_factory.hpp:33-67):
// constant fitness -> uniform over [0, N-1] (UniformPairGenerator), otherwise a
// fitness-proportional draw.  The reference's Boost alias table is replaced by
// an inverse-CDF search over fitness quantised to 2^sc so that every prefix
// sum is an exact FP64 integer below 2^52: idx = min{i : target < C[i]},
// target = floor(k53 * C[N-1] / 2^53).
struct Selector {
    bool uniform = true;
    std::vector<double> cum;
    uint64_t total = 0;
    uint32_t n = 0;

    void build(const double* f, uint32_t count, bool constant) {
        n = count;
        uniform = constant;
        if (uniform || n == 0) {
            uniform = true;
            return;
        }
        double m = 0.0;
        for (uint32_t i = 0; i < n; ++i) m = std::max(m, f[i]);
        if (!(m > 0.0)) {
            uniform = true;
            return;
        }
        int L = 0;
        while (((uint64_t)1 << L) < n) ++L;
        const int sc = 52 - L - (std::ilogb(m) + 1);
        cum.resize(n);
        double c = 0.0;
        for (uint32_t i = 0; i < n; ++i) {
            const double fi = f[i] > 0.0 ? f[i] : 0.0;
            c += std::floor(std::ldexp(fi, sc));
            cum[i] = c;
        }
        total = (uint64_t)c;
    }

    uint32_t pick(uint64_t k) const {
        if (uniform) return (uint32_t)(((unsigned __int128)k * n) >> 53);
        const double target = (double)(uint64_t)(((unsigned __int128)k * total) >> 53);
        return (uint32_t)(std::upper_bound(cum.begin(), cum.end(), target) - cum.begin());
    }
};

bool all_equal(const double* f, uint32_t n) {
    for (uint32_t i = 1; i < n; ++i)
        if (!(f[i] == f[0])) return false;
    return true;
}

struct FitnessOp {
    uint32_t kind = CLOTHO_FIT_CONSTANT;
    double coeff = 1.0, inv = 1.0;
    uint32_t traits = 0;

    // ifitness_generator::generate(N) (src/clotho/fitness/normal_fitness_generator.cpp:36-52,
    // src/clotho/fitness/quadratic_fitness_generator.cpp:42-54): constants per N.
    void generate(const clotho_config& cfg, uint32_t n) {
        kind = cfg.fitness_kind;
        traits = cfg.trait_count;
        if (kind == CLOTHO_FIT_GAUSSIAN) {
            const double sigma = std::sqrt(4.0 * (double)n * cfg.mu);
            coeff = (1.0 / sigma) * 0.398942280401432677939946059934;
            inv = 1.0 / (2.0 * sigma * sigma);
        } else if (kind == CLOTHO_FIT_QUADRATIC) {
            const double s = cfg.fitness_scale * std::sqrt(4.0 * (double)n * cfg.mu);
            inv = 1.0 / s;
        } else if (kind == CLOTHO_FIT_GAUSSIAN_FIXED || kind == CLOTHO_FIT_MV_STABILIZING) {
            inv = 1.0 / (2.0 * cfg.fitness_scale * cfg.fitness_scale);
        }
    }

    // ifitness::operator()(double* first, double* last): normal_fitness_metric
    // (src/clotho/fitness/normal_fitness_metric.cpp:25-31; first trait,
    // include/clotho/fitness/normal_fitness_metric.hpp:60-66),
    // quadratic_fitness_metric (src/clotho/fitness/quadratic_fitness_metric.cpp:23-28),
    // constant_fitness_metric; plus the fixed-sigma and multivariate
    // stabilising metrics that the reference lacks.
    double operator()(const double* z) const {
        const double x = traits ? z[0] : 0.0;
        switch (kind) {
            case CLOTHO_FIT_GAUSSIAN: {
                const double ex = (x * x) * inv;
                return coeff * det_exp(-ex);
            }
            case CLOTHO_FIT_QUADRATIC: {
                double r = x * inv;
                r = r * r;
                return r >= 1.0 ? 0.0 : 1.0 - r;
            }
            case CLOTHO_FIT_GAUSSIAN_FIXED:
                return det_exp(-((x * x) * inv));
            case CLOTHO_FIT_MV_STABILIZING: {
                double ss = 0.0;
                for (uint32_t t = 0; t < traits; ++t) ss = ss + z[t] * z[t];
                return det_exp(-(ss * inv));
            }
            default:
                return 1.0;
        }
    }
};

}  // namespace

extern "C" void or_select(const double* fitness, uint32_t n, uint64_t seed, uint32_t generation,
                          uint32_t replicate, uint32_t* parents) {
    Selector sel;
    sel.build(fitness, n, all_equal(fitness, n));
    for (uint32_t i = 0; i < n; ++i) {
        Draw d = draw(seed, generation, S_PARENTS, replicate, i, 0);
        parents[2 * i] = sel.pick(k53(d.x[0], d.x[1]));
        parents[2 * i + 1] = sel.pick(k53(d.x[2], d.x[3]));
    }
}

// Environmental noise of individual i, trait t in generation gen: env_sigma *
// Phi^-1(u), u the open uniform of Philox stream 4 (no reference counterpart:
// the reference ignores environmental factors, examples/qtl/README).
inline double env_noise(const clotho_config& cfg, uint32_t gen, uint32_t i, uint32_t t) {
    const Draw d = draw(cfg.seed, gen, S_ENV, cfg.replicate, i, t);
    return cfg.env_sigma * normal_ppf(uopen(d.x[0], d.x[1]));
}

extern "C" double or_env_noise(const clotho_config* cfg, uint32_t generation, uint32_t individual,
                               uint32_t trait) {
    return env_noise(*cfg, generation, individual, trait);
}

extern "C" double or_fitness_value(const clotho_config* cfg, uint32_t n, const double* z) {
    FitnessOp op;
    op.generate(*cfg, n);
    return op(z);
}

// ===========================================================================
// whole simulation
// ===========================================================================
struct or_sim {
    clotho_config cfg;
    uint32_t N = 0, H = 0, T = 0;
    uint32_t next_N = 0;  // offspring individuals of the following generations (0: N)
    uint32_t inter = 0;   // position in the intermediate batch, 0 .. cfg.intermediates
    uint32_t gen = 1;  // Engine::init increments m_generation to 1 before the first simulate
    std::vector<uint64_t> parent, child;
    uint32_t wp = 0, wc = 0;
    uint32_t A = 0;
    std::vector<double> pos, eff;
    std::vector<uint32_t> age;
    std::vector<uint32_t> fixed_list, lost_list;
    std::vector<double> fitness;  // of the current child population (next parents)
    bool fitness_constant = true;
    std::vector<double> traits;   // child haplotype traits, H x T
    std::vector<uint32_t> top_row, bottom_row, n_events, n_mut, mut_off;
    uint32_t M = 0;
    bool all_neutral = true;
    std::vector<double> cdf_rho, cdf_mu;
    double eff_scale = 1.0, eff_quantum = 1.0;
    int threads = 1;
    double phase[7] = {0, 0, 0, 0, 0, 0, 0};
    std::vector<uint32_t> given;  // parent individual per gamete of the next generation
};

extern "C" or_sim* or_sim_create(const clotho_config* cfg) {
    if (!cfg || cfg->n_individuals == 0) return nullptr;
    if (cfg->intermediates && cfg->compact_every) return nullptr;
    or_sim* s = new or_sim();
    s->cfg = *cfg;
    s->gen = 1 + cfg->intermediates;  // population i of batch g: g (k + 1) + i
    s->N = cfg->n_individuals;
    s->H = 2 * s->N;
    s->T = cfg->trait_count;
    s->cdf_rho.resize(128);
    s->cdf_mu.resize(128);
    int kr = poisson_table(cfg->rho, s->cdf_rho.data(), 128);
    // host mutations: no per-gamete mutation draws (Poisson(0) counts)
    const double mu_drawn = (cfg->flags & CLOTHO_FLAG_HOST_MUTATIONS) ? 0.0 : cfg->mu;
    int km = poisson_table(mu_drawn, s->cdf_mu.data(), 128);
    if (kr < 0 || km < 0) {
        delete s;
        return nullptr;
    }
    s->cdf_rho.resize(kr);
    s->cdf_mu.resize(km);
    s->eff_scale = std::ldexp(1.0, -cfg->effect_quantum_log2);
    s->eff_quantum = std::ldexp(1.0, cfg->effect_quantum_log2);
    // first generation: GeneralFitness::resize appends 1.0 (DS/fitness/general_fitness.hpp:238-243)
    s->fitness.assign(s->N, 1.0);
    s->fitness_constant = true;
    s->traits.assign((size_t)s->H * s->T, 0.0);
    return s;
}

extern "C" void or_sim_destroy(or_sim* s) { delete s; }
extern "C" void or_sim_set_threads(or_sim* s, int threads) { s->threads = threads < 1 ? 1 : threads; }

extern "C" int or_sim_set_population(or_sim* s, const uint64_t* rows, uint32_t words,
                                     const double* positions, const double* effects,
                                     uint32_t alleles, const double* fitness) {
    if (words % 8 || (size_t)words * 64 < alleles) return -1;
    s->wc = words;
    s->child.assign(rows, rows + (size_t)s->H * words);
    s->A = alleles;
    s->pos.assign(positions, positions + alleles);
    s->eff.assign(effects, effects + (size_t)alleles * s->T);
    s->age.assign(alleles, 0);
    s->all_neutral = true;
    for (double w : s->eff)
        if (w != 0.0) {
            s->all_neutral = false;
            break;
        }
    s->traits.assign((size_t)s->H * s->T, 0.0);
    or_phenotype(s->child.data(), s->H, words, s->eff.data(), alleles, s->T, s->traits.data());
    if (fitness) {
        s->fitness.assign(fitness, fitness + s->N);
    } else {
        s->fitness.assign(s->N, 1.0);
    }
    s->fitness_constant = all_equal(s->fitness.data(), s->N);
    std::vector<uint32_t> fx(alleles + 1), ls(alleles + 1);
    uint32_t nf = 0, nl = 0;
    or_free_space(s->child.data(), s->H, words, alleles, fx.data(), &nf, ls.data(), &nl);
    s->inter = 0;  // an upload is analysed like the end of a batch
    s->fixed_list.assign(fx.begin(), fx.begin() + nf);
    s->lost_list.assign(ls.begin(), ls.begin() + nl);
    return 0;
}

namespace {

// One offspring batch [i0, i1): offspring_generator::operator() on a thread
// (DS/offspring_generator/offspring_generator_population_space_row_alignment.hpp:84-115):
// crossover -> mutate -> phenotype -> fixed, with thread-private AND/OR buffers.
void offspring_batch(or_sim* s, const Selector& sel, const std::vector<uint32_t>& free_slots,
                     uint32_t F, uint32_t i0, uint32_t i1, uint64_t* var, uint64_t* fix,
                     double* tsec) {
    const clotho_config& cfg = s->cfg;
    const uint32_t wp = s->wp, wc = s->wc, T = s->T;
    double ev[128];
    auto t0 = std::chrono::steady_clock::now();
    // crossover(): one pair per offspring, gamete 2i from parent id0, 2i+1 from id1
    // given mate pairs (Crossover::update, DS/crossover/crossover.hpp:49-72):
    // gamete 2i from the first individual of pair i, 2i + 1 from the second
    const bool given = !s->given.empty();
    for (uint32_t i = i0; i < i1; ++i) {
        Draw pd = draw(cfg.seed, s->gen, S_PARENTS, cfg.replicate, i, 0);
        const uint32_t ids[2] = {
            given ? s->given[2 * i] : sel.pick(k53(pd.x[0], pd.x[1])),
            given ? s->given[2 * i + 1] : sel.pick(k53(pd.x[2], pd.x[3]))};
        for (uint32_t g = 0; g < 2; ++g) {
            const uint32_t o = 2 * i + g;
            const uint32_t p = ids[g];
            Draw gd = draw(cfg.seed, s->gen, S_GAMETE, cfg.replicate, o, 0);
            const uint32_t ne = poisson_draw(s->cdf_rho, u53(gd.x[0], gd.x[1]));
            const bool swap = u53(gd.x[2], gd.x[3]) < cfg.bias;
            for (uint32_t e = 0; e < ne; e += 4) {
                Draw ed = draw(cfg.seed, s->gen, S_GAMETE, cfg.replicate, o, 1 + e / 4);
                for (uint32_t q = 0; q < 4 && e + q < ne; ++q) ev[e + q] = (double)ed.x[q] * 0x1p-32;
            }
            // the classifier sorts its events (region_classifier ctor); the
            // parity of #{e <= pos} does not depend on their order.
            std::sort(ev, ev + ne);
            const uint32_t top = swap ? 2 * p + 1 : 2 * p;
            const uint32_t bot = swap ? 2 * p : 2 * p + 1;
            s->top_row[o] = top;
            s->bottom_row[o] = bot;
            s->n_events[o] = ne;
            const uint64_t* tr = wp ? s->parent.data() + (size_t)top * wp : nullptr;
            const uint64_t* br = wp ? s->parent.data() + (size_t)bot * wp : nullptr;
            or_crossover(tr, br, wp, s->child.data() + (size_t)o * wc, wc, s->pos.data(), s->A,
                         ev, ne);
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    // mutate(): row i takes mut_pool[mut_dist[i] .. mut_dist[i+1]) (XOR)
    for (uint32_t o = 2 * i0; o < 2 * i1; ++o) {
        uint64_t* row = s->child.data() + (size_t)o * wc;
        for (uint32_t k = s->mut_off[o]; k < s->mut_off[o] + s->n_mut[o]; ++k) {
            const uint32_t slot = k < F ? free_slots[k] : s->A + (k - F);
            row[slot / 64] ^= (uint64_t)1 << (slot % 64);
        }
    }
    auto t2 = std::chrono::steady_clock::now();
    // phenotype(): accumulate, or constant zeros when every weight is neutral
    if (!s->all_neutral) {
        or_phenotype(s->child.data() + (size_t)2 * i0 * wc, 2 * (i1 - i0), wc, s->eff.data(),
                     (uint32_t)s->pos.size(), T, s->traits.data() + (size_t)2 * i0 * T);
    } else {
        std::fill(s->traits.begin() + (size_t)2 * i0 * T, s->traits.begin() + (size_t)2 * i1 * T,
                  0.0);
    }
    auto t3 = std::chrono::steady_clock::now();
    // fixed(): var |= row; fix &= row
    for (uint32_t o = 2 * i0; o < 2 * i1; ++o) {
        const uint64_t* row = s->child.data() + (size_t)o * wc;
        for (uint32_t k = 0; k < wc; ++k) {
            var[k] |= row[k];
            fix[k] &= row[k];
        }
    }
    auto t4 = std::chrono::steady_clock::now();
    tsec[0] = std::chrono::duration<double>(t1 - t0).count();
    tsec[1] = std::chrono::duration<double>(t2 - t1).count();
    tsec[2] = std::chrono::duration<double>(t3 - t2).count();
    tsec[3] = std::chrono::duration<double>(t4 - t3).count();
}

}  // namespace

// Column and row margins over a row subset: the reference's set-bit walk
// (frequency_evaluator<population_space<row_block_alignment>>::operator(),
// include/clotho/data_spaces/analysis/allele_frequency/frequency_evaluator_population_space_row_alignment.hpp:36-61).
extern "C" void or_allele_counts(const uint64_t* rows, uint32_t W, uint32_t A,
                                 const uint32_t* sample, uint32_t n, uint32_t* column_counts,
                                 uint32_t* row_counts) {
    std::fill(column_counts, column_counts + A, 0u);
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t* row = rows + (size_t)sample[i] * W;
        uint32_t m = 0;
        for (uint32_t w = 0; w < W; ++w) {
            uint64_t b = row[w];
            while (b) {
                const uint32_t j = 64u * w + (uint32_t)ctz64(b);
                b &= b - 1;
                if (j < A) ++column_counts[j];
                ++m;
            }
        }
        row_counts[i] = m;
    }
}

// pairwise_difference<association_matrix>::eval with the row-grouped
// evaluator (include/clotho/data_spaces/analysis/pairwise_difference/association_matrix_pairwise_difference.hpp:78-102,
// row_grouped_pairwise_difference.hpp:25-55): for the sample pairs a < b in
// order, the set bits of row a xor row b over W words, counted one bit at a
// time.
extern "C" void or_pairwise_differences(const uint64_t* rows, uint32_t W, const uint32_t* sample,
                                        uint32_t n, uint32_t* out) {
    size_t p = 0;
    for (uint32_t a = 0; a < n; ++a)
        for (uint32_t b = a + 1; b < n; ++b) {
            const uint64_t* ra = rows + (size_t)sample[a] * W;
            const uint64_t* rb = rows + (size_t)sample[b] * W;
            uint32_t d = 0;
            for (uint32_t w = 0; w < W; ++w)
                for (uint64_t x = ra[w] ^ rb[w]; x; x &= x - 1) ++d;
            out[p++] = d;
        }
}

extern "C" int or_sim_compact(or_sim* s);

// Offspring population size of the following generations: the pN the
// reference's population_growth model returns (pN = growth(parent N,
// generation), examples/QTLSim/engine_parallel_pipeline.hpp:181-184).
extern "C" int or_sim_set_population_size(or_sim* s, uint32_t n) {
    if (!s || n == 0) return CLOTHO_ERR_ARG;
    s->next_N = n;
    return 0;
}

// Engine<parallel_pipeline>::simulate (examples/QTLSim/engine_parallel_pipeline.hpp:172-239).
extern "C" int or_sim_step(or_sim* s) {
    const clotho_config& cfg = s->cfg;
    // parent individuals, then the child population grows to pN (:181-197)
    const uint32_t H_par = s->H;
    const uint32_t N_par = s->N;
    if (s->next_N) {
        s->N = s->next_N;
        s->H = 2 * s->N;
    }
    const uint32_t N = s->N, H = s->H, T = s->T;
    auto c0 = std::chrono::steady_clock::now();

    // 1: std::swap(child, parent)
    std::swap(s->parent, s->child);
    s->wp = s->wc;

    // 2: removeFixedAlleles (:409-447; population_space::remove_fixed_allele,
    //    DS/population_space/population_space_row_block_alignment.hpp:133-153),
    //    with intermediate generations only on the first parent of a batch
    //    (buildProjectedPopulations, QS/engine_parallel_pipeline_batched.hpp:370-382)
    const uint32_t K = cfg.intermediates;
    if (s->inter == 0) {
        for (uint32_t j : s->fixed_list) {
            const uint64_t keep = ~((uint64_t)1 << (j % 64));
            for (uint32_t r = 0; r < H_par; ++r) s->parent[(size_t)r * s->wp + j / 64] &= keep;
        }
    }
    const uint32_t F = (uint32_t)(s->fixed_list.size() + s->lost_list.size());
    std::vector<uint32_t> free_slots(s->fixed_list);
    free_slots.insert(free_slots.end(), s->lost_list.begin(), s->lost_list.end());

    // 3: new mutations.  The reference draws M ~ Poisson(mu*2N) then a uniform
    //    row per mutation (mutation_allocator::allocate, DS/mutation/mutation_allocator.hpp:41-47;
    //    generate_child_mutations :327-367); by Poisson splitting that is the
    //    same law as an independent Poisson(mu) count per gamete, drawn here.
    //    Row o receives mutations [off_o, off_o + n_o) in draw order, as
    //    mutate() reads mut_pool[mut_dist[o] .. mut_dist[o+1]).
    s->n_mut.assign(H, 0);
    s->mut_off.assign(H + 1, 0);
    for (uint32_t o = 0; o < H; ++o) {
        Draw d = draw(cfg.seed, s->gen, S_MUTCOUNT, cfg.replicate, o, 0);
        s->n_mut[o] = poisson_draw(s->cdf_mu, u53(d.x[0], d.x[1]));
        s->mut_off[o + 1] = s->mut_off[o] + s->n_mut[o];
    }
    s->M = s->mut_off[H];

    // 4: child_max_alleles (:393-407) and child->grow (:196)
    const uint32_t all_size = s->A + (s->M > F ? s->M - F : 0);
    if (all_size > cfg.max_alleles) return CLOTHO_ERR_CAPACITY;
    s->wc = ((all_size + 511) / 512) * 8;
    s->child.assign((size_t)H * s->wc, 0);
    s->pos.resize(all_size, 0.0);
    s->age.resize(all_size, 0);
    s->eff.resize((size_t)all_size * T, 0.0);
    s->traits.assign((size_t)H * T, 0.0);

    // 5: per mutation k: slot = free[k] (fixed asc, then lost asc) else append;
    //    AlleleGenerator (position uniform_01, age = generation;
    //    DS/allele_space/allele_generator_vector.hpp:76-83) and
    //    TraitSpaceGenerator2 (neutral ? 0 : Normal(mean, sigma);
    //    DS/phenotype_evaluator/trait_space_generator.hpp:118-129).
    for (uint32_t k = 0; k < s->M; ++k) {
        const uint32_t slot = k < F ? free_slots[k] : s->A + (k - F);
        Draw d = draw(cfg.seed, s->gen, S_MUTATTR, cfg.replicate, k, 0);
        s->pos[slot] = (double)d.x[0] * 0x1p-32;
        s->age[slot] = s->gen;
        for (uint32_t t = 0; t < T; ++t) {
            Draw e = draw(cfg.seed, s->gen, S_MUTATTR, cfg.replicate, k, 1 + t);
            double w = 0.0;
            if (!(u53(e.x[0], e.x[1]) < cfg.neutral_p)) {
                w = cfg.trait_mean + cfg.trait_sigma * normal_ppf(uopen(e.x[2], e.x[3]));
                w = quantize_effect(w, s->eff_scale, s->eff_quantum);
            }
            s->eff[(size_t)slot * T + t] = w;
        }
    }
    // 6: trait_space_vector::isAllNeutral over the whole table
    s->all_neutral = true;
    for (double w : s->eff)
        if (w != 0.0) {
            s->all_neutral = false;
            break;
        }

    // 7: selection from the parent fitness
    Selector sel;
    sel.build(s->fitness.data(), N_par, s->fitness_constant);

    s->top_row.assign(H, 0);
    s->bottom_row.assign(H, 0);
    s->n_events.assign(H, 0);
    auto c1 = std::chrono::steady_clock::now();

    // 8-11: offspring batches over TC threads (thread_pool2 model, :203-226)
    const uint32_t TC = (uint32_t)std::max(1, s->threads);
    const uint32_t batch = (N + TC - 1) / TC;
    std::vector<uint64_t> var((size_t)TC * s->wc, 0), fix((size_t)TC * s->wc, ~(uint64_t)0);
    std::vector<double> tsec((size_t)TC * 4, 0.0);
    std::vector<std::thread> pool;
    for (uint32_t j = 0; j < TC; ++j) {
        const uint32_t i0 = std::min(N, j * batch), i1 = std::min(N, (j + 1) * batch);
        if (i0 >= i1) continue;
        auto job = [=, &sel, &free_slots, &var, &fix, &tsec]() {
            offspring_batch(s, sel, free_slots, F, i0, i1, var.data() + (size_t)j * s->wc,
                            fix.data() + (size_t)j * s->wc, tsec.data() + 4 * j);
        };
        if (j + 1 == TC || TC == 1)
            job();
        else
            pool.emplace_back(job);
    }
    for (auto& th : pool) th.join();
    for (int q = 0; q < 4; ++q) {
        double mx = 0.0;
        for (uint32_t j = 0; j < TC; ++j) mx = std::max(mx, tsec[4 * j + q]);
        s->phase[q] = mx;
    }

    // 12: GeneralFitness::operator()(pop, 0, N) (DS/fitness/general_fitness.hpp:148-177)
    auto c2 = std::chrono::steady_clock::now();
    FitnessOp op;
    op.generate(cfg, N);
    s->fitness.assign(N, 1.0);
    std::vector<double> z(T > 0 ? T : 1);
    for (uint32_t i = 0; i < N; ++i) {
        for (uint32_t t = 0; t < T; ++t)
            z[t] = s->traits[(size_t)(2 * i) * T + t] + s->traits[(size_t)(2 * i + 1) * T + t];
        if (cfg.env_sigma > 0.0)
            for (uint32_t t = 0; t < T; ++t) z[t] = z[t] + env_noise(cfg, s->gen, i, t);
        s->fitness[i] = op(z.data());
    }
    s->fitness_constant = all_equal(s->fitness.data(), N);
    auto c3 = std::chrono::steady_clock::now();

    // 13: free_space.finalize(all_size) (DS/free_space/free_space_accumulator.hpp:142-151),
    //     with intermediate generations only after the last of a batch; before
    //     it the slots the new mutations took leave the front of the free list
    //     (the shared free-space iterator of grow_child_population,
    //     QS/engine_parallel_pipeline_batched.hpp:397-425)
    if (s->inter < K) {
        uint32_t used = std::min(s->M, F);
        const uint32_t from_fixed = std::min<uint32_t>(used, (uint32_t)s->fixed_list.size());
        s->fixed_list.erase(s->fixed_list.begin(), s->fixed_list.begin() + from_fixed);
        used -= from_fixed;
        s->lost_list.erase(s->lost_list.begin(), s->lost_list.begin() + used);
        s->A = all_size;
        ++s->gen;
        ++s->inter;
        s->given.clear();
        auto c4 = std::chrono::steady_clock::now();
        s->phase[4] = std::chrono::duration<double>(c3 - c2).count();
        s->phase[5] = std::chrono::duration<double>(c4 - c3).count();
        s->phase[6] = std::chrono::duration<double>(c1 - c0).count();
        return 0;
    }
    s->inter = 0;
    for (uint32_t j = 1; j < TC; ++j)
        for (uint32_t k = 0; k < s->wc; ++k) {
            var[k] |= var[(size_t)j * s->wc + k];
            fix[k] &= fix[(size_t)j * s->wc + k];
        }
    s->fixed_list.clear();
    s->lost_list.clear();
    for (uint32_t k = 0; k < s->wc; ++k) {
        uint64_t fx = fix[k];
        while (fx) {
            const uint32_t j = 64u * k + (uint32_t)ctz64(fx);
            fx &= fx - 1;
            if (j < all_size) s->fixed_list.push_back(j);
        }
    }
    for (uint32_t k = 0; k < s->wc; ++k) {
        uint64_t ls = ~var[k];
        while (ls) {
            const uint32_t j = 64u * k + (uint32_t)ctz64(ls);
            ls &= ls - 1;
            if (j < all_size) s->lost_list.push_back(j);
        }
    }
    s->A = all_size;
    ++s->gen;
    s->given.clear();
    if (cfg.compact_every) or_sim_compact(s);
    auto c4 = std::chrono::steady_clock::now();
    s->phase[4] = std::chrono::duration<double>(c3 - c2).count();
    s->phase[5] = std::chrono::duration<double>(c4 - c3).count();
    s->phase[6] = std::chrono::duration<double>(c1 - c0).count();
    return 0;
}

// Given mate pairs of the next generation (Crossover<RNG,GS>::update,
// DS/crossover/crossover.hpp:49-72): parents[o] is the parent individual of
// offspring gamete o.
extern "C" int or_sim_set_parents(or_sim* s, const uint32_t* parents, uint32_t gametes) {
    const uint32_t next_n = s->next_N ? s->next_N : s->N;
    if (gametes != 2 * next_n) return -1;
    for (uint32_t o = 0; o < gametes; ++o)
        if (parents[o] >= s->N) return -1;
    s->given.assign(parents, parents + gametes);
    return 0;
}

// MutationGenerator<RNG,GS>::operator()(child, N) (DS/mutation/mutation_generator.hpp:41-52)
// on the current population: mutation k takes the next free slot (fixed
// ascending, lost ascending, then appended), is set in row rows[k] and gets
// its position, effects and the age of the generation that produced the
// population.  A reused fixed column is cleared from every row first (its
// allele leaves the population, as removeFixedAlleles does).  Traits, the
// all-neutral flag and the free lists are recomputed; fitness is not.
extern "C" int or_sim_add_mutations(or_sim* s, const uint32_t* rows, const double* positions,
                                    const double* effects, uint32_t n, uint32_t* slots_out) {
    if (n == 0) return 0;
    const uint32_t T = s->T, H = s->H;
    for (uint32_t k = 0; k < n; ++k)
        if (rows[k] >= H) return CLOTHO_ERR_ARG;
    std::vector<uint32_t> free_slots(s->fixed_list);
    free_slots.insert(free_slots.end(), s->lost_list.begin(), s->lost_list.end());
    const uint32_t F = (uint32_t)free_slots.size(), A = s->A;
    const uint64_t A1 = (uint64_t)A + (n > F ? n - F : 0);
    const uint64_t s_cap = ((uint64_t)s->cfg.max_alleles + 511) / 512 * 512;
    if (A1 > s_cap) return CLOTHO_ERR_CAPACITY;
    const uint32_t W0 = s->wc, W1 = (uint32_t)(((A1 + 511) / 512) * 8);
    if (W1 > W0) {  // re-stride to the wider rows
        std::vector<uint64_t> wide((size_t)H * W1, 0);
        for (uint32_t r = 0; r < H; ++r)
            std::copy(s->child.begin() + (size_t)r * W0, s->child.begin() + (size_t)(r + 1) * W0,
                      wide.begin() + (size_t)r * W1);
        s->child.swap(wide);
        s->wc = W1;
    }
    std::vector<uint32_t> slots(n);
    for (uint32_t k = 0; k < n; ++k) slots[k] = k < F ? free_slots[k] : A + (k - F);
    const uint32_t fixed_reused = std::min(std::min(n, F), (uint32_t)s->fixed_list.size());
    for (uint32_t k = 0; k < fixed_reused; ++k) {
        const uint64_t keep = ~((uint64_t)1 << (slots[k] % 64));
        for (uint32_t r = 0; r < H; ++r) s->child[(size_t)r * W1 + slots[k] / 64] &= keep;
    }
    s->pos.resize(A1, 0.0);
    s->age.resize(A1, 0);
    s->eff.resize((size_t)A1 * T, 0.0);
    const uint32_t age = s->gen > 0 ? s->gen - 1 : 0;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t slot = slots[k];
        s->pos[slot] = positions[k];
        s->age[slot] = age;
        for (uint32_t t = 0; t < T; ++t) s->eff[(size_t)slot * T + t] = effects[(size_t)k * T + t];
        s->child[(size_t)rows[k] * W1 + slot / 64] |= (uint64_t)1 << (slot % 64);
    }
    s->A = (uint32_t)A1;
    s->M = n;
    s->all_neutral = true;
    for (double w : s->eff)
        if (w != 0.0) {
            s->all_neutral = false;
            break;
        }
    s->traits.assign((size_t)H * T, 0.0);
    or_phenotype(s->child.data(), H, W1, s->eff.data(), s->A, T, s->traits.data());
    std::vector<uint32_t> fx(s->A + 1), ls(s->A + 1);
    uint32_t nf = 0, nl = 0;
    or_free_space(s->child.data(), H, W1, s->A, fx.data(), &nf, ls.data(), &nl);
    s->fixed_list.assign(fx.begin(), fx.begin() + nf);
    s->lost_list.assign(ls.begin(), ls.begin() + nl);
    s->inter = 0;
    if (slots_out) std::copy(slots.begin(), slots.end(), slots_out);
    return 0;
}

// Fitness<phenotype_evaluator<...>>::update(pop, eval)
// (DS/fitness/phenotype_evaluator_fitness.hpp:51-69): every individual of the
// current population from its genome traits, with the noise of the generation
// that produced the population.
extern "C" void or_sim_evaluate_fitness(or_sim* s) {
    const clotho_config& cfg = s->cfg;
    const uint32_t N = s->N, T = s->T;
    const uint32_t gen = s->gen > 0 ? s->gen - 1 : 0;
    FitnessOp op;
    op.generate(cfg, N);
    std::vector<double> z(T > 0 ? T : 1);
    s->fitness.assign(N, 1.0);
    for (uint32_t i = 0; i < N; ++i) {
        for (uint32_t t = 0; t < T; ++t)
            z[t] = s->traits[(size_t)(2 * i) * T + t] + s->traits[(size_t)(2 * i + 1) * T + t];
        if (cfg.env_sigma > 0.0)
            for (uint32_t t = 0; t < T; ++t) z[t] = z[t] + env_noise(cfg, gen, i, t);
        s->fitness[i] = op(z.data());
    }
    s->fitness_constant = all_equal(s->fitness.data(), N);
}

// Column compaction (no reference counterpart; the reference reuses free
// slots, QS/engine_parallel_pipeline.hpp:337-348): every slot on the fixed or
// lost list of the current population is removed, the remaining (variable)
// slots keep their order and are renumbered 0..V-1, rows are re-packed to
// W = ceil(V/512)*8 words, positions, effects and ages follow their slots,
// and the free lists become empty.  Genome traits and fitness are unchanged.
extern "C" int or_sim_compact(or_sim* s) {
    const uint32_t A = s->A, W = s->wc, H = s->H, T = s->T;
    std::vector<uint8_t> drop(A, 0);
    for (uint32_t j : s->fixed_list) drop[j] = 1;
    for (uint32_t j : s->lost_list) drop[j] = 1;
    std::vector<uint32_t> kept;
    kept.reserve(A);
    for (uint32_t j = 0; j < A; ++j)
        if (!drop[j]) kept.push_back(j);
    const uint32_t V = (uint32_t)kept.size();
    const uint32_t W2 = ((V + 511) / 512) * 8;
    std::vector<uint64_t> rows((size_t)H * W2, 0);
    for (uint32_t r = 0; r < H; ++r) {
        const uint64_t* in = s->child.data() + (size_t)r * W;
        uint64_t* out = rows.data() + (size_t)r * W2;
        for (uint32_t k = 0; k < V; ++k) {
            const uint32_t j = kept[k];
            if ((in[j / 64] >> (j % 64)) & 1) out[k / 64] |= (uint64_t)1 << (k % 64);
        }
    }
    std::vector<double> pos(V), eff((size_t)V * T);
    std::vector<uint32_t> age(V);
    for (uint32_t k = 0; k < V; ++k) {
        pos[k] = s->pos[kept[k]];
        age[k] = s->age[kept[k]];
        for (uint32_t t = 0; t < T; ++t) eff[(size_t)k * T + t] = s->eff[(size_t)kept[k] * T + t];
    }
    s->child.swap(rows);
    s->wc = W2;
    s->A = V;
    s->pos.swap(pos);
    s->eff.swap(eff);
    s->age.swap(age);
    s->fixed_list.clear();
    s->lost_list.clear();
    s->all_neutral = true;
    for (double w : s->eff)
        if (w != 0.0) {
            s->all_neutral = false;
            break;
        }
    return 0;
}

extern "C" void or_sim_shape(const or_sim* s, clotho_shape* out) {
    std::memset(out, 0, sizeof(*out));
    out->generation = s->gen;
    out->n_individuals = s->N;
    out->haplotypes = s->H;
    out->words = s->wc;
    out->alleles = s->A;
    out->new_mutations = s->M;
    out->fixed_count = (uint32_t)s->fixed_list.size();
    out->lost_count = (uint32_t)s->lost_list.size();
    out->free_count = out->fixed_count + out->lost_count;
    out->variable_count = s->A - out->free_count;
    out->all_neutral = s->all_neutral;
    out->fitness_constant = s->fitness_constant;
}
// Copies of possibly empty vectors (an empty vector may hold no storage at all).
template <class T>
inline void copy_out(void* dst, const std::vector<T>& v, size_t n) {
    if (dst && n) std::memcpy(dst, v.data(), n * sizeof(T));
}
extern "C" void or_sim_rows(const or_sim* s, uint64_t* out) {
    copy_out(out, s->child, s->child.size());
}
extern "C" void or_sim_alleles(const or_sim* s, double* positions, double* effects,
                               uint32_t* ages) {
    copy_out(positions, s->pos, s->A);
    copy_out(effects, s->eff, (size_t)s->A * s->T);
    copy_out(ages, s->age, s->A);
}
extern "C" void or_sim_free_space(const or_sim* s, uint32_t* fixed, uint32_t* lost) {
    copy_out(fixed, s->fixed_list, s->fixed_list.size());
    copy_out(lost, s->lost_list, s->lost_list.size());
}
extern "C" void or_sim_traits(const or_sim* s, double* traits) {
    copy_out(traits, s->traits, s->traits.size());
}
extern "C" void or_sim_fitness(const or_sim* s, double* fitness) {
    copy_out(fitness, s->fitness, s->fitness.size());
}
extern "C" void or_sim_gametes(const or_sim* s, uint32_t* top_row, uint32_t* bottom_row,
                               uint32_t* events, uint32_t* mutations) {
    copy_out(top_row, s->top_row, s->top_row.size());
    copy_out(bottom_row, s->bottom_row, s->bottom_row.size());
    copy_out(events, s->n_events, s->n_events.size());
    copy_out(mutations, s->n_mut, s->n_mut.size());
}
extern "C" void or_sim_phase_seconds(const or_sim* s, double* out7) {
    std::memcpy(out7, s->phase, sizeof(s->phase));
}
