// Device random streams and samplers for the operator-split step.
//
// Counter-based Philox4x32 with 8 rounds (the round count of the generator
// STEPS wraps, src/steps/rng/r123.hpp:76). A stream is addressed by
//   key     = (seed low 32, seed high 32)
//   counter = (tet, period, stream id, channel << 24 | block)
// so every (tet, period, channel) owns an independent sequence and results do
// not depend on launch shape, thread count or batch composition.
//
// All floating-point arithmetic here is + - * / and sqrt (all correctly
// rounded) with contraction disabled (-ffp-contract=off), so results are
// reproducible bit-for-bit on any IEEE-754 machine; log is written out for the
// same reason.
#pragma once

#ifndef __HIPCC_RTC__  // a runtime-compiled build gets these from hipRTC
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace stg {

// Register-pressure fences. The empty asm takes its operands in vector
// registers and hands them back as if changed: no instruction is emitted, and
// the arithmetic is untouched, but nothing that uses the results can be
// scheduled above this point. Used to run a chain of independent pieces one
// after another (instead of all at once, each holding its temporaries) and to
// keep the compiler from holding a value live across a long loop when it is
// cheap to form again.
template <class T>
__device__ __forceinline__ T pin(T x) {
    asm volatile("" : "+v"(x));
    return x;
}
template <class T, class U>
__device__ __forceinline__ void pin2(T &x, U &y) {
    asm volatile("" : "+v"(x), "+v"(y));
}

// Values that are the same in every lane (per-realization constants: the
// blocks of a launch never straddle realizations) taken into scalar registers.
// A load issued after the kernel's first global store cannot be proven
// unclobbered, so the compiler loads it per lane and keeps it, and everything
// formed from it (the Philox round keys of a seed, say), in vector registers.
__device__ __forceinline__ uint32_t uni(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
__device__ __forceinline__ uint64_t uni(uint64_t x) {
    return (uint64_t)uni((uint32_t)x) | ((uint64_t)uni((uint32_t)(x >> 32)) << 32);
}
__device__ __forceinline__ double uni(double x) {
    return __longlong_as_double((long long)uni((uint64_t)__double_as_longlong(x)));
}

struct Philox {
    uint32_t k0, k1;
    uint32_t c0, c1, c2, c3;
    uint32_t w0, w1, w2, w3;
    int left;  // words still unread in w0..w3 (consumed from w0)

    __device__ __forceinline__ void init(uint64_t seed, uint32_t tet, uint32_t period,
                                         uint32_t stream, uint32_t channel) {
        k0 = (uint32_t)seed;
        k1 = (uint32_t)(seed >> 32);
        c0 = tet;
        c1 = period;
        c2 = stream;
        c3 = channel << 24;
        left = 0;
    }

    __device__ __forceinline__ void refill() {
        uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3;
        uint32_t a = k0, b = k1;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            if (r > 0) {
                a += 0x9E3779B9u;
                b += 0xBB67AE85u;
            }
            const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)x0;
            const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)x2;
            const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
            const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
            const uint32_t y0 = hi1 ^ x1 ^ a;
            const uint32_t y2 = hi0 ^ x3 ^ b;
            x0 = y0;
            x1 = lo1;
            x2 = y2;
            x3 = lo0;
        }
        w0 = x0;
        w1 = x1;
        w2 = x2;
        w3 = x3;
        c3 = (c3 & 0xFF000000u) | ((c3 + 1u) & 0x00FFFFFFu);
        left = 4;
    }

    __device__ __forceinline__ uint32_t next() {
        if (left == 0) refill();
        const uint32_t r = w0;
        w0 = w1;
        w1 = w2;
        w2 = w3;
        --left;
        return r;
    }

    // uniform on (0, 1), 32 random bits
    __device__ __forceinline__ double unif() { return ((double)next() + 0.5) * 0x1p-32; }
};

__device__ __forceinline__ double dlog(double x) {
    if (!(x > 0.0)) return -__builtin_inf();
    uint64_t b = (uint64_t)__double_as_longlong(x);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) {
        x = x * 0x1p54;
        b = (uint64_t)__double_as_longlong(x);
        e = (int)((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = __longlong_as_double((long long)((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull));
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e += 1;
    }
    const double f = (m - 1.0) / (m + 1.0);
    const double f2 = f * f;
    // atanh series 1 + f^2/3 + f^4/5 + ... + f^20/21, Horner from the top
    double s = 1.0 / 21.0;
    s = s * f2 + 1.0 / 19.0;
    s = s * f2 + 1.0 / 17.0;
    s = s * f2 + 1.0 / 15.0;
    s = s * f2 + 1.0 / 13.0;
    s = s * f2 + 1.0 / 11.0;
    s = s * f2 + 1.0 / 9.0;
    s = s * f2 + 1.0 / 7.0;
    s = s * f2 + 1.0 / 5.0;
    s = s * f2 + 1.0 / 3.0;
    const double tf = f + f;
    const double lm = tf + tf * (f2 * s);
    const double de = (double)e;
    // ln 2 split so that de * hi is exact
    return de * 0x1.62e42fee00000p-1 + (lm + de * 0x1.a39ef35793c76p-33);
}

// exp(x): x = k ln2 + r with |r| <= ln2/2 (ln2 split so that k * hi is exact),
// e^r by its Taylor series to r^13 (truncation below 5e-18 relative), then
// scaled by 2^k. The oracle evaluates the same operations in the same order.
__device__ __forceinline__ double dexp(double x) {
    if (!(x < 709.782712893384)) return x != x ? x : __builtin_inf();
    if (x < -745.1332191019412) return 0.0;
    const double kd = floor(x * 0x1.71547652b82fep+0 + 0.5);
    const double r = (x - kd * 0x1.62e42fee00000p-1) - kd * 0x1.a39ef35793c76p-33;
    double p = 0x1.6124613a86d09p-33;
    p = p * r + 0x1.1eed8eff8d898p-29;
    p = p * r + 0x1.ae64567f544e4p-26;
    p = p * r + 0x1.27e4fb7789f5cp-22;
    p = p * r + 0x1.71de3a556c734p-19;
    p = p * r + 0x1.a01a01a01a01ap-16;
    p = p * r + 0x1.a01a01a01a01ap-13;
    p = p * r + 0x1.6c16c16c16c17p-10;
    p = p * r + 0x1.1111111111111p-7;
    p = p * r + 0x1.5555555555555p-5;
    p = p * r + 0x1.5555555555555p-3;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return __builtin_ldexp(p, (int)kd);
}

// IEEE-754 square root: correctly rounded on gfx950 as on any conforming host.
__device__ __forceinline__ double dsqrt(double x) { return x > 0.0 ? __builtin_sqrt(x) : 0.0; }

// log(k!) for k < 1024, each entry the correctly rounded double
__device__ __constant__ const double kLfactTable[1024] = {
#include "stg_lfact_table.inc"
};

// log(k!): table below 1024, Stirling series of lgamma(k + 1) above
__device__ __forceinline__ double lfact(int32_t k) {
    if (k < 1024) return kLfactTable[k < 0 ? 0 : k];
    const double x = (double)k + 1.0;
    const double ix = 1.0 / x;
    const double ix2 = ix * ix;
    const double corr = ix * (1.0 / 12.0 - ix2 * (1.0 / 360.0 - ix2 * (1.0 / 1260.0)));
    return (x - 0.5) * dlog(x) - x + 0x1.d67f1c864beb5p-1 + corr;
}

__device__ __forceinline__ double ipow(double base, int32_t n) {
    double r = 1.0;
    while (n > 0) {
        if (n & 1) r = r * base;
        base = base * base;
        n >>= 1;
    }
    return r;
}

// Binomial by sequential inversion (Kachitvichyanukul & Schmeiser 1988, BINV);
// q <= 1/2 and n*q < 30, the switch point of that paper, so the search length
// is n*q + 1 on average and the cap of 150 steps lies more than 20 standard
// deviations above the mean. Written in factorial-scaled form (U_x = u_x x!,
// R_x = f(x) x!, both below 150! < 1e263) so the search loop needs no
// division: U_{x+1} = (U_x - R_x)(x+1), R_{x+1} = R_x (n-x) q/(1-q).
// On a wavefront the loop runs to the largest draw among the lanes; below
// n*q = 30 that costs fewer instructions than the rejection sampler, whose
// wavefront runs to its slowest lane's trial count and exact test.
// The search stops at x = min(n, 150); x and n - x are carried as doubles
// (exact integers), which gives the same operands as converting them each
// step, and x itself is the loop counter (no integer copy per step).
__device__ __forceinline__ int32_t binom_inversion(Philox& g, int32_t n, double q) {
    double r = ipow(1.0 - q, n);
    const double s = q / (1.0 - q);
    double u = g.unif();
    const double lim = (double)(n < 150 ? n : 150);
    double xd = 0.0, nx = (double)n;
    while (!(u <= r) && xd < lim) {
        xd = xd + 1.0;
        u = (u - r) * xd;
        r = r * (nx * s);
        nx = nx - 1.0;
    }
    return (int32_t)xd;
}

// Binomial by transformed rejection with squeeze (Hoermann 1993, BTRS);
// q <= 1/2 and n*q >= 30. Bounded at 1000 trials (acceptance is > 0.8 per
// trial). The branch rarely runs, but its live values set the register budget
// (and so the occupancy) of the whole period kernel, so it is written to hold
// as little as possible across trials:
//   - the setup constants (spq, b, a, c, vr, alpha, m) are formed again in
//     every trial from (n, q), fenced so that they are not hoisted out of the
//     loop;
//   - the log-factorial bound of the exact test (about one trial in seven) is
//     formed when the squeeze fails, its logarithms (up to five, each a
//     Stirling series above 1023) fenced to run one at a time.
// Every value is the one the plain formulas give, in the same order.
__device__ __forceinline__ int32_t binom_btrs(Philox& g, int32_t n, double q) {
    for (int it = 0; it < 1000; ++it) {
        int32_t nn = n;
        double qq = q;
        pin2(nn, qq);
        const double nd = (double)nn;
        const double spq = dsqrt(nd * qq * (1.0 - qq));
        const double b = 1.15 + 2.53 * spq;
        const double ib = 1.0 / b;
        const double a = -0.0873 + 0.0248 * b + 0.01 * qq;
        const double c = nd * qq + 0.5;
        const double vr = 0.92 - 4.2 * ib;
        const double u = g.unif() - 0.5;
        double v = g.unif();
        const double us = 0.5 - fabs(u);
        const double kd = floor((2.0 * a / us + b) * u + c);
        if (kd < 0.0 || kd > nd) continue;
        const int32_t k = (int32_t)kd;
        if (us >= 0.07 && v <= vr) return k;
        // exact test: log v' <= h - log k! - log (n-k)! + (k - m) log(q/(1-q))
        const double alpha = (2.83 + 5.1 * ib) * spq;
        double vv = v * alpha / (a / (us * us) + b);
        const int32_t m = (int32_t)floor((nd + 1.0) * qq);
        pin2(vv, qq);
        double lpq = dlog(qq / (1.0 - qq));
        int32_t a1 = m;
        pin2(lpq, a1);
        double h = lfact(a1);
        int32_t a2 = nn - m;
        pin2(h, a2);
        h = h + lfact(a2);
        pin2(vv, h);
        v = dlog(vv);
        int32_t k1 = k;
        pin2(v, k1);
        double acc = h - lfact(k1);
        int32_t k2 = nn - k;
        pin2(acc, k2);
        acc = acc - lfact(k2);
        if (v <= acc + (double)(k - m) * lpq) return k;
    }
    return (int32_t)floor(((double)n + 1.0) * q);
}

__device__ __forceinline__ int32_t binomial(Philox& g, int32_t n, double p) {
    if (n <= 0 || !(p > 0.0)) return 0;
    if (p >= 1.0) return n;
    const bool flip = p > 0.5;
    const double q = flip ? 1.0 - p : p;
    const int32_t k = ((double)n * q < 30.0) ? binom_inversion(g, n, q) : binom_btrs(g, n, q);
    return flip ? n - k : k;
}

}  // namespace stg
