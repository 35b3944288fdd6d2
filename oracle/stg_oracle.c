/*
 * CPU oracle: operator-split stochastic reaction-diffusion period.
 *
 * TEST INFRASTRUCTURE ONLY -- the checker for the HIP path and the CPU
 * baseline timed by bench.py. Product code never links or calls this file.
 *
 * What it restates (paths relative to the STEPS 5.0.4 source tree):
 *   - per-group Direct SSA with independent kprocs, one group = one tet plus
 *     its membrane triangles:
 *       src/steps/mpi/dist/tetopsplit/operator/ssa_operator.cpp:24-47
 *       src/steps/mpi/dist/tetopsplit/kproc/propensities.hpp:543-586 (drawEvent)
 *   - propensities h*c with falling factorials and volume/area scaling:
 *       src/steps/mpi/dist/tetopsplit/kproc/reactions.cpp:66-119
 *       src/steps/mpi/dist/tetopsplit/kproc/surface_reactions.cpp:176-240
 *   - time-integrated occupancy of species that react and diffuse:
 *       src/steps/mpi/dist/tetopsplit/mol_state.hpp:87-200
 *       src/steps/mpi/dist/tetopsplit/kproc/reactions.cpp:35-40
 *   - diffusion leaving (stochastic rounding, binomial total, categorical
 *     split at <= 10 molecules, sequential conditional binomials above) and
 *     entering (gather of the neighbour's per-face slot):
 *       src/steps/mpi/dist/tetopsplit/operator/diffusion_operator.cpp:41-238
 *       src/steps/mpi/dist/tetopsplit/kproc/diffusions.hpp:211-224
 *       src/steps/rng/rng.hpp:139-156 (stochastic_round)
 *   - split period = 1 / max summed face rate, face rate D*A/(V*d):
 *       src/steps/mpi/tetopsplit/tetopsplit.cpp:5373-5411, diff.cpp:74-79
 *       src/steps/mpi/dist/tetopsplit/simulation.cpp:1288-1337
 *       src/steps/mpi/dist/tetopsplit/simulation_data.hpp:138-148
 *   - period schedules: src/steps/mpi/tetopsplit/tetopsplit.cpp:1528-1560
 *     (TetOpSplit) and src/steps/mpi/dist/tetopsplit/simulation.cpp:1358-1384
 *     (DistTetOpSplit).
 *
 * Random numbers: Philox4x32 with 8 rounds, the generator the reference wraps
 * (src/steps/rng/r123.hpp:76), but keyed by (seed) with counter
 * (tet, period, stream, channel<<24 | block) instead of one sequential stream
 * per rank. Binomials: inversion (Kachitvichyanukul & Schmeiser 1988, BINV)
 * for n*min(p,1-p) < 30 (that paper's switch point), else BTRS (Hoermann 1993, "The generation of binomial
 * random variates", J. Stat. Comput. Simul. 46:101-110). log and sqrt are
 * written out with + - * / only so that the device path reproduces them.
 *
 * Build: oracle/Makefile (-ffp-contract=off: no fused multiply-add anywhere).
 */
#include "stg_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define AVOGADRO 6.02214179e23 /* src/steps/math/constants.hpp:45 */
#define FARADAY 96485.3399     /* src/steps/math/constants.hpp:49 */
#define GAS_CONSTANT 8.314472  /* src/steps/math/constants.hpp:53 */
#define E_CHARGE 1.602176487e-19 /* src/steps/math/constants.hpp:57 */
#define MAX_EVENTS_PER_PERIOD (1 << 22)

/* Threads the OpenMP loops of stgo_run use (for reporting the CPU baseline). */
int32_t stgo_threads(void) { return (int32_t)omp_get_max_threads(); }

/* ------------------------------------------------------------------ Philox */

void stgo_philox4x32_8(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 8; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

typedef struct {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int idx;
} ostream;

static void os_init(ostream *st, uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2,
                    uint32_t channel) {
    st->key[0] = (uint32_t)seed;
    st->key[1] = (uint32_t)(seed >> 32);
    st->ctr[0] = c0;
    st->ctr[1] = c1;
    st->ctr[2] = c2;
    st->ctr[3] = channel << 24;
    st->idx = 4;
}

static uint32_t os_next(ostream *st) {
    if (st->idx == 4) {
        stgo_philox4x32_8(st->ctr, st->key, st->buf);
        st->ctr[3] = (st->ctr[3] & 0xFF000000u) | ((st->ctr[3] + 1u) & 0x00FFFFFFu);
        st->idx = 0;
    }
    return st->buf[st->idx++];
}

/* uniform on the open interval (0, 1) with 32 random bits */
static double os_unif(ostream *st) {
    return ((double)os_next(st) + 0.5) * 0x1p-32;
}

/* ------------------------------------------------------------ scalar math */

static double d_from_bits(uint64_t b) {
    double d;
    memcpy(&d, &b, 8);
    return d;
}
static uint64_t d_to_bits(double d) {
    uint64_t b;
    memcpy(&b, &d, 8);
    return b;
}

/* Natural log of a positive finite double: x = m * 2^e with m in
 * [sqrt(1/2), sqrt(2)), log m = 2 atanh(f), f = (m-1)/(m+1), atanh by its
 * Taylor series to f^21 (|f| <= 0.1716 so the tail is below 1e-17). */
double stgo_log(double x) {
    if (!(x > 0.0)) return -INFINITY;
    uint64_t b = d_to_bits(x);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) { /* subnormal: scale into the normal range */
        x = x * 0x1p54;
        b = d_to_bits(x);
        e = (int)((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = d_from_bits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e += 1;
    }
    double f = (m - 1.0) / (m + 1.0);
    double f2 = f * f;
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
    double tf = f + f;
    double lm = tf + tf * (f2 * s);
    double de = (double)e;
    return de * 0x1.62e42fee00000p-1 + (lm + de * 0x1.a39ef35793c76p-33);
}

/* IEEE-754 square root (correctly rounded, so every conforming machine
 * returns the same double). */
double stgo_sqrt(double x) {
    return x > 0.0 ? sqrt(x) : 0.0;
}

static const double LFACT_TABLE[1024] = {
#include "lfact_table.inc"
};

/* log(k!) : correctly rounded table below 1024, Stirling series of
 * lgamma(k+1) above. */
double stgo_lfact(int32_t k) {
    if (k < 1024) return LFACT_TABLE[k < 0 ? 0 : k];
    double x = (double)k + 1.0;
    double ix = 1.0 / x;
    double ix2 = ix * ix;
    double corr = ix * (1.0 / 12.0 - ix2 * (1.0 / 360.0 - ix2 * (1.0 / 1260.0)));
    return (x - 0.5) * stgo_log(x) - x + 0x1.d67f1c864beb5p-1 + corr;
}

/* exp(x): x = k ln2 + r, |r| <= ln2/2, e^r by Taylor to r^13, times 2^k;
 * the device evaluates the same operations in the same order */
double stgo_exp(double x) {
    if (!(x < 709.782712893384)) return x != x ? x : INFINITY;
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
    return ldexp(p, (int)kd);
}

static double ipow(double base, int32_t n) {
    double r = 1.0;
    while (n > 0) {
        if (n & 1) r = r * base;
        base = base * base;
        n >>= 1;
    }
    return r;
}

/* BINV: sequential inversion from k = 0, q <= 1/2, n*q < 30, written in
 * factorial-scaled form so the search loop has no division: with
 * U_x = u_x x! and R_x = f(x) x!, U_{x+1} = (U_x - R_x)(x+1) and
 * R_{x+1} = R_x (n-x) q/(1-q). */
static int32_t binom_inv(ostream *st, int32_t n, double q) {
    double r = ipow(1.0 - q, n);
    double s = q / (1.0 - q);
    double u = os_unif(st);
    int32_t x = 0;
    for (;;) {
        if (u <= r) return x;
        if (x >= n) return n;
        if (x >= 150) return x;
        x += 1;
        u = (u - r) * (double)x;
        r = r * ((double)(n - x + 1) * s);
    }
}

/* BTRS: transformed rejection with squeeze, q <= 1/2, n*q >= 30. */
static int32_t binom_btrs(ostream *st, int32_t n, double q) {
    double nd = (double)n;
    double spq = stgo_sqrt(nd * q * (1.0 - q));
    double b = 1.15 + 2.53 * spq;
    double ib = 1.0 / b;
    double a = -0.0873 + 0.0248 * b + 0.01 * q;
    double c = nd * q + 0.5;
    double alpha = (2.83 + 5.1 * ib) * spq;
    double vr = 0.92 - 4.2 * ib;
    double lpq = stgo_log(q / (1.0 - q));
    int32_t m = (int32_t)floor((nd + 1.0) * q);
    double h = stgo_lfact(m) + stgo_lfact(n - m);
    for (int it = 0; it < 1000; ++it) {
        double u = os_unif(st) - 0.5;
        double v = os_unif(st);
        double us = 0.5 - fabs(u);
        double kd = floor((2.0 * a / us + b) * u + c);
        if (kd < 0.0 || kd > nd) continue;
        int32_t k = (int32_t)kd;
        if (us >= 0.07 && v <= vr) return k;
        v = stgo_log(v * alpha / (a / (us * us) + b));
        if (v <= h - stgo_lfact(k) - stgo_lfact(n - k) + (double)(k - m) * lpq) return k;
    }
    return m;
}

/* Binomial(n, p); mirrors the role of RNG::getBinom (src/steps/rng/rng.cpp:516-527). */
static int32_t binom(ostream *st, int32_t n, double p) {
    if (n <= 0 || !(p > 0.0)) return 0;
    if (p >= 1.0) return n;
    int flip = p > 0.5;
    double q = flip ? 1.0 - p : p;
    int32_t k = ((double)n * q < 30.0) ? binom_inv(st, n, q) : binom_btrs(st, n, q);
    return flip ? n - k : k;
}

int32_t stgo_binom_stream(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t channel,
                          int32_t n, double p, int32_t reps, int32_t *out) {
    ostream st;
    os_init(&st, seed, c0, c1, c2, channel);
    for (int32_t i = 0; i < reps; ++i) out[i] = binom(&st, n, p);
    return 0;
}

void stgo_uniforms(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t channel,
                   int32_t n, double *out) {
    ostream st;
    os_init(&st, seed, c0, c1, c2, channel);
    for (int32_t i = 0; i < n; ++i) out[i] = os_unif(&st);
}

/* ----------------------------------------------------------------- schedule */

/* src/steps/util/collections.hpp:143-149 */
static int almost_equal(double x, double y) {
    return fabs(x - y) <= DBL_EPSILON * fabs(x + y) * 4 || fabs(x - y) < DBL_MIN;
}

int32_t stgo_schedule(double t0, double t_end, double dt_std, int32_t mode, double *dts,
                      int32_t cap) {
    int32_t n = 0;
    double t = t0;
    if (t_end < t0) return -2;
    if (mode == 0) {
        /* tetopsplit.cpp:1508-1560: periods of dt_std, the last clipped. */
        if (almost_equal(t0, t_end)) return 0;
        int aligned = 0;
        while (t < t_end && !aligned) {
            double per = dt_std;
            if (t + dt_std > t_end) {
                per = t_end - t;
                aligned = 1;
            }
            if (n >= cap) return -1;
            dts[n++] = per;
            t = t + per;
        }
        return n;
    }
    /* simulation.cpp:1358-1384 */
    double nstd_d = floor((t_end - t) / dt_std) - 1.0;
    long long n_std = nstd_d > 0.0 ? (long long)nstd_d : 0;
    for (long long i = 0; i < n_std; ++i) {
        double per = dt_std < t_end - t ? dt_std : t_end - t;
        if (n >= cap) return -1;
        dts[n++] = per;
        t = t + per;
    }
    double next = t + dt_std;
    if (!almost_equal(next, t_end) && next < t_end) {
        if (n + 2 > cap) return -1;
        dts[n++] = dt_std;
        t = t + dt_std;
        dts[n++] = t_end - t;
        t = t_end;
    } else if (!almost_equal(t_end, t)) {
        if (n >= cap) return -1;
        dts[n++] = t_end - t;
    }
    return n;
}

/* ------------------------------------------------------------------- model */

static double face_rate_sum(const double *r4) {
    return ((r4[0] + r4[1]) + r4[2]) + r4[3];
}

/* diffusion constant of species sp in tet t's compartment */
static double tet_dcst(const stgo_system *s, int32_t t, int32_t sp) {
    if (s->tet_comp) return s->comp_dcst[(size_t)s->tet_comp[t] * s->n_specs + sp];
    return s->spec_dcst[sp];
}

/* D_f g_f per face: faces on a diffusion boundary use the boundary's
 * constant (Diff::setDcst / setDirectionDcst, diff.cpp:340-460); a per-face
 * table (Diff::setDirectionDcst per tet, tetopsplit.cpp:4103-4164; inactive
 * rules 0, :4166-4188) replaces both */
static void face_rates(const stgo_system *s, int32_t t, int32_t sp, double r4[4]) {
    const double D = tet_dcst(s, t, sp);
    for (int f = 0; f < 4; ++f) {
        const int32_t d = s->tet_bdir ? s->tet_bdir[4 * t + f] : -1;
        double Df = d < 0 ? D : s->bnd_dcst[(size_t)d * s->n_specs + sp];
        if (s->tet_fD) Df = s->tet_fD[((size_t)sp * s->n_tets + t) * 4 + f];
        r4[f] = Df * s->tet_face_g[4 * t + f];
    }
}

static int tet_clamped(const stgo_system *s, int32_t t, int32_t sp) {
    return s->tet_clamp && ((s->tet_clamp[t] >> sp) & 1u);
}

static int tri_clamped(const stgo_system *s, int32_t tr, int32_t q) {
    return s->tri_clamp && ((s->tri_clamp[tr] >> q) & 1u);
}

/* surface diffusion constant of species q through edge e of triangle tr */
static double edge_dcst(const stgo_system *s, int32_t tr, int32_t q, int e) {
    if (s->tri_eD) return s->tri_eD[((size_t)q * s->n_tris + tr) * 3 + e];
    return s->sspec_dcst[q];
}

/* the species diffuses in some compartment */
static int spec_diffuses(const stgo_system *s, int32_t sp) {
    if (!s->tet_comp) return s->spec_dcst[sp] > 0.0;
    for (int32_t c = 0; c < s->n_comps; ++c)
        if (s->comp_dcst[(size_t)c * s->n_specs + sp] > 0.0) return 1;
    return 0;
}

double stgo_update_period(const stgo_system *s) {
    double mx = 0.0;
    for (int32_t t = 0; t < s->n_tets; ++t) {
        for (int32_t sp = 0; sp < s->n_specs; ++sp) {
            double r4[4];
            face_rates(s, t, sp, r4);
            double sum = face_rate_sum(r4);
            if (sum > mx) mx = sum;
        }
    }
    for (int32_t tr = 0; tr < s->n_tris && s->tri_nbr && s->sspec_dcst; ++tr) {
        for (int32_t q = 0; q < s->n_sspecs; ++q) {
            const double *g = s->tri_edge_g + 3 * tr;
            const double sum = ((edge_dcst(s, tr, q, 0) * g[0]) + (edge_dcst(s, tr, q, 1) * g[1])) +
                               (edge_dcst(s, tr, q, 2) * g[2]);
            if (sum > mx) mx = sum;
        }
    }
    return mx > DBL_EPSILON ? 1.0 / mx : INFINITY;
}

/* c = k * scale^-(order-1): comp_ccst, reac.cpp:44-55 (kcst * pow(vscale, -o1));
 * reactions.cpp:113-119, surface_reactions.cpp:176-203. Negative powers are
 * powers of inv = 1 / scale. */
static double scaled_const(double k, int32_t order, double scale) {
    int32_t o1 = order - 1;
    if (o1 == 0) return k;
    if (o1 < 0) return k * ipow(scale, -o1);
    return k * ipow(1.0 / scale, o1);
}

/* k(V) by linear interpolation in the tabulated rate constant:
 * vdepsreacdef.cpp:179-205 (getVDepK). Returns -1 when V lies outside the
 * table, which the reference reports as a ProgErr. */
static double vdep_const(const stgo_system *s, int32_t tab, double v) {
    const double vmin = s->vdep_vmin[tab], dv = s->vdep_dv[tab];
    const int32_t n = s->vdep_off[tab + 1] - s->vdep_off[tab];
    const double *K = s->vdep_k + s->vdep_off[tab];
    const double v2 = (v - vmin) / dv;
    if (!(v2 >= 0.0) || !(v2 <= (double)(n - 1))) return -1.0;
    const double lv = floor(v2);
    const int32_t l = (int32_t)lv, u = (int32_t)ceil(v2);
    const double r = v2 - lv;
    return (1.0 - r) * K[l] + r * K[u];
}

/* GHK flux of one open channel split for a fixed potential into A = P z F
 * nuFoRT, Bo = c_o e^-nuFoRT, Dn = 1 - e^-nuFoRT (math/ghk.cpp:62-81), so that
 * the flux is (A (c_i - Bo)) / Dn as the reference evaluates it. Returns 0
 * when nuFoRT is outside the range the reference accepts. */
static int ghk_consts(const stgo_system *s, int32_t g, double v, double *A, double *Bo, double *Dn) {
    const double z = (double)s->ghk_valence[g];
    const double nuFoRT = z * (v + s->ghk_vshift[g]) * FARADAY / (GAS_CONSTANT * s->temp);
    if (nuFoRT >= 308 * 2.30 || nuFoRT <= -307 * 2.30) return 0;
    const double e = stgo_exp(-nuFoRT);
    /* a current reading the outer tet, on a triangle without one: 0 */
    const double oconc = s->ghk_voconc[g] < 0.0 ? 0.0 : s->ghk_voconc[g] * 1.0e3;
    if (fabs(nuFoRT) > DBL_EPSILON) {
        *A = s->ghk_perm[g] * z * FARADAY * nuFoRT;
        *Bo = oconc * e;
        *Dn = 1.0 - e;
    } else {
        *A = s->ghk_perm[g] * z * FARADAY;
        *Bo = oconc;
        *Dn = 1.0;
    }
    return 1;
}

/* signed ions per second through one channel, positive = efflux
 * (ghkcurr.cpp:168-208); n_ion / vscale is the inner concentration in M */
static double ghk_rt(const stgo_system *s, int32_t g, double A, double Bo, double Dn, int32_t n_ion,
                     double vscale) {
    const double iconc = (double)n_ion / vscale * 1.0e3;
    const double flux = A * (iconc - Bo) / Dn;
    return flux / (E_CHARGE * (double)s->ghk_valence[g]);
}

/* h = prod over reactants of n(n-1)...(n-lhs+1): reactions.cpp:66-109 */
static double falling(double h, int32_t n, int32_t lhs) {
    for (int32_t i = lhs - 1; i >= 0; --i) h = h * (double)(n - i);
    return h;
}

typedef struct {
    int32_t *tet_grp; /* [n_tets] SSA group of a tet, -1 = its own; null when none */
    int32_t *grp_ptr; /* [groups+1] */
    int32_t *grp_mem; /* members, tet order (the first owns the group) */
    int32_t *tri_ptr; /* [n_tets+1] */
    int32_t *tri_idx; /* [n_tris]   */
    int32_t max_tris;
    uint8_t *tracked; /* [n_specs]  */
    uint8_t *tri_tracked; /* [n_sspecs] surface species that react and diffuse */
    double *tri_corr;     /* [n_sspecs][n_tris] occupancy correction in the period */
} topo;

static int sspec_diffuses(const stgo_system *s, int32_t q) {
    return s->sspec_dcst && s->tri_nbr && s->sspec_dcst[q] > 0.0;
}

static void groups_build(const stgo_system *s, topo *tp);

static void topo_build(const stgo_system *s, topo *tp) {
    tp->tri_ptr = (int32_t *)calloc((size_t)s->n_tets + 1, sizeof(int32_t));
    tp->tri_idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s->n_tris > 0 ? s->n_tris : 1));
    for (int32_t i = 0; i < s->n_tris; ++i) tp->tri_ptr[s->tri_tet[i] + 1]++;
    tp->max_tris = 0;
    for (int32_t t = 0; t < s->n_tets; ++t) {
        if (tp->tri_ptr[t + 1] > tp->max_tris) tp->max_tris = tp->tri_ptr[t + 1];
        tp->tri_ptr[t + 1] += tp->tri_ptr[t];
    }
    int32_t *fill = (int32_t *)calloc((size_t)s->n_tets, sizeof(int32_t));
    for (int32_t i = 0; i < s->n_tris; ++i) {
        int32_t t = s->tri_tet[i];
        tp->tri_idx[tp->tri_ptr[t] + fill[t]++] = i;
    }
    free(fill);
    /* occupancy is tracked for species that diffuse and are changed by a
     * reaction (reactions.cpp:35-40) */
    tp->tracked = (uint8_t *)calloc((size_t)s->n_specs, 1);
    for (int32_t sp = 0; sp < s->n_specs; ++sp) {
        if (!spec_diffuses(s, sp)) continue;
        for (int32_t r = 0; r < s->n_reacs; ++r)
            if (s->reac_upd[r * s->n_specs + sp] != 0) tp->tracked[sp] = 1;
        for (int32_t r = 0; r < s->n_sreacs; ++r)
            if (s->sreac_upd_i[r * s->n_specs + sp] != 0 ||
                (s->sreac_upd_o && s->sreac_upd_o[r * s->n_specs + sp] != 0))
                tp->tracked[sp] = 1;
        for (int32_t g = 0; g < s->n_ghk; ++g)
            if (s->ghk_ion[g] == sp && s->ghk_realflux[g]) tp->tracked[sp] = 1;
    }
    tp->tri_tracked = (uint8_t *)calloc((size_t)(s->n_sspecs > 0 ? s->n_sspecs : 1), 1);
    tp->tri_corr = (double *)calloc((size_t)(s->n_sspecs > 0 ? s->n_sspecs : 1) *
                                        (size_t)(s->n_tris > 0 ? s->n_tris : 1), sizeof(double));
    for (int32_t q = 0; q < s->n_sspecs; ++q) {
        if (!sspec_diffuses(s, q)) continue;
        for (int32_t r = 0; r < s->n_sreacs; ++r)
            if (s->sreac_upd_s[r * s->n_sspecs + q] != 0) tp->tri_tracked[q] = 1;
    }
    groups_build(s, tp);
}

/* does the model reach the outer side of the membrane? */
static int outer_model(const stgo_system *s) {
    for (int32_t i = 0; i < s->n_sreacs * s->n_specs; ++i)
        if ((s->sreac_lhs_o && s->sreac_lhs_o[i]) || (s->sreac_upd_o && s->sreac_upd_o[i])) return 1;
    for (int32_t g = 0; g < s->n_ghk; ++g)
        if (s->ghk_voconc[g] < 0.0) return 1;
    return 0;
}

static int32_t uf_find(int32_t *parent, int32_t x) {
    while (parent[x] != x) x = parent[x] = parent[parent[x]];
    return x;
}

/* SSA groups: a triangle with an outer tet joins it to the inner tet when
 * the model's surface reactions or GHK currents reach the outer side; the
 * reaction system of joined tets is one SSA (the reference's TetOpSplitP runs
 * one SSA over every kproc, tetopsplit.cpp:3217-3354; with diffusion frozen
 * in the period, groups that share no species evolve independently) */
static void groups_build(const stgo_system *s, topo *tp) {
    tp->tet_grp = NULL;
    tp->grp_ptr = NULL;
    tp->grp_mem = NULL;
    if (!s->tri_otet || !outer_model(s)) return;
    const int32_t N = s->n_tets;
    int32_t *parent = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    for (int32_t t = 0; t < N; ++t) parent[t] = t;
    int any = 0;
    for (int32_t tr = 0; tr < s->n_tris; ++tr) {
        const int32_t o = s->tri_otet[tr];
        if (o < 0) continue;
        int32_t a = uf_find(parent, s->tri_tet[tr]), b = uf_find(parent, o);
        if (a != b) parent[a > b ? a : b] = a < b ? a : b;
        any = 1;
    }
    if (!any) {
        free(parent);
        return;
    }
    int32_t *size = (int32_t *)calloc((size_t)N, sizeof(int32_t));
    for (int32_t t = 0; t < N; ++t) size[uf_find(parent, t)]++;
    int32_t *gid = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    int32_t ng = 0, nm = 0;
    for (int32_t t = 0; t < N; ++t) {
        gid[t] = -1;
        const int32_t root = uf_find(parent, t);
        if (size[root] < 2) continue;
        if (root == t) gid[t] = ng++;
        nm++;
    }
    tp->tet_grp = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    tp->grp_ptr = (int32_t *)calloc((size_t)ng + 1, sizeof(int32_t));
    tp->grp_mem = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nm > 0 ? nm : 1));
    for (int32_t t = 0; t < N; ++t) {
        const int32_t root = uf_find(parent, t);
        tp->tet_grp[t] = size[root] < 2 ? -1 : gid[root];
        if (tp->tet_grp[t] >= 0) tp->grp_ptr[tp->tet_grp[t] + 1]++;
    }
    for (int32_t g = 0; g < ng; ++g) tp->grp_ptr[g + 1] += tp->grp_ptr[g];
    int32_t *fill = (int32_t *)calloc((size_t)ng + 1, sizeof(int32_t));
    for (int32_t t = 0; t < N; ++t) {
        const int32_t g = tp->tet_grp[t];
        if (g >= 0) tp->grp_mem[tp->grp_ptr[g] + fill[g]++] = t;
    }
    free(fill);
    free(gid);
    free(size);
    free(parent);
}

static void topo_free(topo *tp) {
    free(tp->tet_grp);
    free(tp->grp_ptr);
    free(tp->grp_mem);
    free(tp->tri_ptr);
    free(tp->tri_idx);
    free(tp->tracked);
    free(tp->tri_tracked);
    free(tp->tri_corr);
}

/* SSA of one tet and its triangles over [0, dt): ssa_operator.cpp:24-47 with
 * the Direct group of propensities.hpp:543-586. Returns events or -1. */
static long ssa_tet(const stgo_system *s, const topo *tp, int32_t t, double dt, uint32_t step,
                    int32_t *cnt, double *corr, double *a, double *ps) {
    const int32_t S = s->n_specs, SS = s->n_sspecs;
    const int32_t R = s->n_reacs, RS = s->n_sreacs;
    const int32_t t0 = tp->tri_ptr[t], nt = tp->tri_ptr[t + 1] - t0;
    const int32_t NG = s->n_ghk, RT = RS + NG; /* entries per triangle */
    const int32_t nk = R + nt * RT;
    if (nk == 0) return 0;
    const double V = s->tet_vol[t];
    const double vscale = 1.0e3 * V * AVOGADRO;
    ostream st;
    const uint32_t ltet = s->tet_comp ? (uint32_t)s->tet_ltet[t] : (uint32_t)t;
    const uint32_t stream = s->tet_comp ? s->comp_stream[s->tet_comp[t]] : s->stream;
    const double *reac_k = s->tet_comp ? s->comp_reac_k + (size_t)s->tet_comp[t] * R : s->reac_k;
    os_init(&st, s->seed, ltet, step, stream, 0u);
    long nev = 0;
    double tcur = 0.0;
    for (;;) {
        /* propensities: volume reactions, then each triangle's sreacs */
        int32_t j = 0;
        for (int32_t r = 0; r < R; ++r, ++j) {
            const int32_t *lhs = s->reac_lhs + r * S;
            int32_t order = 0;
            for (int32_t sp = 0; sp < S; ++sp) order += lhs[sp];
            double h = 1.0;
            int zero = 0;
            for (int32_t sp = 0; sp < S && !zero; ++sp) {
                if (lhs[sp] == 0) continue;
                if (lhs[sp] > cnt[sp]) zero = 1;
                else h = falling(h, cnt[sp], lhs[sp]);
            }
            const double k = s->tet_k ? s->tet_k[(size_t)r * s->n_tets + t] : reac_k[r];
            a[j] = zero ? 0.0 : h * scaled_const(k, order, vscale);
        }
        for (int32_t i = 0; i < nt; ++i) {
            const int32_t tr = tp->tri_idx[t0 + i];
            for (int32_t r = 0; r < RS; ++r, ++j) {
                const int32_t *ls = s->sreac_lhs_s + r * SS;
                const int32_t *li = s->sreac_lhs_i + r * S;
                int32_t order = 0, nvol = 0;
                for (int32_t q = 0; q < SS; ++q) order += ls[q];
                for (int32_t q = 0; q < S; ++q) {
                    order += li[q];
                    nvol += li[q];
                }
                double h = 1.0;
                int zero = 0;
                for (int32_t q = 0; q < SS && !zero; ++q) {
                    if (ls[q] == 0) continue;
                    int32_t n = s->tri_counts[(size_t)q * s->n_tris + tr];
                    if (ls[q] > n) zero = 1;
                    else h = falling(h, n, ls[q]);
                }
                for (int32_t q = 0; q < S && !zero; ++q) {
                    if (li[q] == 0) continue;
                    if (li[q] > cnt[q]) zero = 1;
                    else h = falling(h, cnt[q], li[q]);
                }
                /* outer reactants, on a triangle outside every SSA group: none */
                for (int32_t q = 0; q < S && !zero && s->sreac_lhs_o; ++q)
                    if (s->sreac_lhs_o[r * S + q] > 0) zero = 1;
                if (zero) {
                    a[j] = 0.0;
                } else {
                    double scale = nvol == 0 ? s->tri_area[tr] * AVOGADRO : vscale;
                    /* the triangle's own constant, or a factor 1 / 0 on k(V); an
                     * inactive k(V) reaction reads no table (vdepsreac.cpp:217-230) */
                    const double f = s->tri_sk ? s->tri_sk[(size_t)r * s->n_tris + tr] : 1.0;
                    double k = s->tri_sk ? f : s->sreac_k[r];
                    if (s->sreac_vdep && s->sreac_vdep[r] >= 0) {
                        if (s->tri_sk && f == 0.0) {
                            k = 0.0;
                        } else {
                            k = vdep_const(s, s->sreac_vdep[r], s->tri_v[tr]);
                            if (k < 0.0) return -2;
                            if (s->tri_sk) k = k * f;
                        }
                    }
                    a[j] = h * scaled_const(k, order, scale);
                }
            }
            /* GHK currents: |rt| * open channels (ghkcurr.cpp:168-208) */
            for (int32_t g = 0; g < NG; ++g, ++j) {
                double A, Bo, Dn;
                if (!ghk_consts(s, g, s->tri_v[tr], &A, &Bo, &Dn)) return -3;
                if (s->tri_sk) A = A * s->tri_sk[(size_t)(RS + g) * s->n_tris + tr];
                const int32_t nch = s->tri_counts[(size_t)s->ghk_chan[g] * s->n_tris + tr];
                const double rt = ghk_rt(s, g, A, Bo, Dn, cnt[s->ghk_ion[g]], vscale);
                a[j] = fabs(rt) * (double)nch;
            }
        }
        double acc = 0.0;
        for (int32_t q = 0; q < nk; ++q) {
            acc = acc + a[q];
            ps[q] = acc;
        }
        const double a0 = ps[nk - 1];
        if (a0 < DBL_EPSILON) break;
        const double u1 = os_unif(&st);
        /* -log(u) >= 1 - u: a waiting time certainly past the period end
         * needs no logarithm */
        if (1.0 - u1 > a0 * (dt - tcur) * 1.0000001) break;
        const double u2 = os_unif(&st);
        const double tev = tcur + (-stgo_log(u1) / a0);
        if (tev > dt) break;
        tcur = tev;
        const double x = u2 * a0;
        int32_t sel = -1;
        for (int32_t q = 0; q < nk; ++q) {
            if (ps[q] > x) {
                sel = q;
                break;
            }
        }
        if (sel < 0) {
            for (int32_t q = nk - 1; q >= 0; --q)
                if (a[q] > 0.0) {
                    sel = q;
                    break;
                }
        }
        /* apply: reactions.cpp:124-141 / surface_reactions.cpp:158-172 */
        if (sel < R) {
            /* clamped species keep their counts: Reac::apply, reac.cpp:240-243 */
            const int32_t *upd = s->reac_upd + sel * S;
            for (int32_t sp = 0; sp < S; ++sp) {
                if (upd[sp] == 0 || tet_clamped(s, t, sp)) continue;
                cnt[sp] += upd[sp];
                if (tp->tracked[sp]) corr[sp] -= (double)upd[sp] * tcur;
            }
        } else if ((sel - R) % RT >= RS) {
            /* GHK: one ion in the direction of the flux, charge counted
             * outward positive (ghkcurr.cpp:212-272) */
            const int32_t i = (sel - R) / RT, g = (sel - R) % RT - RS;
            const int32_t tr = tp->tri_idx[t0 + i];
            const int32_t ion = s->ghk_ion[g];
            double A, Bo, Dn;
            ghk_consts(s, g, s->tri_v[tr], &A, &Bo, &Dn);
            if (s->tri_sk) A = A * s->tri_sk[(size_t)(RS + g) * s->n_tris + tr];
            const double rt = ghk_rt(s, g, A, Bo, Dn, cnt[ion], vscale);
            const int efflux = rt >= 0.0;
            if (s->ghk_realflux[g] && !tet_clamped(s, t, ion)) {
                const int32_t dv = efflux ? -1 : 1;
                cnt[ion] += dv;
                if (tp->tracked[ion]) corr[ion] -= (double)dv * tcur;
            }
            s->tri_charge[(size_t)g * s->n_tris + tr] += efflux ? s->ghk_valence[g] : -s->ghk_valence[g];
        } else {
            const int32_t i = (sel - R) / RT, r = (sel - R) % RT;
            const int32_t tr = tp->tri_idx[t0 + i];
            const int32_t *us_ = s->sreac_upd_s + r * SS;
            const int32_t *ui = s->sreac_upd_i + r * S;
            for (int32_t q = 0; q < SS; ++q) {
                if (us_[q] == 0 || tri_clamped(s, tr, q)) continue;
                s->tri_counts[(size_t)q * s->n_tris + tr] += us_[q];
                /* channel-count integral: n_end*dt - sum dn*t_event, the same
                 * value as the reference's piecewise sum (tri.cpp:544-561) */
                if ((s->tri_intg_mask >> q) & 1u)
                    s->tri_intg[(size_t)q * s->n_tris + tr] -= (double)us_[q] * tcur;
                if (tp->tri_tracked[q]) tp->tri_corr[(size_t)q * s->n_tris + tr] -= (double)us_[q] * tcur;
            }
            for (int32_t sp = 0; sp < S; ++sp) {
                if (ui[sp] == 0 || tet_clamped(s, t, sp)) continue;
                cnt[sp] += ui[sp];
                if (tp->tracked[sp]) corr[sp] -= (double)ui[sp] * tcur;
            }
        }
        if (++nev >= MAX_EVENTS_PER_PERIOD) return -1;
    }
    return nev;
}

/* ------------------------------------------------------------ SSA groups */

/* the realization constants of tet t: local index, stream, rate constants */
static uint32_t tet_ltet(const stgo_system *s, int32_t t) {
    return s->tet_comp ? (uint32_t)s->tet_ltet[t] : (uint32_t)t;
}

static uint32_t tet_stream(const stgo_system *s, int32_t t) {
    return s->tet_comp ? s->comp_stream[s->tet_comp[t]] : s->stream;
}

static double tet_reac_k(const stgo_system *s, int32_t t, int32_t r) {
    if (s->tet_k) return s->tet_k[(size_t)r * s->n_tets + t];
    return s->tet_comp ? s->comp_reac_k[(size_t)s->tet_comp[t] * s->n_reacs + r] : s->reac_k[r];
}

/* propensity of volume reaction r in tet t, counts read from the state
 * (reactions.cpp:66-119) */
static double grp_vol_prop(const stgo_system *s, int32_t t, int32_t r) {
    const int32_t S = s->n_specs, N = s->n_tets;
    const int32_t *lhs = s->reac_lhs + r * S;
    int32_t order = 0;
    for (int32_t sp = 0; sp < S; ++sp) order += lhs[sp];
    double h = 1.0;
    for (int32_t sp = 0; sp < S; ++sp) {
        if (lhs[sp] == 0) continue;
        const int32_t n = s->counts[(size_t)sp * N + t];
        if (lhs[sp] > n) return 0.0;
        h = falling(h, n, lhs[sp]);
    }
    return h * scaled_const(tet_reac_k(s, t, r), order, 1.0e3 * s->tet_vol[t] * AVOGADRO);
}

/* GHK constants with E = e^-nuFoRT (1 in the limit) in place of c_o e^-nuFoRT
 * (math/ghk.cpp:62-81), so that c_o can be read from the outer tet */
static int ghk_consts_e(const stgo_system *s, int32_t g, double v, double *A, double *E, double *Dn) {
    const double z = (double)s->ghk_valence[g];
    const double nuFoRT = z * (v + s->ghk_vshift[g]) * FARADAY / (GAS_CONSTANT * s->temp);
    if (nuFoRT >= 308 * 2.30 || nuFoRT <= -307 * 2.30) return 0;
    const double e = stgo_exp(-nuFoRT);
    if (fabs(nuFoRT) > DBL_EPSILON) {
        *A = s->ghk_perm[g] * z * FARADAY * nuFoRT;
        *E = e;
        *Dn = 1.0 - e;
    } else {
        *A = s->ghk_perm[g] * z * FARADAY;
        *E = 1.0;
        *Dn = 1.0;
    }
    return 1;
}

/* signed ions per second through one open channel of GHK current g on
 * triangle tr (inner tet t), the outer concentration from the outer tet when
 * ghk_voconc < 0 (ghkcurr.cpp:168-208); *bad on an out-of-range exponent */
static double grp_ghk_rt(const stgo_system *s, int32_t tr, int32_t t, int32_t g, int *bad) {
    const int32_t N = s->n_tets;
    double A, E, Dn;
    if (!ghk_consts_e(s, g, s->tri_v[tr], &A, &E, &Dn)) {
        *bad = 1;
        return 0.0;
    }
    if (s->tri_sk) A = A * s->tri_sk[(size_t)(s->n_sreacs + g) * s->n_tris + tr];
    const int32_t ion = s->ghk_ion[g];
    double oc = s->ghk_voconc[g] * 1.0e3;
    if (s->ghk_voconc[g] < 0.0) {
        const int32_t o = s->tri_otet ? s->tri_otet[tr] : -1;
        oc = o >= 0 ? (double)s->counts[(size_t)ion * N + o] / (1.0e3 * s->tet_vol[o] * AVOGADRO) * 1.0e3 : 0.0;
    }
    return ghk_rt(s, g, A, oc * E, Dn, s->counts[(size_t)ion * N + t], 1.0e3 * s->tet_vol[t] * AVOGADRO);
}

/* propensity of entry q of triangle tr (surface reaction q < n_sreacs, else
 * GHK current q - n_sreacs) with inner tet t: surface, inner and outer
 * reactants (SReac::rate, sreac.cpp:305-420); constant scaled by the area,
 * the inner or the outer tet's volume (sreac.cpp:150-180) */
static double grp_tri_prop(const stgo_system *s, int32_t tr, int32_t t, int32_t q, int *bad) {
    const int32_t S = s->n_specs, SS = s->n_sspecs, N = s->n_tets, RS = s->n_sreacs;
    if (q >= RS) {
        const int32_t g = q - RS;
        const int32_t nch = s->tri_counts[(size_t)s->ghk_chan[g] * s->n_tris + tr];
        return fabs(grp_ghk_rt(s, tr, t, g, bad)) * (double)nch;
    }
    const int32_t o = s->tri_otet ? s->tri_otet[tr] : -1;
    const int32_t *ls = s->sreac_lhs_s + q * SS;
    const int32_t *li = s->sreac_lhs_i + q * S;
    const int32_t *lo = s->sreac_lhs_o ? s->sreac_lhs_o + q * S : NULL;
    int32_t order = 0, nin = 0, nout = 0;
    double h = 1.0;
    for (int32_t x = 0; x < SS; ++x) {
        order += ls[x];
        if (ls[x] == 0) continue;
        const int32_t n = s->tri_counts[(size_t)x * s->n_tris + tr];
        if (ls[x] > n) return 0.0;
        h = falling(h, n, ls[x]);
    }
    for (int32_t x = 0; x < S; ++x) {
        order += li[x];
        nin += li[x];
        if (li[x] == 0) continue;
        const int32_t n = s->counts[(size_t)x * N + t];
        if (li[x] > n) return 0.0;
        h = falling(h, n, li[x]);
    }
    for (int32_t x = 0; x < S && lo; ++x) {
        order += lo[x];
        nout += lo[x];
        if (lo[x] == 0) continue;
        const int32_t n = o >= 0 ? s->counts[(size_t)x * N + o] : 0;
        if (lo[x] > n) return 0.0;
        h = falling(h, n, lo[x]);
    }
    const double f = s->tri_sk ? s->tri_sk[(size_t)q * s->n_tris + tr] : 1.0;
    double k = s->tri_sk ? f : s->sreac_k[q];
    if (s->sreac_vdep && s->sreac_vdep[q] >= 0) {
        if (s->tri_sk && f == 0.0) {
            k = 0.0;
        } else {
            k = vdep_const(s, s->sreac_vdep[q], s->tri_v[tr]);
            if (k < 0.0) {
                *bad = 2;
                return 0.0;
            }
            if (s->tri_sk) k = k * f;
        }
    }
    double scale;
    if (nin + nout == 0) scale = s->tri_area[tr] * AVOGADRO;
    else if (nout > 0) scale = 1.0e3 * s->tet_vol[o] * AVOGADRO;
    else scale = 1.0e3 * s->tet_vol[t] * AVOGADRO;
    return h * scaled_const(k, order, scale);
}

/* change the count of species sp in tet t by dv unless clamped there, with
 * the occupancy correction of tracked species (mol_state.hpp:87-200) */
static void grp_add(const stgo_system *s, const topo *tp, int32_t t, int32_t sp, int32_t dv, double tcur,
                    double *gcorr) {
    if (tet_clamped(s, t, sp)) return;
    s->counts[(size_t)sp * s->n_tets + t] += dv;
    if (tp->tracked[sp]) gcorr[(size_t)sp * s->n_tets + t] -= (double)dv * tcur;
}

/* Direct SSA of group g over [0, dt): the members' volume reactions in member
 * order, then each member's triangles' entries; counts in the state and
 * occupancy corrections in gcorr [n_specs][n_tets]. Returns events, or a
 * negative code as ssa_tet. */
static long ssa_grp(stgo_system *s, const topo *tp, int32_t g, double dt, uint32_t step, double *gcorr) {
    const int32_t N = s->n_tets, S = s->n_specs, SS = s->n_sspecs, R = s->n_reacs, RS = s->n_sreacs;
    const int32_t RT = RS + s->n_ghk;
    const int32_t m0 = tp->grp_ptr[g], m1 = tp->grp_ptr[g + 1];
    const int32_t owner = tp->grp_mem[m0];
    for (int32_t k = m0; k < m1; ++k) {
        const int32_t t = tp->grp_mem[k];
        for (int32_t sp = 0; sp < S; ++sp) gcorr[(size_t)sp * N + t] = 0.0;
        for (int32_t q = 0; q < SS; ++q)
            if (tp->tri_tracked[q])
                for (int32_t x = tp->tri_ptr[t]; x < tp->tri_ptr[t + 1]; ++x)
                    tp->tri_corr[(size_t)q * s->n_tris + tp->tri_idx[x]] = 0.0;
    }
    ostream st;
    os_init(&st, s->seed, tet_ltet(s, owner), step, tet_stream(s, owner), 0u);
    long nev = 0;
    double tcur = 0.0;
    int bad = 0;
    for (;;) {
        double a0 = 0.0;
        for (int32_t k = m0; k < m1; ++k)
            for (int32_t r = 0; r < R; ++r) a0 = a0 + grp_vol_prop(s, tp->grp_mem[k], r);
        for (int32_t k = m0; k < m1; ++k) {
            const int32_t t = tp->grp_mem[k];
            for (int32_t x = tp->tri_ptr[t]; x < tp->tri_ptr[t + 1]; ++x)
                for (int32_t q = 0; q < RT; ++q) a0 = a0 + grp_tri_prop(s, tp->tri_idx[x], t, q, &bad);
        }
        if (bad) return bad == 2 ? -2 : -3;
        if (a0 < DBL_EPSILON) break;
        const double u1 = os_unif(&st);
        if (1.0 - u1 > a0 * (dt - tcur) * 1.0000001) break;
        const double u2 = os_unif(&st);
        const double tev = tcur + (-stgo_log(u1) / a0);
        if (tev > dt) break;
        tcur = tev;
        const double x = u2 * a0;
        double acc = 0.0;
        int32_t st_t = -1, st_tr = -1, st_q = -1, la_t = -1, la_tr = -1, la_q = -1;
        for (int32_t k = m0; k < m1; ++k) {
            const int32_t t = tp->grp_mem[k];
            for (int32_t r = 0; r < R; ++r) {
                const double p = grp_vol_prop(s, t, r);
                acc = acc + p;
                if (st_t < 0 && acc > x) { st_t = t; st_tr = -1; st_q = r; }
                if (p > 0.0) { la_t = t; la_tr = -1; la_q = r; }
            }
        }
        for (int32_t k = m0; k < m1; ++k) {
            const int32_t t = tp->grp_mem[k];
            for (int32_t y = tp->tri_ptr[t]; y < tp->tri_ptr[t + 1]; ++y)
                for (int32_t q = 0; q < RT; ++q) {
                    const double p = grp_tri_prop(s, tp->tri_idx[y], t, q, &bad);
                    acc = acc + p;
                    if (st_t < 0 && acc > x) { st_t = t; st_tr = tp->tri_idx[y]; st_q = q; }
                    if (p > 0.0) { la_t = t; la_tr = tp->tri_idx[y]; la_q = q; }
                }
        }
        if (st_t < 0) { st_t = la_t; st_tr = la_tr; st_q = la_q; }
        if (st_t < 0) break;
        const int32_t t = st_t;
        if (st_tr < 0) {
            const int32_t *upd = s->reac_upd + st_q * S;
            for (int32_t sp = 0; sp < S; ++sp)
                if (upd[sp] != 0) grp_add(s, tp, t, sp, upd[sp], tcur, gcorr);
        } else {
            const int32_t tr = st_tr, o = s->tri_otet[tr];
            if (st_q >= RS) {
                /* GHK: one ion between the inner and the outer tet
                 * (ghkcurr.cpp:228-272), charge counted outward positive */
                const int32_t gq = st_q - RS, ion = s->ghk_ion[gq];
                const int efflux = grp_ghk_rt(s, tr, t, gq, &bad) >= 0.0;
                if (s->ghk_realflux[gq]) {
                    grp_add(s, tp, t, ion, efflux ? -1 : 1, tcur, gcorr);
                    if (s->ghk_voconc[gq] < 0.0 && o >= 0) grp_add(s, tp, o, ion, efflux ? 1 : -1, tcur, gcorr);
                }
                s->tri_charge[(size_t)gq * s->n_tris + tr] += efflux ? s->ghk_valence[gq] : -s->ghk_valence[gq];
            } else {
                /* SReac::apply, sreac.cpp:440-505 */
                const int32_t *us_ = s->sreac_upd_s + st_q * SS;
                const int32_t *ui = s->sreac_upd_i + st_q * S;
                const int32_t *uo = s->sreac_upd_o ? s->sreac_upd_o + st_q * S : NULL;
                for (int32_t q = 0; q < SS; ++q) {
                    if (us_[q] == 0 || tri_clamped(s, tr, q)) continue;
                    s->tri_counts[(size_t)q * s->n_tris + tr] += us_[q];
                    if ((s->tri_intg_mask >> q) & 1u) s->tri_intg[(size_t)q * s->n_tris + tr] -= (double)us_[q] * tcur;
                    if (tp->tri_tracked[q]) tp->tri_corr[(size_t)q * s->n_tris + tr] -= (double)us_[q] * tcur;
                }
                for (int32_t sp = 0; sp < S; ++sp)
                    if (ui[sp] != 0) grp_add(s, tp, t, sp, ui[sp], tcur, gcorr);
                for (int32_t sp = 0; sp < S && uo && o >= 0; ++sp)
                    if (uo[sp] != 0) grp_add(s, tp, o, sp, uo[sp], tcur, gcorr);
            }
        }
        if (++nev >= MAX_EVENTS_PER_PERIOD) return -1;
    }
    return nev;
}

/* diffusion leaving of one (tet, species): diffusion_operator.cpp:59-130, 222-238 */
static int32_t leave_tet_species(const stgo_system *s, int32_t t, int32_t sp, double dt,
                                 uint32_t step, int32_t *cnt, double corr_sp, int tracked,
                                 int32_t *out4) {
    out4[0] = out4[1] = out4[2] = out4[3] = 0;
    const int32_t n = cnt[sp];
    double r4[4];
    face_rates(s, t, sp, r4);
    const double sum = face_rate_sum(r4);
    if (n <= 0 || !(sum > 0.0)) return 0;
    ostream st;
    const uint32_t ltet = s->tet_comp ? (uint32_t)s->tet_ltet[t] : (uint32_t)t;
    const uint32_t stream = s->tet_comp ? s->comp_stream[s->tet_comp[t]] : s->stream;
    os_init(&st, s->seed, ltet, step, stream, 1u + (uint32_t)sp);
    double occ = (double)n;
    if (tracked && !almost_equal(dt, 0.0)) occ = (double)n + corr_sp / dt;
    const double u = os_unif(&st);
    const double fl = floor(occ);
    int32_t mp = (int32_t)fl + ((occ - fl) > u ? 1 : 0);
    if (mp > n) mp = n;
    if (mp < 0) mp = 0;
    double p = dt * sum;
    if (p > 1.0) p = 1.0;
    if (p < 0.0) p = 0.0;
    const int32_t nl = binom(&st, mp, p);
    if (nl <= 0) return 0;
    /* a clamped source keeps its count (Diff::apply, diff.cpp:476-520, 577) */
    if (!tet_clamped(s, t, sp)) cnt[sp] = n - nl;
    int lastf = -1;
    for (int f = 0; f < 4; ++f)
        if (r4[f] > 0.0) lastf = f;
    if (nl <= s->diff_threshold) {
        for (int32_t m = 0; m < nl; ++m) {
            int chosen = -1;
            for (int tries = 0; tries < 64 && chosen < 0; ++tries) {
                const double selv = os_unif(&st) * sum;
                double cum = 0.0;
                for (int f = 0; f < 4; ++f) {
                    cum = cum + r4[f];
                    if (selv < cum) {
                        chosen = f;
                        break;
                    }
                }
            }
            if (chosen < 0) chosen = lastf;
            out4[chosen] += 1;
        }
    } else {
        int32_t rem = nl;
        double rsum = sum;
        for (int f = 0; f < 4; ++f) {
            if (!(r4[f] > 0.0)) continue;
            int32_t k;
            if (f == lastf) {
                k = rem;
            } else {
                const double pf = r4[f] / rsum;
                k = pf >= 1.0 ? rem : binom(&st, rem, pf);
                rsum = rsum - r4[f];
            }
            out4[f] = k;
            rem -= k;
        }
    }
    return nl;
}

/* surface diffusion out of one (triangle, surface species): the volume's
 * leaving pass with three edges (SDiff block of tetopsplit.cpp:1716-1788) */
static int32_t leave_tri_species(const stgo_system *s, const topo *tp, int32_t tr, int32_t q,
                                 double dt, uint32_t step, int32_t *out3) {
    out3[0] = out3[1] = out3[2] = 0;
    const size_t ix = (size_t)q * s->n_tris + tr;
    const int32_t n = s->tri_counts[ix];
    double r3[3];
    for (int e = 0; e < 3; ++e) r3[e] = edge_dcst(s, tr, q, e) * s->tri_edge_g[3 * tr + e];
    const double sum = (r3[0] + r3[1]) + r3[2];
    if (n <= 0 || !(sum > 0.0)) return 0;
    const int32_t t = s->tri_tet[tr];
    const uint32_t utri = s->tri_ltri ? (uint32_t)s->tri_ltri[tr] : (uint32_t)tr;
    const uint32_t stream = s->tet_comp ? s->comp_stream[s->tet_comp[t]] : s->stream;
    ostream st;
    os_init(&st, s->seed, utri, step, stream, 0x40u + (uint32_t)q);
    double occ = (double)n;
    if (tp->tri_tracked[q] && !almost_equal(dt, 0.0)) occ = (double)n + tp->tri_corr[ix] / dt;
    const double u = os_unif(&st);
    const double fl = floor(occ);
    int32_t mp = (int32_t)fl + ((occ - fl) > u ? 1 : 0);
    if (mp > n) mp = n;
    if (mp < 0) mp = 0;
    double p = dt * sum;
    if (p > 1.0) p = 1.0;
    if (p < 0.0) p = 0.0;
    const int32_t nl = binom(&st, mp, p);
    if (nl <= 0) return 0;
    if (!tri_clamped(s, tr, q)) s->tri_counts[ix] = n - nl;
    int lastf = -1;
    for (int e = 0; e < 3; ++e)
        if (r3[e] > 0.0) lastf = e;
    if (nl <= s->diff_threshold) {
        for (int32_t m = 0; m < nl; ++m) {
            int chosen = -1;
            for (int tries = 0; tries < 64 && chosen < 0; ++tries) {
                const double selv = os_unif(&st) * sum;
                double cum = 0.0;
                for (int e = 0; e < 3; ++e) {
                    cum = cum + r3[e];
                    if (selv < cum) {
                        chosen = e;
                        break;
                    }
                }
            }
            if (chosen < 0) chosen = lastf;
            out3[chosen] += 1;
        }
    } else {
        int32_t rem = nl;
        double rsum = sum;
        for (int e = 0; e < 3; ++e) {
            if (!(r3[e] > 0.0)) continue;
            int32_t k;
            if (e == lastf) {
                k = rem;
            } else {
                const double pf = r3[e] / rsum;
                k = pf >= 1.0 ? rem : binom(&st, rem, pf);
                rsum = rsum - r3[e];
            }
            out3[e] = k;
            rem -= k;
        }
    }
    return nl;
}

/* after the SSA of tet t (counts cnt, occupancy corrections corr): the
 * channel-count integrals of its triangles, then the leaving pass of its
 * species and of its triangles' surface species; writes the counts back and
 * returns the molecules moved */
static uint64_t finish_tet(stgo_system *s, topo *tp, int32_t t, double dt, uint32_t step, int32_t *cnt,
                           const double *corr, int32_t *scratch, int32_t *tout) {
    const int32_t N = s->n_tets, S = s->n_specs, SS = s->n_sspecs, NTR = s->n_tris;
    uint64_t mv = 0;
    if (s->tri_intg_mask != 0) {
        for (int32_t q = 0; q < SS; ++q) {
            if (!((s->tri_intg_mask >> q) & 1u)) continue;
            for (int32_t k = tp->tri_ptr[t]; k < tp->tri_ptr[t + 1]; ++k) {
                const size_t ix = (size_t)q * NTR + tp->tri_idx[k];
                s->tri_intg[ix] += (double)s->tri_counts[ix] * dt;
            }
        }
    }
    for (int32_t sp = 0; sp < S; ++sp) {
        int32_t *o4 = scratch + ((size_t)sp * N + t) * 4;
        if (spec_diffuses(s, sp))
            mv += (uint64_t)leave_tet_species(s, t, sp, dt, step, cnt, corr[sp], tp->tracked[sp], o4);
        else
            o4[0] = o4[1] = o4[2] = o4[3] = 0;
        s->counts[(size_t)sp * N + t] = cnt[sp];
    }
    for (int32_t k = tp->tri_ptr[t]; k < tp->tri_ptr[t + 1]; ++k) {
        const int32_t tr = tp->tri_idx[k];
        for (int32_t q = 0; q < SS; ++q) {
            int32_t *o3 = tout + ((size_t)q * NTR + tr) * 3;
            if (sspec_diffuses(s, q))
                mv += (uint64_t)leave_tri_species(s, tp, tr, q, dt, step, o3);
            else
                o3[0] = o3[1] = o3[2] = 0;
        }
    }
    return mv;
}

int32_t stgo_period(stgo_system *s, double dt, uint32_t step, int32_t *scratch,
                    uint64_t *n_events, uint64_t *n_moved) {
    return stgo_run(s, &dt, 1, step, scratch, n_events, n_moved);
}

int32_t stgo_run(stgo_system *s, const double *dts, int32_t n_periods, uint32_t step0,
                 int32_t *scratch, uint64_t *n_events, uint64_t *n_moved) {
    const int32_t N = s->n_tets, S = s->n_specs;
    topo tp;
    topo_build(s, &tp);
    const int32_t nk_max = s->n_reacs + tp.max_tris * (s->n_sreacs + s->n_ghk) + 1;
    int32_t err = 0;
    uint64_t ev_total = 0, mv_total = 0;
    const int32_t NTR = s->n_tris, SS = s->n_sspecs;
    int32_t *tout = (int32_t *)calloc((size_t)(SS > 0 ? SS : 1) * (size_t)(NTR > 0 ? NTR : 1) * 3,
                                      sizeof(int32_t));
    double *gcorr = tp.tet_grp ? (double *)calloc((size_t)(S > 0 ? S : 1) * (size_t)N, sizeof(double)) : NULL;
    for (int32_t p = 0; p < n_periods && !err; ++p) {
        const double dt = dts[p];
        const uint32_t step = step0 + (uint32_t)p;
        /* pass 1: SSA + leaving, one tet at a time (independent groups) */
#pragma omp parallel reduction(+ : ev_total, mv_total) reduction(| : err)
        {
            int32_t *cnt = (int32_t *)malloc(sizeof(int32_t) * (size_t)S);
            double *corr = (double *)malloc(sizeof(double) * (size_t)S);
            double *a = (double *)malloc(sizeof(double) * (size_t)nk_max);
            double *ps = (double *)malloc(sizeof(double) * (size_t)nk_max);
#pragma omp for schedule(dynamic, 256)
            for (int32_t t = 0; t < N; ++t) {
                const int32_t grp = tp.tet_grp ? tp.tet_grp[t] : -1;
                if (grp >= 0) {
                    /* an SSA group: its first tet runs the SSA and every
                     * member's leaving pass; the other members do nothing */
                    if (tp.grp_mem[tp.grp_ptr[grp]] != t) continue;
                    long nev = ssa_grp(s, &tp, grp, dt, step, gcorr);
                    if (nev < 0) {
                        err |= nev == -2 ? 2 : nev == -3 ? 4 : 1;
                        nev = 0;
                    }
                    ev_total += (uint64_t)nev;
                    for (int32_t k = tp.grp_ptr[grp]; k < tp.grp_ptr[grp + 1]; ++k) {
                        const int32_t mt = tp.grp_mem[k];
                        for (int32_t sp = 0; sp < S; ++sp) {
                            cnt[sp] = s->counts[(size_t)sp * N + mt];
                            corr[sp] = gcorr[(size_t)sp * N + mt];
                        }
                        mv_total += finish_tet(s, &tp, mt, dt, step, cnt, corr, scratch, tout);
                    }
                    continue;
                }
                for (int32_t sp = 0; sp < S; ++sp) {
                    cnt[sp] = s->counts[(size_t)sp * N + t];
                    corr[sp] = 0.0;
                }
                for (int32_t q = 0; q < SS; ++q)
                    if (tp.tri_tracked[q])
                        for (int32_t k = tp.tri_ptr[t]; k < tp.tri_ptr[t + 1]; ++k)
                            tp.tri_corr[(size_t)q * NTR + tp.tri_idx[k]] = 0.0;
                long nev = ssa_tet(s, &tp, t, dt, step, cnt, corr, a, ps);
                if (nev < 0) {
                    err |= nev == -2 ? 2 : nev == -3 ? 4 : 1;
                    nev = 0;
                }
                ev_total += (uint64_t)nev;
                mv_total += finish_tet(s, &tp, t, dt, step, cnt, corr, scratch, tout);
            }
            free(cnt);
            free(corr);
            free(a);
            free(ps);
        }
        /* pass 2: entering, gather of the neighbour's slot (diffusion_operator.cpp:136-220) */
#pragma omp parallel for schedule(static)
        for (int32_t t = 0; t < N; ++t) {
            for (int32_t sp = 0; sp < S; ++sp) {
                if (!spec_diffuses(s, sp)) continue;
                int32_t add = 0;
                for (int f = 0; f < 4; ++f) {
                    const int32_t nb = s->tet_nbr[4 * t + f];
                    if (nb < 0) continue;
                    add += scratch[((size_t)sp * N + nb) * 4 + s->tet_nbr_face[4 * t + f]];
                }
                /* nothing enters a clamped destination (diff.cpp:500-503, 548-551, 570-573) */
                if (!tet_clamped(s, t, sp)) s->counts[(size_t)sp * N + t] += add;
            }
        }
        /* surface molecules entering across triangle edges */
#pragma omp parallel for schedule(static)
        for (int32_t tr = 0; tr < NTR; ++tr) {
            for (int32_t q = 0; q < SS; ++q) {
                if (!sspec_diffuses(s, q)) continue;
                int32_t add = 0;
                for (int e = 0; e < 3; ++e) {
                    const int32_t nb = s->tri_nbr[3 * tr + e];
                    if (nb < 0) continue;
                    add += tout[((size_t)q * NTR + nb) * 3 + s->tri_nbr_edge[3 * tr + e]];
                }
                if (!tri_clamped(s, tr, q)) s->tri_counts[(size_t)q * NTR + tr] += add;
            }
        }
    }
    free(tout);
    free(gcorr);
    topo_free(&tp);
    if (n_events) *n_events += ev_total;
    if (n_moved) *n_moved += mv_total;
    /* -1: event cap hit; -2: a triangle potential outside a k(V) table;
     * -3: GHK exponent out of range */
    return (err & 1) ? -1 : (err & 2) ? -2 : (err & 4) ? -3 : 0;
}
