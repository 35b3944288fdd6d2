/*
 * CPU oracle for the operator-split stochastic reaction-diffusion period.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * algorithm used as the checker for the HIP path (and as the CPU baseline in
 * bench.py). The product never links, loads or calls it.
 *
 * Every function in stg_oracle.c cites the reference file:line it restates.
 * Random draws come from the same keyed Philox4x32-8 streams as the device
 * path, so the oracle and the GPU must agree bit-for-bit on integer counts.
 */
#ifndef STG_ORACLE_H
#define STG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct stgo_system {
    /* mesh (one realization) */
    int32_t n_tets;
    const double  *tet_vol;      /* [n_tets]                 m^3                 */
    const int32_t *tet_nbr;      /* [n_tets*4] neighbour tet, -1 = none          */
    const int32_t *tet_nbr_face; /* [n_tets*4] slot of this face in the nbr      */
    const double  *tet_face_g;   /* [n_tets*4] A_f/(V*d_f) m^-2, 0 = closed face */
    int32_t n_tris;
    const int32_t *tri_tet;      /* [n_tris] inner tet of a membrane triangle    */
    const double  *tri_area;     /* [n_tris] m^2                                 */
    /* model */
    int32_t n_specs, n_reacs, n_sspecs, n_sreacs;
    const int32_t *reac_lhs;     /* [n_reacs*n_specs] reactant multiplicity      */
    const int32_t *reac_upd;     /* [n_reacs*n_specs] net change                 */
    const double  *reac_k;       /* [n_reacs] macroscopic constant (SI)          */
    const double  *spec_dcst;    /* [n_specs] m^2/s, 0 = immobile                */
    const int32_t *sreac_lhs_s;  /* [n_sreacs*n_sspecs]                          */
    const int32_t *sreac_upd_s;  /* [n_sreacs*n_sspecs]                          */
    const int32_t *sreac_lhs_i;  /* [n_sreacs*n_specs] inner-volume reactants    */
    const int32_t *sreac_upd_i;  /* [n_sreacs*n_specs]                           */
    const double  *sreac_k;      /* [n_sreacs]                                   */
    /* state, mutated in place */
    int32_t *counts;             /* [n_specs][n_tets]   species-major SoA        */
    int32_t *tri_counts;         /* [n_sspecs][n_tris]                           */
    /* Philox key and stream id */
    uint64_t seed;
    uint32_t stream;
    /* leaving counts up to this are split one molecule at a time
     * (diffusion_operator.hpp:99 default 10, setDiffApplyThreshold) */
    int32_t diff_threshold;
    /* voltage-dependent surface reactions (may all be null / 0): sreac_vdep[r]
     * is the k(V) table of sreac r or -1; table t spans vdep_k[vdep_off[t] ..
     * vdep_off[t+1]) at V = vdep_vmin[t] + i*vdep_dv[t]; tri_v [n_tris] volts */
    const int32_t *sreac_vdep;
    const double  *vdep_vmin;
    const double  *vdep_dv;
    const int32_t *vdep_off;
    const double  *vdep_k;
    const double  *tri_v;
    /* time integral of the counts of the surface species in tri_intg_mask
     * (bit q = species q), [n_sspecs][n_tris], for time-averaged ohmic
     * channel counts (tri.cpp:544-561, 664-693); may be null when the mask is 0 */
    uint32_t tri_intg_mask;
    double  *tri_intg;
    /* GHK currents (ghkcurr.cpp:168-272) with a virtual outer concentration;
     * tri_charge [n_ghk][n_tris] accumulates the charge each current moved
     * out (elementary charges); may be null when n_ghk is 0 */
    int32_t n_ghk;
    const int32_t *ghk_chan;     /* [n_ghk] surface species */
    const int32_t *ghk_ion;      /* [n_ghk] volume species */
    const int32_t *ghk_valence;
    const double  *ghk_perm;     /* m^3/s */
    const double  *ghk_voconc;   /* M */
    const double  *ghk_vshift;   /* V */
    const int32_t *ghk_realflux;
    double temp;                 /* K */
    int32_t *tri_charge;
    /* compartments (all null / 0 for one compartment): tet t is in
     * compartment tet_comp[t], draws from stream comp_stream[c] with local
     * index tet_ltet[t], reacts with comp_reac_k[c][r] and diffuses with
     * comp_dcst[c][s]; a face f with tet_bdir[4t+f] = d >= 0 lies on a
     * diffusion boundary and uses bnd_dcst[d][s] instead (tetopsplit.cpp:
     * 2745-2931, diff.cpp:340-390) */
    int32_t n_comps;
    const int32_t *tet_comp;     /* [n_tets] */
    const int32_t *tet_ltet;     /* [n_tets] */
    const uint32_t *comp_stream; /* [n_comps] */
    const double *comp_reac_k;   /* [n_comps*n_reacs] */
    const double *comp_dcst;     /* [n_comps*n_specs] */
    const int32_t *tet_bdir;     /* [n_tets*4], or null */
    const double *bnd_dcst;      /* [directions*n_specs] */
    /* surface diffusion (sdiff.cpp:42-90, tetopsplit.cpp:1716-1788), null /
     * zero when no surface species diffuses: edge e of triangle i leads to
     * tri_nbr[3i+e] (its slot tri_nbr_edge[3i+e]) with l_e/(A d_e) =
     * tri_edge_g[3i+e]; triangle i draws with local index tri_ltri[i] (or i) */
    const int32_t *tri_nbr;
    const int32_t *tri_nbr_edge;
    const double  *tri_edge_g;
    const double  *sspec_dcst;   /* [n_sspecs] */
    const int32_t *tri_ltri;
    /* per-element tables, each null when unused (the per-tet and per-triangle
     * setters of TetOpSplitP, tetopsplit.cpp:2189-2437, 3951-4166, 4464):
     * tet_k [n_reacs][n_tets] rate constant of each reaction in each tet (0 =
     * inactive); tet_fD [n_specs][n_tets][4] diffusion constant through each
     * face (replaces D and boundary constants); tet_clamp [n_tets] bit s =
     * species s clamped; tri_sk [n_sreacs + n_ghk][n_tris] constant of each
     * surface reaction on each triangle (a factor on k(V) for a
     * voltage-dependent one) then a factor on each GHK current; tri_eD
     * [n_sspecs][n_tris][3] surface diffusion constant through each edge;
     * tri_clamp [n_tris] bit q = surface species q clamped */
    const double  *tet_k;
    const double  *tet_fD;
    const uint32_t *tet_clamp;
    const double  *tri_sk;
    const double  *tri_eD;
    const uint32_t *tri_clamp;
    /* outer side of the membrane (null when unused): outer-compartment
     * reactants and net changes of each surface reaction [n_sreacs*n_specs]
     * (sreac.cpp:305-505), and the outer tet of each triangle [n_tris] (-1
     * none). A GHK current with ghk_voconc < 0 reads the outer tet's
     * concentration (ghkcurr.cpp:176-182). Tets joined by a triangle whose
     * reactions reach both sides form one SSA group (see stgo_run). */
    const int32_t *sreac_lhs_o;
    const int32_t *sreac_upd_o;
    const int32_t *tri_otet;
} stgo_system;

/* Threads the OpenMP loops of stgo_run use. */
int32_t stgo_threads(void);

/* Philox4x32 with 8 rounds (Random123 philox4x32_R(8, ctr, key)). */
void stgo_philox4x32_8(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Deterministic scalar helpers (exposed for the tests). */
double stgo_log(double x);
double stgo_exp(double x);
double stgo_sqrt(double x);
double stgo_lfact(int32_t k);
/* Binomial(n, p) drawn from the stream (seed, ctr0..2, channel). */
int32_t stgo_binom_stream(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t channel,
                          int32_t n, double p, int32_t reps, int32_t *out);
/* Uniforms in (0,1) from one stream, for KS checks. */
void stgo_uniforms(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t channel,
                   int32_t n, double *out);

/* Split period schedule. mode 0 = TetOpSplit (clip the last period),
 * mode 1 = DistTetOpSplit (n-1 standard periods then one or two remainders).
 * Returns the number of periods written to dts (at most cap), or -1 if cap is
 * too small. */
int32_t stgo_schedule(double t0, double t_end, double dt_std, int32_t mode, double *dts,
                      int32_t cap);

/* Stability-bound split period 1 / max_{tet,species} sum_f D_s g_f
 * (0 if nothing diffuses). */
double stgo_update_period(const stgo_system *s);

/* One operator-split period: per-tet SSA over [0, dt) then multinomial
 * diffusion (leaving pass, then entering pass). scratch must hold
 * n_specs*n_tets*4 int32. Returns 0, or a negative code on a model error. */
int32_t stgo_period(stgo_system *s, double dt, uint32_t step, int32_t *scratch,
                    uint64_t *n_events, uint64_t *n_moved);

/* Run periods from a precomputed schedule; step0 is the first period index. */
int32_t stgo_run(stgo_system *s, const double *dts, int32_t n_periods, uint32_t step0,
                 int32_t *scratch, uint64_t *n_events, uint64_t *n_moved);

#ifdef __cplusplus
}
#endif
#endif
