/*
 * stg -- C ABI of the MI355X operator-split stochastic reaction-diffusion
 * engine (drop-in for the STEPS TetOpSplit / DistTetOpSplit solver engines).
 *
 * Every entry point returns an int status (STG_OK on success); on failure
 * stg_last_error(sim) (or stg_last_error(NULL) before a simulation exists)
 * returns a message. Status codes map one-to-one onto the reference's C++
 * exception types (src/steps/util/error.hpp:161-246) so a host shim can
 * rethrow steps::ArgErr, steps::NotImplErr, steps::ProgErr, steps::SysErr.
 *
 * All arrays are plain host pointers; the library copies what it needs at
 * create time and keeps every per-simulation buffer resident on the current
 * HIP device. One host thread per handle; calls are serialised on the
 * handle's own HIP stream.
 */
#ifndef STG_H
#define STG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STG_OK 0
#define STG_ERR_ARG 1     /* steps::ArgErr     */
#define STG_ERR_NOTIMPL 2 /* steps::NotImplErr */
#define STG_ERR_PROG 3    /* steps::ProgErr    */
#define STG_ERR_SYS 4     /* steps::SysErr     */
#define STG_ERR_DEVICE 5  /* HIP runtime / device failure */

/* Period schedule of stg_run: the last period clipped to the end time as in
 * TetOpSplitP::_runWithoutEField (src/steps/mpi/tetopsplit/tetopsplit.cpp:1528-1560),
 * or n-1 standard periods plus one or two remainders as in
 * OmegaHSimulation::run_rd (src/steps/mpi/dist/tetopsplit/simulation.cpp:1358-1384). */
#define STG_MODE_TETOPSPLIT 0
#define STG_MODE_DISTTETOPSPLIT 1

typedef struct stg_sim stg_sim;

/* Tetrahedral mesh of one realization, whole mesh resident on one GPU.
 * Face f of tet t is the face opposite its f-th vertex. Replaces the
 * neighbour tables of TetOpSplitP::_setup (tetopsplit.cpp:645-690) and the
 * DistMesh neighbour CSR (src/steps/geom/dist/distmesh.cpp:344-420). */
typedef struct stg_mesh {
    int32_t n_tets;
    const double *tet_vol;       /* [n_tets] m^3 */
    const int32_t *tet_nbr;      /* [n_tets*4] neighbour tet, -1 = none */
    const int32_t *tet_nbr_face; /* [n_tets*4] face slot of this face in the neighbour */
    const double *tet_face_g;    /* [n_tets*4] A_f / (V d_f) in m^-2, 0 = closed */
    int32_t n_tris;              /* membrane triangles carrying surface reactions */
    const int32_t *tri_tet;      /* [n_tris] inner tet */
    const double *tri_area;      /* [n_tris] m^2 */
    /* Compartments created with stg_create_comps only (NULL otherwise): */
    const int32_t *tet_nbr_comp; /* [n_tets*4] compartment of the neighbour; tet_nbr indexes
                                    that compartment's mesh */
    const int32_t *tet_face_bdir;/* [n_tets*4] -1, or the diffusion-boundary direction d whose
                                    per-species constant (stg_set_boundary_dcst) opens the face */
    /* Surface diffusion (SDiff, src/steps/mpi/tetopsplit/sdiff.cpp:42-90); NULL when no
     * surface species diffuses. Edge e of triangle i joins it to tri_nbr[3i+e] (-1 none),
     * whose slot for that edge is tri_nbr_edge[3i+e]; tri_edge_g = l_e / (A d_e), m^-2. */
    const int32_t *tri_nbr;      /* [n_tris*3] */
    const int32_t *tri_nbr_edge; /* [n_tris*3] */
    const double *tri_edge_g;    /* [n_tris*3] */
    /* Outer side of each triangle (Tri::oTet, src/steps/mpi/tetopsplit/tri.hpp),
     * NULL when no triangle has one: tri_otet[i] is the tet across the
     * membrane (-1 none) in compartment tri_ocomp[i] (NULL: this one). A
     * triangle whose surface reactions or GHK currents reach the outer
     * compartment joins its inner and outer tets into one SSA group. */
    const int32_t *tri_otet;     /* [n_tris] */
    const int32_t *tri_ocomp;    /* [n_tris] */
    /* Compartment of each edge neighbour (stg_create_comps only; NULL: this
     * one), for surface diffusion across patches of different compartments. */
    const int32_t *tri_nbr_comp; /* [n_tris*3] */
} stg_mesh;

/* Flattened model of one compartment (+ one patch), as the reference's
 * Compdef/Patchdef tables (src/steps/solver/compdef.hpp:252-258, 420). */
typedef struct stg_model {
    int32_t n_specs;             /* volume species */
    int32_t n_reacs;
    const int32_t *reac_lhs;     /* [n_reacs*n_specs] reactant multiplicity, order <= 4 each */
    const int32_t *reac_upd;     /* [n_reacs*n_specs] net change */
    const double *reac_k;        /* [n_reacs] rate constant, SI (M^(1-order)/s) */
    const double *spec_dcst;     /* [n_specs] diffusion constant m^2/s, 0 = immobile */
    int32_t n_sspecs;            /* surface species */
    int32_t n_sreacs;
    const int32_t *sreac_lhs_s;  /* [n_sreacs*n_sspecs] */
    const int32_t *sreac_upd_s;  /* [n_sreacs*n_sspecs] */
    const int32_t *sreac_lhs_i;  /* [n_sreacs*n_specs] inner-compartment reactants */
    const int32_t *sreac_upd_i;  /* [n_sreacs*n_specs] */
    const double *sreac_k;       /* [n_sreacs] */
    /* Voltage-dependent surface reactions (VDepSReac, src/steps/mpi/tetopsplit/
     * vdepsreac.cpp:217-320; VDepSReacdef::getVDepK, src/steps/solver/
     * vdepsreacdef.cpp:194-217): surface reaction r takes its rate constant from
     * table sreac_vdep[r] (-1: the constant sreac_k[r]), linearly interpolated
     * at the triangle's potential. NULL / 0 when the model has none. */
    const int32_t *sreac_vdep;   /* [n_sreacs] */
    int32_t n_vdep;
    const double *vdep_vmin;     /* [n_vdep] V */
    const double *vdep_dv;       /* [n_vdep] V */
    const int32_t *vdep_off;     /* [n_vdep+1] offsets into vdep_k */
    const double *vdep_k;        /* tables of k(vmin + i dv), SI */
    /* GHK currents (GHKcurr::rate/apply, src/steps/mpi/tetopsplit/
     * ghkcurr.cpp:168-272; math::GHKcurrent, src/steps/math/ghk.cpp:62-81):
     * on every membrane triangle, ions of volume species ghk_ion cross at the
     * GHK flux of each channel in surface species ghk_chan. The outer side is
     * a virtual concentration ghk_voconc (M) when >= 0, else the triangle's
     * outer tet; ghk_realflux = 0 moves charge but no ions. Up to 4 currents. */
    int32_t n_ghk;
    const int32_t *ghk_chan;     /* [n_ghk] surface species of the open channel */
    const int32_t *ghk_ion;      /* [n_ghk] volume species of the ion */
    const int32_t *ghk_valence;  /* [n_ghk] nonzero */
    const double *ghk_perm;      /* [n_ghk] single-channel permeability, m^3/s */
    const double *ghk_voconc;    /* [n_ghk] virtual outer concentration, M, or < 0: the outer tet */
    const double *ghk_vshift;    /* [n_ghk] V added to the membrane potential */
    const int32_t *ghk_realflux; /* [n_ghk] 1: the ion count changes per event */
    const double *sspec_dcst;    /* [n_sspecs] surface diffusion constant m^2/s, or NULL */
    /* Outer-compartment reactants and net changes of each surface reaction
     * (SReac olhs / orhs; SReac::rate / apply, sreac.cpp:305-505), NULL when
     * none: [n_sreacs*n_specs]. A reaction has inner or outer volume
     * reactants, not both (model/sreac.cpp:55); with outer ones its constant
     * scales with the outer tet's volume. A GHK current with ghk_voconc < 0
     * reads the outer tet's concentration and moves its ions to and from it
     * (ghkcurr.cpp:176-182, 228-272). */
    const int32_t *sreac_lhs_o;
    const int32_t *sreac_upd_o;
} stg_model;

/* Per-realization parameters of a batch; NULL arrays take the model's values. */
typedef struct stg_realization {
    uint64_t seed;               /* Philox key */
    uint32_t stream;             /* Philox stream id (counter word 2) */
    const double *reac_k;        /* [n_reacs] or NULL */
    const double *sreac_k;       /* [n_sreacs] or NULL */
    const double *spec_dcst;     /* [n_specs] or NULL */
} stg_realization;

/* Engine for one realization: replaces the TetOpSplitP constructor
 * (src/steps/mpi/tetopsplit/tetopsplit.cpp:83-118, pysteps/cysteps_mpi.pyx:80-109)
 * and the DistTetOpSplit constructor (pysteps/cysteps_dist.pyx:1581-1633).
 * At most 536,870,911 tets over all realizations of a handle (each face's
 * gather index, tet x 4 + face, is an int32); more returns STG_ERR_ARG. */
int stg_create(const stg_model *model, const stg_mesh *mesh, uint64_t seed, int32_t mode,
               stg_sim **out);

/* n_real independent realizations (own mesh, constants, key) advanced by the
 * same launches. No reference counterpart: it is the batched ensemble mode. */
int stg_create_batch(const stg_model *model, int32_t n_real, const stg_mesh *meshes,
                     const stg_realization *reals, int32_t mode, stg_sim **out);

/* Several compartments of one simulation (TetOpSplitP over a mesh with
 * several TmComps, tetopsplit.cpp:83-118; Compdef per compartment,
 * src/steps/solver/compdef.hpp). Compartment c is realization c: its mesh
 * (local tet numbering), reaction constants and diffusion constants (zero for
 * reactions and species it does not hold), Philox stream comp_params[c].stream.
 * All compartments advance on one period schedule. Faces between compartments
 * are closed unless they lie on a diffusion boundary (tet_face_bdir). */
int stg_create_comps(const stg_model *model, int32_t n_comps, const stg_mesh *comp_meshes,
                     const stg_realization *comp_params, int32_t mode, stg_sim **out);

/* Diffusion constant of species spec across boundary direction bdir (m^2/s;
 * 0 closes it), as DiffBoundary activation and setDiffBoundarySpecDcst set it
 * per tet direction (tetopsplit.cpp:2745-2931, diff.cpp:307-460). */
int stg_set_boundary_dcst(stg_sim *sim, int32_t bdir, int32_t spec, double dcst);
int stg_get_boundary_dcst(const stg_sim *sim, int32_t bdir, int32_t spec, double *dcst);

int stg_destroy(stg_sim *sim);

/* Message of the last failed call on sim (or of the last failed create when
 * sim is NULL). Never NULL. */
const char *stg_last_error(const stg_sim *sim);

/* TetOpSplitP::reset (tetopsplit.cpp:1436-1503): counts 0, time 0, steps 0.
 * The random streams are not rewound: as the reference's RNG keeps running
 * through a reset, each realization's Philox period counter carries on, so
 * runs after a reset draw fresh numbers. stg_set_seed restarts it. */
int stg_reset(stg_sim *sim);

/* TetOpSplitP::run (tetopsplit.cpp:1507-1526) / OmegaHSimulation::run
 * (simulation.cpp:1418-1453): advance every realization to endtime. When the
 * device reports an error (STG_ERR_PROG: event cap, a potential outside a
 * k(V) table, GHK overflow) the periods have run: time, step counts and the
 * Philox counters stand at endtime, and the counts are those the kernels left;
 * restore a checkpoint or reset to continue from a known state. */
int stg_run(stg_sim *sim, double endtime);

/* TetOpSplitP::advance (tetopsplit.cpp:1956-1966). */
int stg_advance(stg_sim *sim, double adv);

/* Re-key one realization's Philox streams and restart them (period counter
 * 0), as re-seeding the solver's RNG (RNG::initialize, src/steps/rng/rng.hpp)
 * restarts the reference's sequence. */
int stg_set_seed(stg_sim *sim, int32_t real, uint64_t seed);

/* API::getTime / setTime. */
int stg_get_time(const stg_sim *sim, double *t);
int stg_set_time(stg_sim *sim, double t);

/* API::getNSteps: periods executed since reset, per realization. */
int stg_get_nsteps(const stg_sim *sim, int32_t real, uint64_t *n);

/* TetOpSplitP::setNSteps (tetopsplit.cpp:1995-1997): overwrite the period
 * count of one realization. The random streams are keyed by their own counter,
 * so this changes no later result. */
int stg_set_nsteps(stg_sim *sim, int32_t real, uint64_t n);

/* TetOpSplitP::getUpdPeriod (tetopsplit.cpp:5373-5411, getter at :8302-8328). */
int stg_get_update_period(const stg_sim *sim, int32_t real, double *dt);

/* TetOpSplitP::getReacExtent / getDiffExtent (tetopsplit.cpp:8302-8328):
 * SSA events and molecules moved by diffusion since reset. */
int stg_get_extents(const stg_sim *sim, int32_t real, uint64_t *reac_extent,
                    uint64_t *diff_extent);

/* Whole-species count arrays of one realization (n_tets or n_tris long):
 * the bulk form of API::setTetSpecCount / getTetSpecCount and of
 * getBatchTetSpecCountsNP (tetopsplit.cpp:5715-5778). */
int stg_set_tet_counts(stg_sim *sim, int32_t real, int32_t spec, const int32_t *counts);
int stg_get_tet_counts(stg_sim *sim, int32_t real, int32_t spec, int32_t *counts);
int stg_set_tri_counts(stg_sim *sim, int32_t real, int32_t sspec, const int32_t *counts);
int stg_get_tri_counts(stg_sim *sim, int32_t real, int32_t sspec, int32_t *counts);

/* getBatchTetSpecCountsNP (tetopsplit.cpp:5715-5778): counts of selected
 * tets, gathered on the device. */
int stg_get_batch_tet_counts(stg_sim *sim, int32_t real, int32_t spec, const int64_t *tets,
                             size_t n, double *counts);

/* No reference counterpart. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the result of stg_get_batch_tet_counts for at most
 * STG_NOALLOC_MAX_TETS tets (STG_ERR_ARG beyond), with no allocation. The
 * indices and gathered counts are staged in a fixed 1 MiB device array
 * (16 bytes per tet) that the library holds on each GPU it runs on until the
 * process ends. All calls on one GPU share that array: calls that overlap in
 * time on one GPU (other threads or streams) race. */
#define STG_NOALLOC_MAX_TETS 65536
int stg_get_batch_tet_counts_noalloc(stg_sim *sim, int32_t real, int32_t spec, const int64_t *tets,
                                     size_t n, double *counts);

/* Total count of species spec over the listed tets of a realization, reduced
 * on the device (sumBatchTetCountsNP / getROISpecCount, tetopsplit.cpp:5715-5778). */
int stg_sum_batch_tet_counts(stg_sim *sim, int32_t real, int32_t spec, const int64_t *tets, size_t n,
                             int64_t *total);

/* On-device totals of one species over a realization's tets (or triangles):
 * the reduction behind getCompSpecCount / getPatchSpecCount
 * (tetopsplit.cpp:2045-2110). */
int stg_sum_tet_counts(stg_sim *sim, int32_t real, int32_t spec, int64_t *total);
/* Totals of every species in every realization in one device reduction:
 * totals[r * n_specs + s] (stg_sum_all_tri_counts: [r * n_sspecs + q]). Each
 * entry is the reduction of stg_sum_tet_counts (getCompSpecCount,
 * tetopsplit.cpp:2045-2110); the batch of all of them in one call has no
 * reference counterpart (a batched ensemble reads every realization's totals
 * at each save with one device round trip). */
int stg_sum_all_tet_counts(stg_sim *sim, int64_t *totals);
int stg_sum_all_tri_counts(stg_sim *sim, int64_t *totals);
int stg_sum_tri_counts(stg_sim *sim, int32_t real, int32_t sspec, int64_t *total);

/* setCompReacK / setPatchSReacK / setCompDiffD (api.hpp). A changed
 * diffusion constant recomputes the update period, as recomputeUpdPeriod
 * does in TetOpSplitP (tetopsplit.cpp:1515-1517). */
int stg_set_reac_k(stg_sim *sim, int32_t real, int32_t reac, double k);
int stg_set_sreac_k(stg_sim *sim, int32_t real, int32_t sreac, double k);
int stg_set_dcst(stg_sim *sim, int32_t real, int32_t spec, double dcst);
/* Surface diffusion constant of surface species sspec (setPatchSDiffD). */
int stg_set_sdcst(stg_sim *sim, int32_t real, int32_t sspec, double dcst);

/* Per-element tables of one realization (compartment), the device form of
 * the reference's per-tet and per-triangle setters. Each table replaces the
 * realization's constants for its elements; the host composes it from the
 * compartment-level value, per-element overrides and activation flags, as
 * TetOpSplitP keeps one constant and one active flag per kproc. NULL removes
 * the table. A realization with a table runs the general (membrane) variant
 * of the generic period kernel. Triangles are in the realization's input
 * order.
 *
 * Rate constant of every (reaction, tet), [n_reacs][n_tets], SI, 0 for an
 * inactive reaction: _setTetReacK (tetopsplit.cpp:3996), _setTetReacActive
 * (:4045), _setCompReacActive (:2309), Reac::rate (reac.cpp:184-232). */
int stg_set_tet_reac_table(stg_sim *sim, int32_t real, const double *k);
/* Diffusion constant of every (species, tet, face), [n_specs][n_tets][4]
 * m^2/s, face f opposite vertex f, 0 where nothing leaves through the face:
 * _setTetDiffD with and without a direction (tetopsplit.cpp:4103),
 * _setTetDiffActive (:4166), _setCompDiffActive (:2437), and the
 * diffusion-boundary constants of those faces. Recomputes the update period
 * (_computeUpdPeriod, :5373-5411). */
int stg_set_tet_diff_table(stg_sim *sim, int32_t real, const double *d);
/* Clamped species per tet, bit s of mask[t] = species s: reactions and
 * diffusion leave its count unchanged (Reac::apply, reac.cpp:240-243;
 * Diff::apply, diff.cpp:476-577): _setTetSpecClamped (tetopsplit.cpp:3951),
 * _setCompSpecClamped (:2189). */
int stg_set_tet_clamp(stg_sim *sim, int32_t real, const uint32_t *mask);
/* [n_sreacs + n_ghk][n_tris]: each surface reaction's constant on each
 * triangle (SI; for a voltage-dependent reaction a factor 1 or 0 on k(V)),
 * then a factor 1 or 0 on each GHK current: _setTriSReacK (:4464),
 * _setTriSReacActive, _setTriVDepSReacActive, and patches that hold only
 * some of the surface systems. */
int stg_set_tri_sreac_table(stg_sim *sim, int32_t real, const double *k);
/* Surface diffusion constant of every (surface species, triangle, edge),
 * [n_sspecs][n_tris][3]: _setTriSDiffD (tetopsplit.cpp, SDiff::setDirectionDcst,
 * sdiff.cpp) and surface diffusion boundaries. Recomputes the update period. */
int stg_set_tri_sdiff_table(stg_sim *sim, int32_t real, const double *d);
/* Clamped surface species per triangle, bit q = surface species q
 * (_setTriSpecClamped, _setPatchSpecClamped). */
int stg_set_tri_clamp(stg_sim *sim, int32_t real, const uint32_t *mask);

/* Launch shape used by stg_run: threads per block (64..1024, multiple of 64)
 * and whether the launches of one run are replayed from a captured HIP graph.
 * No reference counterpart (device tuning). */
int stg_set_launch(stg_sim *sim, int32_t block, int32_t use_graph);

/* Lanes per tet of the period kernel (no reference counterpart; device
 * tuning): 1 (one lane runs a tet's whole period) or 2, 4, 8 (the tile
 * schedule, for volume-only models on meshes too small to fill the GPU: the
 * lanes of a tet split its species for the entering gather, the leaving pass
 * and the stores, and one of them runs its SSA). A block then holds
 * block / lanes tets. Results are identical for every value. A model that
 * needs the membrane variant (surface reactions, GHK currents, channel
 * integrals, diffusion boundaries, per-element tables) returns
 * STG_ERR_NOTIMPL for lanes > 1, and a run goes back to one lane per tet when
 * the model comes to need it. */
int stg_set_lanes(stg_sim *sim, int32_t lanes);
int stg_get_lanes(const stg_sim *sim, int32_t *lanes);

/* Width of the face slots that carry a period's leaving molecules to the
 * neighbour's entering gather (no reference counterpart; device tuning): 2
 * (16-bit, the default) or 1 (8-bit) bytes per face and diffusing species. A
 * count above the slot's range (32,767 or 127) takes a 32-bit side entry, so
 * results are identical for both widths; 8-bit slots move fewer bytes while
 * few faces pass more than 127 molecules in one period. */
int stg_set_slot_width(stg_sim *sim, int32_t bytes);
int stg_get_slot_width(const stg_sim *sim, int32_t *bytes);

/* Model-specialised kernel (no reference counterpart; device tuning): on = 1
 * compiles, at run time with hipRTC (about a second), a period kernel with
 * this model's reaction network as compile-time constants and uses it for
 * launches of up to 256 threads per block; results are identical to the
 * generic kernel's. min_waves (0..8) asks the compiler to keep at least that
 * many wavefronts per SIMD resident (0: its own choice); each bound is
 * compiled once and kept. The volume reactions become constants; surface
 * reactions, GHK currents and k(V) tables stay run-time tables. on = 0
 * returns to the generic kernel. */
int stg_set_specialize(stg_sim *sim, int32_t on, int32_t min_waves);
int stg_get_specialize(const stg_sim *sim, int32_t *on);

/* TetOpSplitP::setDiffApplyThreshold (tetopsplit.hpp:995,
 * pysteps/cysteps_mpi.pyx:1288; DiffusionOperator default 10,
 * diffusion_operator.hpp:99): leaving counts up to the threshold are split
 * over faces one molecule at a time, larger ones by conditional binomials.
 * Both give the same multinomial distribution. */
int stg_set_diff_apply_threshold(stg_sim *sim, int32_t threshold);

/* TetOpSplitP::checkpoint / restore (tetopsplit.cpp:217-453), own format:
 * time, per-realization period and Philox counters, seeds, extents and
 * constants, all molecule counts, triangle potentials, channel-count
 * integrals and GHK charges. Because every random draw is keyed by (seed, tet, period),
 * a restored simulation continues bit-for-bit as the original would have.
 * stg_restore requires a simulation of the same shape. */
int stg_state_size(const stg_sim *sim, size_t *bytes);
int stg_checkpoint(stg_sim *sim, void *buf, size_t bytes);
int stg_restore(stg_sim *sim, const void *buf, size_t bytes);

/* Membrane potential of every triangle of one realization (volts), used by
 * voltage-dependent surface reactions; TetOpSplitP::setTriV. An attached
 * EField (stg_efield_attach) writes it after every step. Default -65 mV. */
int stg_set_tri_v(stg_sim *sim, int32_t real, const double *v);
int stg_get_tri_v(stg_sim *sim, int32_t real, double *v);

/* Temperature of the GHK flux equation, K (TetOpSplitP::setTemp/getTemp,
 * tetopsplit.cpp:2001-2010; default 293.15, tetopsplit.cpp:1119). */
int stg_set_temp(stg_sim *sim, double kelvin);
int stg_get_temp(const stg_sim *sim, double *kelvin);

/* Number of realizations and tets / triangles of one realization. */
int stg_get_sizes(const stg_sim *sim, int32_t real, int32_t *n_real, int32_t *n_tets,
                  int32_t *n_tris);

/* Average duration (ms) of the period kernel over the launches of the last
 * stg_run, timed with HIP events on the handle's stream; *n = launches. A run
 * of L periods makes L period-kernel launches (each realization steps in the
 * first launches it has periods for) and then one entering-only launch of a
 * separate kernel, which is not included. */
int stg_last_run_kernel_ms(const stg_sim *sim, double *ms_per_launch, int32_t *n);

/* ------------------------------------------------------------------ EField
 * Membrane potential on the vertices of a tet mesh, one membrane. Replaces
 * EField / dVSolverBase (src/steps/solver/efield/efield.cpp:48-110,
 * dVsolver.hpp:126-171) with the BDSystem / PETSc solvers
 * (bdsystem.cpp:40-157, dVsolver_petsc.cpp:163-268) and the DistTetOpSplit
 * EFieldOperator (src/steps/mpi/dist/tetopsplit/operator/efield_operator.cpp:
 * 387-540). Arguments are SI (m, s, V, A, F/m^2, ohm.m, ohm.m^2) as in the
 * reference's setters. */
typedef struct stg_efield stg_efield;

typedef struct stg_efield_mesh {
    int32_t n_verts;
    const double *verts;   /* [n_verts*3] m */
    int32_t n_tets;
    const int32_t *tets;   /* [n_tets*4] */
    int32_t n_tris;
    const int32_t *tris;   /* [n_tris*3] membrane triangles, in the order of the
                              simulation's patch triangles when one is attached */
} stg_efield_mesh;

/* EField::initMesh (efield.cpp:48-110); potential -65 mV, 1 uF/cm^2, 1 ohm.m. */
int stg_efield_create(const stg_efield_mesh *mesh, stg_efield **out);
int stg_efield_destroy(stg_efield *ef);
const char *stg_efield_last_error(const stg_efield *ef);

/* EField::setMembPotential / setMembCapac / setMembVolRes /
 * setSurfaceResistivity (efield.cpp:163-195). */
int stg_efield_set_potential(stg_efield *ef, double v);
int stg_efield_set_capac(stg_efield *ef, double cm);
int stg_efield_set_volres(stg_efield *ef, double ro);
int stg_efield_set_memb_res(stg_efield *ef, double rspec, double vrev);

/* EField::setVertIClamp / setTriIClamp (efield.cpp:213-225, 376-381):
 * positive current depolarises. setVertVClamped (efield.cpp:255-264). */
int stg_efield_set_vert_iclamp(stg_efield *ef, int32_t vert, double amps);
int stg_efield_set_tri_iclamp(stg_efield *ef, int32_t tri, double amps);
int stg_efield_set_vert_clamped(stg_efield *ef, int32_t vert, int32_t clamped);

/* EField::getVertV / setVertV / getTriV (efield.cpp:238-300), all vertices or
 * triangles at once, in volts. */
int stg_efield_set_verts_v(stg_efield *ef, const double *v);
int stg_efield_get_verts_v(stg_efield *ef, double *v);
int stg_efield_get_tris_v(stg_efield *ef, double *v);

/* Ohmic currents (OhmicCurr, Tri::computeI src/steps/mpi/tetopsplit/tri.cpp:
 * 664-693): for each k, channels of surface species sspec[k] of the attached
 * simulation conduct g[k] siemens each with reversal potential erev[k]. */
int stg_efield_set_ohmic(stg_efield *ef, int32_t n, const int32_t *sspec, const double *g,
                         const double *erev);

/* TetOpSplitP::_setTriOhmicErev / _getTriOhmicErev (tetopsplit.cpp:4933-4990,
 * Tri::setOCerev): reversal potential (V) of ohmic current k on one membrane
 * triangle. stg_efield_set_ohmic puts every triangle back to erev[k]. */
int stg_efield_set_tri_ohmic_erev(stg_efield *ef, int32_t k, int32_t tri, double erev);
int stg_efield_get_tri_ohmic_erev(stg_efield *ef, int32_t k, int32_t tri, double *erev);

/* TetOpSplitP::_setTriCapac (tetopsplit.cpp:5187-5203, TetMesh::
 * applyTriCapacitance): specific capacitance (F/m^2) of one membrane triangle;
 * each of its vertices gains (cm - previous) * area / 3. stg_efield_set_capac
 * sets every triangle. */
int stg_efield_set_tri_capac(stg_efield *ef, int32_t tri, double cm);
int stg_efield_get_tri_capac(stg_efield *ef, int32_t tri, double *cm);

/* Tri::getGHKI (tri.cpp:634-660) for every membrane triangle: the charge GHK
 * current g (or every current, g = -1) moved outward over the last EField
 * step, times the elementary charge, over the step (A, outward positive);
 * 0 before the first step after attaching. */
int stg_efield_get_tri_ghk_currents(stg_efield *ef, int32_t g, double *amps);

/* Couple to realization real of a simulation whose patch triangles are the
 * EField's membrane triangles (NULL detaches). */
int stg_efield_attach(stg_efield *ef, stg_sim *sim, int32_t real);

/* EField::advance (efield.cpp:229-233): one implicit step of dt seconds with
 * the current membrane currents. */
int stg_efield_advance(stg_efield *ef, double dt);

/* TetOpSplitP::_runWithEField (tetopsplit.cpp:1851-1954): alternate
 * reaction-diffusion to t+efdt (stg_run) and one EField step, to endtime. */
int stg_efield_run(stg_efield *ef, stg_sim *sim, double endtime, double efdt);

/* EField::checkpoint / restore as TetOpSplitP::checkpoint / restore call
 * them (pEField->checkpoint, tetopsplit.cpp:309; pEField->restore, :446),
 * own format: vertex potentials, the previous step's increment, membrane and
 * clamp parameters (per-triangle capacitance and ohmic reversal potentials
 * included), the start time of the running channel-count integrals and the
 * GHK charges of the last step (the running integrals and charges are in the
 * attached simulation's checkpoint). stg_efield_restore requires the same mesh. */
int stg_efield_state_size(const stg_efield *ef, size_t *bytes);
int stg_efield_checkpoint(stg_efield *ef, void *buf, size_t bytes);
int stg_efield_restore(stg_efield *ef, const void *buf, size_t bytes);

/* Solver choice (no reference counterpart): precision 0 = FP32 conjugate
 * gradients with FP64 dot products inside FP64 iterative refinement, 1 = FP64
 * conjugate gradients; tol = relative FP64 residual ||b - A dV|| / ||b||. */
int stg_efield_set_solver(stg_efield *ef, int32_t precision, double tol, int32_t max_iter);
/* Preconditioner (no reference counterpart; the reference's PETSc path uses
 * point-block Jacobi, efield_operator.cpp:116-120): 1 = one V-cycle of
 * smoothed-aggregation algebraic multigrid for systems above 16,384 vertices
 * (default), set up on the host when the step size or a membrane parameter
 * changes; 0 = Jacobi. Systems of up to 16,384 vertices always use the
 * one-workgroup Jacobi solver. */
int stg_efield_set_precond(stg_efield *ef, int32_t kind);
/* Multigrid levels in use (0 = none), size of the dense coarsest level and
 * the host time of the last set-up in seconds. */
int stg_efield_precond_info(const stg_efield *ef, int32_t *levels, int32_t *coarsest, double *setup_s);
int stg_efield_stats(const stg_efield *ef, int32_t *iters, int32_t *refinements,
                     double *rel_residual);
int stg_efield_sizes(const stg_efield *ef, int32_t *n_verts, int32_t *n_tris, int32_t *nnz);

/* ------------------------------------------------------------- mesh ingest
 * Face topology and measures of a tetrahedral mesh built on the device, the
 * work of the Tetmesh constructor Tetmesh(verts, tets, tris)
 * (src/steps/geom/tetmesh.cpp:78-233) that the mesh loaders feed
 * (pysteps/steps/API_2/geom.py:1466-1592). Triangles are numbered as the
 * reference numbers them: the n_tris supplied triangles first (tris may be
 * NULL when n_tris is 0; repeats dropped), then every other tet face in order
 * of first appearance, tet by tet, faces (v0,v1,v2), (v0,v1,v3), (v0,v2,v3),
 * (v1,v2,v3); a triangle's vertices are stored sorted; tri_tets lists the
 * lowest-numbered tet having the face and the neighbour across it (-1 on the
 * boundary). Coordinates in m; outputs in SI. Errors: STG_ERR_ARG for an index
 * out of range, a degenerate tet, a triangle shared by more than two tets or
 * a supplied triangle that is no tet's face. */
typedef struct stg_meshbuild stg_meshbuild;

int stg_mesh_build(const double *verts, int32_t n_verts, const int32_t *tets, int32_t n_tets,
                   const int32_t *tris, int32_t n_tris, stg_meshbuild **out);
int stg_mesh_build_sizes(const stg_meshbuild *mb, int32_t *n_tets, int32_t *n_tris);
/* Any output pointer may be NULL. Sizes: tet_vol [n_tets], tet_bary
 * [n_tets*3], tet_tri / tet_nbr / tet_nbr_face / tet_face_area /
 * tet_face_dist [n_tets*4] (face f is opposite vertex f), tris [n_tris*3],
 * tri_tets [n_tris*2], tri_area [n_tris], tri_bary [n_tris*3]. */
int stg_mesh_build_fetch(stg_meshbuild *mb, double *tet_vol, double *tet_bary, int32_t *tet_tri,
                         int32_t *tet_nbr, int32_t *tet_nbr_face, double *tet_face_area,
                         double *tet_face_dist, int32_t *tris, int32_t *tri_tets, double *tri_area,
                         double *tri_bary);
int stg_mesh_build_destroy(stg_meshbuild *mb);
/* Message of the last failed stg_mesh_build* call on this thread. */
const char *stg_mesh_build_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
