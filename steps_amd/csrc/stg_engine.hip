// Host side of the C ABI declared in include/stg.h: owns the device buffers of
// one simulation (one or many realizations), builds the period schedule and
// launches the fused period kernel on the handle's HIP stream.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "../../include/stg.h"
#include "stg_kernels.hpp"
#include "stg_rtc_sources.inc"

namespace {

thread_local std::string g_create_error;

struct StgError {
    int code;
    std::string msg;
};

#define HIP_OK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw StgError{STG_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

template <typename T>
T *dalloc(size_t n) {
    void *p = nullptr;
    if (n == 0) n = 1;
    HIP_OK(hipMalloc(&p, n * sizeof(T)));
    return static_cast<T *>(p);
}

// util::almost_equal (src/steps/util/collections.hpp:143-149)
bool almost_equal(double x, double y) {
    return std::fabs(x - y) <= DBL_EPSILON * std::fabs(x + y) * 4 ||
           std::fabs(x - y) < DBL_MIN;
}

// Period lengths from t0 to t_end. mode TETOPSPLIT follows
// TetOpSplitP::_runWithoutEField (tetopsplit.cpp:1528-1560); mode
// DISTTETOPSPLIT follows OmegaHSimulation::run_rd (simulation.cpp:1358-1384).
void schedule(double t0, double t_end, double dt_std, int mode, std::vector<double> &dts) {
    dts.clear();
    double t = t0;
    if (mode == STG_MODE_TETOPSPLIT) {
        if (almost_equal(t0, t_end)) return;
        bool aligned = false;
        while (t < t_end && !aligned) {
            double per = dt_std;
            if (t + dt_std > t_end) {
                per = t_end - t;
                aligned = true;
            }
            dts.push_back(per);
            t = t + per;
        }
        return;
    }
    const double nstd = std::floor((t_end - t) / dt_std) - 1.0;
    const long long n_std = nstd > 0.0 ? (long long)nstd : 0;
    for (long long i = 0; i < n_std; ++i) {
        const double per = std::min(dt_std, t_end - t);
        dts.push_back(per);
        t = t + per;
    }
    const double next = t + dt_std;
    if (!almost_equal(next, t_end) && next < t_end) {
        dts.push_back(dt_std);
        t = t + dt_std;
        dts.push_back(t_end - t);
    } else if (!almost_equal(t_end, t)) {
        dts.push_back(t_end - t);
    }
}

}  // namespace

struct stg_sim {
    int mode = STG_MODE_TETOPSPLIT;
    int n_real = 0;
    int S = 0, R = 0, SS = 0, RS = 0, ND = 0;
    int block = 256;
    int lanes = 1;  // lanes per tet (tile schedule for volume models when > 1)
    int slot_bytes = 2;  // face slot width: 2 (16-bit) or 1 (8-bit)
    bool use_graph = false;
    bool xcd_remap = true;
    int diff_threshold = 10;  // diffusion_operator.hpp:99, tetopsplit.hpp:995
    stg::ModelK mk{};
    std::vector<int32_t> tet_off, tri_off;  // [n_real+1]
    std::vector<int32_t> block_real, block_tet0;
    std::vector<std::vector<int32_t>> tri_perm;  // per realization: user tri -> internal tri
    int32_t NT = 0, NTRI = 0, max_tris = 0;
    int32_t NG = 0;  // GHK currents
    bool outer_model = false;  // surface reactions or GHK currents reach the outer side
    // SSA groups (tets joined across the membrane): host tables, device arrays
    std::vector<int32_t> h_tet_grp, h_grp_ptr, h_grp_mem, h_grp_real;
    int32_t *d_tet_grp = nullptr, *d_grp_ptr = nullptr, *d_grp_mem = nullptr, *d_grp_real = nullptr;
    int32_t *d_tri_otet = nullptr, *d_tri_oreal = nullptr;
    double *d_gcorr = nullptr;
    bool has_groups() const { return d_tet_grp != nullptr; }
    bool linked = false;                   // compartments: one period schedule for all
    int32_t NBD = 0;                       // diffusion-boundary directions
    std::vector<double> h_bnd_D;           // [NBD][S]
    int32_t NSD = 0;                       // diffusing surface species
    std::vector<double> h_sD;              // [n_real][SS] surface diffusion constants
    std::vector<double> h_tri_g;           // [NTRI*3] internal triangle order
    std::vector<uint32_t> h_tet_bdir;      // [NT] 4 x 8-bit direction per face, 0xff = none
    std::vector<double> h_k, h_sk, h_D;  // per-realization tables
    std::vector<double> h_face_g;         // [NT*4], for update-period recomputation
    std::vector<double> vdep_vmin, vdep_dv, vdep_k;
    std::vector<int32_t> vdep_off;
    std::vector<double> dt_std;           // [n_real]
    std::vector<uint64_t> nsteps;         // [n_real] periods since reset (API::getNSteps)
    // [n_real] Philox period counter: advances with nsteps but survives reset,
    // so a reset simulation continues its random streams instead of replaying
    // them (the reference's RNG keeps running through TetOpSplitP::reset,
    // tetopsplit.cpp:1436-1503); re-keying (stg_set_seed) restarts it
    std::vector<uint64_t> rng_steps;
    std::vector<uint64_t> seeds;
    std::vector<uint32_t> streams;
    double time = 0.0;
    // device
    double *d_vol = nullptr;
    int4 *d_nbr = nullptr;  // per face: neighbour's global tet x 4 + its face slot, -1 = none
    double2 *d_g = nullptr;
    int32_t *d_counts = nullptr;
    int16_t *d_out[2] = {nullptr, nullptr};   // face slots (ping-pong), sized for 16-bit ones
    int32_t *d_wide[2] = {nullptr, nullptr};  // escaped slot counts
    int32_t *d_tri_ptr = nullptr;
    int32_t *d_tri_counts = nullptr;
    double *d_tri_area = nullptr;
    int4 *d_block_desc = nullptr;  // [blocks] {block, realization, first tet, end tet}, launch order
    int32_t *d_real_tet0 = nullptr;
    double *d_real_k = nullptr, *d_real_sk = nullptr, *d_real_D = nullptr;
    uint64_t *d_real_seed = nullptr;
    uint32_t *d_real_stream = nullptr;
    // the run's tables in one allocation, uploaded with one copy:
    // [periods per realization, int32][first Philox period, uint32][dt table, double]
    double *d_run_tab = nullptr;
    uint32_t *d_real_step0 = nullptr;  // views into d_run_tab
    int32_t *d_real_nper = nullptr;
    double *d_dt_tab = nullptr;
    size_t dt_tab_cap = 0;
    unsigned long long *d_ext = nullptr;
    double *d_tri_v = nullptr;
    double *d_tri_intg = nullptr;  // [SS][NTRI] time integrals of channel counts, lazily allocated
    int32_t *d_tri_charge = nullptr;  // [NG][NTRI] GHK charge per current, when the model has GHK currents
    uint32_t *d_tet_bdir = nullptr;   // [NT] when there are diffusion boundaries
    double *d_bnd_D = nullptr;        // [NBD][S]
    int4 *d_tri_nbr = nullptr;        // surface diffusion tables
    uint32_t *d_tri_nface = nullptr;
    double *d_tri_g = nullptr;
    int32_t *d_tri_user = nullptr;
    double *d_real_sD = nullptr;
    int32_t *d_tout[2] = {nullptr, nullptr};
    double *d_tri_corr = nullptr;
    // per-element tables (stg_set_tet_reac_table and the like): which tables
    // each realization uses, host copies in device layout, device arrays
    std::vector<int32_t> real_tabs;                    // [n_real] stg::kTab* bits
    std::vector<double> h_tet_k, h_tet_fD, h_tri_sk, h_tri_eD;
    std::vector<uint32_t> h_tet_clamp, h_tri_clamp;
    int32_t *d_real_tabs = nullptr;
    double *d_tet_k = nullptr, *d_tet_fD = nullptr, *d_tri_sk = nullptr, *d_tri_eD = nullptr;
    uint32_t *d_tet_clamp = nullptr, *d_tri_clamp = nullptr;
    double *d_vdep_vmin = nullptr, *d_vdep_dv = nullptr, *d_vdep_k = nullptr;
    int32_t *d_vdep_off = nullptr;
    int32_t *d_err = nullptr;
    unsigned long long *d_sum = nullptr;
    int64_t *d_idx = nullptr;
    // totals of every (realization, species), and the triangle offsets they
    // need, allocated on first use
    long long *d_real_sums = nullptr;
    long long *h_real_sums = nullptr;  // pinned: the read-back of d_real_sums
    unsigned long long *h_sum = nullptr;  // pinned: the read-back of d_sum
    int32_t *h_err = nullptr;             // pinned: the read-back of d_err after a run
    int32_t *d_tri_off = nullptr;
    size_t idx_cap = 0;
    double *d_gather = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // pinned staging of a run's schedule tables (a copy from pageable memory
    // is staged and blocks the host), grown on demand
    unsigned char *h_stage = nullptr;
    size_t h_stage_cap = 0;
    // model-specialised volume kernel (blocks of up to 256), compiled at run time;
    // one compiled module per (membrane variant, waves-per-SIMD bound (0 = the
    // compiler's choice), lanes per tet 1/2/4/8 x face slot width 2/1)
    hipModule_t rtc_mods[2][9][8] = {};
    hipFunction_t rtc_fns[2][9][8] = {};
    hipFunction_t rtc_fn = nullptr;  // the one in use, or null for the generic kernel
    bool rtc_memb = false;           // whether rtc_fn is the membrane variant
    int rtc_waves = 0;               // its waves-per-SIMD bound
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_launches = -1;
    int graph_block = -1;
    std::vector<int32_t> graph_counts;  // blocks per launch the graph was captured with
    // Launch order of the blocks: by periods in the current run, most first, so
    // the blocks still active at launch l are a prefix of launch_blocks[l].
    std::vector<int32_t> desc_np;        // periods per block the uploaded order was made for
    std::vector<int32_t> desc_nper;      // periods per realization the order and launch_blocks were made for
    std::vector<int32_t> launch_blocks;  // [launches] blocks dispatched
    double last_ms = 0.0;
    int32_t last_launches = 0;
    std::string err;

    // the general variant of the period kernel: triangles that can hold a
    // surface reaction, a GHK current or an integrated channel count, or
    // faces on a diffusion boundary
    bool needs_memb() const {
        return (NTRI > 0 && (RS > 0 || NG > 0 || NSD > 0 || mk.tri_intg_mask != 0)) || NBD > 0 || has_tabs() ||
               has_groups();
    }
    // some realization uses a per-element table (read by the membrane
    // variant of the generic kernel only)
    bool has_tabs() const {
        for (int32_t v : real_tabs)
            if (v) return true;
        return false;
    }

    // per-thread LDS rows of the period kernel: counts, occupancy corrections,
    // rate constants, propensities and (membrane variant) GHK constants
    size_t smem_bytes(int B) const {
        const int K = std::max(1, R + max_tris * (RS + NG));
        size_t off = ((size_t)S * B * 4 + 15) & ~(size_t)15;
        const size_t rows = 2 * (size_t)K + (needs_memb() ? 2 * (size_t)max_tris * NG : 0);
        return off + (size_t)S * B * 8 + rows * B * 8;
    }
    // the model-specialised volume kernel keeps only counts and occupancy
    // corrections in LDS (rate constants and propensities are formed in
    // registers); the membrane one adds the triangles' constant rows and the
    // GHK rows
    size_t smem_bytes_lean(int B, bool memb) const {
        const size_t off = ((size_t)S * B * 4 + 15) & ~(size_t)15;
        const size_t rows = memb ? (size_t)max_tris * (RS + NG) + 2 * (size_t)max_tris * NG : 0;
        return off + (size_t)S * B * 8 + rows * B * 8;
    }

    double compute_dt_std(int r) const {
        double mx = 0.0;
        const bool ftab = real_tabs[r] & stg::kTabDiff, etab = real_tabs[r] & stg::kTabTriDiff;
        for (int t = tet_off[r]; t < tet_off[r + 1]; ++t) {
            for (int s = 0; s < S; ++s) {
                const double D = h_D[(size_t)r * S + s];
                const double *g = &h_face_g[4 * (size_t)t];
                double Df[4];
                for (int f = 0; f < 4; ++f) {
                    const uint32_t k = h_tet_bdir.empty() ? 0xffu : (h_tet_bdir[t] >> (8 * f)) & 0xffu;
                    Df[f] = ftab ? h_tet_fD[((size_t)s * NT + t) * 4 + f] : (k == 0xffu ? D : h_bnd_D[(size_t)k * S + s]);
                }
                const double sum = (((Df[0] * g[0]) + (Df[1] * g[1])) + (Df[2] * g[2])) + (Df[3] * g[3]);
                if (sum > mx) mx = sum;
            }
        }
        // surface diffusion (TetOpSplitP::_computeUpdPeriod, tetopsplit.cpp:5373-5411)
        for (int i = tri_off[r]; i < tri_off[r + 1] && NSD > 0; ++i) {
            for (int q = 0; q < SS; ++q) {
                const double D = h_sD[(size_t)r * SS + q];
                const double *g = &h_tri_g[3 * (size_t)i];
                const double *e = etab ? &h_tri_eD[((size_t)q * NTRI + i) * 3] : nullptr;
                const double sum = e ? ((e[0] * g[0]) + (e[1] * g[1])) + (e[2] * g[2])
                                     : ((D * g[0]) + (D * g[1])) + (D * g[2]);
                if (sum > mx) mx = sum;
            }
        }
        return mx > DBL_EPSILON ? 1.0 / mx : INFINITY;
    }

    // split periods of every realization; compartments share the smallest
    void refresh_dt() {
        dt_std.resize(n_real);
        for (int r = 0; r < n_real; ++r) dt_std[r] = compute_dt_std(r);
        if (linked) {
            const double m = *std::min_element(dt_std.begin(), dt_std.end());
            std::fill(dt_std.begin(), dt_std.end(), m);
        }
    }

    void free_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_launches = -1;
    }

    ~stg_sim() {
        free_graph();
        for (auto &plane : rtc_mods)
            for (auto &row : plane)
                for (hipModule_t m : row)
                    if (m) (void)hipModuleUnload(m);
        void *ptrs[] = {d_vol,       d_nbr,         d_g,           d_counts,
                        d_out[0],    d_out[1],      d_wide[0],     d_wide[1],     d_tri_ptr,     d_tri_counts,  d_tri_area,
                        d_block_desc, d_real_tet0,  d_real_k,      d_real_sk,
                        d_real_D,    d_real_seed,   d_real_stream, d_run_tab,
                        d_ext,       d_err,         d_sum,         d_idx,
                        d_gather,    d_tri_v,       d_vdep_vmin,   d_vdep_dv,     d_vdep_k,
                        d_vdep_off,  d_tri_intg,    d_tri_charge,  d_tet_bdir,    d_bnd_D,
                        d_tri_nbr,   d_tri_nface,   d_tri_g,       d_tri_user,    d_real_sD,
                        d_tout[0],   d_tout[1],     d_tri_corr,    d_real_tabs,   d_tet_k,
                        d_tet_fD,    d_tri_sk,      d_tri_eD,      d_tet_clamp,   d_tri_clamp,
                        d_tet_grp,   d_grp_ptr,     d_grp_mem,     d_grp_real,    d_tri_otet,
                        d_tri_oreal, d_gcorr,       d_real_sums,   d_tri_off};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        if (h_stage) (void)hipHostFree(h_stage);
        if (h_real_sums) (void)hipHostFree(h_real_sums);
        if (h_sum) (void)hipHostFree(h_sum);
        if (h_err) (void)hipHostFree(h_err);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void check_real(const stg_sim *s, int32_t real) {
    if (real < 0 || real >= s->n_real)
        throw StgError{STG_ERR_ARG, "realization index " + std::to_string(real) + " out of range"};
}

void build_model(stg_sim *s, const stg_model *m) {
    if (!m) throw StgError{STG_ERR_ARG, "null model"};
    if (m->n_specs < 0 || m->n_specs > stg::kMaxSpecs)
        throw StgError{STG_ERR_ARG, "n_specs must be in [0, 16]"};
    if (m->n_reacs < 0 || m->n_reacs > stg::kMaxReacs)
        throw StgError{STG_ERR_ARG, "n_reacs must be in [0, 32]"};
    if (m->n_sspecs < 0 || m->n_sspecs > stg::kMaxSSpecs)
        throw StgError{STG_ERR_ARG, "n_sspecs must be in [0, 8]"};
    if (m->n_sreacs < 0 || m->n_sreacs > stg::kMaxSReacs)
        throw StgError{STG_ERR_ARG, "n_sreacs must be in [0, 16]"};
    s->S = m->n_specs;
    s->R = m->n_reacs;
    s->SS = m->n_sspecs;
    s->RS = m->n_sreacs;
    s->outer_model = false;
    stg::ModelK &k = s->mk;
    std::memset(&k, 0, sizeof(k));
    k.n_specs = s->S;
    k.n_reacs = s->R;
    k.n_sspecs = s->SS;
    k.n_sreacs = s->RS;
    int ne = 0;
    auto push = [&](int spec, int val) {
        if (val < -128 || val > 127) throw StgError{STG_ERR_ARG, "stoichiometry out of int8 range"};
        if (ne >= stg::kMaxEntries) throw StgError{STG_ERR_ARG, "model tables too large"};
        k.ent[ne] = (int32_t)(((uint32_t)val << 16) | (uint32_t)spec);
        ++ne;
    };
    for (int r = 0; r < s->R; ++r) {
        k.lhs_off[r] = ne;
        int order = 0;
        for (int sp = 0; sp < s->S; ++sp) {
            const int v = m->reac_lhs[r * s->S + sp];
            if (v < 0 || v > 4)
                throw StgError{STG_ERR_ARG, "reactant multiplicity must be in [0, 4]"};
            if (v) push(sp, v);
            order += v;
        }
        k.order[r] = order;
    }
    k.lhs_off[s->R] = ne;
    for (int r = 0; r < s->R; ++r) {
        k.upd_off[r] = ne;
        for (int sp = 0; sp < s->S; ++sp) {
            const int v = m->reac_upd[r * s->S + sp];
            if (v) push(sp, v);
            if (v < 0 && -v > m->reac_lhs[r * s->S + sp])
                throw StgError{STG_ERR_ARG, "a reaction consumes a species it does not react with"};
        }
    }
    k.upd_off[s->R] = ne;
    for (int r = 0; r < s->RS; ++r) {
        k.slhs_off[r] = ne;
        int order = 0, nvol = 0, nin = 0, nout = 0;
        for (int q = 0; q < s->SS; ++q) {
            const int v = m->sreac_lhs_s[r * s->SS + q];
            if (v < 0 || v > 4) throw StgError{STG_ERR_ARG, "reactant multiplicity must be in [0, 4]"};
            if (v) push(stg::kSurfFlag | q, v);
            order += v;
        }
        for (int q = 0; q < s->S; ++q) {
            const int v = m->sreac_lhs_i[r * s->S + q];
            if (v < 0 || v > 4) throw StgError{STG_ERR_ARG, "reactant multiplicity must be in [0, 4]"};
            if (v) push(q, v);
            order += v;
            nvol += v;
            nin += v;
        }
        for (int q = 0; q < s->S && m->sreac_lhs_o; ++q) {
            const int v = m->sreac_lhs_o[r * s->S + q];
            if (v < 0 || v > 4) throw StgError{STG_ERR_ARG, "reactant multiplicity must be in [0, 4]"};
            if (v) push(stg::kOuterFlag | q, v);
            order += v;
            nvol += v;
            nout += v;
        }
        // model/sreac.cpp:55: volume reactants on one side only
        if (nin > 0 && nout > 0)
            throw StgError{STG_ERR_ARG, "a surface reaction can have inner or outer volume reactants, not both"};
        k.sorder[r] = order;
        k.s_nvol[r] = nvol;
        k.s_out[r] = nout > 0 ? 1 : 0;
    }
    k.slhs_off[s->RS] = ne;
    for (int r = 0; r < s->RS; ++r) {
        k.supd_off[r] = ne;
        for (int q = 0; q < s->SS; ++q) {
            const int v = m->sreac_upd_s[r * s->SS + q];
            if (v) push(stg::kSurfFlag | q, v);
            if (v < 0 && -v > m->sreac_lhs_s[r * s->SS + q])
                throw StgError{STG_ERR_ARG, "a surface reaction consumes a species it does not react with"};
        }
        for (int q = 0; q < s->S; ++q) {
            const int v = m->sreac_upd_i[r * s->S + q];
            if (v) push(q, v);
            if (v < 0 && -v > m->sreac_lhs_i[r * s->S + q])
                throw StgError{STG_ERR_ARG, "a surface reaction consumes a species it does not react with"};
        }
        for (int q = 0; q < s->S && m->sreac_upd_o; ++q) {
            const int v = m->sreac_upd_o[r * s->S + q];
            if (v) push(stg::kOuterFlag | q, v);
            if (v < 0 && -v > (m->sreac_lhs_o ? m->sreac_lhs_o[r * s->S + q] : 0))
                throw StgError{STG_ERR_ARG, "a surface reaction consumes a species it does not react with"};
            if (v) s->outer_model = true;
        }
        if (k.s_out[r]) s->outer_model = true;
    }
    k.supd_off[s->RS] = ne;
    // voltage-dependent rate tables
    s->vdep_vmin.clear();
    s->vdep_dv.clear();
    s->vdep_off.assign(1, 0);
    s->vdep_k.clear();
    const int nv = (m->n_vdep > 0 && m->sreac_vdep) ? m->n_vdep : 0;
    for (int t = 0; t < nv; ++t) {
        const int a = m->vdep_off[t], b = m->vdep_off[t + 1];
        if (b - a < 2 || !(m->vdep_dv[t] > 0.0))
            throw StgError{STG_ERR_ARG, "a k(V) table needs two or more points and dv > 0"};
        s->vdep_vmin.push_back(m->vdep_vmin[t]);
        s->vdep_dv.push_back(m->vdep_dv[t]);
        for (int i = a; i < b; ++i) {
            if (!(m->vdep_k[i] >= 0.0)) throw StgError{STG_ERR_ARG, "k(V) must be >= 0"};
            s->vdep_k.push_back(m->vdep_k[i]);
        }
        s->vdep_off.push_back((int32_t)s->vdep_k.size());
    }
    for (int r = 0; r < s->RS; ++r) {
        const int t = (nv > 0) ? m->sreac_vdep[r] : -1;
        if (t < -1 || t >= nv) throw StgError{STG_ERR_ARG, "k(V) table index out of range"};
        k.s_vdep[r] = t;
    }
    // GHK currents
    if (m->n_ghk < 0 || m->n_ghk > stg::kMaxGhk) throw StgError{STG_ERR_ARG, "n_ghk must be in [0, 4]"};
    k.n_ghk = m->n_ghk;
    s->NG = m->n_ghk;
    for (int g = 0; g < m->n_ghk; ++g) {
        if (m->ghk_chan[g] < 0 || m->ghk_chan[g] >= s->SS)
            throw StgError{STG_ERR_ARG, "GHK channel state must be a surface species"};
        if (m->ghk_ion[g] < 0 || m->ghk_ion[g] >= s->S)
            throw StgError{STG_ERR_ARG, "GHK ion must be a volume species"};
        if (m->ghk_valence[g] == 0 || m->ghk_valence[g] < -127 || m->ghk_valence[g] > 127)
            throw StgError{STG_ERR_ARG, "GHK ion valence must be nonzero"};
        if (!(m->ghk_perm[g] > 0.0)) throw StgError{STG_ERR_ARG, "GHK permeability must be > 0"};
        if (std::isnan(m->ghk_voconc[g])) throw StgError{STG_ERR_ARG, "GHK outer concentration is NaN"};
        k.g_outer[g] = m->ghk_voconc[g] < 0.0 ? 1 : 0;  // the outer tet's concentration
        if (k.g_outer[g]) s->outer_model = true;
        k.g_chan[g] = m->ghk_chan[g];
        k.g_ion[g] = m->ghk_ion[g];
        k.g_z[g] = m->ghk_valence[g];
        k.g_real[g] = m->ghk_realflux[g] ? 1 : 0;
        k.g_perm[g] = m->ghk_perm[g];
        // no outer tet (a triangle outside every SSA group): concentration 0
        k.g_oconc[g] = k.g_outer[g] ? 0.0 : m->ghk_voconc[g];
        k.g_vshift[g] = m->ghk_vshift[g];
    }
    k.temp = 293.15;
}

// Diffusing species and occupancy tracking, from the union of all realizations'
// diffusion constants (reactions.cpp:35-40: tracked = reacts and diffuses).
void build_diffusion_sets(stg_sim *s, const stg_model *m) {
    stg::ModelK &k = s->mk;
    k.n_diff = 0;
    k.tracked_mask = 0;
    for (int sp = 0; sp < s->S; ++sp) {
        bool diff = false;
        for (int r = 0; r < s->n_real; ++r) diff = diff || s->h_D[(size_t)r * s->S + sp] > 0.0;
        if (!diff) continue;
        k.dspec[k.n_diff++] = sp;
        bool reacts = false;
        for (int r = 0; r < s->R; ++r) reacts = reacts || m->reac_upd[r * s->S + sp] != 0;
        for (int r = 0; r < s->RS; ++r) reacts = reacts || m->sreac_upd_i[r * s->S + sp] != 0;
        for (int r = 0; r < s->RS && m->sreac_upd_o; ++r) reacts = reacts || m->sreac_upd_o[r * s->S + sp] != 0;
        for (int g = 0; g < m->n_ghk; ++g) reacts = reacts || (m->ghk_ion[g] == sp && m->ghk_realflux[g]);
        if (reacts) k.tracked_mask |= 1 << sp;
    }
    s->ND = k.n_diff;
}

int fail(stg_sim *s, const StgError &e) {
    if (s) s->err = e.msg;
    else g_create_error = e.msg;
    return e.code;
}

template <typename F>
int guard(stg_sim *s, F &&f) {
    try {
        f();
        return STG_OK;
    } catch (const StgError &e) {
        return fail(s, e);
    } catch (const std::bad_alloc &) {
        return fail(s, StgError{STG_ERR_SYS, "host allocation failed"});
    } catch (const std::exception &e) {
        return fail(s, StgError{STG_ERR_PROG, e.what()});
    }
}

void upload_real_tables(stg_sim *s) {
    if (s->d_real_sD && !s->h_sD.empty())
        HIP_OK(hipMemcpyAsync(s->d_real_sD, s->h_sD.data(), sizeof(double) * s->h_sD.size(),
                              hipMemcpyHostToDevice, s->stream));
    if (!s->h_bnd_D.empty())
        HIP_OK(hipMemcpyAsync(s->d_bnd_D, s->h_bnd_D.data(), sizeof(double) * s->h_bnd_D.size(),
                              hipMemcpyHostToDevice, s->stream));
    if (!s->h_k.empty())
        HIP_OK(hipMemcpyAsync(s->d_real_k, s->h_k.data(), sizeof(double) * s->h_k.size(),
                              hipMemcpyHostToDevice, s->stream));
    if (!s->h_sk.empty())
        HIP_OK(hipMemcpyAsync(s->d_real_sk, s->h_sk.data(), sizeof(double) * s->h_sk.size(),
                              hipMemcpyHostToDevice, s->stream));
    if (!s->h_D.empty())
        HIP_OK(hipMemcpyAsync(s->d_real_D, s->h_D.data(), sizeof(double) * s->h_D.size(),
                              hipMemcpyHostToDevice, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
}

// Room in the run tables for a dt table of tab_len entries (periods x
// realizations). A new allocation changes the tables' addresses, so a
// captured launch sequence is dropped; no kernel is in flight between runs.
void reserve_run_tab(stg_sim *s, size_t tab_len) {
    if (s->d_run_tab && tab_len <= s->dt_tab_cap) return;
    const size_t cap = std::max(tab_len * 2, (size_t)16 * s->n_real);
    double *p = dalloc<double>((size_t)s->n_real + cap);  // 8 B per realization for the two int tables
    if (s->d_run_tab) (void)hipFree(s->d_run_tab);
    s->d_run_tab = p;
    s->d_real_nper = reinterpret_cast<int32_t *>(p);
    s->d_real_step0 = reinterpret_cast<uint32_t *>(s->d_real_nper + s->n_real);
    s->d_dt_tab = p + s->n_real;
    s->dt_tab_cap = cap;
    s->free_graph();
}

void create_impl(stg_sim *s, const stg_model *model, int32_t n_real, const stg_mesh *meshes,
                 const stg_realization *reals, int32_t mode, bool linked = false) {
    if (n_real <= 0) throw StgError{STG_ERR_ARG, "need at least one realization"};
    if (!meshes || !reals) throw StgError{STG_ERR_ARG, "null mesh or realization array"};
    if (mode != STG_MODE_TETOPSPLIT && mode != STG_MODE_DISTTETOPSPLIT)
        throw StgError{STG_ERR_ARG, "unknown solver mode"};
    s->mode = mode;
    s->n_real = n_real;
    s->linked = linked;
    build_model(s, model);
    const int S = s->S, R = s->R, RS = s->RS, SS = s->SS;
    s->tet_off.assign(n_real + 1, 0);
    s->tri_off.assign(n_real + 1, 0);
    for (int r = 0; r < n_real; ++r) {
        const stg_mesh &m = meshes[r];
        if (m.n_tets <= 0) throw StgError{STG_ERR_ARG, "mesh has no tetrahedra"};
        if (m.n_tris < 0) throw StgError{STG_ERR_ARG, "negative triangle count"};
        if (!m.tet_vol || !m.tet_nbr || !m.tet_nbr_face || !m.tet_face_g)
            throw StgError{STG_ERR_ARG, "null mesh array"};
        if (m.n_tris > 0 && (!m.tri_tet || !m.tri_area))
            throw StgError{STG_ERR_ARG, "null triangle array"};
        if (!linked && (m.tet_nbr_comp || m.tet_face_bdir))
            throw StgError{STG_ERR_ARG, "compartment links need stg_create_comps"};
        const long long nt = (long long)s->tet_off[r] + m.n_tets;
        const long long ntr = (long long)s->tri_off[r] + m.n_tris;
        if (nt > INT32_MAX / 4 || ntr > INT32_MAX) throw StgError{STG_ERR_ARG, "mesh too large"};
        s->tet_off[r + 1] = (int32_t)nt;
        s->tri_off[r + 1] = (int32_t)ntr;
    }
    s->NT = s->tet_off[n_real];
    s->NTRI = s->tri_off[n_real];
    const int NT = s->NT, NTRI = s->NTRI;

    // host staging of the concatenated mesh
    std::vector<double> vol(NT);
    std::vector<int4> nbr(NT);
    s->h_face_g.assign((size_t)NT * 4, 0.0);
    std::vector<int32_t> tri_ptr(NT + 1, 0);
    std::vector<double> tri_area(std::max(NTRI, 1));
    s->tri_perm.assign(n_real, {});
    s->max_tris = 0;
    int32_t nbd = 0;
    for (int r = 0; r < n_real; ++r)
        if (meshes[r].tet_face_bdir)
            for (int i = 0; i < 4 * meshes[r].n_tets; ++i) nbd = std::max(nbd, meshes[r].tet_face_bdir[i] + 1);
    if (nbd > 255) throw StgError{STG_ERR_ARG, "at most 255 diffusion-boundary directions"};
    s->NBD = nbd;
    if (nbd > 0) s->h_tet_bdir.assign(NT, 0xffffffffu);
    for (int r = 0; r < n_real; ++r) {
        const stg_mesh &m = meshes[r];
        const int32_t t0 = s->tet_off[r];
        for (int t = 0; t < m.n_tets; ++t) {
            vol[t0 + t] = m.tet_vol[t];
            int nb[4];
            for (int f = 0; f < 4; ++f) {
                const int32_t j = m.tet_nbr[4 * t + f];
                const int32_t jf = m.tet_nbr_face[4 * t + f];
                const double g = m.tet_face_g[4 * t + f];
                const int32_t jr = m.tet_nbr_comp ? m.tet_nbr_comp[4 * t + f] : r;
                if (jr < 0 || jr >= n_real) throw StgError{STG_ERR_ARG, "neighbour compartment out of range"};
                const stg_mesh &mj = meshes[jr];
                if (j >= mj.n_tets || j < -1) throw StgError{STG_ERR_ARG, "neighbour index out of range"};
                if (j >= 0 && (jf < 0 || jf > 3)) throw StgError{STG_ERR_ARG, "neighbour face slot out of range"};
                if (j >= 0 && (mj.tet_nbr[4 * j + jf] != t ||
                               (mj.tet_nbr_comp ? mj.tet_nbr_comp[4 * j + jf] : jr) != r))
                    throw StgError{STG_ERR_ARG, "neighbour table is not symmetric"};
                if (!(g >= 0.0) || !std::isfinite(g)) throw StgError{STG_ERR_ARG, "face factor must be finite and >= 0"};
                if (j < 0 && g != 0.0) throw StgError{STG_ERR_ARG, "boundary face with nonzero face factor"};
                const int32_t bd = m.tet_face_bdir ? m.tet_face_bdir[4 * t + f] : -1;
                if (bd < -1 || (bd >= 0 && j < 0))
                    throw StgError{STG_ERR_ARG, "a diffusion-boundary face needs a neighbour"};
                if (jr != r && j >= 0 && bd < 0)
                    throw StgError{STG_ERR_ARG, "a face between compartments must be closed or on a diffusion boundary"};
                if (bd >= 0) {
                    s->h_tet_bdir[t0 + t] &= ~(0xffu << (8 * f));
                    s->h_tet_bdir[t0 + t] |= (uint32_t)bd << (8 * f);
                }
                // the entering gather reads the neighbour's slot for this face directly
                nb[f] = j >= 0 ? 4 * (j + s->tet_off[jr]) + jf : -1;
                s->h_face_g[4 * (size_t)(t0 + t) + f] = g;
            }
            nbr[t0 + t] = make_int4(nb[0], nb[1], nb[2], nb[3]);
        }
        // triangles sorted (stably) by inner tet
        std::vector<int32_t> order(m.n_tris);
        for (int i = 0; i < m.n_tris; ++i) {
            if (m.tri_tet[i] < 0 || m.tri_tet[i] >= m.n_tets)
                throw StgError{STG_ERR_ARG, "triangle inner tet out of range"};
            order[i] = i;
        }
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return m.tri_tet[a] < m.tri_tet[b]; });
        s->tri_perm[r].assign(m.n_tris, 0);
        for (int i = 0; i < m.n_tris; ++i) {
            const int32_t ui = order[i];
            s->tri_perm[r][ui] = i;
            tri_area[s->tri_off[r] + i] = m.tri_area[ui];
            tri_ptr[t0 + m.tri_tet[ui] + 1]++;
        }
    }
    for (int t = 0; t < NT; ++t) {
        s->max_tris = std::max(s->max_tris, tri_ptr[t + 1]);
        tri_ptr[t + 1] += tri_ptr[t];
    }
    if (s->max_tris > 4) throw StgError{STG_ERR_ARG, "a tet has more than four triangles"};

    // surface diffusion: edge neighbours in internal triangle order (in
    // another compartment's patch when tri_nbr_comp says so)
    std::vector<int4> tri_nbr(std::max(NTRI, 1), make_int4(-1, -1, -1, -1));
    std::vector<uint32_t> tri_nface(std::max(NTRI, 1), 0);
    std::vector<int32_t> tri_user(std::max(NTRI, 1), 0);
    s->h_tri_g.assign(3 * (size_t)std::max(NTRI, 1), 0.0);
    for (int r = 0; r < n_real; ++r) {
        const stg_mesh &m = meshes[r];
        const int32_t i0 = s->tri_off[r];
        if (!linked && m.tri_nbr_comp) throw StgError{STG_ERR_ARG, "compartment links need stg_create_comps"};
        for (int u = 0; u < m.n_tris; ++u) {
            const int32_t i = i0 + s->tri_perm[r][u];
            tri_user[i] = u;
            if (!m.tri_nbr) continue;
            int nb[3] = {-1, -1, -1};
            for (int e = 0; e < 3; ++e) {
                const int32_t j = m.tri_nbr[3 * u + e];
                const int32_t je = m.tri_nbr_edge ? m.tri_nbr_edge[3 * u + e] : -1;
                const double g = m.tri_edge_g ? m.tri_edge_g[3 * u + e] : 0.0;
                const int32_t jr = m.tri_nbr_comp ? m.tri_nbr_comp[3 * u + e] : r;
                if (jr < 0 || jr >= n_real) throw StgError{STG_ERR_ARG, "triangle neighbour compartment out of range"};
                const stg_mesh &mj = meshes[jr];
                if (j < -1 || j >= mj.n_tris) throw StgError{STG_ERR_ARG, "triangle neighbour out of range"};
                if (j >= 0 && (je < 0 || je > 2 || !mj.tri_nbr || mj.tri_nbr[3 * j + je] != u ||
                               (mj.tri_nbr_comp ? mj.tri_nbr_comp[3 * j + je] : jr) != r))
                    throw StgError{STG_ERR_ARG, "triangle neighbour table is not symmetric"};
                if (!(g >= 0.0) || !std::isfinite(g) || (j < 0 && g != 0.0))
                    throw StgError{STG_ERR_ARG, "edge factor must be finite, >= 0 and 0 on open edges"};
                nb[e] = j >= 0 ? s->tri_off[jr] + s->tri_perm[jr][j] : -1;
                tri_nface[i] |= (uint32_t)(j >= 0 ? je : 0) << (8 * e);
                s->h_tri_g[3 * (size_t)i + e] = g;
            }
            tri_nbr[i] = make_int4(nb[0], nb[1], nb[2], -1);
        }
    }

    // outer tets of the triangles (internal order, global tets) and the SSA
    // groups they make: union of inner and outer tet over every triangle
    // with an outer tet, when the model reaches the outer side
    std::vector<int32_t> tri_otet(std::max(NTRI, 1), -1), tri_oreal(std::max(NTRI, 1), -1);
    bool any_outer = false;
    for (int r = 0; r < n_real; ++r) {
        const stg_mesh &m = meshes[r];
        if (!m.tri_otet) continue;
        if (!linked && m.tri_ocomp) throw StgError{STG_ERR_ARG, "compartment links need stg_create_comps"};
        for (int u = 0; u < m.n_tris; ++u) {
            const int32_t o = m.tri_otet[u];
            if (o < 0) continue;
            const int32_t orl = m.tri_ocomp ? m.tri_ocomp[u] : r;
            if (orl < 0 || orl >= n_real || o >= meshes[orl].n_tets)
                throw StgError{STG_ERR_ARG, "triangle outer tet out of range"};
            if (orl == r && o == m.tri_tet[u]) throw StgError{STG_ERR_ARG, "a triangle's outer tet is its inner tet"};
            const int32_t i = s->tri_off[r] + s->tri_perm[r][u];
            tri_otet[i] = s->tet_off[orl] + o;
            tri_oreal[i] = orl;
            any_outer = true;
        }
    }
    if (s->outer_model && any_outer) {
        std::vector<int32_t> parent(NT);
        for (int t = 0; t < NT; ++t) parent[t] = t;
        auto find = [&parent](int32_t x) {
            while (parent[x] != x) x = parent[x] = parent[parent[x]];
            return x;
        };
        for (int t = 0; t < NT; ++t)
            for (int i = tri_ptr[t]; i < tri_ptr[t + 1]; ++i)
                if (tri_otet[i] >= 0) {
                    const int32_t a = find(t), b = find(tri_otet[i]);
                    if (a != b) parent[std::max(a, b)] = std::min(a, b);  // the root is the smallest tet
                }
        std::vector<int32_t> size(NT, 0);
        for (int t = 0; t < NT; ++t) size[find(t)] += 1;
        s->h_tet_grp.assign(NT, -1);
        s->h_grp_ptr.assign(1, 0);
        s->h_grp_mem.clear();
        s->h_grp_real.clear();
        std::vector<int32_t> gid(NT, -1);
        for (int t = 0; t < NT; ++t) {  // members in tet order, so the root (smallest) comes first
            const int32_t root = find(t);
            if (size[root] < 2) continue;
            if (gid[root] < 0) {
                gid[root] = (int32_t)s->h_grp_ptr.size() - 1;
                s->h_grp_ptr.push_back(0);
            }
        }
        std::vector<std::vector<int32_t>> mem(s->h_grp_ptr.size() - 1);
        for (int t = 0; t < NT; ++t) {
            const int32_t root = find(t);
            if (size[root] < 2) continue;
            mem[gid[root]].push_back(t);
            s->h_tet_grp[t] = gid[root];
        }
        for (size_t g = 0; g < mem.size(); ++g) {
            for (int32_t t : mem[g]) {
                s->h_grp_mem.push_back(t);
                const int32_t r = (int32_t)(std::upper_bound(s->tet_off.begin(), s->tet_off.end(), t) - s->tet_off.begin()) - 1;
                s->h_grp_real.push_back(r);
            }
            s->h_grp_ptr[g + 1] = (int32_t)s->h_grp_mem.size();
        }
    }

    // per-realization constants
    s->h_k.assign((size_t)n_real * R, 0.0);
    s->h_sk.assign((size_t)n_real * RS, 0.0);
    s->h_D.assign((size_t)n_real * S, 0.0);
    s->seeds.resize(n_real);
    s->streams.resize(n_real);
    for (int r = 0; r < n_real; ++r) {
        const stg_realization &q = reals[r];
        for (int j = 0; j < R; ++j) s->h_k[(size_t)r * R + j] = q.reac_k ? q.reac_k[j] : model->reac_k[j];
        for (int j = 0; j < RS; ++j) s->h_sk[(size_t)r * RS + j] = q.sreac_k ? q.sreac_k[j] : model->sreac_k[j];
        for (int j = 0; j < S; ++j) {
            const double D = q.spec_dcst ? q.spec_dcst[j] : model->spec_dcst[j];
            if (!(D >= 0.0) || !std::isfinite(D)) throw StgError{STG_ERR_ARG, "diffusion constant must be finite and >= 0"};
            s->h_D[(size_t)r * S + j] = D;
        }
        for (int j = 0; j < R; ++j)
            if (!(s->h_k[(size_t)r * R + j] >= 0.0)) throw StgError{STG_ERR_ARG, "rate constant must be >= 0"};
        for (int j = 0; j < RS; ++j)
            if (!(s->h_sk[(size_t)r * RS + j] >= 0.0)) throw StgError{STG_ERR_ARG, "rate constant must be >= 0"};
        s->seeds[r] = q.seed;
        s->streams[r] = q.stream;
    }
    s->h_bnd_D.assign((size_t)s->NBD * S, 0.0);
    s->h_sD.assign((size_t)n_real * SS, 0.0);
    for (int r = 0; r < n_real; ++r)
        for (int q = 0; q < SS; ++q) {
            const double D = model->sspec_dcst ? model->sspec_dcst[q] : 0.0;
            if (!(D >= 0.0) || !std::isfinite(D)) throw StgError{STG_ERR_ARG, "surface diffusion constant must be finite and >= 0"};
            s->h_sD[(size_t)r * SS + q] = D;
        }
    {
        stg::ModelK &k = s->mk;
        k.n_sdiff = 0;
        k.tri_tracked_mask = 0;
        for (int q = 0; q < SS; ++q) {
            if (!(model->sspec_dcst && model->sspec_dcst[q] > 0.0)) continue;
            k.sdspec[k.n_sdiff++] = q;
            for (int j = 0; j < RS; ++j)
                if (model->sreac_upd_s[j * SS + q] != 0) k.tri_tracked_mask |= 1 << q;
        }
        s->NSD = k.n_sdiff;
    }
    build_diffusion_sets(s, model);
    (void)SS;
    s->real_tabs.assign(n_real, 0);
    s->refresh_dt();
    s->nsteps.assign(n_real, 0);
    s->rng_steps.assign(n_real, 0);

    HIP_OK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    HIP_OK(hipEventCreate(&s->ev0));
    HIP_OK(hipEventCreate(&s->ev1));
    s->d_vol = dalloc<double>(NT);
    s->d_nbr = dalloc<int4>(NT);
    s->d_g = dalloc<double2>((size_t)NT * 2);
    s->d_counts = dalloc<int32_t>((size_t)std::max(S, 1) * NT);
    s->d_out[0] = dalloc<int16_t>((size_t)std::max(s->ND, 1) * NT * 4);
    s->d_out[1] = dalloc<int16_t>((size_t)std::max(s->ND, 1) * NT * 4);
    s->d_wide[0] = dalloc<int32_t>((size_t)std::max(s->ND, 1) * NT * 4);
    s->d_wide[1] = dalloc<int32_t>((size_t)std::max(s->ND, 1) * NT * 4);
    s->d_tri_ptr = dalloc<int32_t>(NT + 1);
    s->d_tri_counts = dalloc<int32_t>((size_t)std::max(SS, 1) * std::max(NTRI, 1));
    s->d_tri_area = dalloc<double>(std::max(NTRI, 1));
    s->d_real_tet0 = dalloc<int32_t>(n_real + 1);
    s->d_real_k = dalloc<double>(std::max<size_t>(1, s->h_k.size()));
    s->d_real_sk = dalloc<double>(std::max<size_t>(1, s->h_sk.size()));
    s->d_real_D = dalloc<double>(std::max<size_t>(1, s->h_D.size()));
    if (s->NSD > 0) {
        const size_t nt = std::max(NTRI, 1);
        s->d_tri_nbr = dalloc<int4>(nt);
        s->d_tri_nface = dalloc<uint32_t>(nt);
        s->d_tri_g = dalloc<double>(3 * nt);
        s->d_tri_user = dalloc<int32_t>(nt);
        s->d_real_sD = dalloc<double>(std::max<size_t>(1, s->h_sD.size()));
        s->d_tri_corr = dalloc<double>((size_t)std::max(SS, 1) * nt);
        for (int b = 0; b < 2; ++b) {
            s->d_tout[b] = dalloc<int32_t>((size_t)s->NSD * nt * 4);
            HIP_OK(hipMemset(s->d_tout[b], 0, sizeof(int32_t) * (size_t)s->NSD * nt * 4));
        }
        HIP_OK(hipMemcpy(s->d_tri_nbr, tri_nbr.data(), sizeof(int4) * nt, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->d_tri_nface, tri_nface.data(), sizeof(uint32_t) * nt, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->d_tri_g, s->h_tri_g.data(), sizeof(double) * 3 * nt, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->d_tri_user, tri_user.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice));
    }
    if (s->NBD > 0) {
        s->d_tet_bdir = dalloc<uint32_t>(NT);
        HIP_OK(hipMemcpy(s->d_tet_bdir, s->h_tet_bdir.data(), sizeof(uint32_t) * NT, hipMemcpyHostToDevice));
        s->d_bnd_D = dalloc<double>((size_t)s->NBD * S);
    }
    s->d_real_seed = dalloc<uint64_t>(n_real);
    s->d_real_stream = dalloc<uint32_t>(n_real);
    reserve_run_tab(s, (size_t)16 * n_real);
    s->d_err = dalloc<int32_t>(1);
    s->d_sum = dalloc<unsigned long long>(1);
    HIP_OK(hipMemcpy(s->d_vol, vol.data(), sizeof(double) * NT, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(s->d_nbr, nbr.data(), sizeof(int4) * NT, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(s->d_g, s->h_face_g.data(), sizeof(double) * 4 * (size_t)NT, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(s->d_tri_ptr, tri_ptr.data(), sizeof(int32_t) * (NT + 1), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(s->d_tri_area, tri_area.data(), sizeof(double) * std::max(NTRI, 1), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(s->d_real_tet0, s->tet_off.data(), sizeof(int32_t) * (n_real + 1), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(s->d_real_seed, s->seeds.data(), sizeof(uint64_t) * n_real, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(s->d_real_stream, s->streams.data(), sizeof(uint32_t) * n_real, hipMemcpyHostToDevice));
    HIP_OK(hipMemset(s->d_counts, 0, sizeof(int32_t) * (size_t)std::max(S, 1) * NT));
    HIP_OK(hipMemset(s->d_tri_counts, 0, sizeof(int32_t) * (size_t)std::max(SS, 1) * std::max(NTRI, 1)));
    HIP_OK(hipMemset(s->d_err, 0, sizeof(int32_t)));
    if (s->NG > 0) {
        s->d_tri_charge = dalloc<int32_t>((size_t)s->NG * std::max(NTRI, 1));
        HIP_OK(hipMemset(s->d_tri_charge, 0, sizeof(int32_t) * (size_t)s->NG * std::max(NTRI, 1)));
    }
    s->d_tri_v = dalloc<double>(std::max(NTRI, 1));
    {
        std::vector<double> v0(std::max(NTRI, 1), -65.0e-3);  // DEFAULT_MEMB_POT
        HIP_OK(hipMemcpy(s->d_tri_v, v0.data(), sizeof(double) * v0.size(), hipMemcpyHostToDevice));
    }
    s->d_vdep_vmin = dalloc<double>(std::max<size_t>(1, s->vdep_vmin.size()));
    s->d_vdep_dv = dalloc<double>(std::max<size_t>(1, s->vdep_dv.size()));
    s->d_vdep_off = dalloc<int32_t>(std::max<size_t>(1, s->vdep_off.size()));
    s->d_vdep_k = dalloc<double>(std::max<size_t>(1, s->vdep_k.size()));
    if (!s->vdep_k.empty()) {
        HIP_OK(hipMemcpy(s->d_vdep_vmin, s->vdep_vmin.data(), sizeof(double) * s->vdep_vmin.size(), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->d_vdep_dv, s->vdep_dv.data(), sizeof(double) * s->vdep_dv.size(), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->d_vdep_off, s->vdep_off.data(), sizeof(int32_t) * s->vdep_off.size(), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->d_vdep_k, s->vdep_k.data(), sizeof(double) * s->vdep_k.size(), hipMemcpyHostToDevice));
    }
    if (!s->h_grp_mem.empty()) {
        s->d_tet_grp = dalloc<int32_t>(NT);
        s->d_grp_ptr = dalloc<int32_t>(s->h_grp_ptr.size());
        s->d_grp_mem = dalloc<int32_t>(s->h_grp_mem.size());
        s->d_grp_real = dalloc<int32_t>(s->h_grp_real.size());
        s->d_gcorr = dalloc<double>((size_t)std::max(S, 1) * NT);
        s->d_tri_otet = dalloc<int32_t>(std::max(NTRI, 1));
        s->d_tri_oreal = dalloc<int32_t>(std::max(NTRI, 1));
        HIP_OK(hipMemcpy(s->d_tet_grp, s->h_tet_grp.data(), sizeof(int32_t) * NT, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->d_grp_ptr, s->h_grp_ptr.data(), sizeof(int32_t) * s->h_grp_ptr.size(), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->d_grp_mem, s->h_grp_mem.data(), sizeof(int32_t) * s->h_grp_mem.size(), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->d_grp_real, s->h_grp_real.data(), sizeof(int32_t) * s->h_grp_real.size(),
                         hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->d_tri_otet, tri_otet.data(), sizeof(int32_t) * tri_otet.size(), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->d_tri_oreal, tri_oreal.data(), sizeof(int32_t) * tri_oreal.size(), hipMemcpyHostToDevice));
        HIP_OK(hipMemset(s->d_gcorr, 0, sizeof(double) * (size_t)std::max(S, 1) * NT));
    }
    upload_real_tables(s);
}

// Block decomposition for a block size: every realization starts on a fresh
// block so a block never straddles two realizations.
void build_blocks(stg_sim *s, int B) {
    s->block_real.clear();
    s->block_tet0.clear();
    for (int r = 0; r < s->n_real; ++r) {
        for (int t = s->tet_off[r]; t < s->tet_off[r + 1]; t += B) {
            s->block_real.push_back(r);
            s->block_tet0.push_back(t);
        }
    }
    const size_t nb = s->block_real.size();
    if (s->d_block_desc) (void)hipFree(s->d_block_desc);
    if (s->d_ext) (void)hipFree(s->d_ext);
    s->d_block_desc = dalloc<int4>(nb);
    s->d_ext = dalloc<unsigned long long>(2 * nb);
    HIP_OK(hipMemset(s->d_ext, 0, sizeof(unsigned long long) * 2 * nb));
    s->desc_np.clear();  // the launch order is uploaded at the next run
    s->free_graph();
}

// Orders the blocks for a run in which realization r takes nper[r] periods:
// a block takes part in launch l when its realization has a period to enter
// (1 <= l <= np) or to step (l < np), i.e. when np >= max(l, 1). Blocks sorted
// by np, most first (stable, so realizations keep their tet order), make the
// participants of every launch a prefix: launch l dispatches only those,
// rather than every block of the ensemble with most returning at once.
void order_blocks(stg_sim *s, const std::vector<int32_t> &nper, int L) {
    // the order and the blocks of each launch follow from the blocks'
    // realizations (fixed until the blocks are rebuilt, which clears desc_np)
    // and the periods of each realization: runs of equal length reuse them
    if (!s->desc_np.empty() && nper == s->desc_nper) return;
    s->desc_nper.clear();  // set again once the order is in place
    const size_t nb = s->block_real.size();
    std::vector<int32_t> np(nb);
    for (size_t b = 0; b < nb; ++b) np[b] = nper[s->block_real[b]];
    std::vector<int32_t> at_least((size_t)L + 2, 0);  // blocks with np == v, then >= v
    for (size_t b = 0; b < nb; ++b) at_least[np[b]] += 1;
    for (int v = L - 1; v >= 0; --v) at_least[v] += at_least[v + 1];
    s->launch_blocks.assign((size_t)L + 1, 0);
    for (int l = 0; l <= L; ++l) s->launch_blocks[l] = at_least[std::max(l, 1)];
    if (np == s->desc_np) {
        s->desc_nper = nper;
        return;
    }
    std::vector<int32_t> pos((size_t)L + 2, 0);  // first slot of each np value
    for (int v = L - 1; v >= 0; --v) pos[v] = at_least[v + 1];
    std::vector<int4> desc(nb);
    for (size_t b = 0; b < nb; ++b) {
        const int r = s->block_real[b];
        desc[pos[np[b]]++] = make_int4((int)b, r, s->block_tet0[b], s->tet_off[r + 1]);
    }
    HIP_OK(hipMemcpyAsync(s->d_block_desc, desc.data(), sizeof(int4) * nb, hipMemcpyHostToDevice, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));  // the host vector goes out of scope
    s->desc_np.swap(np);
    s->desc_nper = nper;
}

stg::SysK make_sys(const stg_sim *s, int launch_parity) {
    stg::SysK k{};
    k.n_tets = s->NT;
    k.n_tris = s->NTRI;
    k.n_real = s->n_real;
    k.xcd_remap = s->xcd_remap ? 1 : 0;
    k.n_blocks = (int32_t)s->block_real.size();  // set per launch by launch_all
    k.diff_threshold = s->diff_threshold;
    k.k_max = std::max(1, s->R + s->max_tris * (s->RS + s->NG));
    k.vol = s->d_vol;
    k.nbr = s->d_nbr;
    k.g = s->d_g;
    k.counts = s->d_counts;
    k.out_prev = s->d_out[launch_parity ^ 1];
    k.out_cur = s->d_out[launch_parity];
    k.wide_prev = s->d_wide[launch_parity ^ 1];
    k.wide_cur = s->d_wide[launch_parity];
    k.tri_ptr = s->NTRI > 0 ? s->d_tri_ptr : nullptr;
    k.tri_counts = s->d_tri_counts;
    k.tri_area = s->d_tri_area;
    k.block_desc = s->d_block_desc;
    k.real_tet0 = s->d_real_tet0;
    k.real_k = s->d_real_k;
    k.real_sk = s->d_real_sk;
    k.real_D = s->d_real_D;
    k.real_seed = s->d_real_seed;
    k.real_stream = s->d_real_stream;
    k.real_step0 = s->d_real_step0;
    k.real_nper = s->d_real_nper;
    k.dt_tab = s->d_dt_tab;
    k.ext = s->d_ext;
    k.err = s->d_err;
    k.tri_v = s->d_tri_v;
    k.vdep_vmin = s->d_vdep_vmin;
    k.vdep_dv = s->d_vdep_dv;
    k.vdep_off = s->d_vdep_off;
    k.vdep_k = s->d_vdep_k;
    k.tri_intg = s->d_tri_intg;
    k.tri_charge = s->d_tri_charge;
    k.tet_bdir = s->d_tet_bdir;
    k.bnd_D = s->d_bnd_D;
    k.tri_nbr = s->d_tri_nbr;
    k.tri_nface = s->d_tri_nface;
    k.tri_g = s->d_tri_g;
    k.tri_user = s->d_tri_user;
    k.real_sD = s->d_real_sD;
    k.tout_prev = s->d_tout[launch_parity ^ 1];
    k.tout_cur = s->d_tout[launch_parity];
    k.tri_corr = s->d_tri_corr;
    if (s->has_groups()) {
        k.tet_grp = s->d_tet_grp;
        k.grp_ptr = s->d_grp_ptr;
        k.grp_mem = s->d_grp_mem;
        k.grp_real = s->d_grp_real;
        k.gcorr = s->d_gcorr;
        k.tri_otet = s->d_tri_otet;
        k.tri_oreal = s->d_tri_oreal;
    }
    if (s->has_tabs()) {
        k.real_tabs = s->d_real_tabs;
        k.tet_k = s->d_tet_k;
        k.tet_fD = s->d_tet_fD;
        k.tet_clamp = s->d_tet_clamp;
        k.tri_sk = s->d_tri_sk;
        k.tri_eD = s->d_tri_eD;
        k.tri_clamp = s->d_tri_clamp;
    }
    return k;
}

// the membrane variant of the period kernel is needed when a triangle can
// hold a surface reaction, a GHK current or an integrated channel count
bool needs_memb(const stg_sim *s) { return s->needs_memb(); }

// index of a lanes-per-tet value (1, 2, 4, 8) in the compiled variants
int lanes_index(int lanes) { return lanes == 1 ? 0 : lanes == 2 ? 1 : lanes == 4 ? 2 : 3; }

// index of a (lanes per tet, slot width) pair among the specialised variants
int variant_index(int lanes, int slot_bytes) { return lanes_index(lanes) + (slot_bytes == 1 ? 4 : 0); }

// tets per block of the current launch shape
int tets_per_block(const stg_sim *s) { return s->block / s->lanes; }

using PeriodKernel = void (*)(stg::ModelK, stg::SysK, int);

template <class SlotT>
PeriodKernel period_kernel_w(const stg_sim *s, int B) {
    using stg::DynTables;
    if (needs_memb(s))
        return B <= 256 ? stg::k_period<256, true, DynTables, 1, SlotT> : stg::k_period<1024, true, DynTables, 1, SlotT>;
    static const PeriodKernel vol[2][4] = {
        {stg::k_period<256, false, DynTables, 1, SlotT>, stg::k_period<256, false, DynTables, 2, SlotT>,
         stg::k_period<256, false, DynTables, 4, SlotT>, stg::k_period<256, false, DynTables, 8, SlotT>},
        {stg::k_period<1024, false, DynTables, 1, SlotT>, stg::k_period<1024, false, DynTables, 2, SlotT>,
         stg::k_period<1024, false, DynTables, 4, SlotT>, stg::k_period<1024, false, DynTables, 8, SlotT>}};
    return vol[B <= 256 ? 0 : 1][lanes_index(s->lanes)];
}

PeriodKernel period_kernel(const stg_sim *s, int B) {
    return s->slot_bytes == 1 ? period_kernel_w<int8_t>(s, B) : period_kernel_w<int16_t>(s, B);
}

// ------------------------------------------------------------- specialisation
//
// The volume kernel compiled at run time with this model's reaction network
// as compile-time constants (stg::DynTables describes the fields): the loops
// over reactions, reactants and diffusing species unroll and the table reads
// disappear. Same operations in the same order, so results are identical.

// Code objects compiled in this process, by program source and options: a
// second engine with the same model and bound loads the same code.
std::mutex g_rtc_mutex;
std::map<std::string, std::pair<std::vector<char>, std::string>> g_rtc_cache;

std::string rtc_join(const char *const *parts, size_t n) {
    std::string out;
    for (size_t i = 0; i < n; ++i) out += parts[i];
    return out;
}

// The model's tables as a struct of constants
std::string spec_tables_source(const stg_sim *s) {
    const stg::ModelK &k = s->mk;
    std::ostringstream o;
    auto arr = [&o](const char *name, const std::vector<int> &v) {
        o << "    static constexpr int " << name << "[" << std::max<size_t>(v.size(), 1) << "] = {";
        if (v.empty()) o << "0";
        for (size_t i = 0; i < v.size(); ++i) o << (i ? ", " : "") << v[i];
        o << "};\n";
    };
    const int R = k.n_reacs;
    std::vector<int> order, lhs_off, lhs_sp, lhs_mult, upd_off, upd_sp, upd_dv, dspec;
    for (int q = 0; q < R; ++q) order.push_back(k.order[q]);
    for (int q = 0; q <= R; ++q) {
        lhs_off.push_back(k.lhs_off[q] - k.lhs_off[0]);
        upd_off.push_back(k.upd_off[q] - k.upd_off[0]);
    }
    for (int e = k.lhs_off[0]; e < k.lhs_off[R]; ++e) {
        lhs_sp.push_back(k.ent[e] & 0xffff);
        lhs_mult.push_back(k.ent[e] >> 16);
    }
    for (int e = k.upd_off[0]; e < k.upd_off[R]; ++e) {
        upd_sp.push_back(k.ent[e] & 0xffff);
        upd_dv.push_back(k.ent[e] >> 16);
    }
    for (int d = 0; d < k.n_diff; ++d) dspec.push_back(k.dspec[d]);
    o << "struct StgModelTables {\n    static constexpr bool kStatic = true;\n";
    o << "    static constexpr int S = " << k.n_specs << ", R = " << R << ", n_diff = " << k.n_diff
      << ", tracked_mask = " << k.tracked_mask << ";\n";
    arr("order", order);
    arr("lhs_off", lhs_off);
    arr("lhs_sp", lhs_sp);
    arr("lhs_mult", lhs_mult);
    arr("upd_off", upd_off);
    arr("upd_sp", upd_sp);
    arr("upd_dv", upd_dv);
    arr("dspec", dspec);
    o << "};\n";
    return o.str();
}

void specialize(stg_sim *s, int min_waves) {
    if (min_waves < 0 || min_waves > 8) throw StgError{STG_ERR_ARG, "waves per SIMD bound must be in [0, 8]"};
    const int memb = s->needs_memb() ? 1 : 0;
    // the membrane variant has one lane per tet (a run that needs it goes back to one)
    const int lanes = memb ? 1 : s->lanes;
    const int li = variant_index(lanes, s->slot_bytes);
    if (s->rtc_fns[memb][min_waves][li]) {
        if (s->rtc_fn != s->rtc_fns[memb][min_waves][li]) s->free_graph();
        s->rtc_fn = s->rtc_fns[memb][min_waves][li];
        s->rtc_memb = memb != 0;
        s->rtc_waves = min_waves;
        return;
    }
    const std::string kern = rtc_join(kRtcKernels, sizeof(kRtcKernels) / sizeof(kRtcKernels[0]));
    const std::string rnd = rtc_join(kRtcRandom, sizeof(kRtcRandom) / sizeof(kRtcRandom[0]));
    const std::string lfact = rtc_join(kRtcLfact, sizeof(kRtcLfact) / sizeof(kRtcLfact[0]));
    const std::string name = std::string("stg::k_period<256, ") + (memb ? "true" : "false") + ", StgModelTables, " +
                             std::to_string(lanes) + (s->slot_bytes == 1 ? ", int8_t>" : ", int16_t>");
    std::string src =
        "using __hip_internal::int8_t; using __hip_internal::uint8_t;\n"
        "using __hip_internal::int16_t; using __hip_internal::uint16_t;\n"
        "using __hip_internal::int32_t; using __hip_internal::uint32_t;\n"
        "using __hip_internal::int64_t; using __hip_internal::uint64_t;\n"
        "#include \"stg_kernels.hpp\"\n" +
        spec_tables_source(s) + "template __global__ void " + name +
        "(const stg::ModelK, const stg::SysK, int);\n";
    const char *hdrs[] = {kern.c_str(), rnd.c_str(), lfact.c_str()};
    const char *names[] = {"stg_kernels.hpp", "stg_random.hpp", "stg_lfact_table.inc"};
    std::vector<char> code;
    std::string lowered, log;
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, dev));
    const std::string arch = std::string("--offload-arch=") + prop.gcnArchName;
    // a bound of n waves per SIMD caps the kernel at 512 / n registers per lane
    const std::string waves = std::string(memb ? "-DSTG_PERIOD_WAVES_PER_EU=" : "-DSTG_VOL_WAVES_PER_EU=") +
                              std::to_string(min_waves > 0 ? min_waves : 1);
    // no machine-level loop-invariant motion: it holds constants and addresses
    // in vector registers across the event loop and costs occupancy
    const char *opts[] = {arch.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
                          "-Xclang", "-packed-fp32-ops", "-mllvm", "-disable-machine-licm", waves.c_str()};
    const std::string key = arch + "\n" + waves + "\n" + src;
    {
        std::lock_guard<std::mutex> lk(g_rtc_mutex);
        auto it = g_rtc_cache.find(key);
        if (it != g_rtc_cache.end()) {
            code = it->second.first;
            lowered = it->second.second;
        }
    }
    if (code.empty()) {
        hiprtcProgram prog;
        if (hiprtcCreateProgram(&prog, src.c_str(), "stg_period_spec.hip", 3, hdrs, names) != HIPRTC_SUCCESS)
            throw StgError{STG_ERR_DEVICE, "hiprtcCreateProgram failed"};
        hiprtcAddNameExpression(prog, name.c_str());
        const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
        size_t ls = 0;
        if (hiprtcGetProgramLogSize(prog, &ls) == HIPRTC_SUCCESS && ls > 1) {
            log.resize(ls);
            hiprtcGetProgramLog(prog, &log[0]);
        }
        if (rc == HIPRTC_SUCCESS) {
            const char *low = nullptr;
            size_t cs = 0;
            if (hiprtcGetLoweredName(prog, name.c_str(), &low) == HIPRTC_SUCCESS && low) lowered = low;
            if (hiprtcGetCodeSize(prog, &cs) == HIPRTC_SUCCESS && cs > 0) {
                code.resize(cs);
                hiprtcGetCode(prog, code.data());
            }
        }
        hiprtcDestroyProgram(&prog);
        if (rc != HIPRTC_SUCCESS || code.empty() || lowered.empty())
            throw StgError{STG_ERR_DEVICE, "run-time compilation of the specialised kernel failed: " +
                                               std::string(hiprtcGetErrorString(rc)) + "\n" + log.substr(0, 2000)};
        std::lock_guard<std::mutex> lk(g_rtc_mutex);
        g_rtc_cache[key] = {code, lowered};
    }
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    HIP_OK(hipModuleLoadData(&mod, code.data()));
    if (hipModuleGetFunction(&fn, mod, lowered.c_str()) != hipSuccess) {
        (void)hipModuleUnload(mod);
        throw StgError{STG_ERR_DEVICE, "specialised kernel not found in its module"};
    }
    s->rtc_mods[memb][min_waves][li] = mod;
    s->rtc_fns[memb][min_waves][li] = fn;
    s->rtc_fn = fn;
    s->rtc_memb = memb != 0;
    s->rtc_waves = min_waves;
    s->free_graph();
}

// back to the generic kernel; compiled modules stay loaded for a later switch
void unspecialize(stg_sim *s) {
    if (s->rtc_fn) s->free_graph();
    s->rtc_fn = nullptr;
}

// LDS bytes of the period kernel that launch_all would launch: dynamic rows
// plus the block-reduction array (static)
size_t launch_smem(const stg_sim *s, bool *spec_out) {
    const int B = s->block, BT = tets_per_block(s);
    // per-element tables and SSA groups are read by the generic membrane kernel only
    const bool spec = s->rtc_fn != nullptr && B <= 256 && (s->rtc_memb || !needs_memb(s)) && !s->has_tabs() &&
                      !s->has_groups();
    if (spec_out) *spec_out = spec;
    return spec ? s->smem_bytes_lean(BT, s->rtc_memb) : s->smem_bytes(BT);
}

// The periods of a run: launches 0 .. L-1 of the period kernel (launch l
// steps every realization with more than l periods and enters the molecules
// of the previous launch).
void launch_periods(stg_sim *s, int L) {
    const int B = s->block;
    const PeriodKernel kern = period_kernel(s, B);
    // a volume-only module cannot serve a model that has come to need the
    // membrane variant (e.g. channel integrals requested by an EField)
    bool spec = false;
    const size_t smem = launch_smem(s, &spec);
    for (int l = 0; l < L; ++l) {
        const int nb = s->launch_blocks[l];
        if (nb == 0) continue;
        stg::SysK k = make_sys(s, l & 1);
        k.n_blocks = nb;
        if (spec) {
            int launch = l;
            void *args[] = {(void *)&s->mk, (void *)&k, (void *)&launch};
            HIP_OK(hipModuleLaunchKernel(s->rtc_fn, nb, 1, 1, B, 1, 1, (unsigned)smem, s->stream, args, nullptr));
        } else {
            hipLaunchKernelGGL(kern, dim3(nb), dim3(B), smem, s->stream, s->mk, k, l);
        }
    }
    HIP_OK(hipGetLastError());
}

// Launch L: the molecules that left in the run's last period enter their
// destinations (the realizations whose last period was launch L - 1).
void launch_final_enter(stg_sim *s, int L) {
    const int nb = s->launch_blocks[L];
    if (nb == 0) return;
    stg::SysK k = make_sys(s, L & 1);
    k.n_blocks = nb;
    // one thread per tet of the block
    const int BT = tets_per_block(s);
    const bool memb = needs_memb(s);
    if (s->slot_bytes == 1) {
        if (memb) hipLaunchKernelGGL((stg::k_enter<true, int8_t>), dim3(nb), dim3(BT), 0, s->stream, s->mk, k);
        else hipLaunchKernelGGL((stg::k_enter<false, int8_t>), dim3(nb), dim3(BT), 0, s->stream, s->mk, k);
    } else {
        if (memb) hipLaunchKernelGGL((stg::k_enter<true, int16_t>), dim3(nb), dim3(BT), 0, s->stream, s->mk, k);
        else hipLaunchKernelGGL((stg::k_enter<false, int16_t>), dim3(nb), dim3(BT), 0, s->stream, s->mk, k);
    }
    HIP_OK(hipGetLastError());
}

// A new launch shape (threads per block, lanes per tet): the blocks are laid
// out again for the new tets per block, and the per-block extents are folded
// into realization totals kept in each realization's first block.
void relayout(stg_sim *s, int block, int lanes) {
    if (block == s->block && lanes == s->lanes && !s->block_real.empty()) return;
    std::vector<unsigned long long> keep;
    const std::vector<int32_t> old_real = s->block_real;
    if (!old_real.empty()) {
        keep.resize(2 * old_real.size());
        HIP_OK(hipMemcpy(keep.data(), s->d_ext, sizeof(unsigned long long) * keep.size(), hipMemcpyDeviceToHost));
    }
    // a specialised volume kernel is compiled per lane count: switch to this
    // count's variant, or to the generic kernel until it is compiled
    if (s->rtc_fn && !s->rtc_memb && lanes != s->lanes)
        s->rtc_fn = s->rtc_fns[0][s->rtc_waves][variant_index(lanes, s->slot_bytes)];
    s->block = block;
    s->lanes = lanes;
    build_blocks(s, tets_per_block(s));
    if (keep.empty()) return;
    std::vector<unsigned long long> ext(2 * s->block_real.size(), 0);
    std::vector<unsigned long long> tot(2 * s->n_real, 0);
    for (size_t b = 0; b < old_real.size(); ++b) {
        tot[2 * old_real[b]] += keep[2 * b];
        tot[2 * old_real[b] + 1] += keep[2 * b + 1];
    }
    std::vector<bool> done(s->n_real, false);
    for (size_t b = 0; b < s->block_real.size(); ++b) {
        const int r = s->block_real[b];
        if (done[r]) continue;
        done[r] = true;
        ext[2 * b] = tot[2 * r];
        ext[2 * b + 1] = tot[2 * r + 1];
    }
    HIP_OK(hipMemcpy(s->d_ext, ext.data(), sizeof(unsigned long long) * ext.size(), hipMemcpyHostToDevice));
}

void run_impl(stg_sim *s, double endtime) {
    if (!(endtime >= s->time) && !almost_equal(s->time, endtime))
        throw StgError{STG_ERR_ARG, "Endtime is before current simulation time"};
    if (almost_equal(s->time, endtime)) {
        s->time = endtime;
        s->last_launches = 0;
        return;
    }
    // the tile schedule serves volume-only models: a model that has come to
    // need the membrane variant (an EField asking for channel integrals) goes
    // back to one lane per tet
    if (s->lanes > 1 && needs_memb(s)) relayout(s, s->block, 1);
    if (s->block_real.empty()) build_blocks(s, tets_per_block(s));
    // per-realization schedules
    std::vector<std::vector<double>> dts(s->n_real);
    int L = 0;
    for (int r = 0; r < s->n_real; ++r) {
        schedule(s->time, endtime, s->dt_std[r], s->mode, dts[r]);
        if (dts[r].size() > (size_t)INT32_MAX / 2) throw StgError{STG_ERR_ARG, "too many periods"};
        L = std::max<int>(L, (int)dts[r].size());
    }
    if (L > 0) {
        std::vector<double> tab((size_t)L * s->n_real, 0.0);
        std::vector<int32_t> nper(s->n_real);
        std::vector<uint32_t> step0(s->n_real);
        for (int r = 0; r < s->n_real; ++r) {
            nper[r] = (int32_t)dts[r].size();
            step0[r] = (uint32_t)s->rng_steps[r];
            for (size_t i = 0; i < dts[r].size(); ++i) tab[i * s->n_real + r] = dts[r][i];
        }
        reserve_run_tab(s, tab.size());
        // the three tables go through pinned staging in the device layout, so
        // one asynchronous copy moves them; the run ends with a stream
        // synchronisation, so the staging is free again for the next run
        const size_t b_tab = sizeof(double) * tab.size(), b_int = sizeof(int32_t) * (size_t)s->n_real;
        const size_t need = b_tab + 2 * b_int;
        if (need > s->h_stage_cap) {
            if (s->h_stage) HIP_OK(hipHostFree(s->h_stage));
            s->h_stage = nullptr;
            s->h_stage_cap = 0;
            HIP_OK(hipHostMalloc((void **)&s->h_stage, 2 * need, hipHostMallocDefault));
            s->h_stage_cap = 2 * need;
        }
        std::memcpy(s->h_stage, nper.data(), b_int);
        std::memcpy(s->h_stage + b_int, step0.data(), b_int);
        std::memcpy(s->h_stage + 2 * b_int, tab.data(), b_tab);
        HIP_OK(hipMemcpyAsync(s->d_run_tab, s->h_stage, need, hipMemcpyHostToDevice, s->stream));
        order_blocks(s, nper, L);
        // the LDS footprint can grow after stg_set_launch checked it (an
        // EField asking for channel integrals switches to the membrane variant)
        const size_t smem = launch_smem(s, nullptr);
        if (smem + sizeof(unsigned long long) * 32 > 160 * 1024)
            throw StgError{STG_ERR_ARG, "the model's LDS footprint exceeds 160 KiB at " + std::to_string(s->block) +
                                            " threads per block; choose a smaller block with stg_set_launch"};
        if (smem > 64 * 1024)
            HIP_OK(hipFuncSetAttribute((const void *)period_kernel(s, s->block),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        HIP_OK(hipEventRecord(s->ev0, s->stream));
        if (s->use_graph) {
            if (s->graph_launches != L || s->graph_block != s->block || s->graph_counts != s->launch_blocks) {
                s->free_graph();
                HIP_OK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
                launch_periods(s, L);
                HIP_OK(hipStreamEndCapture(s->stream, &s->graph));
                HIP_OK(hipGraphInstantiate(&s->graph_exec, s->graph, nullptr, nullptr, 0));
                s->graph_launches = L;
                s->graph_block = s->block;
                s->graph_counts = s->launch_blocks;
            }
            HIP_OK(hipGraphLaunch(s->graph_exec, s->stream));
        } else {
            launch_periods(s, L);
        }
        HIP_OK(hipEventRecord(s->ev1, s->stream));
        launch_final_enter(s, L);
        // the device error flags come back with the run's synchronisation
        if (!s->h_err) HIP_OK(hipHostMalloc((void **)&s->h_err, sizeof(int32_t), hipHostMallocDefault));
        HIP_OK(hipMemcpyAsync(s->h_err, s->d_err, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        s->last_ms = ms;
        s->last_launches = L;
        // the launches have run: the clock, the period counts and the Philox
        // counters move on even when a device error is reported below, so a
        // later run never repeats this run's draws
        for (int r = 0; r < s->n_real; ++r) {
            s->nsteps[r] += dts[r].size();
            s->rng_steps[r] += dts[r].size();
        }
        s->time = endtime;
        const int32_t err = *s->h_err;
        if (err) {
            HIP_OK(hipMemset(s->d_err, 0, sizeof(int32_t)));
            if (err & 1) throw StgError{STG_ERR_PROG, "SSA exceeded 4194304 events in one tet and period"};
            if (err & 2) throw StgError{STG_ERR_PROG, "triangle potential outside a voltage-dependent rate table"};
            throw StgError{STG_ERR_PROG, "GHK flux overflow: |z V F / (R T)| too large"};
        }
    }
    s->time = endtime;
}

}  // namespace

extern "C" {

int stg_create(const stg_model *model, const stg_mesh *mesh, uint64_t seed, int32_t mode,
               stg_sim **out) {
    stg_realization r{};
    r.seed = seed;
    r.stream = 0;
    return stg_create_batch(model, 1, mesh, &r, mode, out);
}

int stg_create_batch(const stg_model *model, int32_t n_real, const stg_mesh *meshes,
                     const stg_realization *reals, int32_t mode, stg_sim **out) {
    if (!out) {
        g_create_error = "null output handle";
        return STG_ERR_ARG;
    }
    *out = nullptr;
    stg_sim *s = new (std::nothrow) stg_sim();
    if (!s) {
        g_create_error = "host allocation failed";
        return STG_ERR_SYS;
    }
    const int rc = guard(nullptr, [&] { create_impl(s, model, n_real, meshes, reals, mode); });
    if (rc != STG_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return STG_OK;
}

int stg_create_comps(const stg_model *model, int32_t n_comps, const stg_mesh *comp_meshes,
                     const stg_realization *comp_params, int32_t mode, stg_sim **out) {
    if (!out) {
        g_create_error = "null output handle";
        return STG_ERR_ARG;
    }
    *out = nullptr;
    stg_sim *s = new (std::nothrow) stg_sim();
    if (!s) {
        g_create_error = "host allocation failed";
        return STG_ERR_SYS;
    }
    const int rc = guard(nullptr, [&] { create_impl(s, model, n_comps, comp_meshes, comp_params, mode, true); });
    if (rc != STG_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return STG_OK;
}

int stg_set_boundary_dcst(stg_sim *s, int32_t bdir, int32_t spec, double dcst) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        if (bdir < 0 || bdir >= s->NBD) throw StgError{STG_ERR_ARG, "boundary direction out of range"};
        if (spec < 0 || spec >= s->S) throw StgError{STG_ERR_ARG, "species index out of range"};
        if (!(dcst >= 0.0) || !std::isfinite(dcst))
            throw StgError{STG_ERR_ARG, "diffusion constant must be finite and >= 0"};
        const bool diffuses = std::find(s->mk.dspec, s->mk.dspec + s->mk.n_diff, spec) != s->mk.dspec + s->mk.n_diff;
        if (dcst > 0.0 && !diffuses)
            throw StgError{STG_ERR_NOTIMPL, "species has no diffusion rule in any compartment"};
        s->h_bnd_D[(size_t)bdir * s->S + spec] = dcst;
        s->refresh_dt();
        upload_real_tables(s);
        HIP_OK(hipStreamSynchronize(s->stream));
    });
}

int stg_get_boundary_dcst(const stg_sim *s, int32_t bdir, int32_t spec, double *dcst) {
    if (!s || !dcst || bdir < 0 || bdir >= s->NBD || spec < 0 || spec >= s->S) return STG_ERR_ARG;
    *dcst = s->h_bnd_D[(size_t)bdir * s->S + spec];
    return STG_OK;
}

int stg_destroy(stg_sim *sim) {
    delete sim;
    return STG_OK;
}

const char *stg_last_error(const stg_sim *sim) {
    return sim ? sim->err.c_str() : g_create_error.c_str();
}

int stg_reset(stg_sim *s) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        HIP_OK(hipMemsetAsync(s->d_counts, 0, sizeof(int32_t) * (size_t)std::max(s->S, 1) * s->NT, s->stream));
        HIP_OK(hipMemsetAsync(s->d_tri_counts, 0,
                              sizeof(int32_t) * (size_t)std::max(s->SS, 1) * std::max(s->NTRI, 1), s->stream));
        if (s->d_ext)
            HIP_OK(hipMemsetAsync(s->d_ext, 0, sizeof(unsigned long long) * 2 * s->block_real.size(), s->stream));
        if (s->d_tri_intg)  // Tri::reset clears the ohmic-current integrals (tri.cpp:441)
            HIP_OK(hipMemsetAsync(s->d_tri_intg, 0, sizeof(double) * (size_t)s->SS * s->NTRI, s->stream));
        if (s->d_tri_charge)  // ... and the GHK charges (tri.cpp:435)
            HIP_OK(hipMemsetAsync(s->d_tri_charge, 0, sizeof(int32_t) * (size_t)s->NG * s->NTRI, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
        s->time = 0.0;
        // the Philox period counters (rng_steps) keep running
        std::fill(s->nsteps.begin(), s->nsteps.end(), 0);
    });
}

int stg_run(stg_sim *s, double endtime) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] { run_impl(s, endtime); });
}

int stg_advance(stg_sim *s, double adv) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        if (adv < 0.0) throw StgError{STG_ERR_ARG, "Time to advance cannot be negative"};
        run_impl(s, s->time + adv);
    });
}

int stg_set_seed(stg_sim *s, int32_t real, uint64_t seed) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        s->seeds[real] = seed;
        s->rng_steps[real] = 0;  // a re-keyed generator starts its streams afresh
        HIP_OK(hipMemcpy(s->d_real_seed + real, &seed, sizeof(uint64_t), hipMemcpyHostToDevice));
    });
}

int stg_get_time(const stg_sim *s, double *t) {
    if (!s || !t) return STG_ERR_ARG;
    *t = s->time;
    return STG_OK;
}

int stg_set_time(stg_sim *s, double t) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        if (!(t >= 0.0)) throw StgError{STG_ERR_ARG, "time must be >= 0"};
        s->time = t;
    });
}

int stg_get_nsteps(const stg_sim *s, int32_t real, uint64_t *n) {
    if (!s || !n) return STG_ERR_ARG;
    return guard(const_cast<stg_sim *>(s), [&] {
        check_real(s, real);
        *n = s->nsteps[real];
    });
}

int stg_set_nsteps(stg_sim *s, int32_t real, uint64_t n) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        s->nsteps[real] = n;
    });
}

int stg_get_update_period(const stg_sim *s, int32_t real, double *dt) {
    if (!s || !dt) return STG_ERR_ARG;
    return guard(const_cast<stg_sim *>(s), [&] {
        check_real(s, real);
        *dt = s->dt_std[real];
    });
}

int stg_get_extents(const stg_sim *cs, int32_t real, uint64_t *reac, uint64_t *diff) {
    stg_sim *s = const_cast<stg_sim *>(cs);
    if (!s || !reac || !diff) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        *reac = 0;
        *diff = 0;
        if (s->block_real.empty()) return;
        std::vector<unsigned long long> ext(2 * s->block_real.size());
        HIP_OK(hipMemcpy(ext.data(), s->d_ext, sizeof(unsigned long long) * ext.size(), hipMemcpyDeviceToHost));
        for (size_t b = 0; b < s->block_real.size(); ++b) {
            if (s->block_real[b] != real) continue;
            *reac += ext[2 * b];
            *diff += ext[2 * b + 1];
        }
    });
}

int stg_set_tet_counts(stg_sim *s, int32_t real, int32_t spec, const int32_t *counts) {
    if (!s || !counts) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        if (spec < 0 || spec >= s->S) throw StgError{STG_ERR_ARG, "species index out of range"};
        const int n = s->tet_off[real + 1] - s->tet_off[real];
        for (int i = 0; i < n; ++i)
            if (counts[i] < 0) throw StgError{STG_ERR_ARG, "molecule counts must be >= 0"};
        HIP_OK(hipMemcpy(s->d_counts + (size_t)spec * s->NT + s->tet_off[real], counts, sizeof(int32_t) * n,
                         hipMemcpyHostToDevice));
    });
}

int stg_get_tet_counts(stg_sim *s, int32_t real, int32_t spec, int32_t *counts) {
    if (!s || !counts) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        if (spec < 0 || spec >= s->S) throw StgError{STG_ERR_ARG, "species index out of range"};
        const int n = s->tet_off[real + 1] - s->tet_off[real];
        HIP_OK(hipMemcpy(counts, s->d_counts + (size_t)spec * s->NT + s->tet_off[real], sizeof(int32_t) * n,
                         hipMemcpyDeviceToHost));
    });
}

int stg_set_tri_counts(stg_sim *s, int32_t real, int32_t sspec, const int32_t *counts) {
    if (!s || !counts) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        if (sspec < 0 || sspec >= s->SS) throw StgError{STG_ERR_ARG, "surface species index out of range"};
        const int n = s->tri_off[real + 1] - s->tri_off[real];
        std::vector<int32_t> internal(std::max(n, 1));
        for (int i = 0; i < n; ++i) {
            if (counts[i] < 0) throw StgError{STG_ERR_ARG, "molecule counts must be >= 0"};
            internal[s->tri_perm[real][i]] = counts[i];
        }
        if (n > 0)
            HIP_OK(hipMemcpy(s->d_tri_counts + (size_t)sspec * s->NTRI + s->tri_off[real], internal.data(),
                             sizeof(int32_t) * n, hipMemcpyHostToDevice));
    });
}

int stg_get_tri_counts(stg_sim *s, int32_t real, int32_t sspec, int32_t *counts) {
    if (!s || !counts) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        if (sspec < 0 || sspec >= s->SS) throw StgError{STG_ERR_ARG, "surface species index out of range"};
        const int n = s->tri_off[real + 1] - s->tri_off[real];
        if (n == 0) return;
        std::vector<int32_t> internal(n);
        HIP_OK(hipMemcpy(internal.data(), s->d_tri_counts + (size_t)sspec * s->NTRI + s->tri_off[real],
                         sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) counts[i] = internal[s->tri_perm[real][i]];
    });
}

static void check_indices(const stg_sim *s, int32_t real, const int64_t *tets, size_t n) {
    const int64_t nt = s->tet_off[real + 1] - s->tet_off[real];
    for (size_t i = 0; i < n; ++i)
        if (tets[i] < 0 || tets[i] >= nt) throw StgError{STG_ERR_ARG, "tetrahedron index out of range"};
}

// device copy of a tet index list, checked against the realization's size
static void stage_indices(stg_sim *s, int32_t real, const int64_t *tets, size_t n) {
    check_indices(s, real, tets, n);
    if (n > s->idx_cap) {
        if (s->d_idx) (void)hipFree(s->d_idx);
        if (s->d_gather) (void)hipFree(s->d_gather);
        s->d_idx = nullptr;
        s->d_gather = nullptr;
        s->idx_cap = n;
        s->d_idx = dalloc<int64_t>(n);
        s->d_gather = dalloc<double>(n);
    }
    if (n) HIP_OK(hipMemcpyAsync(s->d_idx, tets, sizeof(int64_t) * n, hipMemcpyHostToDevice, s->stream));
}

int stg_get_batch_tet_counts(stg_sim *s, int32_t real, int32_t spec, const int64_t *tets, size_t n,
                             double *counts) {
    if (!s || (n > 0 && (!tets || !counts))) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        if (spec < 0 || spec >= s->S) throw StgError{STG_ERR_ARG, "species index out of range"};
        stage_indices(s, real, tets, n);
        if (n == 0) return;
        const int B = 256;
        const int nb = (int)((n + B - 1) / B);
        hipLaunchKernelGGL(stg::k_gather_counts, dim3(nb), dim3(B), 0, s->stream, s->d_counts + (size_t)spec * s->NT,
                           s->d_idx, n, s->tet_off[real], s->d_gather);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(counts, s->d_gather, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
    });
}

}  // extern "C"

// Staging of the allocation-free batch getter: up to STG_NOALLOC_MAX_TETS tet
// indices (int64) followed by as many gathered counts (double), 1 MiB. Each
// GPU the library runs on holds its own copy until the process ends. Host
// code reaches it only through the address hipGetSymbolAddress returns for
// the device (scratch_address), never through the array's name.
static __device__ __align__(256) unsigned char scratch[STG_NOALLOC_MAX_TETS * (sizeof(int64_t) + sizeof(double))];

namespace {
std::mutex g_scratch_mutex;
std::map<int, unsigned char *> g_scratch_addr;  // device -> its copy of scratch

// the current device's copy, looked up once per device with that device current
unsigned char *scratch_address() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    auto it = g_scratch_addr.find(dev);
    if (it != g_scratch_addr.end()) return it->second;
    unsigned char *p = nullptr;
    HIP_OK(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
    if (!p) throw StgError{STG_ERR_DEVICE, "no device address for the staging array"};
    g_scratch_addr[dev] = p;
    return p;
}
}  // namespace

extern "C" {

int stg_get_batch_tet_counts_noalloc(stg_sim *s, int32_t real, int32_t spec, const int64_t *tets, size_t n,
                                     double *counts) {
    if (!s || (n > 0 && (!tets || !counts))) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        if (spec < 0 || spec >= s->S) throw StgError{STG_ERR_ARG, "species index out of range"};
        if (n > STG_NOALLOC_MAX_TETS)
            throw StgError{STG_ERR_ARG, "at most " + std::to_string(STG_NOALLOC_MAX_TETS) +
                                            " tets per call of the allocation-free batch getter"};
        check_indices(s, real, tets, n);
        if (n == 0) return;
        unsigned char *base = scratch_address();
        int64_t *d_idx = reinterpret_cast<int64_t *>(base);
        double *d_out = reinterpret_cast<double *>(base + STG_NOALLOC_MAX_TETS * sizeof(int64_t));
        HIP_OK(hipMemcpyAsync(d_idx, tets, sizeof(int64_t) * n, hipMemcpyHostToDevice, s->stream));
        const int B = 256;
        const int nb = (int)((n + B - 1) / B);
        hipLaunchKernelGGL(stg::k_gather_counts, dim3(nb), dim3(B), 0, s->stream, s->d_counts + (size_t)spec * s->NT,
                           d_idx, n, s->tet_off[real], d_out);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(counts, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
    });
}

static void sum_range(stg_sim *s, const int32_t *ptr, int32_t n, int64_t *total) {
    if (!s->h_sum) HIP_OK(hipHostMalloc((void **)&s->h_sum, sizeof(unsigned long long), hipHostMallocDefault));
    HIP_OK(hipMemsetAsync(s->d_sum, 0, sizeof(unsigned long long), s->stream));
    if (n > 0) {
        const int B = 256;
        int nb = (n + B - 1) / B;
        nb = std::min(nb, 2048);
        hipLaunchKernelGGL(stg::k_sum_counts, dim3(nb), dim3(B), 0, s->stream, ptr, n, s->d_sum);
        HIP_OK(hipGetLastError());
    }
    HIP_OK(hipMemcpyAsync(s->h_sum, s->d_sum, sizeof(unsigned long long), hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    *total = (int64_t)*s->h_sum;
}

int stg_sum_batch_tet_counts(stg_sim *s, int32_t real, int32_t spec, const int64_t *tets, size_t n,
                             int64_t *total) {
    if (!s || !total || (n > 0 && !tets)) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        if (spec < 0 || spec >= s->S) throw StgError{STG_ERR_ARG, "species index out of range"};
        stage_indices(s, real, tets, n);
        if (!s->h_sum) HIP_OK(hipHostMalloc((void **)&s->h_sum, sizeof(unsigned long long), hipHostMallocDefault));
        HIP_OK(hipMemsetAsync(s->d_sum, 0, sizeof(unsigned long long), s->stream));
        if (n > 0) {
            const int B = 256;
            const int nb = (int)std::min<size_t>((n + B - 1) / B, 2048);
            hipLaunchKernelGGL(stg::k_sum_indexed, dim3(nb), dim3(B), 0, s->stream,
                               s->d_counts + (size_t)spec * s->NT, s->d_idx, (int64_t)n, s->tet_off[real],
                               s->d_sum);
            HIP_OK(hipGetLastError());
        }
        HIP_OK(hipMemcpyAsync(s->h_sum, s->d_sum, sizeof(unsigned long long), hipMemcpyDeviceToHost, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
        *total = (int64_t)*s->h_sum;
    });
}

int stg_sum_tet_counts(stg_sim *s, int32_t real, int32_t spec, int64_t *total) {
    if (!s || !total) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        if (spec < 0 || spec >= s->S) throw StgError{STG_ERR_ARG, "species index out of range"};
        sum_range(s, s->d_counts + (size_t)spec * s->NT + s->tet_off[real], s->tet_off[real + 1] - s->tet_off[real],
                  total);
    });
}

// Totals of every (realization, species) in one reduction over a
// species-major [S][n] count array with realization offsets off (device).
static void sum_all(stg_sim *s, const int32_t *counts, int32_t n, int32_t S, const int32_t *off, int64_t *totals) {
    const size_t nout = (size_t)s->n_real * S;
    if (nout == 0) return;
    const size_t cap = (size_t)s->n_real * std::max(std::max(s->S, s->SS), 1);
    if (!s->d_real_sums) s->d_real_sums = dalloc<long long>(cap);
    if (!s->h_real_sums) HIP_OK(hipHostMalloc((void **)&s->h_real_sums, sizeof(long long) * cap, hipHostMallocDefault));
    hipLaunchKernelGGL(stg::k_sum_real_counts, dim3((unsigned)nout), dim3(256), 0, s->stream, counts, n, S, off,
                       s->d_real_sums);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(s->h_real_sums, s->d_real_sums, sizeof(long long) * nout, hipMemcpyDeviceToHost,
                          s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    std::memcpy(totals, s->h_real_sums, sizeof(long long) * nout);
}

int stg_sum_all_tet_counts(stg_sim *s, int64_t *totals) {
    if (!s || !totals) return STG_ERR_ARG;
    return guard(s, [&] { sum_all(s, s->d_counts, s->NT, s->S, s->d_real_tet0, totals); });
}

int stg_sum_all_tri_counts(stg_sim *s, int64_t *totals) {
    if (!s || !totals) return STG_ERR_ARG;
    return guard(s, [&] {
        if (!s->d_tri_off) {
            s->d_tri_off = dalloc<int32_t>(s->tri_off.size());
            HIP_OK(hipMemcpy(s->d_tri_off, s->tri_off.data(), sizeof(int32_t) * s->tri_off.size(),
                             hipMemcpyHostToDevice));
        }
        sum_all(s, s->d_tri_counts, s->NTRI, s->SS, s->d_tri_off, totals);
    });
}

int stg_sum_tri_counts(stg_sim *s, int32_t real, int32_t sspec, int64_t *total) {
    if (!s || !total) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        if (sspec < 0 || sspec >= s->SS) throw StgError{STG_ERR_ARG, "surface species index out of range"};
        sum_range(s, s->d_tri_counts + (size_t)sspec * s->NTRI + s->tri_off[real],
                  s->tri_off[real + 1] - s->tri_off[real], total);
    });
}

int stg_set_reac_k(stg_sim *s, int32_t real, int32_t reac, double k) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        if (reac < 0 || reac >= s->R) throw StgError{STG_ERR_ARG, "reaction index out of range"};
        if (!(k >= 0.0) || !std::isfinite(k)) throw StgError{STG_ERR_ARG, "rate constant must be finite and >= 0"};
        s->h_k[(size_t)real * s->R + reac] = k;
        upload_real_tables(s);
    });
}

int stg_set_sreac_k(stg_sim *s, int32_t real, int32_t sreac, double k) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        if (sreac < 0 || sreac >= s->RS) throw StgError{STG_ERR_ARG, "surface reaction index out of range"};
        if (!(k >= 0.0) || !std::isfinite(k)) throw StgError{STG_ERR_ARG, "rate constant must be finite and >= 0"};
        s->h_sk[(size_t)real * s->RS + sreac] = k;
        upload_real_tables(s);
    });
}

int stg_set_dcst(stg_sim *s, int32_t real, int32_t spec, double dcst) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        if (spec < 0 || spec >= s->S) throw StgError{STG_ERR_ARG, "species index out of range"};
        if (!(dcst >= 0.0) || !std::isfinite(dcst))
            throw StgError{STG_ERR_ARG, "diffusion constant must be finite and >= 0"};
        const bool was_diff = (std::find(s->mk.dspec, s->mk.dspec + s->mk.n_diff, spec) != s->mk.dspec + s->mk.n_diff);
        if (dcst > 0.0 && !was_diff)
            throw StgError{STG_ERR_NOTIMPL, "enabling diffusion of a species that was immobile at creation"};
        s->h_D[(size_t)real * s->S + spec] = dcst;
        s->refresh_dt();
        upload_real_tables(s);
    });
}

int stg_set_sdcst(stg_sim *s, int32_t real, int32_t sspec, double dcst) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        if (sspec < 0 || sspec >= s->SS) throw StgError{STG_ERR_ARG, "surface species index out of range"};
        if (!(dcst >= 0.0) || !std::isfinite(dcst))
            throw StgError{STG_ERR_ARG, "diffusion constant must be finite and >= 0"};
        const bool was = std::find(s->mk.sdspec, s->mk.sdspec + s->mk.n_sdiff, sspec) != s->mk.sdspec + s->mk.n_sdiff;
        if (dcst > 0.0 && !was)
            throw StgError{STG_ERR_NOTIMPL, "enabling surface diffusion of a species that was immobile at creation"};
        s->h_sD[(size_t)real * s->SS + sspec] = dcst;
        s->refresh_dt();
        upload_real_tables(s);
        HIP_OK(hipStreamSynchronize(s->stream));
    });
}

// ------------------------------------------------------ per-element tables
}  // extern "C"

namespace {

// Copies realization real's rows of a table into its host copy (layout
// [rows][n_total][width], the realization's elements at offset off..off+n)
// and uploads the whole table; src == nullptr drops the realization's use.
template <typename T>
void put_table(stg_sim *s, int32_t real, int32_t bit, const T *src, int rows, int32_t n_total, int32_t off,
               int32_t n, int width, std::vector<T> &host, T *&dev, const std::vector<int32_t> *perm = nullptr) {
    if (src) {
        const size_t sz = (size_t)std::max(rows, 1) * std::max(n_total, 1) * width;
        if (host.size() != sz) host.assign(sz, T{});
        for (int q = 0; q < rows; ++q)
            for (int32_t i = 0; i < n; ++i) {
                const int32_t j = perm ? (*perm)[i] : i;  // user order -> internal order
                for (int w = 0; w < width; ++w)
                    host[((size_t)q * n_total + off + j) * width + w] = src[((size_t)q * n + i) * width + w];
            }
        if (!dev) dev = dalloc<T>(sz);
        HIP_OK(hipMemcpy(dev, host.data(), sz * sizeof(T), hipMemcpyHostToDevice));
        s->real_tabs[real] |= bit;
    } else {
        s->real_tabs[real] &= ~bit;
    }
    if (!s->d_real_tabs) s->d_real_tabs = dalloc<int32_t>(s->n_real);
    HIP_OK(hipMemcpy(s->d_real_tabs, s->real_tabs.data(), sizeof(int32_t) * s->n_real, hipMemcpyHostToDevice));
    s->free_graph();  // the kernel variant and its arguments change
}

}  // namespace

extern "C" {

int stg_set_tet_reac_table(stg_sim *s, int32_t real, const double *k) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        const int32_t n = s->tet_off[real + 1] - s->tet_off[real];
        if (k)
            for (size_t i = 0; i < (size_t)s->R * n; ++i)
                if (!(k[i] >= 0.0) || !std::isfinite(k[i]))
                    throw StgError{STG_ERR_ARG, "rate constants must be finite and >= 0"};
        if (k && s->R == 0) return;
        put_table(s, real, stg::kTabReac, k, s->R, s->NT, s->tet_off[real], n, 1, s->h_tet_k, s->d_tet_k);
    });
}

int stg_set_tet_diff_table(stg_sim *s, int32_t real, const double *d) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        const int32_t n = s->tet_off[real + 1] - s->tet_off[real];
        if (d) {
            for (int sp = 0; sp < s->S; ++sp) {
                const bool diffuses = std::find(s->mk.dspec, s->mk.dspec + s->mk.n_diff, sp) != s->mk.dspec + s->mk.n_diff;
                for (size_t i = 0; i < (size_t)n * 4; ++i) {
                    const double v = d[(size_t)sp * n * 4 + i];
                    if (!(v >= 0.0) || !std::isfinite(v))
                        throw StgError{STG_ERR_ARG, "diffusion constants must be finite and >= 0"};
                    if (v > 0.0 && !diffuses)
                        throw StgError{STG_ERR_NOTIMPL, "enabling diffusion of a species that was immobile at creation"};
                }
            }
        }
        put_table(s, real, stg::kTabDiff, d, s->S, s->NT, s->tet_off[real], n, 4, s->h_tet_fD, s->d_tet_fD);
        s->refresh_dt();
    });
}

int stg_set_tet_clamp(stg_sim *s, int32_t real, const uint32_t *mask) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        const int32_t n = s->tet_off[real + 1] - s->tet_off[real];
        put_table(s, real, stg::kTabClamp, mask, 1, s->NT, s->tet_off[real], n, 1, s->h_tet_clamp, s->d_tet_clamp);
    });
}

int stg_set_tri_sreac_table(stg_sim *s, int32_t real, const double *k) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        const int32_t n = s->tri_off[real + 1] - s->tri_off[real];
        const int rows = s->RS + s->NG;
        if (k)
            for (size_t i = 0; i < (size_t)rows * n; ++i)
                if (!(k[i] >= 0.0) || !std::isfinite(k[i]))
                    throw StgError{STG_ERR_ARG, "surface constants and factors must be finite and >= 0"};
        if (k && (rows == 0 || n == 0)) return;
        put_table(s, real, stg::kTabTriReac, k, rows, s->NTRI, s->tri_off[real], n, 1, s->h_tri_sk, s->d_tri_sk,
                  &s->tri_perm[real]);
    });
}

int stg_set_tri_sdiff_table(stg_sim *s, int32_t real, const double *d) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        const int32_t n = s->tri_off[real + 1] - s->tri_off[real];
        if (d) {
            if (s->NSD == 0) throw StgError{STG_ERR_NOTIMPL, "no surface species diffuses in this model"};
            for (int q = 0; q < s->SS; ++q) {
                const bool diffuses = std::find(s->mk.sdspec, s->mk.sdspec + s->mk.n_sdiff, q) != s->mk.sdspec + s->mk.n_sdiff;
                for (size_t i = 0; i < (size_t)n * 3; ++i) {
                    const double v = d[(size_t)q * n * 3 + i];
                    if (!(v >= 0.0) || !std::isfinite(v))
                        throw StgError{STG_ERR_ARG, "diffusion constants must be finite and >= 0"};
                    if (v > 0.0 && !diffuses)
                        throw StgError{STG_ERR_NOTIMPL, "enabling surface diffusion of a species that was immobile at creation"};
                }
            }
        }
        if (d && n == 0) return;
        put_table(s, real, stg::kTabTriDiff, d, s->SS, s->NTRI, s->tri_off[real], n, 3, s->h_tri_eD, s->d_tri_eD,
                  &s->tri_perm[real]);
        s->refresh_dt();
    });
}

int stg_set_tri_clamp(stg_sim *s, int32_t real, const uint32_t *mask) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        const int32_t n = s->tri_off[real + 1] - s->tri_off[real];
        if (mask && n == 0) return;
        put_table(s, real, stg::kTabTriClamp, mask, 1, s->NTRI, s->tri_off[real], n, 1, s->h_tri_clamp,
                  s->d_tri_clamp, &s->tri_perm[real]);
    });
}

int stg_set_specialize(stg_sim *s, int32_t on, int32_t min_waves) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        if (on) specialize(s, min_waves);
        else unspecialize(s);
    });
}

int stg_get_specialize(const stg_sim *s, int32_t *on) {
    if (!s || !on) return STG_ERR_ARG;
    *on = s->rtc_fn != nullptr ? 1 : 0;
    return STG_OK;
}

int stg_set_launch(stg_sim *s, int32_t block, int32_t use_graph) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        if (block < 64 || block > 1024 || block % 64 != 0)
            throw StgError{STG_ERR_ARG, "block must be a multiple of 64 in [64, 1024]"};
        if (s->smem_bytes(block / s->lanes) > 160 * 1024)
            throw StgError{STG_ERR_ARG, "block too large for the model's LDS footprint"};
        relayout(s, block, s->lanes);
        s->use_graph = use_graph != 0;
        s->free_graph();
    });
}

int stg_set_lanes(stg_sim *s, int32_t lanes) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8)
            throw StgError{STG_ERR_ARG, "lanes per tet must be 1, 2, 4 or 8"};
        if (lanes > 1 && needs_memb(s))
            throw StgError{STG_ERR_NOTIMPL, "several lanes per tet serve volume-only models"};
        const bool spec_on = s->rtc_fn != nullptr;
        relayout(s, s->block, lanes);
        // specialisation stays on: compile this lane count's variant if needed
        if (spec_on && !s->rtc_fn) specialize(s, s->rtc_waves);
    });
}

int stg_get_lanes(const stg_sim *s, int32_t *lanes) {
    if (!s || !lanes) return STG_ERR_ARG;
    *lanes = s->lanes;
    return STG_OK;
}

int stg_set_slot_width(stg_sim *s, int32_t bytes) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        if (bytes != 1 && bytes != 2) throw StgError{STG_ERR_ARG, "face slot width must be 1 or 2 bytes"};
        if (bytes == s->slot_bytes) return;
        // slots live only within a run (its trailing entering pass empties
        // them), so the width can change between runs
        const bool spec_on = s->rtc_fn != nullptr;
        s->slot_bytes = bytes;
        s->rtc_fn = nullptr;
        s->free_graph();
        if (spec_on) specialize(s, s->rtc_waves);
    });
}

int stg_get_slot_width(const stg_sim *s, int32_t *bytes) {
    if (!s || !bytes) return STG_ERR_ARG;
    *bytes = s->slot_bytes;
    return STG_OK;
}

// ------------------------------------------------------------ checkpoint
// Layout (all little-endian, host byte order): header, per-realization time
// counters and constants, then the count arrays as stored on the device.
namespace {
// "TGSTATS3": counts, constants, triangle V, Philox period counters, channel
// integrals and GHK charges
constexpr uint64_t kStateMagic = 0x33534154534754ull;
struct StateHeader {
    uint64_t magic;
    int32_t n_real, n_tets, n_tris, n_specs, n_sspecs, n_reacs, n_sreacs, diff_threshold;
    double time;
    int32_t intg_mask, has_charge;
};
constexpr size_t kWordsPerReal = 5;  // nsteps, rng_steps, seed, reaction extent, diffusion extent

size_t state_bytes(const stg_sim *s) {
    return sizeof(StateHeader) + s->n_real * (sizeof(uint64_t) * kWordsPerReal) +
           sizeof(double) * (s->h_k.size() + s->h_sk.size() + s->h_D.size() + s->h_bnd_D.size() +
                             s->h_sD.size()) +
           sizeof(int32_t) * ((size_t)s->S * s->NT + (size_t)s->SS * s->NTRI) +
           sizeof(double) * (size_t)s->NTRI +
           (s->d_tri_intg ? sizeof(double) * (size_t)s->SS * s->NTRI : 0) +
           (s->d_tri_charge ? sizeof(int32_t) * (size_t)s->NG * s->NTRI : 0);
}
}  // namespace

int stg_state_size(const stg_sim *s, size_t *bytes) {
    if (!s || !bytes) return STG_ERR_ARG;
    *bytes = state_bytes(s);
    return STG_OK;
}

int stg_checkpoint(stg_sim *s, void *buf, size_t bytes) {
    if (!s || !buf) return STG_ERR_ARG;
    return guard(s, [&] {
        if (bytes < state_bytes(s)) throw StgError{STG_ERR_ARG, "checkpoint buffer too small"};
        char *p = static_cast<char *>(buf);
        StateHeader h{kStateMagic, s->n_real, s->NT, s->NTRI, s->S, s->SS, s->R, s->RS, s->diff_threshold,
                      s->time, s->d_tri_intg ? s->mk.tri_intg_mask : 0, s->d_tri_charge ? 1 : 0};
        std::memcpy(p, &h, sizeof(h));
        p += sizeof(h);
        std::vector<unsigned long long> ext(2 * s->block_real.size(), 0);
        if (!ext.empty())
            HIP_OK(hipMemcpy(ext.data(), s->d_ext, sizeof(unsigned long long) * ext.size(), hipMemcpyDeviceToHost));
        for (int r = 0; r < s->n_real; ++r) {
            uint64_t ev = 0, mv = 0;
            for (size_t b = 0; b < s->block_real.size(); ++b)
                if (s->block_real[b] == r) {
                    ev += ext[2 * b];
                    mv += ext[2 * b + 1];
                }
            const uint64_t v[kWordsPerReal] = {s->nsteps[r], s->rng_steps[r], s->seeds[r], ev, mv};
            std::memcpy(p, v, sizeof(v));
            p += sizeof(v);
        }
        for (const auto *vec : {&s->h_k, &s->h_sk, &s->h_D, &s->h_bnd_D, &s->h_sD}) {
            std::memcpy(p, vec->data(), vec->size() * sizeof(double));
            p += vec->size() * sizeof(double);
        }
        const size_t nc = (size_t)s->S * s->NT, nt = (size_t)s->SS * s->NTRI;
        if (nc) HIP_OK(hipMemcpy(p, s->d_counts, nc * sizeof(int32_t), hipMemcpyDeviceToHost));
        p += nc * sizeof(int32_t);
        if (nt) HIP_OK(hipMemcpy(p, s->d_tri_counts, nt * sizeof(int32_t), hipMemcpyDeviceToHost));
        p += nt * sizeof(int32_t);
        if (s->NTRI) HIP_OK(hipMemcpy(p, s->d_tri_v, (size_t)s->NTRI * sizeof(double), hipMemcpyDeviceToHost));
        p += (size_t)s->NTRI * sizeof(double);
        if (s->d_tri_intg) {
            HIP_OK(hipMemcpy(p, s->d_tri_intg, (size_t)s->SS * s->NTRI * sizeof(double), hipMemcpyDeviceToHost));
            p += (size_t)s->SS * s->NTRI * sizeof(double);
        }
        if (s->d_tri_charge)
            HIP_OK(hipMemcpy(p, s->d_tri_charge, (size_t)s->NG * s->NTRI * sizeof(int32_t), hipMemcpyDeviceToHost));
    });
}

int stg_restore(stg_sim *s, const void *buf, size_t bytes) {
    if (!s || !buf) return STG_ERR_ARG;
    return guard(s, [&] {
        if (bytes < sizeof(StateHeader)) throw StgError{STG_ERR_ARG, "checkpoint too short"};
        const char *p = static_cast<const char *>(buf);
        StateHeader h;
        std::memcpy(&h, p, sizeof(h));
        if (h.magic != kStateMagic) throw StgError{STG_ERR_ARG, "not an stg checkpoint"};
        if (h.n_real != s->n_real || h.n_tets != s->NT || h.n_tris != s->NTRI || h.n_specs != s->S ||
            h.n_sspecs != s->SS || h.n_reacs != s->R || h.n_sreacs != s->RS)
            throw StgError{STG_ERR_ARG, "checkpoint was written by a simulation of another shape"};
        p += sizeof(h);
        std::vector<unsigned long long> ext(2 * s->block_real.size(), 0);
        std::vector<bool> placed(s->n_real, false);
        const bool want_intg = h.intg_mask != 0, want_charge = h.has_charge != 0;
        if (want_charge && !s->d_tri_charge) throw StgError{STG_ERR_ARG, "checkpoint holds GHK charges this model lacks"};
        if (want_intg && (s->SS == 0 || s->NTRI == 0))
            throw StgError{STG_ERR_ARG, "checkpoint holds channel integrals this model lacks"};
        const size_t need = state_bytes(s) + (want_intg && !s->d_tri_intg ? sizeof(double) * (size_t)s->SS * s->NTRI : 0) -
                            (!want_intg && s->d_tri_intg ? sizeof(double) * (size_t)s->SS * s->NTRI : 0);
        if (bytes < need) throw StgError{STG_ERR_ARG, "checkpoint too short"};
        for (int r = 0; r < s->n_real; ++r) {
            uint64_t v[kWordsPerReal];
            std::memcpy(v, p, sizeof(v));
            p += sizeof(v);
            s->nsteps[r] = v[0];
            s->rng_steps[r] = v[1];
            s->seeds[r] = v[2];
            for (size_t b = 0; b < s->block_real.size(); ++b)
                if (s->block_real[b] == r && !placed[r]) {
                    ext[2 * b] = v[3];
                    ext[2 * b + 1] = v[4];
                    placed[r] = true;
                }
        }
        for (auto *vec : {&s->h_k, &s->h_sk, &s->h_D, &s->h_bnd_D, &s->h_sD}) {
            std::memcpy(vec->data(), p, vec->size() * sizeof(double));
            p += vec->size() * sizeof(double);
        }
        const size_t nc = (size_t)s->S * s->NT, nt = (size_t)s->SS * s->NTRI;
        if (nc) HIP_OK(hipMemcpy(s->d_counts, p, nc * sizeof(int32_t), hipMemcpyHostToDevice));
        p += nc * sizeof(int32_t);
        if (nt) HIP_OK(hipMemcpy(s->d_tri_counts, p, nt * sizeof(int32_t), hipMemcpyHostToDevice));
        p += nt * sizeof(int32_t);
        if (s->NTRI) HIP_OK(hipMemcpy(s->d_tri_v, p, (size_t)s->NTRI * sizeof(double), hipMemcpyHostToDevice));
        p += (size_t)s->NTRI * sizeof(double);
        if (want_intg) {
            const size_t n = (size_t)s->SS * s->NTRI;
            if (!s->d_tri_intg) s->d_tri_intg = dalloc<double>(n);
            HIP_OK(hipMemcpy(s->d_tri_intg, p, n * sizeof(double), hipMemcpyHostToDevice));
            p += n * sizeof(double);
            if ((s->mk.tri_intg_mask | h.intg_mask) != s->mk.tri_intg_mask) s->mk.tri_intg_mask |= h.intg_mask;
        } else if (s->d_tri_intg) {
            HIP_OK(hipMemset(s->d_tri_intg, 0, (size_t)s->SS * s->NTRI * sizeof(double)));
        }
        if (want_charge)
            HIP_OK(hipMemcpy(s->d_tri_charge, p, (size_t)s->NG * s->NTRI * sizeof(int32_t), hipMemcpyHostToDevice));
        if (!ext.empty())
            HIP_OK(hipMemcpy(s->d_ext, ext.data(), sizeof(unsigned long long) * ext.size(), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->d_real_seed, s->seeds.data(), sizeof(uint64_t) * s->n_real, hipMemcpyHostToDevice));
        s->time = h.time;
        s->diff_threshold = h.diff_threshold;
        s->refresh_dt();
        upload_real_tables(s);
        s->free_graph();
    });
}

int stg_set_diff_apply_threshold(stg_sim *s, int32_t threshold) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        if (threshold < 0) throw StgError{STG_ERR_ARG, "threshold must be >= 0"};
        s->diff_threshold = threshold;
        s->free_graph();
    });
}

int stg_set_tri_v(stg_sim *s, int32_t real, const double *v) {
    if (!s || !v) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        const int n = s->tri_off[real + 1] - s->tri_off[real];
        std::vector<double> internal(std::max(n, 1));
        for (int i = 0; i < n; ++i) internal[s->tri_perm[real][i]] = v[i];
        if (n > 0)
            HIP_OK(hipMemcpy(s->d_tri_v + s->tri_off[real], internal.data(), sizeof(double) * n,
                             hipMemcpyHostToDevice));
    });
}

int stg_get_tri_v(stg_sim *s, int32_t real, double *v) {
    if (!s || !v) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        const int n = s->tri_off[real + 1] - s->tri_off[real];
        if (n == 0) return;
        std::vector<double> internal(n);
        HIP_OK(hipMemcpy(internal.data(), s->d_tri_v + s->tri_off[real], sizeof(double) * n,
                         hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) v[i] = internal[s->tri_perm[real][i]];
    });
}

// Device pointer of the triangle potentials of one realization (internal
// triangle order), for the EField (not part of the public header).
int stg_internal_tri_v_device(stg_sim *s, int32_t real, double **ptr) {
    if (!s || !ptr) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        *ptr = s->d_tri_v + s->tri_off[real];
    });
}

int stg_set_temp(stg_sim *s, double kelvin) {
    if (!s) return STG_ERR_ARG;
    return guard(s, [&] {
        if (!(kelvin >= 0.0)) throw StgError{STG_ERR_ARG, "temperature must be >= 0 K"};
        if (s->mk.temp != kelvin) {
            s->mk.temp = kelvin;
            s->free_graph();
        }
    });
}

int stg_get_temp(const stg_sim *s, double *kelvin) {
    if (!s || !kelvin) return STG_ERR_ARG;
    *kelvin = s->mk.temp;
    return STG_OK;
}

// Device pointer of the GHK charge of each current on each triangle of one
// realization (internal order; current g at ptr[g * stride + i]), null when
// the model has no GHK currents; for the EField (not part of the public header).
int stg_internal_tri_charge_device(stg_sim *s, int32_t real, int32_t **ptr, int32_t *n_ghk, int32_t *stride) {
    if (!s || !ptr || !n_ghk || !stride) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        *ptr = s->d_tri_charge ? s->d_tri_charge + s->tri_off[real] : nullptr;
        *n_ghk = s->d_tri_charge ? s->NG : 0;
        *stride = s->NTRI;
    });
}

// Start accumulating the time integral of the counts of the surface species
// in mask (bit q = species q) on every triangle, and return the [SS][NTRI]
// integral array (internal triangle order, all realizations). For the
// EField's time-averaged channel counts (not part of the public header).
int stg_internal_tri_intg_device(stg_sim *s, uint32_t mask, double **ptr) {
    if (!s || !ptr) return STG_ERR_ARG;
    return guard(s, [&] {
        if (s->SS == 0 || s->NTRI == 0) throw StgError{STG_ERR_ARG, "the simulation has no membrane"};
        if (mask == 0 || (s->SS < 32 && (mask >> s->SS) != 0))
            throw StgError{STG_ERR_ARG, "surface species mask out of range"};
        if (!s->d_tri_intg) {
            const size_t n = (size_t)s->SS * s->NTRI;
            s->d_tri_intg = dalloc<double>(n);
            HIP_OK(hipMemsetAsync(s->d_tri_intg, 0, n * sizeof(double), s->stream));
            HIP_OK(hipStreamSynchronize(s->stream));
        }
        if ((s->mk.tri_intg_mask | (int32_t)mask) != s->mk.tri_intg_mask) {
            s->mk.tri_intg_mask |= (int32_t)mask;
            s->free_graph();  // the model tables are baked into a captured graph
        }
        *ptr = s->d_tri_intg;
    });
}

// Device pointer of the triangle counts and the realization's triangle
// offset, for the EField's membrane currents (not part of the public header).
int stg_internal_tri_counts_device(stg_sim *s, int32_t real, const int32_t **ptr,
                                   int32_t *n_tris_total, int32_t *tri0) {
    if (!s || !ptr || !n_tris_total || !tri0) return STG_ERR_ARG;
    return guard(s, [&] {
        check_real(s, real);
        *ptr = s->d_tri_counts;
        *n_tris_total = s->NTRI;
        *tri0 = s->tri_off[real];
    });
}

// User triangle order -> internal (inner-tet sorted) slot of one realization.
int stg_internal_tri_perm(const stg_sim *s, int32_t real, const int32_t **perm, int32_t *n) {
    if (!s || !perm || !n) return STG_ERR_ARG;
    return guard(const_cast<stg_sim *>(s), [&] {
        check_real(s, real);
        *perm = s->tri_perm[real].data();
        *n = (int32_t)s->tri_perm[real].size();
    });
}

int stg_get_sizes(const stg_sim *s, int32_t real, int32_t *n_real, int32_t *n_tets, int32_t *n_tris) {
    if (!s) return STG_ERR_ARG;
    return guard(const_cast<stg_sim *>(s), [&] {
        check_real(s, real);
        if (n_real) *n_real = s->n_real;
        if (n_tets) *n_tets = s->tet_off[real + 1] - s->tet_off[real];
        if (n_tris) *n_tris = s->tri_off[real + 1] - s->tri_off[real];
    });
}

// mean over the full-period launches of the last run (the trailing
// entering-only k_enter is timed outside them)
int stg_last_run_kernel_ms(const stg_sim *s, double *ms_per_launch, int32_t *n) {
    if (!s || !ms_per_launch || !n) return STG_ERR_ARG;
    *n = s->last_launches;
    *ms_per_launch = s->last_launches > 0 ? s->last_ms / s->last_launches : 0.0;
    return STG_OK;
}

}  // extern "C"
