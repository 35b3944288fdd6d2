// Device kernels of the operator-split period (gfx950).
//
// One launch = one split period for every realization that still has periods
// left in the current run, fused as
//   entering(previous period)  ->  SSA over [0, dt)  ->  leaving(this period)
// One thread owns one tetrahedron (and the membrane triangles whose inner tet
// it is). Counts are read once and written once per period; the time-averaged
// occupancy used by the leaving pass lives in LDS for the whole period, so it
// never touches HBM. Leaving molecules go to a per-(diffusing species, tet,
// face) int32 slot; the owner of the neighbouring tet gathers them at the
// start of the next launch from the other half of a ping-pong buffer, so no
// atomics are needed and the result is independent of scheduling.
//
// Reference semantics:
//   SSA         src/steps/mpi/dist/tetopsplit/operator/ssa_operator.cpp:24-47,
//               kproc/propensities.hpp:543-586 (Direct group draw)
//   propensity  kproc/reactions.cpp:66-119, kproc/surface_reactions.cpp:176-240
//   occupancy   mol_state.hpp:87-200
//   diffusion   operator/diffusion_operator.cpp:41-238, kproc/diffusions.hpp:211-224
#pragma once

#ifndef __HIPCC_RTC__  // a runtime-compiled build gets these from hipRTC
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "stg_random.hpp"

namespace stg {

constexpr int kMaxSpecs = 16;
constexpr int kMaxReacs = 32;
constexpr int kMaxSSpecs = 8;
constexpr int kMaxSReacs = 16;
constexpr int kMaxEntries = 384;
constexpr int kMaxGhk = 4;
constexpr int kSurfFlag = 0x80;   // entry species with this bit: surface species (idx & 0x7f)
constexpr int kOuterFlag = 0x100; // entry species with this bit: outer-compartment species (idx & 0xff)
constexpr long kMaxEventsPerPeriod = 1L << 22;
constexpr double kAvogadro = 6.02214179e23;  // src/steps/math/constants.hpp:45
constexpr double kDblEpsilon = 2.220446049250313080847e-16;
constexpr double kFaraday = 96485.3399;        // src/steps/math/constants.hpp:49
constexpr double kGasConstant = 8.314472;      // src/steps/math/constants.hpp:53
constexpr double kECharge = 1.602176487e-19;   // src/steps/math/constants.hpp:57

// Model tables, passed by value in the kernel-argument segment. Every field is
// a 32-bit word so that wave-uniform reads are scalar loads (sub-dword fields
// would be fetched with per-lane vector loads).
struct ModelK {
    int32_t n_specs, n_reacs, n_sspecs, n_sreacs, n_diff;
    int32_t tracked_mask;  // bit s: occupancy tracked for species s
    int32_t tri_intg_mask; // bit q: time integral of surface species q is accumulated
    int32_t dspec[kMaxSpecs];
    int32_t order[kMaxReacs];
    int32_t sorder[kMaxSReacs];
    int32_t s_nvol[kMaxSReacs];  // number of volume reactants of a surface reaction
    int32_t s_vdep[kMaxSReacs];  // k(V) table of a surface reaction, -1 = constant k
    int32_t lhs_off[kMaxReacs + 1], upd_off[kMaxReacs + 1];
    int32_t slhs_off[kMaxSReacs + 1], supd_off[kMaxSReacs + 1];
    int32_t ent[kMaxEntries];  // species (low 16 bits) | signed value << 16
    // GHK currents: channel-state surface species, inner ion, valence, whether
    // ions move; permeability (m/s), virtual outer concentration (M), V shift (V)
    int32_t n_sdiff;            // surface species that diffuse
    int32_t sdspec[kMaxSSpecs];
    int32_t tri_tracked_mask;   // bit q: occupancy tracked for surface species q
    int32_t n_ghk;
    int32_t g_chan[kMaxGhk], g_ion[kMaxGhk], g_z[kMaxGhk], g_real[kMaxGhk];
    double g_perm[kMaxGhk], g_oconc[kMaxGhk], g_vshift[kMaxGhk];
    double temp;  // K
    // outer side of the membrane: s_out[q] = 1 when surface reaction q's
    // volume reactants are outer species (its constant scales with the outer
    // tet's volume, sreac.cpp:150-180); g_outer[g] = 1 when GHK current g
    // reads the outer tet's concentration (ghkcurr.cpp:176-182)
    int32_t s_out[kMaxSReacs];
    int32_t g_outer[kMaxGhk];
};

// Where the period kernel reads the volume reaction network from. DynTables: the
// ModelK tables at run time (the library's own build). A model-specialised
// build (compiled at run time for one model) supplies a struct with
// kStatic = true and the same tables as compile-time constants:
//   S, R, order[R], lhs_off[R+1], lhs_sp[], lhs_mult[], upd_off[R+1],
//   upd_sp[], upd_dv[], tracked_mask, n_diff, dspec[n_diff]
// so the loops over reactions, reactants and species unroll and the
// constants fold; every operation and its order stay the same, so both
// builds produce identical results.
struct DynTables {
    static constexpr bool kStatic = false;
};

template <class MT>
__device__ __forceinline__ int tab_S(const ModelK &m) {
    if constexpr (MT::kStatic) return MT::S;
    else return m.n_specs;
}
template <class MT>
__device__ __forceinline__ int tab_R(const ModelK &m) {
    if constexpr (MT::kStatic) return MT::R;
    else return m.n_reacs;
}
template <class MT>
__device__ __forceinline__ int tab_ndiff(const ModelK &m) {
    if constexpr (MT::kStatic) return MT::n_diff;
    else return m.n_diff;
}
template <class MT>
__device__ __forceinline__ int tab_dspec(const ModelK &m, int d) {
    if constexpr (MT::kStatic) return MT::dspec[d];
    else return m.dspec[d];
}
template <class MT>
__device__ __forceinline__ int tab_tracked(const ModelK &m) {
    if constexpr (MT::kStatic) return MT::tracked_mask;
    else return m.tracked_mask;
}

// Mesh, state and per-realization tables (device pointers).
struct SysK {
    int32_t n_tets;        // all realizations
    int32_t n_tris;
    int32_t n_real;
    int32_t xcd_remap;     // nonzero: contiguous block ranges per XCD
    int32_t n_blocks;          // blocks in this launch
    int32_t diff_threshold;      // leaving counts up to this are split per molecule
    int32_t k_max;               // most reactions of one tet group (LDS rows per thread)
    const double *vol;           // [NT]
    const int4 *nbr;             // [NT] gather index of each face: the neighbour's global tet
                                 // x 4 + this face's slot in it, -1 = no neighbour
    const double2 *g;            // [NT*2] face factors (f0,f1),(f2,f3)
    int32_t *counts;             // [S][NT]
    const void *out_prev;        // [ND][NT][4] molecules leaving through each face, 8- or 16-bit
    void *out_cur;               // (the kernel's SlotT); a count above the slot's maximum is
                                 // its escape value here and lives in wide_*
    const int32_t *wide_prev;    // [ND][NT][4] escaped face counts (touched only on escape)
    int32_t *wide_cur;
    const int32_t *tri_ptr;      // [NT+1] or null when there are no triangles
    int32_t *tri_counts;         // [SS][NTRI]
    const double *tri_area;      // [NTRI]
    const int4 *block_desc;      // [launched blocks] {block id (extents slot), realization,
                                 // first tet, end tet of the realization}
    const int32_t *real_tet0;    // [R+1]
    const double *real_k;        // [R][n_reacs]
    const double *real_sk;       // [R][n_sreacs]
    const double *real_D;        // [R][n_specs]
    const uint64_t *real_seed;   // [R]
    const uint32_t *real_stream; // [R]
    const uint32_t *real_step0;  // [R] period counter at launch 0
    const int32_t *real_nper;    // [R] periods in this run
    const double *dt_tab;        // [L][R]
    unsigned long long *ext;     // [blocks][2] events, molecules moved, by block id
    int32_t *err;                // error flags: 1 event cap, 2 potential outside a k(V) table,
                                 // 4 GHK exponent out of range
    const double *tri_v;         // [NTRI] membrane potential of each triangle (V)
    const double *vdep_vmin;     // [n_vdep]
    const double *vdep_dv;       // [n_vdep]
    const int32_t *vdep_off;     // [n_vdep+1]
    const double *vdep_k;        // concatenated k(V) tables
    double *tri_intg;            // [SS][NTRI] time integral of counts (tri_intg_mask species)
    int32_t *tri_charge;         // [n_ghk][NTRI] GHK charge moved out per current (elementary charges), or null
    const uint32_t *tet_bdir;    // [NT] 4 x 8-bit diffusion-boundary direction per face
                                 // (0xff = none), or null
    const double *bnd_D;         // [directions][n_specs] diffusion constant across a boundary
    // surface diffusion
    const int4 *tri_nbr;         // [NTRI] neighbour triangle per edge (global), -1 none
    const uint32_t *tri_nface;   // [NTRI] 3 x 8-bit edge slot in the neighbour
    const double *tri_g;         // [NTRI*3] l_e / (A d_e)
    const int32_t *tri_user;     // [NTRI] triangle index in its realization's input order
    const double *real_sD;       // [R][n_sspecs]
    const int32_t *tout_prev;    // [n_sdiff][NTRI][4]
    int32_t *tout_cur;           // [n_sdiff][NTRI][4]
    double *tri_corr;            // [n_sspecs][NTRI] occupancy correction within the period
    // Per-element tables, null when no realization has one; bit kTab* of
    // real_tabs[r] says which ones realization r uses (membrane variant only)
    const int32_t *real_tabs;    // [R]
    const double *tet_k;         // [n_reacs][NT] rate constant of each reaction in each tet
    const double *tet_fD;        // [n_specs][NT][4] diffusion constant through each face
    const uint32_t *tet_clamp;   // [NT] bit s: species s clamped in the tet
    const double *tri_sk;        // [n_sreacs + n_ghk][NTRI] surface constant / k(V) and GHK factor
    const double *tri_eD;        // [n_sspecs][NTRI][3] surface diffusion constant through each edge
    const uint32_t *tri_clamp;   // [NTRI] bit q: surface species q clamped on the triangle
    // SSA groups (membrane models coupling both sides): tets joined by
    // triangles whose reactions or GHK currents reach both sides; null when
    // there are none. tet_grp[t] = group of t or -1; members of group g are
    // grp_mem[grp_ptr[g] .. grp_ptr[g+1]) (global tets, first = owner) with
    // realizations grp_real[]; gcorr holds their occupancy corrections
    const int32_t *tet_grp;      // [NT]
    const int32_t *grp_ptr;      // [groups + 1]
    const int32_t *grp_mem;      // [members]
    const int32_t *grp_real;     // [members]
    double *gcorr;               // [n_specs][NT]
    const int32_t *tri_otet;     // [NTRI] outer tet (global) of each triangle, -1 none
    const int32_t *tri_oreal;    // [NTRI] its realization
};

// bits of SysK::real_tabs
constexpr int32_t kTabReac = 1, kTabDiff = 2, kTabClamp = 4, kTabTriReac = 8, kTabTriDiff = 16,
                  kTabTriClamp = 32;

// both structures travel in the 4 KiB kernel-argument segment
static_assert(sizeof(ModelK) + sizeof(SysK) + 16 <= 4096, "kernel arguments exceed 4 KiB");

// Face slots are 8- or 16-bit (SlotT = int8_t or int16_t, chosen per engine):
// a period's molecules through one face are mostly a few, and narrow slots cut
// the only per-face traffic (written by the owner, gathered by the neighbour).
// A count above slot_max<SlotT> is stored as slot_escape<SlotT> with the count
// itself in a 32-bit side table at the same index, which is read only when the
// escape is seen. 8-bit slots escape above 127 molecules per face and period,
// 16-bit ones above 32,767: the narrower the slots, the fewer bytes while
// escapes are rare.
template <class SlotT>
constexpr int32_t slot_max = (1 << (8 * sizeof(SlotT) - 1)) - 1;
template <class SlotT>
constexpr int32_t slot_escape = -(1 << (8 * sizeof(SlotT) - 1));

// the four slots of one (tet, diffusing species), one store
template <class SlotT>
__device__ __forceinline__ void put_slots(void *slots, int32_t *wide, size_t base, int4 o) {
    constexpr int bits = 8 * (int)sizeof(SlotT);
    constexpr uint32_t mask = (1u << bits) - 1u;
    const int32_t v[4] = {o.x, o.y, o.z, o.w};
    uint32_t h[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        if (v[f] > slot_max<SlotT>) {
            wide[base + f] = v[f];
            h[f] = (uint32_t)slot_escape<SlotT> & mask;
        } else {
            h[f] = (uint32_t)v[f] & mask;
        }
    }
    SlotT *p = static_cast<SlotT *>(slots) + base;
    if constexpr (sizeof(SlotT) == 2)
        *reinterpret_cast<uint2 *>(p) = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
    else
        *reinterpret_cast<uint32_t *>(p) = h[0] | (h[1] << 8) | (h[2] << 16) | (h[3] << 24);
}

// c = k scale^-(order-1) (comp_ccst, src/steps/mpi/tetopsplit/reac.cpp:44-55),
// as a product with powers of scale or of inv = 1 / scale: no division per call
__device__ __forceinline__ double scaled_const(double k, int order, double scale, double inv) {
    const int o1 = order - 1;
    if (o1 == 0) return k;
    if (o1 < 0) return k * ipow(scale, -o1);
    return k * ipow(inv, o1);
}

// h *= n (n-1) ... (n-L+1), multiplied from the smallest factor up
__device__ __forceinline__ double falling(double h, int32_t n, int32_t L) {
    for (int32_t i = L - 1; i >= 0; --i) h = h * (double)(n - i);
    return h;
}

// Per-thread LDS arrays are laid out [index][blockDim] so that a wavefront
// touching the same index hits 64 consecutive banks.
struct Lds {
    int32_t *cnt;   // [S][B]
    double *corr;   // [S][B]
    double *a;      // [K][B] propensities
    double *kc;     // [K][B] scaled rate constants of this tet's reactions
    double *gx;     // [2 * triangles * GHK currents][B] GHK flux constants
    int B, tid;
    __device__ __forceinline__ int32_t &c(int s) const { return cnt[s * B + tid]; }
    __device__ __forceinline__ double &o(int s) const { return corr[s * B + tid]; }
    __device__ __forceinline__ double &p(int j) const { return a[j * B + tid]; }
    __device__ __forceinline__ double &c(int j, int) const { return kc[j * B + tid]; }
    __device__ __forceinline__ double &g(int j, int w) const { return gx[(2 * j + w) * B + tid]; }
};

// GHK flux of one open channel (math/ghk.cpp:62-81), split for a fixed
// potential into P z F nuFoRT, c_o e^-nuFoRT and 1 - e^-nuFoRT so that the
// flux is (A (c_i - Bo)) / Dn, the reference's left-to-right evaluation.
// Returns false when nuFoRT is outside the range the reference accepts.
__device__ __forceinline__ bool ghk_consts(const ModelK &m, int g, double v, double &A, double &Bo,
                                           double &Dn) {
    const double z = (double)m.g_z[g];
    const double nuFoRT = z * (v + m.g_vshift[g]) * kFaraday / (kGasConstant * m.temp);
    if (nuFoRT >= 308 * 2.30 || nuFoRT <= -307 * 2.30) return false;
    const double e = dexp(-nuFoRT);
    const double oconc = m.g_oconc[g] * 1.0e3;  // mol/m^3
    if (fabs(nuFoRT) > kDblEpsilon) {
        A = m.g_perm[g] * z * kFaraday * nuFoRT;
        Bo = oconc * e;
        Dn = 1.0 - e;
    } else {  // the limit nuFoRT -> 0
        A = m.g_perm[g] * z * kFaraday;
        Bo = oconc;
        Dn = 1.0;
    }
    return true;
}

// ions per second through one channel, signed: positive is an efflux
// (ghkcurr.cpp:168-208); n_ion / vscale is the inner concentration in M
__device__ __forceinline__ double ghk_rt(const ModelK &m, int g, double A, double Bo, double Dn,
                                         int32_t n_ion, double vscale) {
    const double iconc = (double)n_ion / vscale * 1.0e3;
    const double flux = A * (iconc - Bo) / Dn;
    return flux / (kECharge * (double)m.g_z[g]);
}

// k(V) by linear interpolation in a table (VDepSReacdef::getVDepK,
// src/steps/solver/vdepsreacdef.cpp:194-217); sets error bit 2 outside it.
__device__ __forceinline__ double vdep_k(const SysK &sys, int tab, double v) {
    const double vmin = sys.vdep_vmin[tab], dv = sys.vdep_dv[tab];
    const int n = sys.vdep_off[tab + 1] - sys.vdep_off[tab];
    const double *K = sys.vdep_k + sys.vdep_off[tab];
    const double v2 = (v - vmin) / dv;
    if (!(v2 >= 0.0) || !(v2 <= (double)(n - 1))) {
        atomicOr(sys.err, 2);
        return 0.0;
    }
    const double lv = floor(v2);
    const int l = (int)lv, u = (int)ceil(v2);
    const double r = v2 - lv;
    return (1.0 - r) * K[l] + r * K[u];
}

// combinatorial factor h of a surface reaction on triangle tri (0 if a
// reactant is short)
__device__ __forceinline__ double surf_h(const ModelK &m, const SysK &sys, const Lds &L, int r, int tri) {
    double h = 1.0;
    for (int e = m.slhs_off[r]; e < m.slhs_off[r + 1]; ++e) {
        const int sp = m.ent[e] & 0xffff;
        const int mult = m.ent[e] >> 16;
        int32_t n;
        if (sp & kOuterFlag) return 0.0;  // a triangle outside every SSA group has no outer tet
        if (sp & kSurfFlag) n = sys.tri_counts[(size_t)(sp & 0x7f) * sys.n_tris + tri];
        else n = L.c(sp);
        if (mult > n) return 0.0;
        h = falling(h, n, mult);
    }
    return h;
}

// Direct-method SSA of one tet group over [0, dt). Returns events, or -1 when
// the per-period event cap is hit.
// MEMB = false compiles a variant without triangles (no surface reactions,
// GHK currents, k(V) tables or channel integrals) for volume-only models.
template <bool MEMB, class MT>
__device__ __forceinline__ long ssa_group(const ModelK &m, const SysK &sys, const Lds &L, int r, int t,
                          uint32_t ltet, uint32_t period, double dt, uint64_t seed, uint32_t stream,
                          int rtab = 0, uint32_t clampm = 0u) {
    const int R = tab_R<MT>(m), RS = m.n_sreacs, NG = m.n_ghk;
    const int RT = RS + NG;  // entries per triangle: surface reactions, then GHK currents
    int tp0 = 0, ntri = 0;
    if (MEMB && sys.tri_ptr != nullptr && RT > 0) {
        tp0 = sys.tri_ptr[t];
        ntri = sys.tri_ptr[t + 1] - tp0;
    }
    const int K = R + ntri * RT;
    if (K == 0) return 0;
    const double vscale = 1.0e3 * sys.vol[t] * kAvogadro;
    const double ivscale = 1.0 / vscale;
    const double *kr = sys.real_k + (size_t)r * R;
    const double *ksr = sys.real_sk + (size_t)r * RS;
    // rate constants scaled by volume / area, formed once per period
    {
        int j = 0;
        if constexpr (MT::kStatic) {
#pragma unroll
            for (int q = 0; q < MT::R; ++q) L.c(q, 0) = scaled_const(kr[q], MT::order[q], vscale, ivscale);
            j = MT::R;
        } else {
            for (int q = 0; q < R; ++q, ++j) {
                // a per-tet table replaces the realization's constant
                const double kq = (MEMB && (rtab & kTabReac)) ? sys.tet_k[(size_t)q * sys.n_tets + t] : kr[q];
                L.c(j, 0) = scaled_const(kq, m.order[q], vscale, ivscale);
            }
        }
        for (int i = 0; i < ntri; ++i) {
            const int tri = tp0 + i;
            const double ascale = sys.tri_area[tri] * kAvogadro;
            const double iascale = 1.0 / ascale;
            const bool ttab = MEMB && (rtab & kTabTriReac);
            for (int q = 0; q < RS; ++q, ++j) {
                double k;
                if (ttab) {
                    // the triangle's constant, or its factor (1 active, 0 not) on k(V);
                    // an inactive k(V) reaction reads no table
                    const double f = sys.tri_sk[(size_t)q * sys.n_tris + tri];
                    k = m.s_vdep[q] < 0 ? f : (f == 0.0 ? 0.0 : vdep_k(sys, m.s_vdep[q], sys.tri_v[tri]) * f);
                } else {
                    k = m.s_vdep[q] < 0 ? ksr[q] : vdep_k(sys, m.s_vdep[q], sys.tri_v[tri]);
                }
                L.c(j, 0) = m.s_nvol[q] == 0 ? scaled_const(k, m.sorder[q], ascale, iascale)
                                             : scaled_const(k, m.sorder[q], vscale, ivscale);
            }
            for (int q = 0; q < NG; ++q, ++j) {
                if (!ghk_consts(m, q, sys.tri_v[tri], L.c(j, 0), L.g(i * NG + q, 0), L.g(i * NG + q, 1))) {
                    atomicOr(sys.err, 4);
                    L.c(j, 0) = 0.0;
                    L.g(i * NG + q, 0) = 0.0;
                    L.g(i * NG + q, 1) = 1.0;
                } else if (ttab) {
                    // factor 1 or 0 on the current's P z F nuFoRT term
                    L.c(j, 0) = L.c(j, 0) * sys.tri_sk[(size_t)(RS + q) * sys.n_tris + tri];
                }
            }
        }
    }
    Philox g;
    g.init(seed, ltet, period, stream, 0u);
    long nev = 0;
    double tcur = 0.0;
    while (true) {
        // With compile-time reaction tables the per-tet constants in LDS are
        // loop-invariant and the compiler would hold all of them in registers
        // across the event loop (137 VGPRs for four species, spilling under a
        // 4-wave bound); this compiler-only fence keeps them as LDS reads inside
        // each event (117 VGPRs, no scratch). No instruction is emitted.
        if constexpr (MT::kStatic) asm volatile("" ::: "memory");
        double a0 = 0.0;
        int j = 0;
        if constexpr (MT::kStatic) {
            // the same products in the same order; a short reactant zeroes the
            // propensity whatever the later factors are
#pragma unroll
            for (int q = 0; q < MT::R; ++q) {
                double h = 1.0;
                bool zero = false;
#pragma unroll
                for (int e = MT::lhs_off[q]; e < MT::lhs_off[q + 1]; ++e) {
                    const int32_t n = L.c(MT::lhs_sp[e]);
                    if (MT::lhs_mult[e] > n) zero = true;
                    else h = falling(h, n, MT::lhs_mult[e]);
                }
                const double av = zero ? 0.0 : h * L.c(q, 0);
                L.p(q) = av;
                a0 = a0 + av;
            }
            j = MT::R;
        } else {
            for (int q = 0; q < R; ++q, ++j) {
                double h = 1.0;
                bool zero = false;
                for (int e = m.lhs_off[q]; e < m.lhs_off[q + 1]; ++e) {
                    const int sp = m.ent[e] & 0xffff;
                    const int mult = m.ent[e] >> 16;
                    const int32_t n = L.c(sp);
                    if (mult > n) {
                        zero = true;
                        break;
                    }
                    h = falling(h, n, mult);
                }
                const double av = zero ? 0.0 : h * L.c(j, 0);
                L.p(j) = av;
                a0 = a0 + av;
            }
        }
        for (int i = 0; i < ntri; ++i) {
            for (int q = 0; q < RS; ++q, ++j) {
                const double h = surf_h(m, sys, L, q, tp0 + i);
                const double av = h == 0.0 ? 0.0 : h * L.c(j, 0);
                L.p(j) = av;
                a0 = a0 + av;
            }
            for (int q = 0; q < NG; ++q, ++j) {
                const int32_t nch = sys.tri_counts[(size_t)m.g_chan[q] * sys.n_tris + tp0 + i];
                const double rt = ghk_rt(m, q, L.c(j, 0), L.g(i * NG + q, 0), L.g(i * NG + q, 1),
                                         L.c(m.g_ion[q]), vscale);
                const double av = fabs(rt) * (double)nch;
                L.p(j) = av;
                a0 = a0 + av;
            }
        }
        if (a0 < kDblEpsilon) break;
        const double u1 = g.unif();
        // -log(u) >= 1 - u: a waiting time certainly past the period end
        // needs no logarithm
        if (1.0 - u1 > a0 * (dt - tcur) * 1.0000001) break;
        const double u2 = g.unif();
        const double tev = tcur + (-dlog(u1) / a0);
        if (tev > dt) break;
        tcur = tev;
        // first q whose running sum exceeds x, else the last nonzero propensity
        const double x = u2 * a0;
        int sel = -1, last = -1;
        double acc = 0.0;
        for (int q = 0; q < K; ++q) {
            const double pq = L.p(q);
            acc = acc + pq;
            if (sel < 0 && acc > x) sel = q;
            if (pq > 0.0) last = q;
        }
        if (sel < 0) sel = last;
        if (!MEMB || sel < R) {
            if constexpr (MT::kStatic) {
                // the selected reaction's entries, in table order
#pragma unroll
                for (int q = 0; q < MT::R; ++q) {
                    if (sel != q) continue;
#pragma unroll
                    for (int e = MT::upd_off[q]; e < MT::upd_off[q + 1]; ++e) {
                        const int sp = MT::upd_sp[e], dv = MT::upd_dv[e];
                        L.c(sp) += dv;
                        if ((MT::tracked_mask >> sp) & 1) L.o(sp) -= (double)dv * tcur;
                    }
                }
            } else {
                // clamped species keep their counts (Reac::apply, reac.cpp:240-243)
                for (int e = m.upd_off[sel]; e < m.upd_off[sel + 1]; ++e) {
                    const int sp = m.ent[e] & 0xffff;
                    const int dv = m.ent[e] >> 16;
                    if (MEMB && ((clampm >> sp) & 1u)) continue;
                    L.c(sp) += dv;
                    if ((m.tracked_mask >> sp) & 1) L.o(sp) -= (double)dv * tcur;
                }
            }
        } else if ((sel - R) % RT >= RS) {
            // GHK current: one ion crosses in the direction of the flux
            // (ghkcurr.cpp:212-272); the triangle's charge counts it outward positive
            const int i = (sel - R) / RT;
            const int q = (sel - R) - i * RT - RS;
            const int ion = m.g_ion[q];
            const double rt = ghk_rt(m, q, L.c(sel, 0), L.g(i * NG + q, 0), L.g(i * NG + q, 1), L.c(ion),
                                     vscale);
            const bool efflux = rt >= 0.0;
            if (m.g_real[q] && !((clampm >> ion) & 1u)) {
                const int dv = efflux ? -1 : 1;
                L.c(ion) += dv;
                if ((m.tracked_mask >> ion) & 1) L.o(ion) -= (double)dv * tcur;
            }
            sys.tri_charge[(size_t)q * sys.n_tris + tp0 + i] += efflux ? m.g_z[q] : -m.g_z[q];
        } else {
            const int i = (sel - R) / RT;
            const int qs = (sel - R) - i * RT;
            const int tri = tp0 + i;
            const uint32_t tclamp = (MEMB && (rtab & kTabTriClamp)) ? sys.tri_clamp[tri] : 0u;
            for (int e = m.supd_off[qs]; e < m.supd_off[qs + 1]; ++e) {
                const int sp = m.ent[e] & 0xffff;
                const int dv = m.ent[e] >> 16;
                if (sp & kOuterFlag) continue;  // no outer tet outside an SSA group
                if (sp & kSurfFlag) {
                    if ((tclamp >> (sp & 0x7f)) & 1u) continue;
                    const size_t ix = (size_t)(sp & 0x7f) * sys.n_tris + tri;
                    sys.tri_counts[ix] += dv;
                    // integral of n over the period = n_end * dt - sum dn * t_event
                    if ((m.tri_intg_mask >> (sp & 0x7f)) & 1) sys.tri_intg[ix] -= (double)dv * tcur;
                    if ((m.tri_tracked_mask >> (sp & 0x7f)) & 1) sys.tri_corr[ix] -= (double)dv * tcur;
                } else {
                    if ((clampm >> sp) & 1u) continue;
                    L.c(sp) += dv;
                    if ((m.tracked_mask >> sp) & 1) L.o(sp) -= (double)dv * tcur;
                }
            }
        }
        if (++nev >= kMaxEventsPerPeriod) return -1;
    }
    return nev;
}

// Direct-method SSA of one tet for the model-specialised volume kernel. The
// draws and the floating-point operations are those of ssa_group, in the same
// order; what differs is where the intermediate values live:
//   - a reaction's rate constant c = k scale^-(order-1) is formed when it is
//     used, from k in scalar registers and the tet's volume scale, instead of
//     being stored per tet;
//   - the selection forms the propensities again (the counts have not moved
//     since the sum was taken) instead of reading them back.
// Its only per-thread LDS is then the counts and occupancy corrections (48 B
// for four species, against 144 B with the two K-row tables), so LDS no longer
// caps the kernel at four wavefronts per SIMD.
template <class MT>
__device__ long ssa_lean(const SysK &sys, const Lds &L, int r, int t, uint32_t ltet, uint32_t period,
                         double dt, uint64_t seed, uint32_t stream) {
    constexpr int R = MT::R;
    if constexpr (R == 0) {
        return 0;
    } else {
        const double vscale = 1.0e3 * sys.vol[t] * kAvogadro;
        const double ivscale = 1.0 / vscale;
        double kq[R];
#pragma unroll
        for (int q = 0; q < R; ++q) kq[q] = uni(sys.real_k[(size_t)r * R + q]);
        Philox g;
        g.init(seed, ltet, period, stream, 0u);
        long nev = 0;
        double tcur = 0.0;
        while (true) {
            asm volatile("" ::: "memory");  // counts are read from LDS in every event
            double a0 = 0.0;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                double h = 1.0;
                bool zero = false;
#pragma unroll
                for (int e = MT::lhs_off[q]; e < MT::lhs_off[q + 1]; ++e) {
                    const int32_t n = L.c(MT::lhs_sp[e]);
                    if (MT::lhs_mult[e] > n) zero = true;
                    else h = falling(h, n, MT::lhs_mult[e]);
                }
                a0 = a0 + (zero ? 0.0 : h * scaled_const(kq[q], MT::order[q], vscale, ivscale));
            }
            if (a0 < kDblEpsilon) break;
            const double u1 = g.unif();
            if (1.0 - u1 > a0 * (dt - tcur) * 1.0000001) break;
            const double u2 = g.unif();
            const double tev = tcur + (-dlog(u1) / a0);
            if (tev > dt) break;
            tcur = tev;
            const double x = u2 * a0;
            int sel = -1, last = -1;
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                double h = 1.0;
                bool zero = false;
#pragma unroll
                for (int e = MT::lhs_off[q]; e < MT::lhs_off[q + 1]; ++e) {
                    const int32_t n = L.c(MT::lhs_sp[e]);
                    if (MT::lhs_mult[e] > n) zero = true;
                    else h = falling(h, n, MT::lhs_mult[e]);
                }
                const double pq = zero ? 0.0 : h * scaled_const(kq[q], MT::order[q], vscale, ivscale);
                acc = acc + pq;
                if (sel < 0 && acc > x) sel = q;
                if (pq > 0.0) last = q;
            }
            if (sel < 0) sel = last;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if (sel != q) continue;
#pragma unroll
                for (int e = MT::upd_off[q]; e < MT::upd_off[q + 1]; ++e) {
                    const int sp = MT::upd_sp[e], dv = MT::upd_dv[e];
                    L.c(sp) += dv;
                    if ((MT::tracked_mask >> sp) & 1) L.o(sp) -= (double)dv * tcur;
                }
            }
            if (++nev >= kMaxEventsPerPeriod) return -1;
        }
        return nev;
    }
}

// Propensity of volume reaction q of a model-specialised kernel (the value
// ssa_group forms: h times the rate constant scaled by the tet's volume)
template <class MT>
__device__ __forceinline__ double vol_prop(const Lds &L, int q, double k, double vscale, double ivscale) {
    double h = 1.0;
    bool zero = false;
#pragma unroll
    for (int e = MT::lhs_off[q]; e < MT::lhs_off[q + 1]; ++e) {
        const int32_t n = L.c(MT::lhs_sp[e]);
        if (MT::lhs_mult[e] > n) zero = true;
        else h = falling(h, n, MT::lhs_mult[e]);
    }
    return zero ? 0.0 : h * scaled_const(k, MT::order[q], vscale, ivscale);
}

// Propensity of triangle entry e = i * RT + q (surface reaction q < RS, else GHK
// current q - RS) of the lean membrane SSA; its constants are LDS row e.
__device__ __forceinline__ double tri_prop(const ModelK &m, const SysK &sys, const Lds &L, int e, int i, int q,
                                           int tp0, double vscale) {
    const int RS = m.n_sreacs, NG = m.n_ghk;
    if (q < RS) {
        const double h = surf_h(m, sys, L, q, tp0 + i);
        return h == 0.0 ? 0.0 : h * L.c(e, 0);
    }
    const int qg = q - RS;
    const int32_t nch = sys.tri_counts[(size_t)m.g_chan[qg] * sys.n_tris + tp0 + i];
    const double rt = ghk_rt(m, qg, L.c(e, 0), L.g(i * NG + qg, 0), L.g(i * NG + qg, 1), L.c(m.g_ion[qg]), vscale);
    return fabs(rt) * (double)nch;
}

// Direct-method SSA of one tet group for the model-specialised membrane
// kernel: ssa_group's draws and floating-point operations in the same order,
// with the volume reactions' rate constants formed when used (as ssa_lean) and
// every propensity formed again for the selection instead of stored. LDS holds
// only the triangles' rate constants and GHK constants (rows i * RT + q), so a
// thread needs K - R fewer constant rows and no propensity rows: 164 -> 84 B
// per thread for the calcium membrane model.
template <class MT>
__device__ long ssa_memb_lean(const ModelK &m, const SysK &sys, const Lds &L, int r, int t, uint32_t ltet,
                              uint32_t period, double dt, uint64_t seed, uint32_t stream) {
    constexpr int R = MT::R;
    const int RS = m.n_sreacs, NG = m.n_ghk;
    const int RT = RS + NG;
    int tp0 = 0, ntri = 0;
    if (sys.tri_ptr != nullptr && RT > 0) {
        tp0 = sys.tri_ptr[t];
        ntri = sys.tri_ptr[t + 1] - tp0;
    }
    if (R + ntri * RT == 0) return 0;
    const double vscale = 1.0e3 * sys.vol[t] * kAvogadro;
    const double ivscale = 1.0 / vscale;
    double kq[R > 0 ? R : 1];
#pragma unroll
    for (int q = 0; q < R; ++q) kq[q] = uni(sys.real_k[(size_t)r * R + q]);
    const double *ksr = sys.real_sk + (size_t)r * RS;
    for (int i = 0; i < ntri; ++i) {
        const double ascale = sys.tri_area[tp0 + i] * kAvogadro;
        const double iascale = 1.0 / ascale;
        for (int q = 0; q < RS; ++q) {
            const double k = m.s_vdep[q] < 0 ? ksr[q] : vdep_k(sys, m.s_vdep[q], sys.tri_v[tp0 + i]);
            L.c(i * RT + q, 0) = m.s_nvol[q] == 0 ? scaled_const(k, m.sorder[q], ascale, iascale)
                                                  : scaled_const(k, m.sorder[q], vscale, ivscale);
        }
        for (int q = 0; q < NG; ++q) {
            const int e = i * RT + RS + q;
            if (!ghk_consts(m, q, sys.tri_v[tp0 + i], L.c(e, 0), L.g(i * NG + q, 0), L.g(i * NG + q, 1))) {
                atomicOr(sys.err, 4);
                L.c(e, 0) = 0.0;
                L.g(i * NG + q, 0) = 0.0;
                L.g(i * NG + q, 1) = 1.0;
            }
        }
    }
    Philox g;
    g.init(seed, ltet, period, stream, 0u);
    long nev = 0;
    double tcur = 0.0;
    while (true) {
        asm volatile("" ::: "memory");
        double a0 = 0.0;
#pragma unroll
        for (int q = 0; q < R; ++q) a0 = a0 + vol_prop<MT>(L, q, kq[q], vscale, ivscale);
        for (int i = 0; i < ntri; ++i)
            for (int q = 0; q < RT; ++q) a0 = a0 + tri_prop(m, sys, L, i * RT + q, i, q, tp0, vscale);
        if (a0 < kDblEpsilon) break;
        const double u1 = g.unif();
        if (1.0 - u1 > a0 * (dt - tcur) * 1.0000001) break;
        const double u2 = g.unif();
        const double tev = tcur + (-dlog(u1) / a0);
        if (tev > dt) break;
        tcur = tev;
        const double x = u2 * a0;
        int sel = -1, last = -1;
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const double pq = vol_prop<MT>(L, q, kq[q], vscale, ivscale);
            acc = acc + pq;
            if (sel < 0 && acc > x) sel = q;
            if (pq > 0.0) last = q;
        }
        for (int i = 0; i < ntri; ++i) {
            for (int q = 0; q < RT; ++q) {
                const int j = R + i * RT + q;
                const double pq = tri_prop(m, sys, L, i * RT + q, i, q, tp0, vscale);
                acc = acc + pq;
                if (sel < 0 && acc > x) sel = j;
                if (pq > 0.0) last = j;
            }
        }
        if (sel < 0) sel = last;
        if (sel < R) {
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if (sel != q) continue;
#pragma unroll
                for (int e = MT::upd_off[q]; e < MT::upd_off[q + 1]; ++e) {
                    const int sp = MT::upd_sp[e], dv = MT::upd_dv[e];
                    L.c(sp) += dv;
                    if ((MT::tracked_mask >> sp) & 1) L.o(sp) -= (double)dv * tcur;
                }
            }
        } else if ((sel - R) % RT >= RS) {
            // GHK current: one ion crosses in the direction of the flux
            // (ghkcurr.cpp:212-272); the triangle's charge counts it outward positive
            const int i = (sel - R) / RT;
            const int q = (sel - R) - i * RT - RS;
            const int ion = m.g_ion[q];
            const double rt = ghk_rt(m, q, L.c(sel - R, 0), L.g(i * NG + q, 0), L.g(i * NG + q, 1), L.c(ion),
                                     vscale);
            const bool efflux = rt >= 0.0;
            if (m.g_real[q]) {
                const int dv = efflux ? -1 : 1;
                L.c(ion) += dv;
                if ((m.tracked_mask >> ion) & 1) L.o(ion) -= (double)dv * tcur;
            }
            sys.tri_charge[(size_t)q * sys.n_tris + tp0 + i] += efflux ? m.g_z[q] : -m.g_z[q];
        } else {
            const int i = (sel - R) / RT;
            const int qs = (sel - R) - i * RT;
            const int tri = tp0 + i;
            for (int e = m.supd_off[qs]; e < m.supd_off[qs + 1]; ++e) {
                const int sp = m.ent[e] & 0xffff;
                const int dv = m.ent[e] >> 16;
                if (sp & kOuterFlag) continue;  // no outer tet outside an SSA group
                if (sp & kSurfFlag) {
                    const size_t ix = (size_t)(sp & 0x7f) * sys.n_tris + tri;
                    sys.tri_counts[ix] += dv;
                    if ((m.tri_intg_mask >> (sp & 0x7f)) & 1) sys.tri_intg[ix] -= (double)dv * tcur;
                    if ((m.tri_tracked_mask >> (sp & 0x7f)) & 1) sys.tri_corr[ix] -= (double)dv * tcur;
                } else {
                    L.c(sp) += dv;
                    if ((m.tracked_mask >> sp) & 1) L.o(sp) -= (double)dv * tcur;
                }
            }
        }
        if (++nev >= kMaxEventsPerPeriod) return -1;
    }
    return nev;
}

__device__ __forceinline__ bool almost_zero_dt(double dt) {
    // util::almost_equal(dt, 0.0) of src/steps/util/collections.hpp:143-149
    const double ad = fabs(dt);
    return ad <= kDblEpsilon * ad * 4 || ad < 2.2250738585072014e-308;
}

// Leaving pass of one (tet, species) holding n molecules with occupancy
// correction corr: returns the molecules that leave (the caller takes them
// off the count unless the species is clamped there: a clamped source keeps
// its count, Diff::apply, diff.cpp:476-520, 577) and writes their split over
// the 4 faces. Faces on a diffusion boundary (bdir byte != 0xff) take the
// boundary's constant for this species instead of D (0 when it is closed); a
// per-face table fD (non-null) replaces both.
// This function, tri_leave, ssa_group and group_period are always inlined:
// the membrane kernel calls them from more than one place, and a real call
// takes the kernel-argument structures by reference, which makes every lane
// copy them into its private (scratch) memory at each call (up to 3 KiB per
// lane; the generic 6M-tet membrane period took 1.23 ms that way, 0.32 ms
// inlined).
__device__ __forceinline__ int32_t leave(const SysK &sys, int32_t n, double corr, int sp, bool tracked, double D,
                         const double r4in[4], uint32_t bdir, int S, double dt, uint64_t seed,
                         uint32_t ltet, uint32_t period, uint32_t stream, int4 &o,
                         const double *fD = nullptr) {
    o = make_int4(0, 0, 0, 0);
    double r4[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const uint32_t k = (bdir >> (8 * f)) & 0xffu;
        const double Df = fD != nullptr ? fD[f] : (k == 0xffu ? D : sys.bnd_D[(size_t)k * S + sp]);
        r4[f] = Df * r4in[f];
    }
    const double sum = ((r4[0] + r4[1]) + r4[2]) + r4[3];
    if (n <= 0 || !(sum > 0.0)) return 0;
    Philox g;
    g.init(seed, ltet, period, stream, 1u + (uint32_t)sp);
    double occ = (double)n;
    if (tracked && !almost_zero_dt(dt)) occ = (double)n + corr / dt;
    const double u = g.unif();
    const double fl = floor(occ);
    int32_t mp = (int32_t)fl + ((occ - fl) > u ? 1 : 0);
    mp = mp > n ? n : mp;
    mp = mp < 0 ? 0 : mp;
    double p = dt * sum;
    p = p > 1.0 ? 1.0 : p;
    p = p < 0.0 ? 0.0 : p;
    const int32_t nl = binomial(g, mp, p);
    if (nl <= 0) return 0;
    int lastf = -1;
#pragma unroll
    for (int f = 0; f < 4; ++f)
        if (r4[f] > 0.0) lastf = f;
    int32_t k4[4] = {0, 0, 0, 0};
    if (nl <= sys.diff_threshold) {
        for (int32_t mol = 0; mol < nl; ++mol) {
            int chosen = -1;
            for (int tries = 0; tries < 64 && chosen < 0; ++tries) {
                const double selv = g.unif() * sum;
                double cum = 0.0;
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    cum = cum + r4[f];
                    if (chosen < 0 && selv < cum) chosen = f;
                }
            }
            if (chosen < 0) chosen = lastf;
#pragma unroll
            for (int f = 0; f < 4; ++f) k4[f] += (chosen == f) ? 1 : 0;
        }
    } else {
        int32_t rem = nl;
        double rsum = sum;
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            if (!(r4[f] > 0.0)) continue;
            int32_t k;
            if (f == lastf) {
                k = rem;
            } else {
                const double pf = r4[f] / rsum;
                k = pf >= 1.0 ? rem : binomial(g, rem, pf);
                rsum = rsum - r4[f];
            }
            k4[f] = k;
            rem -= k;
        }
    }
    o = make_int4(k4[0], k4[1], k4[2], k4[3]);
    return nl;
}

// Surface diffusion out of one (triangle, surface species): the leaving pass
// of the volume with three edge directions (SDiff, sdiff.cpp:42-90, and the
// SDiff block of tetopsplit.cpp:1716-1788). Writes three edge slots.
// A per-edge table eD (non-null) replaces D; a clamped source keeps its count.
__device__ __forceinline__ int32_t tri_leave(const SysK &sys, int q, bool tracked, double D, int tri, double dt,
                             uint64_t seed, uint32_t period, uint32_t stream, int4 &o,
                             const double *eD = nullptr, bool clamped = false) {
    o = make_int4(0, 0, 0, 0);
    const size_t ix = (size_t)q * sys.n_tris + tri;
    const int32_t n = sys.tri_counts[ix];
    double r3[3];
#pragma unroll
    for (int e = 0; e < 3; ++e) r3[e] = (eD != nullptr ? eD[e] : D) * sys.tri_g[3 * (size_t)tri + e];
    const double sum = (r3[0] + r3[1]) + r3[2];
    if (n <= 0 || !(sum > 0.0)) return 0;
    Philox g;
    g.init(seed, (uint32_t)sys.tri_user[tri], period, stream, 0x40u + (uint32_t)q);
    double occ = (double)n;
    if (tracked && !almost_zero_dt(dt)) occ = (double)n + sys.tri_corr[ix] / dt;
    const double u = g.unif();
    const double fl = floor(occ);
    int32_t mp = (int32_t)fl + ((occ - fl) > u ? 1 : 0);
    mp = mp > n ? n : mp;
    mp = mp < 0 ? 0 : mp;
    double p = dt * sum;
    p = p > 1.0 ? 1.0 : p;
    p = p < 0.0 ? 0.0 : p;
    const int32_t nl = binomial(g, mp, p);
    if (nl <= 0) return 0;
    if (!clamped) sys.tri_counts[ix] = n - nl;
    int lastf = -1;
#pragma unroll
    for (int e = 0; e < 3; ++e)
        if (r3[e] > 0.0) lastf = e;
    int32_t k3[3] = {0, 0, 0};
    if (nl <= sys.diff_threshold) {
        for (int32_t mol = 0; mol < nl; ++mol) {
            int chosen = -1;
            for (int tries = 0; tries < 64 && chosen < 0; ++tries) {
                const double selv = g.unif() * sum;
                double cum = 0.0;
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    cum = cum + r3[e];
                    if (chosen < 0 && selv < cum) chosen = e;
                }
            }
            if (chosen < 0) chosen = lastf;
#pragma unroll
            for (int e = 0; e < 3; ++e) k3[e] += (chosen == e) ? 1 : 0;
        }
    } else {
        int32_t rem = nl;
        double rsum = sum;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            if (!(r3[e] > 0.0)) continue;
            int32_t k;
            if (e == lastf) {
                k = rem;
            } else {
                const double pf = r3[e] / rsum;
                k = pf >= 1.0 ? rem : binomial(g, rem, pf);
                rsum = rsum - r3[e];
            }
            k3[e] = k;
            rem -= k;
        }
    }
    o = make_int4(k3[0], k3[1], k3[2], 0);
    return nl;
}

// ---------------------------------------------------------------- SSA groups
//
// Membrane triangles whose surface reactions or GHK currents reach the outer
// compartment couple their inner and outer tets into one SSA group: the
// reaction system of the group is one Direct SSA over the period (diffusion
// is frozen within it, as in the reference's operator splitting), run by the
// thread of the group's first tet. Members' counts stay in HBM and their
// occupancy corrections in gcorr; the other members' threads do nothing.

// GHK constants as ghk_consts, with the factor E = e^-nuFoRT (1 in the
// nuFoRT -> 0 limit) in place of c_o e^-nuFoRT, so that the outer
// concentration c_o can be the outer tet's at each evaluation: Bo = c_o E.
__device__ __forceinline__ bool ghk_consts_e(const ModelK &m, int g, double v, double &A, double &E, double &Dn) {
    const double z = (double)m.g_z[g];
    const double nuFoRT = z * (v + m.g_vshift[g]) * kFaraday / (kGasConstant * m.temp);
    if (nuFoRT >= 308 * 2.30 || nuFoRT <= -307 * 2.30) return false;
    const double e = dexp(-nuFoRT);
    if (fabs(nuFoRT) > kDblEpsilon) {
        A = m.g_perm[g] * z * kFaraday * nuFoRT;
        E = e;
        Dn = 1.0 - e;
    } else {
        A = m.g_perm[g] * z * kFaraday;
        E = 1.0;
        Dn = 1.0;
    }
    return true;
}

__device__ __forceinline__ int real_tab(const SysK &sys, int r) {
    return sys.real_tabs != nullptr ? sys.real_tabs[r] : 0;
}

__device__ __forceinline__ uint32_t tet_clamp_of(const SysK &sys, int t, int rtab) {
    return (rtab & kTabClamp) ? sys.tet_clamp[t] : 0u;
}

// propensity of volume reaction q in group member t (realization r)
__device__ double grp_vol_prop(const ModelK &m, const SysK &sys, int t, int r, int rtab, int q) {
    const int NT = sys.n_tets;
    double h = 1.0;
    for (int e = m.lhs_off[q]; e < m.lhs_off[q + 1]; ++e) {
        const int sp = m.ent[e] & 0xffff;
        const int mult = m.ent[e] >> 16;
        const int32_t n = sys.counts[(size_t)sp * NT + t];
        if (mult > n) return 0.0;
        h = falling(h, n, mult);
    }
    const double vscale = 1.0e3 * sys.vol[t] * kAvogadro;
    const double k = (rtab & kTabReac) ? sys.tet_k[(size_t)q * NT + t] : sys.real_k[(size_t)r * m.n_reacs + q];
    return h * scaled_const(k, m.order[q], vscale, 1.0 / vscale);
}

// GHK current g on triangle tri (inner tet t): signed ions per second
// through one open channel (0 and error bit 4 when nuFoRT is out of range)
__device__ double grp_ghk_rt(const ModelK &m, const SysK &sys, int tri, int t, int rtab, int g) {
    const int NT = sys.n_tets;
    double A, E, Dn;
    if (!ghk_consts_e(m, g, sys.tri_v[tri], A, E, Dn)) {
        atomicOr(sys.err, 4);
        return 0.0;
    }
    if (rtab & kTabTriReac) A = A * sys.tri_sk[(size_t)(m.n_sreacs + g) * sys.n_tris + tri];
    const int ion = m.g_ion[g];
    double oc = m.g_oconc[g] * 1.0e3;
    if (m.g_outer[g]) {
        const int o = sys.tri_otet[tri];
        oc = o >= 0 ? (double)sys.counts[(size_t)ion * NT + o] / (1.0e3 * sys.vol[o] * kAvogadro) * 1.0e3 : 0.0;
    }
    const double vscale = 1.0e3 * sys.vol[t] * kAvogadro;
    return ghk_rt(m, g, A, oc * E, Dn, sys.counts[(size_t)ion * NT + t], vscale);
}

// propensity of triangle entry q (surface reaction q < RS, else GHK current
// q - RS) on triangle tri of group member t
__device__ double grp_tri_prop(const ModelK &m, const SysK &sys, int tri, int t, int r, int rtab, int q) {
    const int NT = sys.n_tets, NTR = sys.n_tris, RS = m.n_sreacs;
    if (q >= RS) {
        const int g = q - RS;
        const int32_t nch = sys.tri_counts[(size_t)m.g_chan[g] * NTR + tri];
        return fabs(grp_ghk_rt(m, sys, tri, t, rtab, g)) * (double)nch;
    }
    const int o = sys.tri_otet[tri];
    double h = 1.0;
    for (int e = m.slhs_off[q]; e < m.slhs_off[q + 1]; ++e) {
        const int sp = m.ent[e] & 0xffff;
        const int mult = m.ent[e] >> 16;
        int32_t n;
        if (sp & kSurfFlag) n = sys.tri_counts[(size_t)(sp & 0x7f) * NTR + tri];
        else if (sp & kOuterFlag) n = o >= 0 ? sys.counts[(size_t)(sp & 0xff) * NT + o] : 0;
        else n = sys.counts[(size_t)sp * NT + t];
        if (mult > n) return 0.0;
        h = falling(h, n, mult);
    }
    double k;
    if (rtab & kTabTriReac) {
        const double f = sys.tri_sk[(size_t)q * NTR + tri];
        k = m.s_vdep[q] < 0 ? f : (f == 0.0 ? 0.0 : vdep_k(sys, m.s_vdep[q], sys.tri_v[tri]) * f);
    } else {
        k = m.s_vdep[q] < 0 ? sys.real_sk[(size_t)r * RS + q] : vdep_k(sys, m.s_vdep[q], sys.tri_v[tri]);
    }
    double scale;
    if (m.s_nvol[q] == 0) scale = sys.tri_area[tri] * kAvogadro;
    else if (m.s_out[q]) scale = 1.0e3 * sys.vol[o] * kAvogadro;   // o >= 0: its reactants are present
    else scale = 1.0e3 * sys.vol[t] * kAvogadro;
    return h * scaled_const(k, m.sorder[q], scale, 1.0 / scale);
}

// One period of group g (see above); returns through nev / moved.
template <class SlotT>
__device__ __forceinline__ void group_period(const ModelK &m, const SysK &sys, int g, int launch, bool do_enter, bool do_step,
                             long &nev_out, int32_t &moved_out) {
    const int NT = sys.n_tets, NTR = sys.n_tris, S = m.n_specs, R = m.n_reacs, RS = m.n_sreacs;
    const int RT = RS + m.n_ghk;
    const int k0 = sys.grp_ptr[g], k1 = sys.grp_ptr[g + 1];
    if (do_enter) {
        for (int k = k0; k < k1; ++k) {
            const int t = sys.grp_mem[k], r = sys.grp_real[k];
            const int rtab = real_tab(sys, r);
            const uint32_t clampm = tet_clamp_of(sys, t, rtab);
            const int4 gi4 = sys.nbr[t];
            const int gi[4] = {gi4.x, gi4.y, gi4.z, gi4.w};
            for (int d = 0; d < m.n_diff; ++d) {
                const int sp = m.dspec[d];
                int32_t add = 0;
                for (int f = 0; f < 4; ++f) {
                    if (gi[f] < 0) continue;
                    const size_t ix = (size_t)d * NT * 4 + (size_t)gi[f];
                    int32_t v = static_cast<const SlotT *>(sys.out_prev)[ix];
                    if (v == slot_escape<SlotT>) v = sys.wide_prev[ix];
                    add += v;
                }
                if (!((clampm >> sp) & 1u)) sys.counts[(size_t)sp * NT + t] += add;
            }
            if (m.n_sdiff > 0 && sys.tri_ptr != nullptr) {
                for (int tri = sys.tri_ptr[t]; tri < sys.tri_ptr[t + 1]; ++tri) {
                    const int4 nb = sys.tri_nbr[tri];
                    const uint32_t ne = sys.tri_nface[tri];
                    const int tnb[3] = {nb.x, nb.y, nb.z};
                    const uint32_t tclamp = (rtab & kTabTriClamp) ? sys.tri_clamp[tri] : 0u;
                    for (int d = 0; d < m.n_sdiff; ++d) {
                        const int32_t *op = sys.tout_prev + (size_t)d * NTR * 4;
                        int32_t add = 0;
                        for (int e = 0; e < 3; ++e)
                            if (tnb[e] >= 0) add += op[(size_t)tnb[e] * 4 + ((ne >> (8 * e)) & 0xff)];
                        if (!((tclamp >> m.sdspec[d]) & 1u)) sys.tri_counts[(size_t)m.sdspec[d] * NTR + tri] += add;
                    }
                }
            }
        }
    }
    if (!do_step) return;
    const int ro = sys.grp_real[k0], to = sys.grp_mem[k0];
    const uint32_t period = sys.real_step0[ro] + (uint32_t)launch;
    const double dt = sys.dt_tab[(size_t)launch * sys.n_real + ro];
    // occupancy corrections of the period start at zero
    for (int k = k0; k < k1; ++k) {
        const int t = sys.grp_mem[k];
        for (int sp = 0; sp < S; ++sp)
            if ((m.tracked_mask >> sp) & 1) sys.gcorr[(size_t)sp * NT + t] = 0.0;
        if (m.tri_tracked_mask != 0 && sys.tri_ptr != nullptr)
            for (int q = 0; q < m.n_sspecs; ++q)
                if ((m.tri_tracked_mask >> q) & 1)
                    for (int tri = sys.tri_ptr[t]; tri < sys.tri_ptr[t + 1]; ++tri) sys.tri_corr[(size_t)q * NTR + tri] = 0.0;
    }
    // Direct SSA over the group: the members' volume reactions in member
    // order, then each member's triangles' entries
    Philox gph;
    gph.init(sys.real_seed[ro], (uint32_t)(to - sys.real_tet0[ro]), period, sys.real_stream[ro], 0u);
    long nev = 0;
    double tcur = 0.0;
    while (true) {
        double a0 = 0.0;
        for (int k = k0; k < k1; ++k) {
            const int t = sys.grp_mem[k], r = sys.grp_real[k], rtab = real_tab(sys, r);
            for (int q = 0; q < R; ++q) a0 = a0 + grp_vol_prop(m, sys, t, r, rtab, q);
        }
        for (int k = k0; k < k1 && sys.tri_ptr != nullptr; ++k) {
            const int t = sys.grp_mem[k], r = sys.grp_real[k], rtab = real_tab(sys, r);
            for (int tri = sys.tri_ptr[t]; tri < sys.tri_ptr[t + 1]; ++tri)
                for (int q = 0; q < RT; ++q) a0 = a0 + grp_tri_prop(m, sys, tri, t, r, rtab, q);
        }
        if (a0 < kDblEpsilon) break;
        const double u1 = gph.unif();
        if (1.0 - u1 > a0 * (dt - tcur) * 1.0000001) break;
        const double u2 = gph.unif();
        const double tev = tcur + (-dlog(u1) / a0);
        if (tev > dt) break;
        tcur = tev;
        // selection: the first entry whose running sum exceeds x, else the
        // last nonzero one; an entry is (member k, triangle or -1, index q)
        const double x = u2 * a0;
        int sk = -1, stri = -1, sq = -1, lk = -1, ltri = -1, lq = -1;
        double acc = 0.0;
        for (int k = k0; k < k1; ++k) {
            const int t = sys.grp_mem[k], r = sys.grp_real[k], rtab = real_tab(sys, r);
            for (int q = 0; q < R; ++q) {
                const double pq = grp_vol_prop(m, sys, t, r, rtab, q);
                acc = acc + pq;
                if (sk < 0 && acc > x) { sk = k; stri = -1; sq = q; }
                if (pq > 0.0) { lk = k; ltri = -1; lq = q; }
            }
        }
        for (int k = k0; k < k1 && sys.tri_ptr != nullptr; ++k) {
            const int t = sys.grp_mem[k], r = sys.grp_real[k], rtab = real_tab(sys, r);
            for (int tri = sys.tri_ptr[t]; tri < sys.tri_ptr[t + 1]; ++tri)
                for (int q = 0; q < RT; ++q) {
                    const double pq = grp_tri_prop(m, sys, tri, t, r, rtab, q);
                    acc = acc + pq;
                    if (sk < 0 && acc > x) { sk = k; stri = tri; sq = q; }
                    if (pq > 0.0) { lk = k; ltri = tri; lq = q; }
                }
        }
        if (sk < 0) { sk = lk; stri = ltri; sq = lq; }
        if (sk < 0) break;  // every propensity is zero (a0 came from rounding)
        const int t = sys.grp_mem[sk], r = sys.grp_real[sk], rtab = real_tab(sys, r);
        const uint32_t clampm = tet_clamp_of(sys, t, rtab);
        if (stri < 0) {
            for (int e = m.upd_off[sq]; e < m.upd_off[sq + 1]; ++e) {
                const int sp = m.ent[e] & 0xffff;
                const int dv = m.ent[e] >> 16;
                if ((clampm >> sp) & 1u) continue;
                sys.counts[(size_t)sp * NT + t] += dv;
                if ((m.tracked_mask >> sp) & 1) sys.gcorr[(size_t)sp * NT + t] -= (double)dv * tcur;
            }
        } else {
            const int o = sys.tri_otet[stri];
            const uint32_t oclamp = o >= 0 ? tet_clamp_of(sys, o, real_tab(sys, sys.tri_oreal[stri])) : 0u;
            if (sq >= RS) {
                // GHK current: one ion crosses in the direction of the flux,
                // from the inner tet to the outer one or back
                // (ghkcurr.cpp:228-272); charge counted outward positive
                const int gq = sq - RS;
                const int ion = m.g_ion[gq];
                const bool efflux = grp_ghk_rt(m, sys, stri, t, rtab, gq) >= 0.0;
                if (m.g_real[gq]) {
                    const int dv = efflux ? -1 : 1;
                    if (!((clampm >> ion) & 1u)) {
                        sys.counts[(size_t)ion * NT + t] += dv;
                        if ((m.tracked_mask >> ion) & 1) sys.gcorr[(size_t)ion * NT + t] -= (double)dv * tcur;
                    }
                    if (m.g_outer[gq] && o >= 0 && !((oclamp >> ion) & 1u)) {
                        sys.counts[(size_t)ion * NT + o] -= dv;
                        if ((m.tracked_mask >> ion) & 1) sys.gcorr[(size_t)ion * NT + o] += (double)dv * tcur;
                    }
                }
                sys.tri_charge[(size_t)gq * sys.n_tris + stri] += efflux ? m.g_z[gq] : -m.g_z[gq];
            } else {
                // surface reaction: triangle, inner and outer tet counts
                // (SReac::apply, sreac.cpp:440-505)
                const uint32_t tclamp = (rtab & kTabTriClamp) ? sys.tri_clamp[stri] : 0u;
                for (int e = m.supd_off[sq]; e < m.supd_off[sq + 1]; ++e) {
                    const int sp = m.ent[e] & 0xffff;
                    const int dv = m.ent[e] >> 16;
                    if (sp & kSurfFlag) {
                        const int qs = sp & 0x7f;
                        if ((tclamp >> qs) & 1u) continue;
                        const size_t ix = (size_t)qs * NTR + stri;
                        sys.tri_counts[ix] += dv;
                        if ((m.tri_intg_mask >> qs) & 1) sys.tri_intg[ix] -= (double)dv * tcur;
                        if ((m.tri_tracked_mask >> qs) & 1) sys.tri_corr[ix] -= (double)dv * tcur;
                    } else if (sp & kOuterFlag) {
                        const int so = sp & 0xff;
                        if (o < 0 || ((oclamp >> so) & 1u)) continue;
                        sys.counts[(size_t)so * NT + o] += dv;
                        if ((m.tracked_mask >> so) & 1) sys.gcorr[(size_t)so * NT + o] -= (double)dv * tcur;
                    } else {
                        if ((clampm >> sp) & 1u) continue;
                        sys.counts[(size_t)sp * NT + t] += dv;
                        if ((m.tracked_mask >> sp) & 1) sys.gcorr[(size_t)sp * NT + t] -= (double)dv * tcur;
                    }
                }
            }
        }
        if (++nev >= kMaxEventsPerPeriod) {
            atomicOr(sys.err, 1);
            nev = 0;
            break;
        }
    }
    nev_out += nev;
    // channel-count integrals, then every member's leaving pass
    int32_t moved = 0;
    for (int k = k0; k < k1; ++k) {
        const int t = sys.grp_mem[k], r = sys.grp_real[k], rtab = real_tab(sys, r);
        const uint32_t clampm = tet_clamp_of(sys, t, rtab);
        const int tp0 = sys.tri_ptr != nullptr ? sys.tri_ptr[t] : 0;
        const int tp1 = sys.tri_ptr != nullptr ? sys.tri_ptr[t + 1] : 0;
        if (m.tri_intg_mask != 0)
            for (int q = 0; q < m.n_sspecs; ++q)
                if ((m.tri_intg_mask >> q) & 1)
                    for (int tri = tp0; tri < tp1; ++tri) {
                        const size_t ix = (size_t)q * NTR + tri;
                        sys.tri_intg[ix] += (double)sys.tri_counts[ix] * dt;
                    }
        const uint32_t ltet = (uint32_t)(t - sys.real_tet0[r]);
        const uint32_t per = sys.real_step0[r] + (uint32_t)launch;
        const double dtr = sys.dt_tab[(size_t)launch * sys.n_real + r];
        const uint64_t seed = sys.real_seed[r];
        const uint32_t stream = sys.real_stream[r];
        const uint32_t bdir = sys.tet_bdir != nullptr ? sys.tet_bdir[t] : 0xffffffffu;
        for (int d = 0; d < m.n_diff; ++d) {
            const int sp = m.dspec[d];
            const double2 g01 = sys.g[2 * (size_t)t];
            const double2 g23 = sys.g[2 * (size_t)t + 1];
            const double r4[4] = {g01.x, g01.y, g23.x, g23.y};
            const double *fD = (rtab & kTabDiff) ? sys.tet_fD + ((size_t)sp * NT + t) * 4 : nullptr;
            const size_t ci = (size_t)sp * NT + t;
            const int32_t n = sys.counts[ci];
            int4 o;
            const int32_t nl = leave(sys, n, sys.gcorr[ci], sp, (m.tracked_mask >> sp) & 1,
                                     sys.real_D[(size_t)r * S + sp], r4, bdir, S, dtr, seed, ltet, per, stream, o, fD);
            if (nl > 0 && !((clampm >> sp) & 1u)) sys.counts[ci] = n - nl;
            moved += nl;
            put_slots<SlotT>(sys.out_cur, sys.wide_cur, ((size_t)d * NT + t) * 4, o);
        }
        if (m.n_sdiff > 0) {
            int4 *toc = reinterpret_cast<int4 *>(sys.tout_cur);
            for (int tri = tp0; tri < tp1; ++tri) {
                const uint32_t tclamp = (rtab & kTabTriClamp) ? sys.tri_clamp[tri] : 0u;
                for (int d = 0; d < m.n_sdiff; ++d) {
                    const int q = m.sdspec[d];
                    int4 o;
                    const double *eD = (rtab & kTabTriDiff) ? sys.tri_eD + ((size_t)q * NTR + tri) * 3 : nullptr;
                    moved += tri_leave(sys, q, (m.tri_tracked_mask >> q) & 1, sys.real_sD[(size_t)r * m.n_sspecs + q],
                                       tri, dtr, seed, per, stream, o, eD, (tclamp >> q) & 1u);
                    toc[(size_t)d * NTR + tri] = o;
                }
            }
        }
    }
    moved_out += moved;
}

#ifndef STG_PERIOD_WAVES_PER_EU
#define STG_PERIOD_WAVES_PER_EU 1
#endif
#ifndef STG_VOL_WAVES_PER_EU
#define STG_VOL_WAVES_PER_EU 1
#endif

// Per-block extents (events, molecules moved): wavefront reduction, then one
// slot per block, added by thread 0. Every thread of the block calls it.
__device__ __forceinline__ void block_extents(const SysK &sys, int slot, long nev, int32_t moved) {
    unsigned long long ev = (unsigned long long)nev, mv = (unsigned long long)moved;
    for (int off = 32; off > 0; off >>= 1) {
        ev += __shfl_xor(ev, off, 64);
        mv += __shfl_xor(mv, off, 64);
    }
    __shared__ unsigned long long red[2][16];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (lane == 0) {
        red[0][wave] = ev;
        red[1][wave] = mv;
    }
    __syncthreads();
    if (tid == 0) {
        unsigned long long se = 0, sm = 0;
        for (int w = 0; w < ((int)blockDim.x + 63) / 64; ++w) {
            se += red[0][w];
            sm += red[1][w];
        }
        sys.ext[2 * (size_t)slot] += se;
        sys.ext[2 * (size_t)slot + 1] += sm;
    }
}

// Volume-only models with G > 1 lanes per tet (the tile schedule, for meshes
// too small to fill the GPU, where a period takes one wavefront's critical
// path): the G lanes of a tet gather, leave and store its species s with
// s = lane mod G, and lane 0 runs the tet's SSA. Each species keeps its own
// Philox channel, so the draws, and the results, are those of one lane per
// tet; a wavefront holds 64 / G tets, which shortens the SSA's slowest lane
// and splits the leaving pass over the species.
template <int G, class MT, class SlotT>
__device__ void period_tile(const ModelK &m, const SysK &sys, unsigned char *smem, const int4 bd, int r, int launch,
                            bool do_enter, bool do_step, long &nev, int32_t &moved) {
    const int B = blockDim.x;
    const int BT = B / G;  // tets per block
    const int tid = threadIdx.x;
    const int slot = tid / G, sub = tid % G;
    const int S = tab_S<MT>(m);
    const int ND = tab_ndiff<MT>(m);
    Lds L;
    L.B = BT;
    L.tid = slot;
    L.cnt = reinterpret_cast<int32_t *>(smem);
    L.corr = reinterpret_cast<double *>(smem + (((size_t)S * BT * 4 + 15) & ~(size_t)15));
    L.kc = L.corr + (size_t)S * BT;
    L.a = L.kc + (size_t)sys.k_max * BT;
    L.gx = nullptr;
    const int t = bd.z + slot;
    const bool valid = t < bd.w;
    const int NT = sys.n_tets;
    if (valid) {
        for (int s = sub; s < S; s += G) {
            L.c(s) = sys.counts[(size_t)s * NT + t];
            L.o(s) = 0.0;
        }
        if (do_enter) {
            const int4 gi4 = sys.nbr[t];
            const int nbv[4] = {gi4.x, gi4.y, gi4.z, gi4.w};
            size_t off[4];
#pragma unroll
            for (int f = 0; f < 4; ++f) off[f] = nbv[f] >= 0 ? (size_t)nbv[f] : 0;
            for (int d = sub; d < ND; d += G) {
                const SlotT *op = static_cast<const SlotT *>(sys.out_prev) + (size_t)d * NT * 4;
                int32_t v[4];
#pragma unroll
                for (int f = 0; f < 4; ++f) v[f] = op[off[f]];
                int32_t add = 0;
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    if (v[f] == slot_escape<SlotT>) v[f] = sys.wide_prev[(size_t)d * NT * 4 + off[f]];
                    add += nbv[f] >= 0 ? v[f] : 0;
                }
                L.c(tab_dspec<MT>(m, d)) += add;
            }
        }
    }
    __syncthreads();  // every species of the tet has entered
    if (do_step) {
        const int tet0 = sys.real_tet0[r];
        const uint32_t ltet = (uint32_t)(t - tet0);
        const uint32_t period = uni(sys.real_step0[r]) + (uint32_t)launch;
        const double dt = uni(sys.dt_tab[(size_t)launch * sys.n_real + r]);
        const uint64_t seed = uni(sys.real_seed[r]);
        const uint32_t stream = uni(sys.real_stream[r]);
        if (valid && sub == 0) {
            if constexpr (MT::kStatic) nev = ssa_lean<MT>(sys, L, r, t, ltet, period, dt, seed, stream);
            else nev = ssa_group<false, MT>(m, sys, L, r, t, ltet, period, dt, seed, stream);
            if (nev < 0) {
                atomicOr(sys.err, 1);
                nev = 0;
            }
        }
        __syncthreads();  // the SSA's counts and occupancy corrections are in LDS
        if (valid) {
            const double *Dr = sys.real_D + (size_t)r * S;
            for (int d = sub; d < ND; d += G) {
                const int sp = tab_dspec<MT>(m, d);
                const int tg = pin(t);
                const double2 g01 = sys.g[2 * (size_t)tg];
                const double2 g23 = sys.g[2 * (size_t)tg + 1];
                const double r4[4] = {g01.x, g01.y, g23.x, g23.y};
                int4 o;
                const int32_t n = L.c(sp);
                // the lanes of a wavefront hold different species: D is read
                // per lane, not taken from the first lane
                const int32_t nl = leave(sys, n, L.o(sp), sp, (tab_tracked<MT>(m) >> sp) & 1, Dr[sp], r4,
                                         0xffffffffu, S, dt, seed, ltet, period, stream, o);
                if (nl > 0) L.c(sp) = n - nl;
                moved += nl;
                put_slots<SlotT>(sys.out_cur, sys.wide_cur, ((size_t)d * NT + pin(t)) * 4, o);
            }
        }
        __syncthreads();  // a species may be stored by another lane than the one that moved it
    }
    if (valid)
        for (int s = sub; s < S; s += G) sys.counts[(size_t)s * NT + t] = L.c(s);
}

template <int MAXB, bool MEMB, class MT = DynTables, int G = 1, class SlotT = int16_t>
__global__ void __launch_bounds__(MAXB, MEMB ? STG_PERIOD_WAVES_PER_EU : STG_VOL_WAVES_PER_EU)
    k_period(const ModelK m, const SysK sys, int launch) {
    static_assert(G == 1 || (!MEMB && G <= 8 && (G & (G - 1)) == 0), "lanes per tet: 1, or 2/4/8 for volume models");
    extern __shared__ __align__(16) unsigned char smem[];
    const int B = blockDim.x;
    const int tid = threadIdx.x;
    int b = blockIdx.x;
    if (sys.xcd_remap) {
        // Hardware dispatches consecutive blocks round-robin over 8 XCDs; give
        // each XCD a contiguous range of tets so face neighbours share an L2.
        const int nb = sys.n_blocks;
        const int per = (nb + 7) / 8;
        const int x = b & 7, i = b >> 3;
        const int full = nb - 8 * (per - 1);  // XCDs that get 'per' blocks
        const int lb = (x < full) ? x * per + i : full * per + (x - full) * (per - 1) + i;
        b = lb;
    }
    const int4 bd = sys.block_desc[b];
    const int r = bd.y;
    const int np = sys.real_nper[r];
    const bool do_enter = launch >= 1 && launch <= np;
    const bool do_step = launch < np;
    if (!do_enter && !do_step) return;  // uniform over the block

    long nev = 0;
    int32_t moved = 0;
    if constexpr (G > 1) {
        period_tile<G, MT, SlotT>(m, sys, smem, bd, r, launch, do_enter, do_step, nev, moved);
        block_extents(sys, bd.x, nev, moved);
        return;
    }
    const int S = tab_S<MT>(m);
    Lds L;
    L.B = B;
    L.tid = tid;
    L.cnt = reinterpret_cast<int32_t *>(smem);
    L.corr = reinterpret_cast<double *>(smem + (((size_t)S * B * 4 + 15) & ~(size_t)15));
    L.kc = L.corr + (size_t)S * B;
    if constexpr (MT::kStatic && MEMB) {
        // lean membrane SSA: only the triangle rows, then the GHK rows
        L.a = nullptr;
        L.gx = L.kc + (size_t)(sys.k_max - MT::R) * B;
    } else {
        L.a = L.kc + (size_t)sys.k_max * B;
        L.gx = MEMB ? L.a + (size_t)sys.k_max * B : nullptr;
    }

    const int t = bd.z + tid;
    const int tend = bd.w;
    const bool valid = t < tend;
    // per-element tables of this realization (uniform over the block); the
    // engine runs the generic kernel whenever a realization has one
    const int rtab = (MEMB && !MT::kStatic && sys.real_tabs != nullptr) ? sys.real_tabs[r] : 0;
    // a tet of a multi-tet SSA group: the group's first tet runs it all
    bool in_group = false;
    if constexpr (MEMB && !MT::kStatic) {
        if (valid && sys.tet_grp != nullptr) {
            const int grp = sys.tet_grp[t];
            if (grp >= 0) {
                in_group = true;
                if (sys.grp_mem[sys.grp_ptr[grp]] == t)
                    group_period<SlotT>(m, sys, grp, launch, do_enter, do_step, nev, moved);
            }
        }
    }
    if (valid && !in_group) {
        const int NT = sys.n_tets;
        // clamped species: nothing enters, leaves or reacts away (Diff::apply,
        // diff.cpp:476-577; Reac::apply, reac.cpp:240-243)
        const uint32_t clampm = (MEMB && !MT::kStatic && (rtab & kTabClamp)) ? sys.tet_clamp[t] : 0u;
        for (int s = 0; s < S; ++s) {
            L.c(s) = sys.counts[(size_t)s * NT + t];
            L.o(s) = 0.0;
        }
        const int4 gi4 = sys.nbr[t];
        if (do_enter) {
            // every slot is read unconditionally (a missing neighbour reads
            // slot 0 and is masked) so the four loads are in flight together
            const int nbv[4] = {gi4.x, gi4.y, gi4.z, gi4.w};
            size_t off[4];
#pragma unroll
            for (int f = 0; f < 4; ++f) off[f] = nbv[f] >= 0 ? (size_t)nbv[f] : 0;
            // the slots of up to four diffusing species are loaded together and
            // the (rare) escapes resolved once all of them have arrived: an
            // escape test per species would wait on that species' loads before
            // the next species' loads are issued
            const int ND = tab_ndiff<MT>(m);
            for (int d0 = 0; d0 < ND; d0 += 4) {
                int32_t v[4][4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const SlotT *op = static_cast<const SlotT *>(sys.out_prev) + (size_t)(d0 + j) * NT * 4;
#pragma unroll
                    for (int f = 0; f < 4; ++f) v[j][f] = d0 + j < ND ? (int32_t)op[off[f]] : 0;
                }
                bool esc = false;
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int f = 0; f < 4; ++f) esc = esc || v[j][f] == slot_escape<SlotT>;
                if (esc) {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
#pragma unroll
                        for (int f = 0; f < 4; ++f)
                            if (v[j][f] == slot_escape<SlotT>) v[j][f] = sys.wide_prev[(size_t)(d0 + j) * NT * 4 + off[f]];
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (d0 + j >= ND) break;
                    int32_t add = 0;
#pragma unroll
                    for (int f = 0; f < 4; ++f) add += nbv[f] >= 0 ? v[j][f] : 0;
                    const int dsp = tab_dspec<MT>(m, d0 + j);
                    if (!MEMB || !((clampm >> dsp) & 1u)) L.c(dsp) += add;
                }
            }
        }
        int tp0 = 0, tp1 = 0;
        if (MEMB && sys.tri_ptr != nullptr) {
            tp0 = sys.tri_ptr[t];
            tp1 = sys.tri_ptr[t + 1];
        }
        if (MEMB && do_enter && m.n_sdiff > 0) {
            // surface molecules entering this tet's triangles across their edges
            for (int tri = tp0; tri < tp1; ++tri) {
                const int4 nb = sys.tri_nbr[tri];
                const uint32_t ne = sys.tri_nface[tri];
                const int nbv[3] = {nb.x, nb.y, nb.z};
                const uint32_t tclamp = (rtab & kTabTriClamp) ? sys.tri_clamp[tri] : 0u;
                for (int d = 0; d < m.n_sdiff; ++d) {
                    const int32_t *op = sys.tout_prev + (size_t)d * sys.n_tris * 4;
                    int32_t add = 0;
#pragma unroll
                    for (int e = 0; e < 3; ++e)
                        if (nbv[e] >= 0) add += op[(size_t)nbv[e] * 4 + ((ne >> (8 * e)) & 0xff)];
                    if (!((tclamp >> m.sdspec[d]) & 1u)) sys.tri_counts[(size_t)m.sdspec[d] * sys.n_tris + tri] += add;
                }
            }
        }
        if (do_step) {
            const int tet0 = sys.real_tet0[r];
            const uint32_t ltet = (uint32_t)(t - tet0);
            // per-realization constants, in scalar registers
            const uint32_t period = uni(sys.real_step0[r]) + (uint32_t)launch;
            const double dt = uni(sys.dt_tab[(size_t)launch * sys.n_real + r]);
            const uint64_t seed = uni(sys.real_seed[r]);
            const uint32_t stream = uni(sys.real_stream[r]);
            if (MEMB && m.tri_tracked_mask != 0)
                for (int q = 0; q < m.n_sspecs; ++q)
                    if ((m.tri_tracked_mask >> q) & 1)
                        for (int tri = tp0; tri < tp1; ++tri) sys.tri_corr[(size_t)q * sys.n_tris + tri] = 0.0;
            if constexpr (MT::kStatic && !MEMB) nev = ssa_lean<MT>(sys, L, r, t, ltet, period, dt, seed, stream);
            else if constexpr (MT::kStatic && MEMB)
                nev = ssa_memb_lean<MT>(m, sys, L, r, t, ltet, period, dt, seed, stream);
            else nev = ssa_group<MEMB, MT>(m, sys, L, r, t, ltet, period, dt, seed, stream, rtab, clampm);
            if (nev < 0) {
                atomicOr(sys.err, 1);
                nev = 0;
            }
            if (MEMB && m.tri_intg_mask != 0) {
                for (int q = 0; q < m.n_sspecs; ++q) {
                    if (!((m.tri_intg_mask >> q) & 1)) continue;
                    for (int tri = tp0; tri < tp1; ++tri) {
                        const size_t ix = (size_t)q * sys.n_tris + tri;
                        sys.tri_intg[ix] += (double)sys.tri_counts[ix] * dt;
                    }
                }
            }
            if (tab_ndiff<MT>(m) > 0) {
                const uint32_t bdir = (MEMB && sys.tet_bdir != nullptr) ? sys.tet_bdir[t] : 0xffffffffu;
                const double *Dr = sys.real_D + (size_t)r * S;
                for (int d = 0; d < tab_ndiff<MT>(m); ++d) {
                    const int sp = tab_dspec<MT>(m, d);
                    // the face factors are read again for every species (cache
                    // hits) from a fenced index: held across the species loop
                    // they would cost 8 vector registers at its peak
                    const int tg = pin(t);
                    const double2 g01 = sys.g[2 * (size_t)tg];
                    const double2 g23 = sys.g[2 * (size_t)tg + 1];
                    const double r4[4] = {g01.x, g01.y, g23.x, g23.y};
                    int4 o;
                    const double *fD = (MEMB && (rtab & kTabDiff)) ? sys.tet_fD + ((size_t)sp * NT + t) * 4 : nullptr;
                    const int32_t n = L.c(sp);
                    const int32_t nl = leave(sys, n, L.o(sp), sp, (tab_tracked<MT>(m) >> sp) & 1, uni(Dr[sp]), r4,
                                             bdir, S, dt, seed, ltet, period, stream, o, fD);
                    if (nl > 0 && !(MEMB && ((clampm >> sp) & 1u))) L.c(sp) = n - nl;
                    moved += nl;
                    put_slots<SlotT>(sys.out_cur, sys.wide_cur, ((size_t)d * NT + pin(t)) * 4, o);
                }
            }
            if (MEMB && m.n_sdiff > 0) {
                const double *sDr = sys.real_sD + (size_t)r * m.n_sspecs;
                int4 *toc = reinterpret_cast<int4 *>(sys.tout_cur);
                for (int tri = tp0; tri < tp1; ++tri) {
                    const uint32_t tclamp = (rtab & kTabTriClamp) ? sys.tri_clamp[tri] : 0u;
                    for (int d = 0; d < m.n_sdiff; ++d) {
                        const int q = m.sdspec[d];
                        int4 o;
                        const double *eD = (rtab & kTabTriDiff) ? sys.tri_eD + ((size_t)q * sys.n_tris + tri) * 3 : nullptr;
                        moved += tri_leave(sys, q, (m.tri_tracked_mask >> q) & 1, uni(sDr[q]), tri, dt, seed,
                                           period, stream, o, eD, (tclamp >> q) & 1u);
                        toc[(size_t)d * sys.n_tris + tri] = o;
                    }
                }
            }
        }
        // store addresses formed again rather than held from the loads
        const int tw = pin(t);
        for (int s = 0; s < S; ++s) sys.counts[(size_t)s * NT + tw] = L.c(s);
    }
    block_extents(sys, bd.x, nev, moved);
}

// The entering pass alone, for the launch after the last period of a run: the
// same gather as the head of k_period, for the blocks whose realization's
// last period was the previous launch (the first launch_blocks[L] block
// descriptors). A kernel of its own so that every k_period dispatch of a
// single-realization run is a full period.
template <bool MEMB, class SlotT>
__global__ void __launch_bounds__(1024) k_enter(const ModelK m, const SysK sys) {
    const int4 bd = sys.block_desc[blockIdx.x];
    const int t = bd.z + (int)threadIdx.x;
    if (t >= bd.w) return;
    const int NT = sys.n_tets;
    const int rtab = (MEMB && sys.real_tabs != nullptr) ? sys.real_tabs[bd.y] : 0;
    const uint32_t clampm = (rtab & kTabClamp) ? sys.tet_clamp[t] : 0u;
    const int4 gi4 = sys.nbr[t];
    const int nbv[4] = {gi4.x, gi4.y, gi4.z, gi4.w};
    size_t off[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) off[f] = nbv[f] >= 0 ? (size_t)nbv[f] : 0;
    for (int d = 0; d < m.n_diff; ++d) {
        const SlotT *op = static_cast<const SlotT *>(sys.out_prev) + (size_t)d * NT * 4;
        int32_t add = 0;
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            int32_t v = op[off[f]];
            if (v == slot_escape<SlotT>) v = sys.wide_prev[(size_t)d * NT * 4 + off[f]];
            add += nbv[f] >= 0 ? v : 0;
        }
        if (!((clampm >> m.dspec[d]) & 1u)) sys.counts[(size_t)m.dspec[d] * NT + t] += add;
    }
    if (MEMB && m.n_sdiff > 0 && sys.tri_ptr != nullptr) {
        for (int tri = sys.tri_ptr[t]; tri < sys.tri_ptr[t + 1]; ++tri) {
            const int4 nb = sys.tri_nbr[tri];
            const uint32_t ne = sys.tri_nface[tri];
            const int tnb[3] = {nb.x, nb.y, nb.z};
            const uint32_t tclamp = (rtab & kTabTriClamp) ? sys.tri_clamp[tri] : 0u;
            for (int d = 0; d < m.n_sdiff; ++d) {
                const int32_t *op = sys.tout_prev + (size_t)d * sys.n_tris * 4;
                int32_t add = 0;
#pragma unroll
                for (int e = 0; e < 3; ++e)
                    if (tnb[e] >= 0) add += op[(size_t)tnb[e] * 4 + ((ne >> (8 * e)) & 0xff)];
                if (!((tclamp >> m.sdspec[d]) & 1u)) sys.tri_counts[(size_t)m.sdspec[d] * sys.n_tris + tri] += add;
            }
        }
    }
}

// getBatchTetSpecCountsNP gather
__global__ void k_gather_counts(const int32_t *__restrict__ counts, const int64_t *__restrict__ idx,
                                size_t n, int32_t base, double *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)counts[base + idx[i]];
}

// Sum of n int32 counts into one int64 (block partials + one atomic per block).
__global__ void k_sum_counts(const int32_t *__restrict__ v, int32_t n,
                             unsigned long long *__restrict__ total) {
    long long acc = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        acc += v[i];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    __shared__ long long red[16];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long s = 0;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) s += red[w];
        atomicAdd(total, (unsigned long long)s);
    }
}

// Totals of every (realization, species): block b = r * S + s reduces
// species s over realization r's elements [off[r], off[r + 1]) of a
// species-major [S][n] array; no atomics, so the sums are the same every call.
__global__ void k_sum_real_counts(const int32_t *__restrict__ v, int32_t n, int32_t S,
                                  const int32_t *__restrict__ off, long long *__restrict__ out) {
    const int r = blockIdx.x / S, s = blockIdx.x % S;
    const int32_t *p = v + (size_t)s * n;
    long long acc = 0;
    for (int i = off[r] + (int)threadIdx.x; i < off[r + 1]; i += blockDim.x) acc += p[i];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    __shared__ long long red[16];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long tot = 0;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) tot += red[w];
        out[blockIdx.x] = tot;
    }
}

// Sum of counts[base + idx[i]] over i < n into one int64 (an ROI total).
__global__ void k_sum_indexed(const int32_t *__restrict__ counts, const int64_t *__restrict__ idx,
                              int64_t n, int32_t base, unsigned long long *__restrict__ total) {
    long long acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc += counts[base + idx[i]];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    __shared__ long long red[16];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long s = 0;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) s += red[w];
        atomicAdd(total, (unsigned long long)s);
    }
}

}  // namespace stg
