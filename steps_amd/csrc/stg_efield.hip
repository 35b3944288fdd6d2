// Membrane-potential (EField) step on the GPU: vertex system assembly and a
// mixed-precision conjugate-gradient solve.
//
// The system of one step (dVSolverBase::_advance,
// src/steps/solver/efield/dVsolver.hpp:126-171) in microns / ms / mV / pF / nS / pA:
//   (C_i/dt + G_i + sum_j c_ij) dV_i - sum_j c_ij dV_j
//       = I_i + G_i (V_ext - V_i) + sum_j c_ij (V_j - V_i)
// with c_ij = sigma * w_ij, w_ij the coupling of a piecewise-linear potential
// over the barycentric dual cells (TetCoupler, tetcoupler.cpp:48-290), which is
// the P1 stiffness -sum_tets V grad(phi_i).grad(phi_j) assembled here directly.
// The matrix is symmetric positive definite; the reference solves it by banded
// LU (bdsystem.cpp) or PETSc CG (dVsolver_petsc.cpp:163-268). Here: CG with the
// matrix and vectors in FP32 and every dot product accumulated in FP64, wrapped
// in FP64 iterative refinement (residual b - A x in FP64) until the FP64
// residual meets the tolerance. An all-FP64 CG mode gives the reference
// solution on the device. The preconditioner is one V-cycle of smoothed-
// aggregation algebraic multigrid (set up on the host once per time step size
// and membrane parameters, applied on the device) for systems above 16,384
// vertices, and Jacobi below that or when chosen.
//
// Reductions are deterministic and atomic-free: each block writes its partial,
// then one workgroup sums the partials in a fixed order and updates the CG
// scalars on the device, so an iteration needs no host round trip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/stg.h"

namespace {

struct EfError {
    int code;
    std::string msg;
};

#define EF_HIP(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw EfError{STG_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

template <typename T>
T *ef_alloc(size_t n) {
    void *p = nullptr;
    EF_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return static_cast<T *>(p);
}

template <typename T>
void ef_put(T *d, const std::vector<T> &h) {
    if (!h.empty()) EF_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

constexpr int kB = 256;
constexpr int kCoarsest = 400;  // levels at or below this size get a dense inverse

// ------------------------------------------------------------- reductions

// Sum of one double per thread of the block into partial[blockIdx.x].
__device__ __forceinline__ void block_partial(double v, double *partial) {
    __shared__ double red[kB / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < (int)(blockDim.x + 63) / 64; ++i) s += red[i];
        partial[blockIdx.x] = s;
    }
}

// One workgroup sums the nb partials in a fixed order and applies `op`:
//   0: scal[0] = sum, scal[1] = 0            (CG start: rz, beta)
//   1: scal[3] = sum                          (p.q of this iteration)
//   2: scal[1] = sum / scal[0], scal[0] = sum (new rz and beta), out[1] = sum
//   3: out[0] = sum                           (a plain reduction)
__global__ void __launch_bounds__(1024) k_reduce(const double *partial, int nb, int op, double *scal,
                                                 double *out) {
    __shared__ double red[16];
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partial[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) red[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < (int)(blockDim.x + 63) / 64; ++i) t += red[i];
        if (op == 0) {
            scal[0] = t;
            scal[1] = 0.0;
            out[0] = t;  // kept: scal[0] moves on with the iterations
        } else if (op == 1) {
            scal[3] = t;
        } else if (op == 2) {
            scal[1] = scal[0] > 0.0 ? t / scal[0] : 0.0;
            scal[0] = t;
            out[1] = t;  // beside the start's rz in out[0]: one read-back for both
        } else {
            out[0] = t;
        }
    }
}

// ------------------------------------------------------------- kernels

// The off-diagonal couplings are stored ELLPACK-style, slot-major
// ([width][n], rows padded with column i and coupling 0), so the 64 rows of a
// wavefront read each slot with one coalesced load.
struct EfDev {
    int n;                 // vertices
    int width;             // slots per row
    const int *col;        // [width][n] off-diagonal columns
    const double *w;       // [width][n] geometric coupling (um)
    const double *wsum;    // [n]   row sums of w
    const float *a32;      // [width][n] -c_ij in FP32 (0 for clamped rows/cols)
    const float *d32;      // [n]   diagonal in FP32
    const float *m32;      // [n]   row sum in FP32: C/dt + G + sigma * (couplings to clamped vertices)
    const double *d64;     // [n]   diagonal in FP64
    const unsigned char *clamped;
};

// FP32 products use the difference form (A x)_i = m_i x_i + sum_j a_ij (x_j - x_i),
// with m_i the row sum of the FP64 matrix rounded once. Rounding a_ij and the
// diagonal separately would leave each Laplacian row (zero sum) off by ~1e-7
// of its diagonal, a spurious mass term that swamps the membrane capacitance on
// the smooth modes and costs refinements; the difference form is exact on
// constants, like the FP64 operator.

// membrane current of one triangle (pA, outward positive) from ohmic channels
// (Tri::computeI, src/steps/mpi/tetopsplit/tri.cpp:664-693)
struct OhmicK {
    int n;
    int sspec[8];
    double g[8];      // S per channel
    double erev[8];   // V
};

// The open-channel count is the time average over the last span seconds
// (integral / span, Tri::computeI) when intg is given, else the present count.
// GHK charges: current g of EField triangle t at charge[g * charge_stride + tri_map[t]].
// oc_erev: [oc.n][n_tris] reversal potential per triangle, or null for oc.erev.
__global__ void k_vertex_current(int n, const int *vt_ptr, const int *vt_tri, const int *tri_v,
                                 const double *V, const double *tri_clamp, const double *vert_clamp,
                                 const int *tri_counts, const double *intg, double span,
                                 const int *charge, int n_ghk, int charge_stride, const int *tri_map,
                                 int n_tris_sim, OhmicK oc, const double *oc_erev, int n_tris, double *I) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double cur = vert_clamp[i];
    for (int e = vt_ptr[i]; e < vt_ptr[i + 1]; ++e) {
        const int t = vt_tri[e];
        double c = tri_clamp[t];
        // GHK charge moved out over the span, as a mean current (Tri::computeI)
        if (charge != nullptr) {
            int q = 0;
            for (int g = 0; g < n_ghk; ++g) q += charge[(size_t)g * charge_stride + tri_map[t]];
            c -= (double)q * 1.602176487e-19 / span * 1e12;
        }
        if (oc.n > 0 && tri_counts != nullptr) {
            const double vt = (V[tri_v[3 * t]] + V[tri_v[3 * t + 1]] + V[tri_v[3 * t + 2]]) / 3.0;
            double im = 0.0;
            for (int k = 0; k < oc.n; ++k) {
                const size_t ix = (size_t)oc.sspec[k] * n_tris_sim + tri_map[t];
                const double nopen = intg ? intg[ix] / span : (double)tri_counts[ix];
                const double erev = oc_erev ? oc_erev[(size_t)k * n_tris + t] : oc.erev[k];
                im += nopen * oc.g[k] * (vt * 1e-3 - erev);
            }
            c -= im * 1e12;
        }
        cur += c / 3.0;
    }
    I[i] = cur;
}

// b = I + G (Vext - V) + sum_j c_ij (V_j - V_i); diagonal C/dt + G + sigma*wsum
__global__ void k_assemble(EfDev m, double sigma, double oodt, const double *cap, const double *gext,
                           double vext, const double *V, const double *I, double *b, double *d64,
                           float *d32, float *m32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n) return;
    if (m.clamped[i]) {
        b[i] = 0.0;
        d64[i] = 1.0;
        d32[i] = 1.0f;
        m32[i] = 1.0f;
        return;
    }
    const double vi = V[i];
    double rhs = I[i] + gext[i] * (vext - vi);
    double wcl = 0.0;  // couplings to clamped vertices: a diagonal term only
    for (int k = 0; k < m.width; ++k) {
        const int e = k * m.n + i;
        const int j = m.col[e];
        rhs += sigma * m.w[e] * (V[j] - vi);
        if (m.clamped[j]) wcl += m.w[e];
    }
    b[i] = rhs;
    const double d = cap[i] * oodt + gext[i] + sigma * m.wsum[i];
    d64[i] = d;
    d32[i] = (float)d;
    m32[i] = (float)(cap[i] * oodt + gext[i] + sigma * wcl);
}

__global__ void k_offdiag32(EfDev m, double sigma, float *a32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n) return;
    for (int k = 0; k < m.width; ++k) {
        const int e = k * m.n + i;
        const bool off = m.clamped[i] || m.clamped[m.col[e]];
        a32[e] = off ? 0.0f : (float)(-sigma * m.w[e]);
    }
}

// (A x)_i of the fine matrix for x given element-wise by xv(j), in slot order.
// The slots are taken kEllChunk at a time: their column and coupling loads are
// issued together, then their gathers, then the products are summed in slot
// order, so a row costs about one dependent round trip per chunk instead of
// one per slot, and every sum is the one the plain slot loop forms.
constexpr int kEllChunk = 8;

template <typename T, typename F>
__device__ __forceinline__ double ell_row(const EfDev &m, double sigma, int i, F xv) {
    constexpr int U = kEllChunk;
    if constexpr (sizeof(T) == 4) {
        const float xi = (float)xv(i);
        float acc = m.m32[i] * xi;
        for (int k0 = 0; k0 < m.width; k0 += U) {
            int c[U];
            float a[U], v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool on = k0 + u < m.width;
                const int e = (k0 + u) * m.n + i;
                c[u] = on ? m.col[e] : i;
                a[u] = on ? m.a32[e] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = k0 + u < m.width ? (float)xv(c[u]) : xi;
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (k0 + u < m.width) acc += a[u] * (v[u] - xi);
        }
        return acc;
    } else {
        double acc = m.d64[i] * (double)xv(i);
        if (!m.clamped[i])
            for (int k0 = 0; k0 < m.width; k0 += U) {
                int c[U];
                double w[U], v[U];
                bool use[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const bool on = k0 + u < m.width;
                    const int e = (k0 + u) * m.n + i;
                    c[u] = on ? m.col[e] : i;
                    w[u] = on ? m.w[e] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const bool on = k0 + u < m.width;
                    use[u] = on && !m.clamped[c[u]];
                    v[u] = on ? (double)xv(c[u]) : 0.0;
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (use[u]) acc -= sigma * w[u] * v[u];
            }
        return acc;
    }
}

// r = b - A x in FP64, and ||r||^2, ||b||^2
__global__ void k_residual64(EfDev m, double sigma, const double *b, const double *x, double *r,
                             double *part_r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double rr = 0.0;
    if (i < m.n) {
        const double ax = ell_row<double>(m, sigma, i, [&](int j) { return x[j]; });
        const double ri = b[i] - ax;
        r[i] = ri;
        rr = ri * ri;
    }
    block_partial(rr, part_r);
}

// PCG iteration, kernel A: p_new = z + beta p_old; q = A p_new; partial (p_new, q).
// Neighbour values of p_new are formed on the fly from z and p_old, so p_old
// is only read and p_new only written (ping-pong buffers).
template <typename T>
__global__ void k_cg_a(EfDev m, double sigma, const T *z, const T *p, T *pn, T *q, const double *scal,
                       int first, double *part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double pq = 0.0;
    if (i < m.n) {
        // the first iteration never reads p_old (it is uninitialised memory)
        const double beta = first ? 0.0 : scal[1];
        auto pv = [&](int j) { return first ? z[j] : (T)((double)z[j] + beta * (double)p[j]); };
        const T pi = pv(i);
        const T qi = (T)ell_row<T>(m, sigma, i, pv);
        q[i] = qi;
        pn[i] = pi;
        pq = (double)pi * (double)qi;
    }
    block_partial(pq, part);
}

// kernel B: alpha = rz / pq; x += alpha p; r -= alpha q; z = r / diag; partial (r, z)
template <typename T>
__global__ void k_cg_b(EfDev m, const double *scal, T *x, T *r, T *z, const T *pnew, const T *q,
                       double *part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double rz = 0.0;
    if (i < m.n) {
        const double alpha = scal[0] / scal[3];
        const double xi = (double)x[i] + alpha * (double)pnew[i];
        const double ri = (double)r[i] - alpha * (double)q[i];
        const double dg = sizeof(T) == 4 ? (double)m.d32[i] : m.d64[i];
        const T zi = (T)(ri / dg);
        x[i] = (T)xi;
        r[i] = (T)ri;
        z[i] = zi;
        rz = (double)r[i] * (double)zi;
    }
    block_partial(rz, part);
}

// initial: r = residual (as T), z = r / diag, partial (r, z)
template <typename T>
__global__ void k_cg_init(EfDev m, const double *r64, T *x, T *r, T *z, double *part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double rz = 0.0;
    if (i < m.n) {
        const T ri = (T)r64[i];
        const double dg = sizeof(T) == 4 ? (double)m.d32[i] : m.d64[i];
        const T zi = (T)((double)ri / dg);
        x[i] = (T)0;
        r[i] = ri;
        z[i] = zi;
        rz = (double)ri * (double)zi;
    }
    block_partial(rz, part);
}

// ------------------------------------------------------------- multigrid
//
// One V-cycle z = M r of smoothed-aggregation AMG (Vanek, Mandel & Brezina
// 1996). Level 0 is the ELL matrix above; coarser levels and the transfer
// operators are CSR. Smoothing is damped Jacobi from a zero guess, one sweep
// down and one up, which keeps M symmetric for CG.

template <typename T>
__device__ __forceinline__ T fine_diag(const EfDev &m, int i) {
    if constexpr (sizeof(T) == 4) return m.d32[i];
    else return m.d64[i];
}

// down sweep on level 0: x = om D^-1 b, r = b - A x (x of the neighbours formed on the fly)
template <typename T>
__global__ void k_amg0_down(EfDev m, double sigma, double om, const T *b, T *x, T *r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n) return;
    auto xv = [&](int j) { return (T)(om * (double)b[j] / (double)fine_diag<T>(m, j)); };
    const T xi = xv(i);
    x[i] = xi;
    r[i] = (T)((double)b[i] - ell_row<T>(m, sigma, i, xv));
}

// up sweep on level 0: out = x + om D^-1 (b - A x); with part given, also the
// block partials of (b, out), which CG needs next (the preconditioned rz)
template <typename T>
__global__ void k_amg0_up(EfDev m, double sigma, double om, const T *b, const T *x, T *out, double *part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double bz = 0.0;
    if (i < m.n) {
        auto xv = [&](int j) { return x[j]; };
        const double ax = ell_row<T>(m, sigma, i, xv);
        const T zi = (T)((double)x[i] + om * ((double)b[i] - ax) / (double)fine_diag<T>(m, i));
        out[i] = zi;
        bz = (double)b[i] * (double)zi;
    }
    if (part) block_partial(bz, part);
}

// CSR kernels give each row a group of G lanes (G a power of two sized to the
// operator's mean row length): the group reads the row with stride G, so the
// loads of a wavefront are coalesced, and sums it with a fixed shuffle tree,
// so results do not depend on timing. Each lane takes kCsrChunk of its entries
// at a time (loads first, then gathers, then the sum in entry order), so a
// lane waits on one dependent round trip per chunk rather than per entry.
constexpr int kCsrChunk = 4;

// sum over the lane's entries e of f(val[e], x[col[e]]), in entry order
template <typename T, int G, typename F>
__device__ __forceinline__ double csr_lane_sum(const int *ptr, const int *col, const T *val, const T *x, int row,
                                               int lane, F f) {
    constexpr int U = kCsrChunk;
    const int end = ptr[row + 1];
    double acc = 0.0;
    for (int e0 = ptr[row] + lane; e0 < end; e0 += U * G) {
        int c[U];
        T a[U], v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + u * G;
            c[u] = e < end ? col[e] : row;
            a[u] = e < end ? val[e] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = e0 + u * G < end ? x[c[u]] : (T)0;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (e0 + u * G < end) acc += f(a[u], v[u]);
    }
    return acc;
}

template <typename T, int G>
__device__ __forceinline__ double csr_row_dot(const int *ptr, const int *col, const T *val, const T *x, int row,
                                              bool valid, int lane) {
    double acc = 0.0;
    if (valid)
        acc = csr_lane_sum<T, G>(ptr, col, val, x, row, lane,
                                 [](T a, T v) { return (double)a * (double)v; });
    for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
    return acc;
}

// (A x)_row of a level operator. FP32 levels use the difference form with the
// row sums rs of the FP64 Galerkin operator (the diagonal entry adds nothing);
// FP64 levels the plain product.
template <typename T, int G>
__device__ __forceinline__ double csr_row_apply(const int *ptr, const int *col, const T *val, const float *rs,
                                                const T *x, int row, bool valid, int lane) {
    if constexpr (sizeof(T) == 4) {
        double acc = 0.0;
        if (valid) {
            const float xr = (float)x[row];
            acc = csr_lane_sum<T, G>(ptr, col, val, x, row, lane,
                                     [xr](T a, T v) { return (double)a * (double)((float)v - xr); });
        }
        for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
        return valid ? acc + (double)rs[row] * (double)x[row] : 0.0;
    } else {
        return csr_row_dot<T, G>(ptr, col, val, x, row, valid, lane);
    }
}

// r = b - A x on a CSR level (diagonal stored in the rows)
template <typename T, int G>
__global__ void k_csr_resid(int n, const int *ptr, const int *col, const T *val, const float *rs, const T *b,
                            const T *x, T *r) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = t / G, lane = t & (G - 1);
    const bool valid = row < n;
    const double ax = csr_row_apply<T, G>(ptr, col, val, rs, x, row, valid, lane);
    if (valid && lane == 0) r[row] = (T)((double)b[row] - ax);
}

// out = x + om D^-1 (b - A x) on a CSR level (post-smoothing)
template <typename T, int G>
__global__ void k_csr_up(int n, const int *ptr, const int *col, const T *val, const float *rs, const T *dinv,
                         double om, const T *b, const T *x, T *out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = t / G, lane = t & (G - 1);
    const bool valid = row < n;
    const double ax = csr_row_apply<T, G>(ptr, col, val, rs, x, row, valid, lane);
    if (valid && lane == 0) out[row] = (T)((double)x[row] + om * ((double)b[row] - ax) * (double)dinv[row]);
}

// out (+)= M in for a CSR operator (restriction R = P^T, or prolongation P).
// With xs given, a restriction also starts the coarse level's down sweep:
// xs = om D^-1 out (damped Jacobi from a zero guess).
template <typename T, int G>
__global__ void k_csr_apply(int n, const int *ptr, const int *col, const T *val, const T *in, T *out, int add,
                            const T *dinv, double om, T *xs) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = t / G, lane = t & (G - 1);
    const bool valid = row < n;
    const double acc = csr_row_dot<T, G>(ptr, col, val, in, row, valid, lane);
    if (valid && lane == 0) {
        const T o = (T)((add ? (double)out[row] : 0.0) + acc);
        out[row] = o;
        if (xs) xs[row] = (T)(om * (double)o * (double)dinv[row]);
    }
}

// Coarsest level: x_c = A_c^-1 b_c with the dense inverse, 16 lanes per row
// (coalesced reads of the row, fixed shuffle tree)
template <typename T>
__global__ void k_dense_apply(int nc, const T *inv, const T *b, T *xc) {
    constexpr int H = 16;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / H, hl = t & (H - 1);
    double acc = 0.0;
    if (i < nc)
        for (int j = hl; j < nc; j += H) acc += (double)inv[(size_t)i * nc + j] * (double)b[j];
    for (int off = H / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, H);
    if (i < nc && hl == 0) xc[i] = (T)acc;
}

// PCG pieces around the multigrid preconditioner
template <typename T>
__global__ void k_cg_update(int n, const double *scal, T *x, T *r, const T *pnew, const T *q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double alpha = scal[0] / scal[3];
    x[i] = (T)((double)x[i] + alpha * (double)pnew[i]);
    r[i] = (T)((double)r[i] - alpha * (double)q[i]);
}

template <typename T>
__global__ void k_cg_start(int n, const double *r64, T *x, T *r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = (T)0;
    r[i] = (T)r64[i];
}


// Sum over the block of one double per thread; every thread gets the total.
__device__ __forceinline__ double block_allsum(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < (int)(blockDim.x + 63) / 64; ++i) s += red[i];
        red[16] = s;
    }
    __syncthreads();
    return red[16];
}

// Whole Jacobi-PCG solve of a small system in one workgroup: no launches or
// host round trips per iteration. Adds the solution to x64; writes the
// iteration count. The loop condition depends only on block-wide sums, so
// every thread leaves the loop together.
template <typename T>
__global__ void __launch_bounds__(1024) k_cg_small(EfDev m, double sigma, const double *r64, T *x, T *r,
                                                   T *z, T *p, T *q, double rel_tol, int max_iter,
                                                   double *x64, int *iters) {
    __shared__ double red[17];
    const int n = m.n, B = blockDim.x, tid = threadIdx.x;
    const bool f32 = sizeof(T) == 4;
    double loc = 0.0;
    for (int i = tid; i < n; i += B) {
        const T ri = (T)r64[i];
        const T zi = (T)((double)ri / (f32 ? (double)m.d32[i] : m.d64[i]));
        x[i] = (T)0;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        loc += (double)ri * (double)zi;
    }
    double rz = block_allsum(loc, red);
    const double stop = rel_tol * rel_tol * rz;
    int it = 0;
    while (it < max_iter && rz > stop && rz > 0.0) {
        loc = 0.0;
        for (int i = tid; i < n; i += B) {
            double qi;
            if (f32) {
                const float pi = (float)p[i];
                float acc = m.m32[i] * pi;
                for (int k = 0; k < m.width; ++k) {
                    const int e = k * n + i;
                    acc += m.a32[e] * ((float)p[m.col[e]] - pi);
                }
                qi = acc;
            } else {
                double acc = m.d64[i] * (double)p[i];
                if (!m.clamped[i])
                    for (int k = 0; k < m.width; ++k) {
                        const int e = k * n + i;
                        if (!m.clamped[m.col[e]]) acc -= sigma * m.w[e] * (double)p[m.col[e]];
                    }
                qi = acc;
            }
            q[i] = (T)qi;
            loc += (double)p[i] * (double)(T)qi;
        }
        const double pq = block_allsum(loc, red);
        const double alpha = rz / pq;
        loc = 0.0;
        for (int i = tid; i < n; i += B) {
            const double ri = (double)r[i] - alpha * (double)q[i];
            x[i] = (T)((double)x[i] + alpha * (double)p[i]);
            r[i] = (T)ri;
            const T zi = (T)((double)(T)ri / (f32 ? (double)m.d32[i] : m.d64[i]));
            z[i] = zi;
            loc += (double)(T)ri * (double)zi;
        }
        const double rzn = block_allsum(loc, red);
        const double beta = rzn / rz;
        rz = rzn;
        for (int i = tid; i < n; i += B) p[i] = (T)((double)z[i] + beta * (double)p[i]);
        __syncthreads();
        ++it;
    }
    for (int i = tid; i < n; i += B) x64[i] += (double)x[i];
    if (tid == 0) *iters = it;
}

template <typename T>
__global__ void k_accumulate(int n, double *x64, const T *d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x64[i] += (double)d[i];
}

// potential of each membrane triangle (mean of its vertices, volts) into the
// attached simulation's slot for that triangle
__global__ void k_tri_v(int nt, const int *tri_v, const double *V, const int *tri_map, double *out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nt) out[tri_map[t]] = (V[tri_v[3 * t]] + V[tri_v[3 * t + 1]] + V[tri_v[3 * t + 2]]) / 3.0 * 1e-3;
}

// Move the GHK charges of each current into last ([n_ghk][nt], when given) and
// clear them for the next step (Tri::resetECharge, tri.cpp:449-468).
__global__ void k_take_charge(int nt, const int *tri_map, int n_ghk, int stride, int *charge, int *last) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    for (int g = 0; g < n_ghk; ++g) {
        int *c = charge + (size_t)g * stride + tri_map[t];
        if (last) last[(size_t)g * nt + t] = *c;
        *c = 0;
    }
}

// restart the time integrals of the channel species on this EField's triangles
__global__ void k_zero_intg(int nt, const int *tri_map, int n_tris_sim, uint32_t mask, double *intg) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    for (int q = 0; q < 32; ++q)
        if ((mask >> q) & 1u) intg[(size_t)q * n_tris_sim + tri_map[t]] = 0.0;
}

__global__ void k_apply_dv(int n, const unsigned char *clamped, const double *dv, double *V) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !clamped[i]) V[i] += dv[i];
}

__global__ void k_sumsq(int n, const double *v, double *part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double x = i < n ? v[i] : 0.0;
    block_partial(x * x, part);
}

}  // namespace

// Multigrid hierarchy on the device. Level l < nlev-1 holds the operator of
// that level (level 0: the fine ELL matrix, not stored here), its prolongation
// P_l to level l+1 and restriction R_l = P_l^T; the last level is dense.
struct AmgLevel {
    int n = 0;
    double om = 0.0;                                    // Jacobi weight 4 / (3 rho(D^-1 A))
    int ga = 1, gp = 1, gr = 1;                         // lanes per row of A_l, P_l, R_l
    int *a_ptr = nullptr, *a_col = nullptr;             // A_l (l >= 1), diagonal included
    float *a32 = nullptr, *dinv32 = nullptr, *rs32 = nullptr;  // rs32: row sums of A_l
    double *a64 = nullptr, *dinv64 = nullptr;
    int *p_ptr = nullptr, *p_col = nullptr;             // P_l: n rows
    float *p32 = nullptr;
    double *p64 = nullptr;
    int *r_ptr = nullptr, *r_col = nullptr;             // R_l: n_{l+1} rows
    float *r32 = nullptr;
    double *r64 = nullptr;
    double *b = nullptr, *x = nullptr, *r = nullptr, *out = nullptr;  // work vectors, FP64-sized
};

struct AmgDev {
    bool built = false, usable = false;
    double key_oodt = -1.0, key_sigma = -1.0;
    long key_ver = -1;
    std::vector<AmgLevel> lev;   // all but the coarsest
    int nc = 0;                  // coarsest size
    float *inv32 = nullptr;
    double *inv64 = nullptr;
    double *cb = nullptr, *cx = nullptr;
    double setup_s = 0.0;
    // two CG iterations (V-cycles included) captured once per hierarchy and
    // precision: the launch sequence holds this hierarchy's pointers and weights
    hipGraphExec_t pair32 = nullptr, pair64 = nullptr;

    void release() {
        if (pair32) (void)hipGraphExecDestroy(pair32);
        if (pair64) (void)hipGraphExecDestroy(pair64);
        pair32 = pair64 = nullptr;
        for (AmgLevel &L : lev) {
            void *ptrs[] = {L.a_ptr, L.a_col, L.a32, L.dinv32, L.rs32, L.a64, L.dinv64, L.p_ptr, L.p_col, L.p32,
                            L.p64, L.r_ptr, L.r_col, L.r32, L.r64, L.b, L.x, L.r, L.out};
            for (void *q : ptrs)
                if (q) (void)hipFree(q);
        }
        lev.clear();
        void *ptrs[] = {inv32, inv64, cb, cx};
        for (void *q : ptrs)
            if (q) (void)hipFree(q);
        inv32 = nullptr;
        inv64 = nullptr;
        cb = cx = nullptr;
        built = usable = false;
        nc = 0;
    }
};

constexpr int kEfMaxGhk = 4;  // GHK currents a simulation can carry

struct stg_efield {
    int n = 0, n_tris = 0, nnz = 0, width = 0;
    std::vector<int> tri_v;                 // [n_tris*3]
    std::vector<double> vert_area, tri_area, cap_h, gext_h;
    std::vector<double> vert_clamp_h, tri_clamp_h;
    std::vector<double> tri_cm_h;           // [n_tris] specific capacitance of each triangle
    std::vector<double> oc_erev_h;          // [8][n_tris] per-triangle reversal potentials, empty = oc.erev
    std::vector<unsigned char> clamped_h;
    double sigma = 1.0 / 1.0e-3;            // nS/um, 1 ohm.m
    double vext = 0.0;
    double cm = 0.01;
    int precision = 0;                      // 0 = FP32 CG + FP64 refinement, 1 = FP64 CG
    double tol = 1e-8;                      // relative FP64 residual target
    double floor_tol = 1e-9;                // accepted when refinement stagnates above tol
    double inner_tol = 1e-6;                // relative preconditioned residual per FP32 solve
    int inner_max = 4000;                   // FP32 iterations per refinement
    int max_iter = 20000;
    int last_iters = 0, last_refine = 0;
    double last_rel = 0.0;
    int precond = 1;                        // 1 = multigrid above 16,384 vertices, 0 = Jacobi
    long params_ver = 0;                    // bumped whenever the matrix parameters change
    std::vector<int> col_h;                 // host copy of the coupling table ([width][n])
    std::vector<double> w_h, wsum_h;
    AmgDev amg;
    OhmicK oc{};
    stg_sim *sim = nullptr;
    int sim_real = 0;
    double *intg = nullptr;                 // attached simulation's channel-count integrals
    uint32_t intg_mask = 0;
    int32_t *charge = nullptr;              // attached simulation's GHK charges (this realization)
    int32_t n_ghk = 0, charge_stride = 0;   // GHK currents; distance between their charge rows
    int32_t *ghk_last = nullptr;            // [n_ghk][n_tris] charges of the last step
    double ghk_dt = 0.0;                    // its length in seconds, 0 before the first step
    double intg_t0 = 0.0;                   // simulation time the integrals started from
    std::string err;
    hipStream_t stream = nullptr;
    // device
    int *col = nullptr, *vt_ptr = nullptr, *vt_tri = nullptr, *tri_vd = nullptr;
    int *tri_map = nullptr;  // EField triangle -> the attached simulation's triangle slot
    double *w = nullptr, *wsum = nullptr, *d64 = nullptr, *V = nullptr, *I = nullptr, *b = nullptr;
    double *cap = nullptr, *gext = nullptr, *vclamp = nullptr, *tclamp = nullptr, *oc_erev = nullptr;
    double *x64 = nullptr, *r64 = nullptr, *dv_prev = nullptr;
    float *a32 = nullptr, *d32 = nullptr, *m32 = nullptr;
    unsigned char *clamped = nullptr;
    void *xs = nullptr, *rs = nullptr, *zs = nullptr, *ps = nullptr, *qs = nullptr;  // T-typed work
    double *part = nullptr, *scal = nullptr, *outs = nullptr;
    double *h_fetch = nullptr;  // pinned host words the convergence checks read back
    unsigned *cnt = nullptr;
    bool a32_dirty = true;
    int nb = 1;

    EfDev dev() const {
        return EfDev{n, width, col, w, wsum, a32, d32, m32, d64, clamped};
    }

    ~stg_efield() {
        amg.release();
        void *ptrs[] = {col, vt_ptr, vt_tri, tri_vd, tri_map, w, wsum, d64, V, I, b, cap, gext, vclamp,
                        tclamp, x64, r64, dv_prev, a32, d32, m32, clamped, xs, rs, zs, ps, qs, part,
                        scal, outs, cnt, oc_erev, ghk_last};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        if (h_fetch) (void)hipHostFree(h_fetch);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// stg_sim accessors defined in stg_engine.hip
extern "C" int stg_internal_tri_counts_device(stg_sim *sim, int32_t real, const int32_t **ptr,
                                              int32_t *n_tris_total, int32_t *tri0);
extern "C" int stg_internal_tri_perm(const stg_sim *sim, int32_t real, const int32_t **perm,
                                     int32_t *n);
extern "C" int stg_internal_tri_v_device(stg_sim *sim, int32_t real, double **ptr);
extern "C" int stg_internal_tri_intg_device(stg_sim *sim, uint32_t mask, double **ptr);
extern "C" int stg_internal_tri_charge_device(stg_sim *sim, int32_t real, int32_t **ptr, int32_t *n_ghk,
                                              int32_t *stride);

namespace {

thread_local std::string g_ef_error;

template <typename F>
int ef_guard(stg_efield *e, F &&f) {
    try {
        f();
        return STG_OK;
    } catch (const EfError &x) {
        if (e) e->err = x.msg;
        else g_ef_error = x.msg;
        return x.code;
    } catch (const std::bad_alloc &) {
        if (e) e->err = "host allocation failed";
        else g_ef_error = "host allocation failed";
        return STG_ERR_SYS;
    } catch (const std::exception &x) {
        if (e) e->err = x.what();
        else g_ef_error = x.what();
        return STG_ERR_PROG;
    }
}

void upload_params(stg_efield *e) {
    ++e->params_ver;
    ef_put(e->cap, e->cap_h);
    ef_put(e->gext, e->gext_h);
    ef_put(e->vclamp, e->vert_clamp_h);
    ef_put(e->tclamp, e->tri_clamp_h);
    ef_put(e->clamped, e->clamped_h);
    if (!e->oc_erev_h.empty()) {
        if (!e->oc_erev) e->oc_erev = ef_alloc<double>(e->oc_erev_h.size());
        ef_put(e->oc_erev, e->oc_erev_h);
    }
    e->a32_dirty = true;
}

void build(stg_efield *e, const stg_efield_mesh *m) {
    if (!m || m->n_verts <= 0 || m->n_tets <= 0 || !m->verts || !m->tets)
        throw EfError{STG_ERR_ARG, "empty or null EField mesh"};
    if (m->n_tris < 0 || (m->n_tris > 0 && !m->tris)) throw EfError{STG_ERR_ARG, "bad membrane triangles"};
    const int n = m->n_verts;
    e->n = n;
    e->n_tris = m->n_tris;
    // P1 stiffness: w_ij = -sum_tets V grad(phi_i).grad(phi_j), coordinates in microns.
    // Rows are short (a vertex touches a few tens of others): linear search.
    std::vector<std::vector<std::pair<int, double>>> rows(n);
    auto add = [&rows](int i, int j, double v) {
        auto &row = rows[i];
        for (auto &kv : row)
            if (kv.first == j) {
                kv.second += v;
                return;
            }
        row.emplace_back(j, v);
    };
    for (int t = 0; t < m->n_tets; ++t) {
        int v[4];
        double P[4][3];
        for (int a = 0; a < 4; ++a) {
            v[a] = m->tets[4 * t + a];
            if (v[a] < 0 || v[a] >= n) throw EfError{STG_ERR_ARG, "tet vertex index out of range"};
            for (int k = 0; k < 3; ++k) P[a][k] = m->verts[3 * v[a] + k] * 1e6;
        }
        double E[3][3];
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) E[a][k] = P[a + 1][k] - P[0][k];
        const double det = E[0][0] * (E[1][1] * E[2][2] - E[1][2] * E[2][1]) -
                           E[0][1] * (E[1][0] * E[2][2] - E[1][2] * E[2][0]) +
                           E[0][2] * (E[1][0] * E[2][1] - E[1][1] * E[2][0]);
        if (!(std::fabs(det) > 0.0)) throw EfError{STG_ERR_ARG, "degenerate tetrahedron"};
        // inverse of E; gradients of phi_1..3 are the columns of inv(E)
        double inv[3][3];
        inv[0][0] = (E[1][1] * E[2][2] - E[1][2] * E[2][1]) / det;
        inv[0][1] = (E[0][2] * E[2][1] - E[0][1] * E[2][2]) / det;
        inv[0][2] = (E[0][1] * E[1][2] - E[0][2] * E[1][1]) / det;
        inv[1][0] = (E[1][2] * E[2][0] - E[1][0] * E[2][2]) / det;
        inv[1][1] = (E[0][0] * E[2][2] - E[0][2] * E[2][0]) / det;
        inv[1][2] = (E[0][2] * E[1][0] - E[0][0] * E[1][2]) / det;
        inv[2][0] = (E[1][0] * E[2][1] - E[1][1] * E[2][0]) / det;
        inv[2][1] = (E[0][1] * E[2][0] - E[0][0] * E[2][1]) / det;
        inv[2][2] = (E[0][0] * E[1][1] - E[0][1] * E[1][0]) / det;
        double g[4][3];
        for (int k = 0; k < 3; ++k) {
            for (int a = 1; a < 4; ++a) g[a][k] = inv[k][a - 1];
            g[0][k] = -(g[1][k] + g[2][k] + g[3][k]);
        }
        const double vol = std::fabs(det) / 6.0;
        for (int a = 0; a < 4; ++a)
            for (int c = 0; c < 4; ++c) {
                if (a == c) continue;
                const double kij = vol * (g[a][0] * g[c][0] + g[a][1] * g[c][1] + g[a][2] * g[c][2]);
                add(v[a], v[c], -kij);
            }
    }
    // Couplings that cancel to exactly zero over their tets carry nothing and
    // are dropped (removing an exact zero term leaves every sum bit-identical).
    // On Kuhn (Freudenthal) meshes the P1 stiffness is the 7-point stencil:
    // 8 of an interior vertex's 14 edges cancel, and the table width falls from
    // 14 to 6, which every fine-level pass reads.
    size_t width = 1, nnz = 0;
    for (int i = 0; i < n; ++i) {
        auto &r = rows[i];
        r.erase(std::remove_if(r.begin(), r.end(), [](const std::pair<int, double> &kv) { return kv.second == 0.0; }),
                r.end());
        width = std::max(width, rows[i].size());
        nnz += rows[i].size();
    }
    std::vector<int> col(width * (size_t)n);
    std::vector<double> w(width * (size_t)n, 0.0), wsum(n, 0.0);
    for (int i = 0; i < n; ++i) {
        std::sort(rows[i].begin(), rows[i].end());
        for (size_t k = 0; k < width; ++k) {
            const bool have = k < rows[i].size();
            col[k * n + i] = have ? rows[i][k].first : i;
            w[k * n + i] = have ? rows[i][k].second : 0.0;
            if (have) wsum[i] += rows[i][k].second;
        }
        std::vector<std::pair<int, double>>().swap(rows[i]);
    }
    e->nnz = (int)nnz;
    e->width = (int)width;
    // membrane: triangle areas, one third to each vertex (tetmesh.cpp:231-260)
    e->tri_v.assign(m->tris, m->tris + 3 * (size_t)m->n_tris);
    e->tri_area.resize(m->n_tris);
    e->vert_area.assign(n, 0.0);
    std::vector<std::vector<int>> vt(n);
    for (int t = 0; t < m->n_tris; ++t) {
        double P[3][3];
        for (int a = 0; a < 3; ++a) {
            const int vi = e->tri_v[3 * t + a];
            if (vi < 0 || vi >= n) throw EfError{STG_ERR_ARG, "triangle vertex index out of range"};
            for (int k = 0; k < 3; ++k) P[a][k] = m->verts[3 * vi + k] * 1e6;
            vt[vi].push_back(t);
        }
        double u[3], v[3];
        for (int k = 0; k < 3; ++k) {
            u[k] = P[1][k] - P[0][k];
            v[k] = P[2][k] - P[0][k];
        }
        const double cx = u[1] * v[2] - u[2] * v[1], cy = u[2] * v[0] - u[0] * v[2],
                     cz = u[0] * v[1] - u[1] * v[0];
        const double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
        e->tri_area[t] = area;
        for (int a = 0; a < 3; ++a) e->vert_area[e->tri_v[3 * t + a]] += area / 3.0;
    }
    std::vector<int> vt_ptr(n + 1, 0), vt_tri;
    for (int i = 0; i < n; ++i) {
        for (int t : vt[i]) vt_tri.push_back(t);
        vt_ptr[i + 1] = (int)vt_tri.size();
    }
    e->cap_h.resize(n);
    for (int i = 0; i < n; ++i) e->cap_h[i] = e->cm * e->vert_area[i];
    e->tri_cm_h.assign(m->n_tris, e->cm);
    e->gext_h.assign(n, 0.0);
    e->vert_clamp_h.assign(n, 0.0);
    e->tri_clamp_h.assign(std::max(m->n_tris, 1), 0.0);
    e->clamped_h.assign(n, 0);

    EF_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    e->nb = (n + kB - 1) / kB;
    e->col = ef_alloc<int>(col.size());
    e->w = ef_alloc<double>(w.size());
    e->wsum = ef_alloc<double>(n);
    e->a32 = ef_alloc<float>(w.size());
    e->d32 = ef_alloc<float>(n);
    e->m32 = ef_alloc<float>(n);
    e->d64 = ef_alloc<double>(n);
    e->V = ef_alloc<double>(n);
    e->I = ef_alloc<double>(n);
    e->b = ef_alloc<double>(n);
    e->cap = ef_alloc<double>(n);
    e->gext = ef_alloc<double>(n);
    e->vclamp = ef_alloc<double>(n);
    e->tclamp = ef_alloc<double>(std::max(m->n_tris, 1));
    e->clamped = ef_alloc<unsigned char>(n);
    e->x64 = ef_alloc<double>(n);
    e->r64 = ef_alloc<double>(n);
    e->dv_prev = ef_alloc<double>(n);
    e->xs = ef_alloc<double>(n);
    e->rs = ef_alloc<double>(n);
    e->zs = ef_alloc<double>(n);
    e->ps = ef_alloc<double>(2 * (size_t)n);
    e->qs = ef_alloc<double>(n);
    e->part = ef_alloc<double>(e->nb);
    e->scal = ef_alloc<double>(4);
    e->outs = ef_alloc<double>(8);
    EF_HIP(hipHostMalloc((void **)&e->h_fetch, 2 * sizeof(double), hipHostMallocDefault));
    e->cnt = ef_alloc<unsigned>(8);
    e->vt_ptr = ef_alloc<int>(n + 1);
    e->vt_tri = ef_alloc<int>(std::max<size_t>(vt_tri.size(), 1));
    e->tri_vd = ef_alloc<int>(std::max<size_t>(e->tri_v.size(), 1));
    ef_put(e->col, col);
    ef_put(e->w, w);
    ef_put(e->wsum, wsum);
    e->col_h = std::move(col);
    e->w_h = std::move(w);
    e->wsum_h = std::move(wsum);
    ef_put(e->vt_ptr, vt_ptr);
    ef_put(e->vt_tri, vt_tri);
    ef_put(e->tri_vd, e->tri_v);
    EF_HIP(hipMemset(e->cnt, 0, 8 * sizeof(unsigned)));
    EF_HIP(hipMemset(e->dv_prev, 0, n * sizeof(double)));
    std::vector<double> v0(n, -65.0);
    ef_put(e->V, v0);
    upload_params(e);
}

// Jacobi PCG on A d = r (r given in r64), result added to x64. Returns iterations.
// Device scalars read back by the host's convergence decisions, through
// pinned memory (a copy to pageable memory goes through a staging buffer,
// which adds to every round trip). fetch2 reads two adjacent words at once.
double fetch(stg_efield *e, const double *d) {
    EF_HIP(hipMemcpyAsync(e->h_fetch, d, sizeof(double), hipMemcpyDeviceToHost, e->stream));
    EF_HIP(hipStreamSynchronize(e->stream));
    return e->h_fetch[0];
}
void fetch2(stg_efield *e, const double *d, double &a, double &b) {
    EF_HIP(hipMemcpyAsync(e->h_fetch, d, 2 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    EF_HIP(hipStreamSynchronize(e->stream));
    a = e->h_fetch[0];
    b = e->h_fetch[1];
}

constexpr int kSmallSystem = 16384;  // solved by one persistent workgroup

template <typename T>
int pcg_amg(stg_efield *e, double rel_tol, int max_iter);

template <typename T>
int pcg(stg_efield *e, double rel_tol, int max_iter) {
    const EfDev m = e->dev();
    const int nb = e->nb, n = e->n;
    T *x = (T *)e->xs, *r = (T *)e->rs, *z = (T *)e->zs, *p = (T *)e->ps, *q = (T *)e->qs;
    if (n <= kSmallSystem) {
        hipLaunchKernelGGL(k_cg_small<T>, dim3(1), dim3(1024), 0, e->stream, m, e->sigma, e->r64, x, r, z,
                           p, q, rel_tol, max_iter, e->x64, (int *)e->cnt + 7);
        int it = 0;
        EF_HIP(hipMemcpyAsync(&it, (int *)e->cnt + 7, sizeof(int), hipMemcpyDeviceToHost, e->stream));
        EF_HIP(hipStreamSynchronize(e->stream));
        return it;
    }
    if (e->precond == 1 && e->amg.usable) return pcg_amg<T>(e, rel_tol, max_iter);
    hipLaunchKernelGGL(k_cg_init<T>, dim3(nb), dim3(kB), 0, e->stream, m, e->r64, x, r, z, e->part);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, e->stream, e->part, nb, 0, e->scal, e->outs);
    const double rz0 = fetch(e, e->scal);
    int it = 0;
    if (!(rz0 > 0.0)) return 0;
    const int check = 16;
    while (it < max_iter) {
        for (int k = 0; k < check && it < max_iter; ++k, ++it) {
            T *pold = p + (size_t)(it & 1) * n, *pnew = p + (size_t)((it + 1) & 1) * n;
            hipLaunchKernelGGL(k_cg_a<T>, dim3(nb), dim3(kB), 0, e->stream, m, e->sigma, z, pold, pnew, q,
                               e->scal, it == 0 ? 1 : 0, e->part);
            hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, e->stream, e->part, nb, 1, e->scal, e->outs);
            hipLaunchKernelGGL(k_cg_b<T>, dim3(nb), dim3(kB), 0, e->stream, m, e->scal, x, r, z, pnew, q,
                               e->part);
            hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, e->stream, e->part, nb, 2, e->scal, e->outs);
        }
        const double rz = fetch(e, e->scal);
        if (!(rz > rel_tol * rel_tol * rz0)) break;
    }
    EF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_accumulate<T>, dim3(nb), dim3(kB), 0, e->stream, n, e->x64, x);
    return it;
}

// ------------------------------------------------------------- multigrid setup (host, FP64)

struct Csr {
    int n = 0, m = 0;  // rows, columns
    std::vector<int> ptr, col;
    std::vector<double> val;
};

// The fine matrix of a step with 1/dt = oodt (1/ms), exactly as k_assemble and
// k_offdiag32 form it: diagonal C/dt + G + sigma sum_j w_ij, couplings
// -sigma w_ij between unclamped vertices, identity rows for clamped ones.
Csr fine_matrix(const stg_efield *e, double oodt) {
    const int n = e->n, W = e->width;
    Csr A;
    A.n = A.m = n;
    A.ptr.assign(n + 1, 0);
    A.col.reserve((size_t)n * (W + 1));
    A.val.reserve((size_t)n * (W + 1));
    for (int i = 0; i < n; ++i) {
        if (e->clamped_h[i]) {
            A.col.push_back(i);
            A.val.push_back(1.0);
        } else {
            A.col.push_back(i);
            A.val.push_back(e->cap_h[i] * oodt + e->gext_h[i] + e->sigma * e->wsum_h[i]);
            for (int k = 0; k < W; ++k) {
                const int j = e->col_h[(size_t)k * n + i];
                const double w = e->w_h[(size_t)k * n + i];
                if (j == i || w == 0.0 || e->clamped_h[j]) continue;
                A.col.push_back(j);
                A.val.push_back(-e->sigma * w);
            }
        }
        A.ptr[i + 1] = (int)A.col.size();
    }
    return A;
}

double diag_of(const Csr &A, int i) {
    for (int e = A.ptr[i]; e < A.ptr[i + 1]; ++e)
        if (A.col[e] == i) return A.val[e];
    return 0.0;
}

// Upper estimate of the largest eigenvalue of D^-1 A: the Rayleigh quotient
// (x, A x) / (x, D x) after 20 power steps, with a 10% margin, capped by the
// Gershgorin bound max_i sum_j |a_ij| / a_ii.
double rho_dinv_a(const Csr &A, const std::vector<double> &dinv) {
    const int n = A.n;
    std::vector<double> x(n), y(n);
    uint64_t h = 0x9E3779B97F4A7C15ull;
    for (int i = 0; i < n; ++i) {
        h ^= h << 13;
        h ^= h >> 7;
        h ^= h << 17;
        x[i] = 0.5 + (double)(h >> 11) * 0x1p-53;
    }
    double lam = 0.0;
    for (int it = 0; it < 20; ++it) {
        double num = 0.0, den = 0.0, nrm = 0.0;
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int e = A.ptr[i]; e < A.ptr[i + 1]; ++e) s += A.val[e] * x[A.col[e]];
            num += x[i] * s;
            den += x[i] * x[i] / dinv[i];
            y[i] = dinv[i] * s;
            nrm += y[i] * y[i];
        }
        lam = num / den;
        nrm = std::sqrt(nrm);
        if (!(nrm > 0.0)) break;
        for (int i = 0; i < n; ++i) x[i] = y[i] / nrm;
    }
    double g = 0.0;
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int e = A.ptr[i]; e < A.ptr[i + 1]; ++e) s += std::fabs(A.val[e]);
        g = std::max(g, s * dinv[i]);
    }
    return std::min(1.1 * lam, g);
}

// Greedy aggregation on the strength graph (|a_ij| >= theta sqrt(a_ii a_jj)):
// 1) a vertex whose strong neighbours are all free seeds an aggregate with
// them; 2) a free vertex joins the first-pass aggregate of its strongest
// neighbour; 3) what is left seeds aggregates with its free strong
// neighbours. Vertices without strong neighbours (clamped or membrane rows the
// capacitance dominates) stay outside every aggregate: -1.
std::vector<int> aggregate(const Csr &A, const std::vector<double> &diag, double theta, int &nagg) {
    const int n = A.n;
    std::vector<int> sp(n + 1, 0), sc;
    std::vector<double> sv;
    sc.reserve(A.col.size());
    sv.reserve(A.col.size());
    for (int i = 0; i < n; ++i) {
        for (int e = A.ptr[i]; e < A.ptr[i + 1]; ++e) {
            const int j = A.col[e];
            const double a = A.val[e];
            if (j == i || a == 0.0) continue;
            if (a * a >= theta * theta * std::fabs(diag[i] * diag[j])) {
                sc.push_back(j);
                sv.push_back(std::fabs(a));
            }
        }
        sp[i + 1] = (int)sc.size();
    }
    std::vector<int> agg(n, -1);
    nagg = 0;
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0 || sp[i] == sp[i + 1]) continue;
        bool free = true;
        for (int k = sp[i]; k < sp[i + 1] && free; ++k) free = agg[sc[k]] < 0;
        if (!free) continue;
        agg[i] = nagg;
        for (int k = sp[i]; k < sp[i + 1]; ++k) agg[sc[k]] = nagg;
        ++nagg;
    }
    const std::vector<int> agg1 = agg;
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        int best = -1;
        double bw = 0.0;
        for (int k = sp[i]; k < sp[i + 1]; ++k)
            if (agg1[sc[k]] >= 0 && sv[k] > bw) {
                bw = sv[k];
                best = agg1[sc[k]];
            }
        if (best >= 0) agg[i] = best;
    }
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0 || sp[i] == sp[i + 1]) continue;
        agg[i] = nagg;
        for (int k = sp[i]; k < sp[i + 1]; ++k)
            if (agg[sc[k]] < 0) agg[sc[k]] = nagg;
        ++nagg;
    }
    return agg;
}

// P = (I - om D^-1 A) P_tent with P_tent the aggregate indicator
Csr smoothed_prolongator(const Csr &A, const std::vector<double> &dinv, const std::vector<int> &agg, int nagg,
                         double om) {
    Csr P;
    P.n = A.n;
    P.m = nagg;
    P.ptr.assign(A.n + 1, 0);
    std::vector<std::pair<int, double>> row;
    for (int i = 0; i < A.n; ++i) {
        row.clear();
        auto addv = [&row](int c, double v) {
            for (auto &kv : row)
                if (kv.first == c) {
                    kv.second += v;
                    return;
                }
            row.emplace_back(c, v);
        };
        if (agg[i] >= 0) addv(agg[i], 1.0);
        for (int e = A.ptr[i]; e < A.ptr[i + 1]; ++e) {
            const int j = A.col[e];
            if (agg[j] >= 0) addv(agg[j], -om * dinv[i] * A.val[e]);
        }
        std::sort(row.begin(), row.end());
        for (const auto &kv : row)
            if (kv.second != 0.0) {
                P.col.push_back(kv.first);
                P.val.push_back(kv.second);
            }
        P.ptr[i + 1] = (int)P.col.size();
    }
    return P;
}

Csr transpose(const Csr &P) {
    Csr R;
    R.n = P.m;
    R.m = P.n;
    R.ptr.assign(P.m + 1, 0);
    for (int c : P.col) ++R.ptr[c + 1];
    for (int c = 0; c < P.m; ++c) R.ptr[c + 1] += R.ptr[c];
    R.col.resize(P.col.size());
    R.val.resize(P.val.size());
    std::vector<int> pos(R.ptr.begin(), R.ptr.end() - 1);
    for (int i = 0; i < P.n; ++i)
        for (int e = P.ptr[i]; e < P.ptr[i + 1]; ++e) {
            const int k = pos[P.col[e]]++;
            R.col[k] = i;
            R.val[k] = P.val[e];
        }
    return R;
}

// C = X Y with a dense accumulator per row (columns sorted)
Csr spgemm(const Csr &X, const Csr &Y) {
    Csr C;
    C.n = X.n;
    C.m = Y.m;
    C.ptr.assign(X.n + 1, 0);
    std::vector<double> acc(Y.m, 0.0);
    std::vector<int> mark(Y.m, -1), touched;
    for (int i = 0; i < X.n; ++i) {
        touched.clear();
        for (int e = X.ptr[i]; e < X.ptr[i + 1]; ++e) {
            const int j = X.col[e];
            const double a = X.val[e];
            for (int f = Y.ptr[j]; f < Y.ptr[j + 1]; ++f) {
                const int c = Y.col[f];
                if (mark[c] != i) {
                    mark[c] = i;
                    acc[c] = 0.0;
                    touched.push_back(c);
                }
                acc[c] += a * Y.val[f];
            }
        }
        std::sort(touched.begin(), touched.end());
        for (int c : touched) {
            C.col.push_back(c);
            C.val.push_back(acc[c]);
        }
        C.ptr[i + 1] = (int)C.col.size();
    }
    return C;
}

// Inverse of a small dense matrix by Gauss-Jordan elimination with partial pivoting
std::vector<double> dense_inverse(const Csr &A) {
    const int n = A.n;
    std::vector<double> M((size_t)n * n, 0.0), I((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        for (int e = A.ptr[i]; e < A.ptr[i + 1]; ++e) M[(size_t)i * n + A.col[e]] = A.val[e];
        I[(size_t)i * n + i] = 1.0;
    }
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(M[(size_t)r * n + c]) > std::fabs(M[(size_t)piv * n + c])) piv = r;
        const double pv = M[(size_t)piv * n + c];
        if (!(std::fabs(pv) > 0.0)) throw EfError{STG_ERR_PROG, "singular coarsest multigrid level"};
        if (piv != c)
            for (int k = 0; k < n; ++k) {
                std::swap(M[(size_t)piv * n + k], M[(size_t)c * n + k]);
                std::swap(I[(size_t)piv * n + k], I[(size_t)c * n + k]);
            }
        const double ip = 1.0 / pv;
        for (int k = 0; k < n; ++k) {
            M[(size_t)c * n + k] *= ip;
            I[(size_t)c * n + k] *= ip;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = M[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                M[(size_t)r * n + k] -= f * M[(size_t)c * n + k];
                I[(size_t)r * n + k] -= f * I[(size_t)c * n + k];
            }
        }
    }
    return I;
}

int *up_int(const std::vector<int> &v) {
    int *d = ef_alloc<int>(v.size());
    ef_put(d, v);
    return d;
}

double *up_f64(const std::vector<double> &v) {
    double *d = ef_alloc<double>(v.size());
    ef_put(d, v);
    return d;
}

float *up_f32(const std::vector<double> &v) {
    std::vector<float> f(v.begin(), v.end());
    float *d = ef_alloc<float>(f.size());
    ef_put(d, f);
    return d;
}


// lanes per row: the power of two nearest the mean row length, 1..64
// Lanes per row: about one chunk of kCsrChunk entries per lane. A group as
// wide as the row (one entry per lane) costs more threads and shuffle steps:
// on the 1 M-vertex cube's hierarchy the V-cycle's CSR kernels took 196 us
// per CG iteration that way against 165 us with this choice.
int group_size(const Csr &M) {
    const double mean = M.n > 0 ? (double)M.col.size() / M.n : 1.0;
    int g = 1;
    while (g < 64 && g * 1.5 * kCsrChunk < mean) g *= 2;
    return g;
}

// Build the hierarchy for the matrix of a step with 1/dt = oodt. Leaves
// amg.usable false when coarsening stalls above a size a dense inverse can take.
void amg_build(stg_efield *e, double oodt) {
    const auto t0 = std::chrono::steady_clock::now();
    AmgDev &G = e->amg;
    G.release();
    G.built = true;
    G.key_oodt = oodt;
    G.key_sigma = e->sigma;
    G.key_ver = e->params_ver;
    Csr A = fine_matrix(e, oodt);
    for (int l = 0; l < 12 && A.n > kCoarsest; ++l) {
        std::vector<double> diag(A.n), dinv(A.n);
        for (int i = 0; i < A.n; ++i) {
            diag[i] = diag_of(A, i);
            if (!(diag[i] > 0.0)) throw EfError{STG_ERR_PROG, "multigrid level with a non-positive diagonal"};
            dinv[i] = 1.0 / diag[i];
        }
        const double om = 4.0 / (3.0 * rho_dinv_a(A, dinv));
        int nagg = 0;
        // every coupling counts as strong: on these P1 stiffness matrices a
        // threshold leaves small aggregates and dense coarse operators
        const std::vector<int> agg = aggregate(A, diag, 0.0, nagg);
        if (nagg == 0 || nagg > 0.85 * A.n) break;  // coarsening stalled
        Csr P = smoothed_prolongator(A, dinv, agg, nagg, om);
        Csr R = transpose(P);
        Csr Ac = spgemm(R, spgemm(A, P));
        AmgLevel L;
        L.n = A.n;
        L.om = om;
        L.ga = group_size(A);
        L.gp = group_size(P);
        L.gr = group_size(R);
        if (l > 0) {
            L.a_ptr = up_int(A.ptr);
            L.a_col = up_int(A.col);
            L.a32 = up_f32(A.val);
            L.a64 = up_f64(A.val);
            L.dinv32 = up_f32(dinv);
            L.dinv64 = up_f64(dinv);
            std::vector<double> rsum(A.n, 0.0);
            for (int i = 0; i < A.n; ++i)
                for (int e = A.ptr[i]; e < A.ptr[i + 1]; ++e) rsum[i] += A.val[e];
            L.rs32 = up_f32(rsum);
            L.b = ef_alloc<double>(A.n);
            L.out = ef_alloc<double>(A.n);
        }
        L.p_ptr = up_int(P.ptr);
        L.p_col = up_int(P.col);
        L.p32 = up_f32(P.val);
        L.p64 = up_f64(P.val);
        L.r_ptr = up_int(R.ptr);
        L.r_col = up_int(R.col);
        L.r32 = up_f32(R.val);
        L.r64 = up_f64(R.val);
        L.x = ef_alloc<double>(A.n);
        L.r = ef_alloc<double>(A.n);
        G.lev.push_back(L);
        A = std::move(Ac);
    }
    if (G.lev.empty() || A.n > 5 * kCoarsest) {
        G.release();
        G.built = true;  // remember the key: Jacobi for this matrix
        G.key_oodt = oodt;
        G.key_sigma = e->sigma;
        G.key_ver = e->params_ver;
        return;
    }
    const std::vector<double> inv = dense_inverse(A);
    G.nc = A.n;
    G.inv32 = up_f32(inv);
    G.inv64 = up_f64(inv);
    G.cb = ef_alloc<double>(A.n);
    G.cx = ef_alloc<double>(A.n);
    G.usable = true;
    G.setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// launch K<T, G> with G lanes per row over n rows
template <typename T, template <typename, int> class K>
struct GroupLaunch {
    template <typename... A>
    static void run(int g, int n, hipStream_t st, A... args) {
        const long threads = (long)n * g;
        const dim3 grid((unsigned)((threads + kB - 1) / kB));
        switch (g) {
            case 1: hipLaunchKernelGGL((K<T, 1>::fn), grid, dim3(kB), 0, st, args...); break;
            case 2: hipLaunchKernelGGL((K<T, 2>::fn), grid, dim3(kB), 0, st, args...); break;
            case 4: hipLaunchKernelGGL((K<T, 4>::fn), grid, dim3(kB), 0, st, args...); break;
            case 8: hipLaunchKernelGGL((K<T, 8>::fn), grid, dim3(kB), 0, st, args...); break;
            case 16: hipLaunchKernelGGL((K<T, 16>::fn), grid, dim3(kB), 0, st, args...); break;
            case 32: hipLaunchKernelGGL((K<T, 32>::fn), grid, dim3(kB), 0, st, args...); break;
            default: hipLaunchKernelGGL((K<T, 64>::fn), grid, dim3(kB), 0, st, args...); break;
        }
    }
};
template <typename T, int G> struct ResidK { static constexpr auto fn = k_csr_resid<T, G>; };
template <typename T, int G> struct UpK { static constexpr auto fn = k_csr_up<T, G>; };
template <typename T, int G> struct ApplyK { static constexpr auto fn = k_csr_apply<T, G>; };


template <typename T>
const T *pick(const float *f, const double *d) {
    if constexpr (sizeof(T) == 4) return reinterpret_cast<const T *>(f);
    else return reinterpret_cast<const T *>(d);
}

// z = M b0: one V-cycle on the device; the block partials of (b0, z0) go to
// part (one per block of the fine level), for k_reduce
template <typename T>
void vcycle(stg_efield *e, const T *b0, T *z0, double *part) {
    AmgDev &G = e->amg;
    const int L = (int)G.lev.size();
    const EfDev m = e->dev();
    auto blocks = [](int n) { return dim3((n + kB - 1) / kB); };
    for (int l = 0; l < L; ++l) {
        AmgLevel &V = G.lev[l];
        const T *b = l == 0 ? b0 : reinterpret_cast<const T *>(V.b);
        T *x = reinterpret_cast<T *>(V.x), *r = reinterpret_cast<T *>(V.r);
        if (l == 0) {
            hipLaunchKernelGGL(k_amg0_down<T>, blocks(V.n), dim3(kB), 0, e->stream, m, e->sigma, V.om, b, x, r);
        } else {  // x = om D^-1 b was written with b by the restriction
            GroupLaunch<T, ResidK>::run(V.ga, V.n, e->stream, V.n, (const int *)V.a_ptr, (const int *)V.a_col,
                                        pick<T>(V.a32, V.a64), (const float *)V.rs32, b, (const T *)x, r);
        }
        if (l + 1 < L) {
            AmgLevel &C = G.lev[l + 1];
            GroupLaunch<T, ApplyK>::run(V.gr, C.n, e->stream, C.n, (const int *)V.r_ptr, (const int *)V.r_col,
                                        pick<T>(V.r32, V.r64), (const T *)r, reinterpret_cast<T *>(C.b), 0,
                                        pick<T>(C.dinv32, C.dinv64), C.om, reinterpret_cast<T *>(C.x));
        } else {
            GroupLaunch<T, ApplyK>::run(V.gr, G.nc, e->stream, G.nc, (const int *)V.r_ptr, (const int *)V.r_col,
                                        pick<T>(V.r32, V.r64), (const T *)r, reinterpret_cast<T *>(G.cb), 0,
                                        (const T *)nullptr, 0.0, (T *)nullptr);
        }
    }
    hipLaunchKernelGGL(k_dense_apply<T>, blocks(G.nc * 16), dim3(kB), 0, e->stream, G.nc,
                       pick<T>(G.inv32, G.inv64), (const T *)G.cb, reinterpret_cast<T *>(G.cx));
    for (int l = L - 1; l >= 0; --l) {
        AmgLevel &V = G.lev[l];
        const T *xc = reinterpret_cast<const T *>(l + 1 < L ? G.lev[l + 1].out : G.cx);
        T *x = reinterpret_cast<T *>(V.x);
        GroupLaunch<T, ApplyK>::run(V.gp, V.n, e->stream, V.n, (const int *)V.p_ptr, (const int *)V.p_col,
                                    pick<T>(V.p32, V.p64), xc, x, 1, (const T *)nullptr, 0.0, (T *)nullptr);
        const T *b = l == 0 ? b0 : reinterpret_cast<const T *>(V.b);
        T *out = l == 0 ? z0 : reinterpret_cast<T *>(V.out);
        if (l == 0)
            hipLaunchKernelGGL(k_amg0_up<T>, blocks(V.n), dim3(kB), 0, e->stream, m, e->sigma, V.om, b,
                               (const T *)x, out, part);
        else
            GroupLaunch<T, UpK>::run(V.ga, V.n, e->stream, V.n, (const int *)V.a_ptr, (const int *)V.a_col,
                                     pick<T>(V.a32, V.a64), (const float *)V.rs32, pick<T>(V.dinv32, V.dinv64),
                                     V.om, b, (const T *)x, out);
    }
}

// CG on A d = r (r given in r64) preconditioned by one V-cycle per iteration;
// result added to x64. Returns iterations.
template <typename T>
int pcg_amg(stg_efield *e, double rel_tol, int max_iter) {
    const EfDev m = e->dev();
    const int nb = e->nb, n = e->n;
    T *x = (T *)e->xs, *r = (T *)e->rs, *z = (T *)e->zs, *p = (T *)e->ps, *q = (T *)e->qs;
    hipLaunchKernelGGL(k_cg_start<T>, dim3(nb), dim3(kB), 0, e->stream, n, e->r64, x, r);
    vcycle<T>(e, r, z, e->part);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, e->stream, e->part, nb, 0, e->scal, e->outs);
    // iteration it: the search directions alternate between two buffers, and
    // the first iteration starts from p = z
    auto iteration = [&](int it) {
        T *pold = p + (size_t)(it & 1) * n, *pnew = p + (size_t)((it + 1) & 1) * n;
        hipLaunchKernelGGL(k_cg_a<T>, dim3(nb), dim3(kB), 0, e->stream, m, e->sigma, z, pold, pnew, q, e->scal,
                           it == 0 ? 1 : 0, e->part);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, e->stream, e->part, nb, 1, e->scal, e->outs);
        hipLaunchKernelGGL(k_cg_update<T>, dim3(nb), dim3(kB), 0, e->stream, n, e->scal, x, r, (const T *)pnew,
                           (const T *)q);
        vcycle<T>(e, r, z, e->part);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, e->stream, e->part, nb, 2, e->scal, e->outs);
    };
    // Iterations run in pairs between convergence checks (each check is a
    // host round trip). After the first pair, a pair (even it, odd it) is the
    // same launch sequence every time, so it is captured once as a graph and
    // replayed: about 20 launches per iteration cost their launch gaps once.
    // r.z of the start is read back together with the first pair's (one
    // round trip). It is positive whenever CG runs (the residual is nonzero and
    // the V-cycle is SPD); were it not, the first pair's correction is dropped.
    hipGraphExec_t &pair = sizeof(T) == 4 ? e->amg.pair32 : e->amg.pair64;
    int it = 0;
    double rz0 = 0.0;
    {
        while (it < max_iter) {
            if (it > 0 && (it & 1) == 0 && it + 2 <= max_iter) {
                if (!pair) {
                    hipGraph_t g = nullptr;
                    EF_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
                    iteration(it);
                    iteration(it + 1);
                    EF_HIP(hipStreamEndCapture(e->stream, &g));
                    const hipError_t rc = hipGraphInstantiate(&pair, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    EF_HIP(rc);
                }
                EF_HIP(hipGraphLaunch(pair, e->stream));
                it += 2;
            } else {
                for (int k = 0; k < 2 && it < max_iter; ++k, ++it) iteration(it);
            }
            double rz = 0.0;
            if (it <= 2) {
                fetch2(e, e->outs, rz0, rz);
                if (!(rz0 > 0.0)) return 0;
            } else {
                rz = fetch(e, e->scal);
            }
            if (!(rz > rel_tol * rel_tol * rz0)) break;
        }
    }
    EF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_accumulate<T>, dim3(nb), dim3(kB), 0, e->stream, n, e->x64, x);
    return it;
}

// refresh the attached simulation's triangle potentials (TetOpSplitP::_refreshEFTrisV,
// tetopsplit.cpp:1931)
void publish_tri_v(stg_efield *e) {
    if (!e->sim || e->n_tris == 0) return;
    double *dst = nullptr;
    const int rc = stg_internal_tri_v_device(e->sim, e->sim_real, &dst);
    if (rc != STG_OK) throw EfError{rc, "attached simulation rejected the triangle potentials"};
    hipLaunchKernelGGL(k_tri_v, dim3((e->n_tris + kB - 1) / kB), dim3(kB), 0, e->stream, e->n_tris,
                       e->tri_vd, e->V, e->tri_map, dst);
    EF_HIP(hipGetLastError());
}

// Ohmic currents average the channel counts over the time the simulation
// advanced since the previous EField step (Tri::computeI, tri.cpp:664-693):
// start the integrals of the channel species from the present time.
void start_integrals(stg_efield *e) {
    e->intg = nullptr;
    e->intg_mask = 0;
    e->charge = nullptr;
    e->n_ghk = 0;
    e->ghk_dt = 0.0;  // no step yet: every GHK current reads 0 (Tri::getGHKI)
    if (!e->sim) return;
    int rc0 = stg_get_time(e->sim, &e->intg_t0);
    if (rc0 != STG_OK) throw EfError{rc0, stg_last_error(e->sim)};
    rc0 = stg_internal_tri_charge_device(e->sim, e->sim_real, &e->charge, &e->n_ghk, &e->charge_stride);
    if (rc0 != STG_OK) throw EfError{rc0, stg_last_error(e->sim)};
    if (e->charge && e->n_tris > 0) {
        if (!e->ghk_last) e->ghk_last = ef_alloc<int32_t>((size_t)kEfMaxGhk * e->n_tris);
        hipLaunchKernelGGL(k_take_charge, dim3((e->n_tris + kB - 1) / kB), dim3(kB), 0, e->stream, e->n_tris,
                           e->tri_map, e->n_ghk, e->charge_stride, e->charge, nullptr);
        EF_HIP(hipGetLastError());
        EF_HIP(hipStreamSynchronize(e->stream));
    }
    if (e->oc.n == 0) return;
    uint32_t mask = 0;
    for (int k = 0; k < e->oc.n; ++k) mask |= 1u << e->oc.sspec[k];
    int rc = stg_internal_tri_intg_device(e->sim, mask, &e->intg);
    if (rc != STG_OK) throw EfError{rc, stg_last_error(e->sim)};
    const int32_t *tc = nullptr;
    int32_t ntot = 0, tri0 = 0;
    rc = stg_internal_tri_counts_device(e->sim, e->sim_real, &tc, &ntot, &tri0);
    if (rc != STG_OK) throw EfError{rc, "attached simulation has no triangle counts"};
    e->intg_mask = mask;
    hipLaunchKernelGGL(k_zero_intg, dim3((e->n_tris + kB - 1) / kB), dim3(kB), 0, e->stream, e->n_tris,
                       e->tri_map, ntot, mask, e->intg + tri0);
    EF_HIP(hipGetLastError());
    EF_HIP(hipStreamSynchronize(e->stream));
}

void advance_impl(stg_efield *e, double dt_s) {
    if (!(dt_s > 0.0)) throw EfError{STG_ERR_ARG, "EField time step must be positive"};
    const int n = e->n, nb = e->nb;
    const EfDev m = e->dev();
    const int32_t *tri_counts = nullptr;
    const double *intg = nullptr;
    const int32_t *charge = nullptr;
    double span = 0.0, now = 0.0;
    int32_t ntot = 0, tri0 = 0;
    if (e->sim) {
        int rc = stg_get_time(e->sim, &now);
        if (rc != STG_OK) throw EfError{rc, stg_last_error(e->sim)};
        span = now - e->intg_t0;
        // GHK charge is turned into a current only over simulated time
        if (e->charge && span > 0.0) charge = e->charge;
        if (e->oc.n > 0) {
            rc = stg_internal_tri_counts_device(e->sim, e->sim_real, &tri_counts, &ntot, &tri0);
            if (rc != STG_OK) throw EfError{rc, "attached simulation has no triangle counts"};
            tri_counts += tri0;
            // no simulated time since the last step (or the clock was reset):
            // the present counts stand in for the average
            if (e->intg && span > 0.0) intg = e->intg + tri0;
        }
    }
    hipLaunchKernelGGL(k_vertex_current, dim3(nb), dim3(kB), 0, e->stream, n, e->vt_ptr, e->vt_tri,
                       e->tri_vd, e->V, e->tclamp, e->vclamp, tri_counts, intg, span, charge, e->n_ghk,
                       e->charge_stride, e->tri_map, ntot, e->oc, e->oc_erev_h.empty() ? nullptr : e->oc_erev,
                       e->n_tris, e->I);
    if (e->sim) {
        if (e->intg)
            hipLaunchKernelGGL(k_zero_intg, dim3((e->n_tris + kB - 1) / kB), dim3(kB), 0, e->stream,
                               e->n_tris, e->tri_map, ntot, e->intg_mask, e->intg + tri0);
        if (e->charge && e->n_tris > 0) {
            // the step's charges become the currents the getters report
            hipLaunchKernelGGL(k_take_charge, dim3((e->n_tris + kB - 1) / kB), dim3(kB), 0, e->stream,
                               e->n_tris, e->tri_map, e->n_ghk, e->charge_stride, e->charge,
                               span > 0.0 ? e->ghk_last : nullptr);
            if (span > 0.0) e->ghk_dt = span;
        }
        e->intg_t0 = now;
    }
    const double oodt = 1.0 / (dt_s * 1e3);
    // The hierarchy only preconditions: CG itself runs on the matrix assembled
    // for this exact step. It is rebuilt when a membrane or bulk parameter
    // changed, or when 1/dt moved by more than a quarter (step sizes from
    // accumulated times differ in their last bits from step to step).
    if (e->precond == 1 && n > kSmallSystem) {
        const AmgDev &G = e->amg;
        if (!G.built || std::fabs(G.key_oodt - oodt) > 0.25 * oodt || G.key_sigma != e->sigma ||
            G.key_ver != e->params_ver)
            amg_build(e, oodt);
    }
    hipLaunchKernelGGL(k_assemble, dim3(nb), dim3(kB), 0, e->stream, m, e->sigma, oodt, e->cap, e->gext,
                       e->vext, e->V, e->I, e->b, e->d64, e->d32, e->m32);
    if (e->a32_dirty) {
        hipLaunchKernelGGL(k_offdiag32, dim3(nb), dim3(kB), 0, e->stream, e->dev(), e->sigma, e->a32);
        e->a32_dirty = false;
    }
    // initial guess: the previous step's increment
    EF_HIP(hipMemcpyAsync(e->x64, e->dv_prev, n * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
    // b.b and the first residual's r.r come back in one round trip
    hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(kB), 0, e->stream, n, e->b, e->part);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, e->stream, e->part, nb, 3, e->scal, e->outs + 3);
    hipLaunchKernelGGL(k_residual64, dim3(nb), dim3(kB), 0, e->stream, m, e->sigma, e->b, e->x64, e->r64,
                       e->part);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, e->stream, e->part, nb, 3, e->scal, e->outs + 4);
    double bb = 0.0, rr0 = 0.0;
    fetch2(e, e->outs + 3, bb, rr0);
    int iters = 0, refine = 0;
    double rel = 0.0, prev = INFINITY;
    bool ok = true;
    if (bb > 0.0) {
        for (refine = 0;; ++refine) {
            double rr = rr0;
            if (refine > 0) {
                hipLaunchKernelGGL(k_residual64, dim3(nb), dim3(kB), 0, e->stream, m, e->sigma, e->b, e->x64,
                                   e->r64, e->part);
                hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, e->stream, e->part, nb, 3, e->scal,
                                   e->outs + 4);
                rr = fetch(e, e->outs + 4);
            }
            rel = std::sqrt(rr / bb);
            if (rel <= e->tol) break;
            // the FP64 residual has a rounding floor near eps * cond(A); stop
            // once a refinement no longer halves it
            if (refine >= 2 && rel > 0.5 * prev) {
                ok = rel <= e->floor_tol;
                break;
            }
            if (refine >= 12) {
                ok = false;
                break;
            }
            prev = rel;
            if (e->precision == 1) iters += pcg<double>(e, std::max(e->tol / rel, 1e-15), e->max_iter);
            else  // the inner solve only needs to close the remaining gap to tol
                iters += pcg<float>(e, std::max(e->inner_tol, 0.1 * e->tol / rel),
                                    std::min(e->inner_max, e->max_iter));
        }
    } else {
        EF_HIP(hipMemsetAsync(e->x64, 0, n * sizeof(double), e->stream));
    }
    e->last_iters = iters;
    e->last_refine = refine;
    e->last_rel = rel;
    if (!ok) throw EfError{STG_ERR_PROG, "EField solve did not reach the residual tolerance"};
    hipLaunchKernelGGL(k_apply_dv, dim3(nb), dim3(kB), 0, e->stream, n, e->clamped, e->x64, e->V);
    publish_tri_v(e);
    EF_HIP(hipMemcpyAsync(e->dv_prev, e->x64, n * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
    EF_HIP(hipGetLastError());
    EF_HIP(hipStreamSynchronize(e->stream));
}

void check_v(const stg_efield *e, int32_t v) {
    if (v < 0 || v >= e->n) throw EfError{STG_ERR_ARG, "vertex index out of range"};
}

void check_t(const stg_efield *e, int32_t t) {
    if (t < 0 || t >= e->n_tris) throw EfError{STG_ERR_ARG, "triangle index out of range"};
}

}  // namespace

extern "C" {

int stg_efield_create(const stg_efield_mesh *mesh, stg_efield **out) {
    if (!out) {
        g_ef_error = "null output handle";
        return STG_ERR_ARG;
    }
    *out = nullptr;
    stg_efield *e = new (std::nothrow) stg_efield();
    if (!e) return STG_ERR_SYS;
    const int rc = ef_guard(nullptr, [&] { build(e, mesh); });
    if (rc != STG_OK) {
        delete e;
        return rc;
    }
    *out = e;
    return STG_OK;
}

int stg_efield_destroy(stg_efield *e) {
    delete e;
    return STG_OK;
}

const char *stg_efield_last_error(const stg_efield *e) {
    return e ? e->err.c_str() : g_ef_error.c_str();
}

int stg_efield_set_potential(stg_efield *e, double v) {
    if (!e) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        std::vector<double> h(e->n, v * 1e3);
        ef_put(e->V, h);
        EF_HIP(hipMemset(e->dv_prev, 0, e->n * sizeof(double)));
        publish_tri_v(e);
        EF_HIP(hipStreamSynchronize(e->stream));
    });
}

int stg_efield_set_capac(stg_efield *e, double cm) {
    if (!e) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        if (!(cm >= 0.0)) throw EfError{STG_ERR_ARG, "capacitance must be >= 0"};
        e->cm = cm;  // F/m^2 = pF/um^2 (efield.cpp:151-160)
        for (int i = 0; i < e->n; ++i) e->cap_h[i] = cm * e->vert_area[i];
        std::fill(e->tri_cm_h.begin(), e->tri_cm_h.end(), cm);  // TetMesh::applySurfaceCapacitance
        upload_params(e);
    });
}

int stg_efield_set_volres(stg_efield *e, double ro) {
    if (!e) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        if (!(ro > 0.0)) throw EfError{STG_ERR_ARG, "volume resistivity must be > 0"};
        e->sigma = 1.0 / (ro * 1.0e-3);  // efield.cpp:183-186
        e->a32_dirty = true;
    });
}

int stg_efield_set_memb_res(stg_efield *e, double rspec, double vrev) {
    if (!e) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        if (!(rspec > 0.0)) throw EfError{STG_ERR_ARG, "membrane resistivity must be > 0"};
        const double rs = 1.0e3 * rspec;  // efield.cpp:188-195
        for (int i = 0; i < e->n; ++i) e->gext_h[i] = e->vert_area[i] / rs;
        e->vext = vrev * 1.0e3;
        upload_params(e);
    });
}

int stg_efield_set_vert_iclamp(stg_efield *e, int32_t v, double amps) {
    if (!e) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        check_v(e, v);
        e->vert_clamp_h[v] = amps * 1e12;
        upload_params(e);
    });
}

int stg_efield_set_tri_iclamp(stg_efield *e, int32_t t, double amps) {
    if (!e) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        check_t(e, t);
        e->tri_clamp_h[t] = amps * 1e12;
        upload_params(e);
    });
}

int stg_efield_set_tri_capac(stg_efield *e, int32_t t, double cm) {
    if (!e) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        check_t(e, t);
        if (!(cm >= 0.0)) throw EfError{STG_ERR_ARG, "capacitance must be >= 0"};
        // TetMesh::applyTriCapacitance: each vertex's share of the area changes
        // by (cm - previous cm) * area / 3
        const double d = (cm - e->tri_cm_h[t]) * e->tri_area[t] / 3.0;
        e->tri_cm_h[t] = cm;
        for (int a = 0; a < 3; ++a) e->cap_h[e->tri_v[3 * t + a]] += d;
        upload_params(e);
    });
}

int stg_efield_get_tri_capac(stg_efield *e, int32_t t, double *cm) {
    if (!e || !cm) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        check_t(e, t);
        *cm = e->tri_cm_h[t];
    });
}

int stg_efield_set_tri_ohmic_erev(stg_efield *e, int32_t k, int32_t t, double erev) {
    if (!e) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        check_t(e, t);
        if (k < 0 || k >= e->oc.n) throw EfError{STG_ERR_ARG, "ohmic current index out of range"};
        if (e->oc_erev_h.empty()) {
            e->oc_erev_h.resize((size_t)8 * e->n_tris);
            for (int j = 0; j < 8; ++j)
                std::fill(e->oc_erev_h.begin() + (size_t)j * e->n_tris,
                          e->oc_erev_h.begin() + (size_t)(j + 1) * e->n_tris, j < e->oc.n ? e->oc.erev[j] : 0.0);
        }
        e->oc_erev_h[(size_t)k * e->n_tris + t] = erev;
        upload_params(e);
    });
}

int stg_efield_get_tri_ohmic_erev(stg_efield *e, int32_t k, int32_t t, double *erev) {
    if (!e || !erev) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        check_t(e, t);
        if (k < 0 || k >= e->oc.n) throw EfError{STG_ERR_ARG, "ohmic current index out of range"};
        *erev = e->oc_erev_h.empty() ? e->oc.erev[k] : e->oc_erev_h[(size_t)k * e->n_tris + t];
    });
}

int stg_efield_get_tri_ghk_currents(stg_efield *e, int32_t g, double *amps) {
    if (!e || !amps) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        if (g < -1 || g >= std::max(e->n_ghk, 1)) throw EfError{STG_ERR_ARG, "GHK current index out of range"};
        const int nt = e->n_tris;
        if (e->ghk_dt == 0.0 || e->n_ghk == 0 || !e->ghk_last) {
            std::fill(amps, amps + nt, 0.0);
            return;
        }
        std::vector<int32_t> q((size_t)e->n_ghk * nt);
        EF_HIP(hipStreamSynchronize(e->stream));
        EF_HIP(hipMemcpy(q.data(), e->ghk_last, q.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (int t = 0; t < nt; ++t) {
            int c = 0;  // Tri::getGHKI sums the charges before converting
            if (g >= 0) c = q[(size_t)g * nt + t];
            else
                for (int k = 0; k < e->n_ghk; ++k) c += q[(size_t)k * nt + t];
            amps[t] = ((double)c * 1.602176487e-19) / e->ghk_dt;
        }
    });
}

int stg_efield_set_vert_clamped(stg_efield *e, int32_t v, int32_t clamped) {
    if (!e) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        check_v(e, v);
        e->clamped_h[v] = clamped ? 1 : 0;
        upload_params(e);
    });
}

int stg_efield_set_verts_v(stg_efield *e, const double *v) {
    if (!e || !v) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        std::vector<double> h(v, v + e->n);
        for (auto &x : h) x *= 1e3;
        ef_put(e->V, h);
        publish_tri_v(e);
        EF_HIP(hipStreamSynchronize(e->stream));
    });
}

int stg_efield_get_verts_v(stg_efield *e, double *v) {
    if (!e || !v) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        EF_HIP(hipMemcpy(v, e->V, e->n * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < e->n; ++i) v[i] *= 1e-3;
    });
}

int stg_efield_get_tris_v(stg_efield *e, double *v) {
    if (!e || !v) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        std::vector<double> h(e->n);
        EF_HIP(hipMemcpy(h.data(), e->V, e->n * sizeof(double), hipMemcpyDeviceToHost));
        for (int t = 0; t < e->n_tris; ++t)
            v[t] = (h[e->tri_v[3 * t]] + h[e->tri_v[3 * t + 1]] + h[e->tri_v[3 * t + 2]]) / 3.0 * 1e-3;
    });
}

int stg_efield_set_ohmic(stg_efield *e, int32_t n, const int32_t *sspec, const double *g,
                         const double *erev) {
    if (!e) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        if (n < 0 || n > 8) throw EfError{STG_ERR_ARG, "between 0 and 8 ohmic currents"};
        e->oc.n = n;
        e->oc_erev_h.clear();  // every triangle starts from the current's own reversal potential
        for (int k = 0; k < n; ++k) {
            if (sspec[k] < 0 || sspec[k] >= 8) throw EfError{STG_ERR_ARG, "bad channel species index"};
            e->oc.sspec[k] = sspec[k];
            e->oc.g[k] = g[k];
            e->oc.erev[k] = erev[k];
        }
        start_integrals(e);
    });
}

int stg_efield_attach(stg_efield *e, stg_sim *sim, int32_t real) {
    if (!e) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        if (sim) {
            const int32_t *perm = nullptr;
            int32_t nt = 0;
            const int rc = stg_internal_tri_perm(sim, real, &perm, &nt);
            if (rc != STG_OK) throw EfError{rc, "bad realization"};
            if (nt != e->n_tris)
                throw EfError{STG_ERR_ARG, "simulation and EField membranes differ in triangle count"};
            std::vector<int> map(perm, perm + nt);
            if (!e->tri_map) e->tri_map = ef_alloc<int>(std::max(nt, 1));
            ef_put(e->tri_map, map);
        }
        e->sim = sim;
        e->sim_real = real;
        start_integrals(e);
        publish_tri_v(e);
        EF_HIP(hipStreamSynchronize(e->stream));
    });
}

int stg_efield_advance(stg_efield *e, double dt) {
    if (!e) return STG_ERR_ARG;
    return ef_guard(e, [&] { advance_impl(e, dt); });
}

int stg_efield_run(stg_efield *e, stg_sim *sim, double endtime, double efdt) {
    if (!e || !sim) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        if (!(efdt > 0.0)) throw EfError{STG_ERR_ARG, "EField dt must be positive"};
        double t = 0.0;
        int rc = stg_get_time(sim, &t);
        if (rc != STG_OK) throw EfError{rc, stg_last_error(sim)};
        if (endtime < t) throw EfError{STG_ERR_ARG, "Endtime is before current simulation time"};
        // TetOpSplitP::_runWithEField (tetopsplit.cpp:1851-1954)
        while (t < endtime) {
            const double t0 = t;
            double t1 = std::min(t0 + efdt, endtime);
            if (std::fabs(t1 - endtime) <= DBL_EPSILON * std::fabs(t1 + endtime) * 4) t1 = endtime;
            rc = stg_run(sim, t1);
            if (rc != STG_OK) throw EfError{rc, stg_last_error(sim)};
            advance_impl(e, t1 - t0);
            t = t1;
        }
    });
}

// ------------------------------------------------------------ checkpoint
// Own format: header, vertex potentials and the previous step's increment
// (the CG starting guess), the membrane and clamp parameters, the time the
// present channel-count integrals started from, and the GHK charges of the
// last step with its length (Tri::checkpoint keeps pECharge_last and its dt,
// tri.cpp:133-157).
}  // extern "C"

namespace {
constexpr uint64_t kEfStateMagic = 0x3253464547544ull;  // "TGEFS2"
struct EfStateHeader {
    uint64_t magic;
    int32_t n, n_tris;
    double intg_t0, sigma, vext, cm;
    int32_t has_oc_erev, n_ghk;
    double ghk_dt;
};
std::vector<std::vector<double> *> ef_state_vectors(stg_efield *e) {
    std::vector<std::vector<double> *> v{&e->cap_h, &e->gext_h, &e->vert_clamp_h, &e->tri_clamp_h, &e->tri_cm_h};
    if (!e->oc_erev_h.empty()) v.push_back(&e->oc_erev_h);
    return v;
}
size_t ef_state_bytes(stg_efield *e, bool oc_erev, int n_ghk) {
    size_t b = sizeof(EfStateHeader) + 2 * sizeof(double) * (size_t)e->n;
    for (auto *v : {&e->cap_h, &e->gext_h, &e->vert_clamp_h, &e->tri_clamp_h, &e->tri_cm_h})
        b += v->size() * sizeof(double);
    if (oc_erev) b += (size_t)8 * e->n_tris * sizeof(double);
    return b + e->clamped_h.size() + (size_t)n_ghk * e->n_tris * sizeof(int32_t);
}
int ef_saved_ghk(const stg_efield *e) { return e->ghk_last ? e->n_ghk : 0; }
}  // namespace

extern "C" {

int stg_efield_state_size(const stg_efield *e, size_t *bytes) {
    if (!e || !bytes) return STG_ERR_ARG;
    auto *m = const_cast<stg_efield *>(e);
    *bytes = ef_state_bytes(m, !e->oc_erev_h.empty(), ef_saved_ghk(e));
    return STG_OK;
}

int stg_efield_checkpoint(stg_efield *e, void *buf, size_t bytes) {
    if (!e || !buf) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        const int ng = ef_saved_ghk(e);
        if (bytes < ef_state_bytes(e, !e->oc_erev_h.empty(), ng))
            throw EfError{STG_ERR_ARG, "checkpoint buffer too small"};
        char *p = static_cast<char *>(buf);
        const EfStateHeader h{kEfStateMagic, e->n,  e->n_tris, e->intg_t0, e->sigma, e->vext, e->cm,
                              e->oc_erev_h.empty() ? 0 : 1, ng, e->ghk_dt};
        std::memcpy(p, &h, sizeof(h));
        p += sizeof(h);
        const size_t nb = (size_t)e->n * sizeof(double);
        EF_HIP(hipStreamSynchronize(e->stream));
        EF_HIP(hipMemcpy(p, e->V, nb, hipMemcpyDeviceToHost));
        p += nb;
        EF_HIP(hipMemcpy(p, e->dv_prev, nb, hipMemcpyDeviceToHost));
        p += nb;
        for (const auto *v : ef_state_vectors(e)) {
            std::memcpy(p, v->data(), v->size() * sizeof(double));
            p += v->size() * sizeof(double);
        }
        std::memcpy(p, e->clamped_h.data(), e->clamped_h.size());
        p += e->clamped_h.size();
        if (ng > 0)
            EF_HIP(hipMemcpy(p, e->ghk_last, (size_t)ng * e->n_tris * sizeof(int32_t), hipMemcpyDeviceToHost));
    });
}

int stg_efield_restore(stg_efield *e, const void *buf, size_t bytes) {
    if (!e || !buf) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        if (bytes < sizeof(EfStateHeader)) throw EfError{STG_ERR_ARG, "checkpoint too short"};
        const char *p = static_cast<const char *>(buf);
        EfStateHeader h;
        std::memcpy(&h, p, sizeof(h));
        if (h.magic != kEfStateMagic) throw EfError{STG_ERR_ARG, "not an EField checkpoint"};
        if (h.n != e->n || h.n_tris != e->n_tris)
            throw EfError{STG_ERR_ARG, "checkpoint was written by an EField of another mesh"};
        if (h.n_ghk != 0 && (h.n_ghk != e->n_ghk || !e->ghk_last))
            throw EfError{STG_ERR_ARG, "checkpoint holds GHK currents the attached simulation lacks"};
        if (bytes < ef_state_bytes(e, h.has_oc_erev != 0, h.n_ghk)) throw EfError{STG_ERR_ARG, "checkpoint too short"};
        p += sizeof(h);
        const size_t nb = (size_t)e->n * sizeof(double);
        EF_HIP(hipStreamSynchronize(e->stream));
        EF_HIP(hipMemcpy(e->V, p, nb, hipMemcpyHostToDevice));
        p += nb;
        EF_HIP(hipMemcpy(e->dv_prev, p, nb, hipMemcpyHostToDevice));
        p += nb;
        if (h.has_oc_erev) e->oc_erev_h.resize((size_t)8 * e->n_tris);
        else e->oc_erev_h.clear();
        for (auto *v : ef_state_vectors(e)) {
            std::memcpy(v->data(), p, v->size() * sizeof(double));
            p += v->size() * sizeof(double);
        }
        std::memcpy(e->clamped_h.data(), p, e->clamped_h.size());
        p += e->clamped_h.size();
        if (h.n_ghk > 0)
            EF_HIP(hipMemcpy(e->ghk_last, p, (size_t)h.n_ghk * e->n_tris * sizeof(int32_t), hipMemcpyHostToDevice));
        e->ghk_dt = h.n_ghk > 0 ? h.ghk_dt : 0.0;
        e->intg_t0 = h.intg_t0;
        e->sigma = h.sigma;
        e->vext = h.vext;
        e->cm = h.cm;
        upload_params(e);
        publish_tri_v(e);
        EF_HIP(hipStreamSynchronize(e->stream));
    });
}

int stg_efield_set_solver(stg_efield *e, int32_t precision, double tol, int32_t max_iter) {
    if (!e) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        if (precision != 0 && precision != 1) throw EfError{STG_ERR_ARG, "precision is 0 (mixed) or 1 (FP64)"};
        if (!(tol > 0.0) || max_iter <= 0) throw EfError{STG_ERR_ARG, "bad tolerance or iteration cap"};
        e->precision = precision;
        e->tol = tol;
        e->max_iter = max_iter;
    });
}

int stg_efield_set_precond(stg_efield *e, int32_t kind) {
    if (!e) return STG_ERR_ARG;
    return ef_guard(e, [&] {
        if (kind != 0 && kind != 1) throw EfError{STG_ERR_ARG, "preconditioner is 0 (Jacobi) or 1 (multigrid)"};
        e->precond = kind;
    });
}

int stg_efield_precond_info(const stg_efield *e, int32_t *levels, int32_t *coarsest, double *setup_s) {
    if (!e || !levels || !coarsest || !setup_s) return STG_ERR_ARG;
    *levels = e->amg.usable ? (int32_t)e->amg.lev.size() + 1 : 0;
    *coarsest = e->amg.usable ? e->amg.nc : 0;
    *setup_s = e->amg.usable ? e->amg.setup_s : 0.0;
    return STG_OK;
}

int stg_efield_stats(const stg_efield *e, int32_t *iters, int32_t *refinements, double *rel_residual) {
    if (!e || !iters || !refinements || !rel_residual) return STG_ERR_ARG;
    *iters = e->last_iters;
    *refinements = e->last_refine;
    *rel_residual = e->last_rel;
    return STG_OK;
}

int stg_efield_sizes(const stg_efield *e, int32_t *n_verts, int32_t *n_tris, int32_t *nnz) {
    if (!e) return STG_ERR_ARG;
    if (n_verts) *n_verts = e->n;
    if (n_tris) *n_tris = e->n_tris;
    if (nnz) *nnz = e->nnz;
    return STG_OK;
}

}  // extern "C"
