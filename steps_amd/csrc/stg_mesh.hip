// Face topology and measures of a tetrahedral mesh, built on the device: the
// work of the Tetmesh constructor (src/steps/geom/tetmesh.cpp:78-233) that the
// mesh loaders (pysteps/steps/API_2/geom.py:1466-1592) hand their vertex, tet
// and triangle arrays to, for meshes of millions of tets.
//
// The faces are the supplied triangles (positions 0 .. ns-1) followed by the
// four faces of every tet in the reference's order, (v0,v1,v2), (v0,v1,v3),
// (v0,v2,v3), (v1,v2,v3) (positions ns + 4 t + j; face j is the one opposite
// vertex 3 - j). Each is keyed by its sorted vertex triple (a, b, c). Two
// stable rocPRIM radix sorts (by c, then by a * n_verts + b) bring equal keys
// together, in position order; each run is one triangle, and its first
// position is where the triangle first appears. Triangles are numbered in
// order of first appearance (an exclusive scan over "first appearance" marks
// in position order), which is the reference's numbering. All measures are
// formed with the same operations, in the same order, as the host builder.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "../../include/stg.h"

namespace {

struct MbError {
    int code;
    std::string msg;
};

#define MB_HIP(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw MbError{STG_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

constexpr int kB = 256;

dim3 grid_for(long n) { return dim3((unsigned)((n + kB - 1) / kB)); }

// volume |e1 . (e2 x e3)| / 6 and barycentre of each tet
__global__ void k_tet_geom(int nt, const double *verts, const int32_t *tets, double *vol, double *bary,
                           int32_t *err) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    double p[4][3];
    for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 3; ++k) p[a][k] = verts[3 * (size_t)tets[4 * (size_t)t + a] + k];
    double e1[3], e2[3], e3[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = p[1][k] - p[0][k];
        e2[k] = p[2][k] - p[0][k];
        e3[k] = p[3][k] - p[0][k];
    }
    const double cx = e2[1] * e3[2] - e2[2] * e3[1];
    const double cy = e2[2] * e3[0] - e2[0] * e3[2];
    const double cz = e2[0] * e3[1] - e2[1] * e3[0];
    const double v = fabs(e1[0] * cx + e1[1] * cy + e1[2] * cz) / 6.0;
    vol[t] = v;
    if (!(v > 0.0)) atomicOr(err, 1);
    for (int k = 0; k < 3; ++k) bary[3 * (size_t)t + k] = (((p[0][k] + p[1][k]) + p[2][k]) + p[3][k]) / 4.0;
}

// sorted vertex triple of face position i: a supplied triangle (i < ns) or
// face j = (i - ns) & 3 of tet (i - ns) >> 2 in the reference's face order;
// hi = a * nv + b, lo = c, id = i
__global__ void k_face_keys(long nf, int ns, int nv, const int32_t *tets, const int32_t *tris, uint64_t *hi,
                            uint32_t *lo, int32_t *id) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    int v[3];
    if (i < ns) {
        for (int k = 0; k < 3; ++k) v[k] = tris[3 * i + k];
    } else {
        const long t = (i - ns) >> 2;
        const int j = (int)((i - ns) & 3);
        // (0,1,2), (0,1,3), (0,2,3), (1,2,3)
        const int a0 = j == 3 ? 1 : 0;
        const int a1 = j <= 1 ? 1 : 2;
        const int a2 = j == 0 ? 2 : 3;
        v[0] = tets[4 * t + a0];
        v[1] = tets[4 * t + a1];
        v[2] = tets[4 * t + a2];
    }
    if (v[0] > v[1]) { const int s = v[0]; v[0] = v[1]; v[1] = s; }
    if (v[1] > v[2]) { const int s = v[1]; v[1] = v[2]; v[2] = s; }
    if (v[0] > v[1]) { const int s = v[0]; v[0] = v[1]; v[1] = s; }
    hi[i] = (uint64_t)v[0] * (uint64_t)nv + (uint64_t)v[1];
    lo[i] = (uint32_t)v[2];
    id[i] = (int32_t)i;
}

template <typename T>
__global__ void k_gather(long n, const T *src, const int32_t *idx, T *dst) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

// 1 where a sorted position starts a new key; that position's face is the
// key's first appearance, marked in position order
__global__ void k_starts(long nf, const uint64_t *hi, const uint32_t *lo, const int32_t *pos, int32_t *start,
                         int32_t *first_mark) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const bool s = i == 0 || hi[i] != hi[i - 1] || lo[i] != lo[i - 1];
    start[i] = s ? 1 : 0;
    if (s) first_mark[pos[i]] = 1;
}

__global__ void k_fill(long n, int32_t *p, int32_t v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// one thread per key run: the triangle (numbered by its first appearance),
// its vertices, the tets having it (the first two tet faces of the run, in
// position order) and the neighbour relation across it. A run with more than
// two tet faces is a non-manifold triangle (error 2); one with none is a
// supplied triangle that is no tet's face (error 4).
__global__ void k_tri_build(long nf, int ns, int nv, const uint64_t *hi, const uint32_t *lo, const int32_t *pos,
                            const int32_t *start, const int32_t *tri_no, int32_t *tris, int32_t *tri_tets,
                            int32_t *tet_tri, int32_t *nbr, int32_t *nbr_face, int32_t *err) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf || !start[i]) return;
    const int tri = tri_no[pos[i]];
    int32_t fa = -1, fb = -1;
    int nface = 0;
    for (long k = i; k < nf && (k == i || !start[k]); ++k) {
        const int32_t p = pos[k];
        if (p < ns) continue;  // a supplied copy of the triangle
        const int32_t slot = 4 * ((p - ns) >> 2) + (3 - ((p - ns) & 3));  // tet * 4 + face opposite vertex
        if (nface == 0) fa = slot;
        else if (nface == 1) fb = slot;
        ++nface;
        tet_tri[slot] = tri;
    }
    if (nface > 2) atomicOr(err, 2);
    if (nface == 0) atomicOr(err, 4);
    tris[3 * (size_t)tri + 0] = (int32_t)(hi[i] / (uint64_t)nv);
    tris[3 * (size_t)tri + 1] = (int32_t)(hi[i] % (uint64_t)nv);
    tris[3 * (size_t)tri + 2] = (int32_t)lo[i];
    tri_tets[2 * (size_t)tri] = fa >= 0 ? (fa >> 2) : -1;
    tri_tets[2 * (size_t)tri + 1] = fb >= 0 ? (fb >> 2) : -1;
    if (fa >= 0 && fb >= 0 && nface == 2) {
        nbr[fa] = fb >> 2;
        nbr_face[fa] = fb & 3;
        nbr[fb] = fa >> 2;
        nbr_face[fb] = fa & 3;
    }
}

// area 0.5 |(p1 - p0) x (p2 - p0)| and barycentre of each triangle
__global__ void k_tri_geom(int ntri, const double *verts, const int32_t *tris, double *area, double *bary) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    double p[3][3];
    for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 3; ++k) p[a][k] = verts[3 * (size_t)tris[3 * (size_t)t + a] + k];
    double u[3], w[3];
    for (int k = 0; k < 3; ++k) {
        u[k] = p[1][k] - p[0][k];
        w[k] = p[2][k] - p[0][k];
    }
    const double cx = u[1] * w[2] - u[2] * w[1];
    const double cy = u[2] * w[0] - u[0] * w[2];
    const double cz = u[0] * w[1] - u[1] * w[0];
    area[t] = 0.5 * sqrt((cx * cx + cy * cy) + cz * cz);
    for (int k = 0; k < 3; ++k) bary[3 * (size_t)t + k] = ((p[0][k] + p[1][k]) + p[2][k]) / 3.0;
}

// per tet face: the triangle's area and the distance between the barycentres
// of the tet and its neighbour (0 on the boundary)
__global__ void k_face_measures(long nf, const int32_t *tet_tri, const double *tri_area, const int32_t *nbr,
                                const double *bary, double *face_area, double *face_dist) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    face_area[i] = tri_area[tet_tri[i]];
    const int32_t j = nbr[i];
    double d = 0.0;
    if (j >= 0) {
        const long t = i >> 2;
        const double dx = bary[3 * t] - bary[3 * (size_t)j];
        const double dy = bary[3 * t + 1] - bary[3 * (size_t)j + 1];
        const double dz = bary[3 * t + 2] - bary[3 * (size_t)j + 2];
        d = sqrt((dx * dx + dy * dy) + dz * dz);
    }
    face_dist[i] = d;
}

template <typename T>
T *mb_alloc(size_t n) {
    void *p = nullptr;
    MB_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return static_cast<T *>(p);
}

}  // namespace

struct stg_meshbuild {
    int nt = 0, nv = 0, ntri = 0;
    std::string err;
    std::vector<void *> owned;
    double *tet_vol = nullptr, *tet_bary = nullptr, *face_area = nullptr, *face_dist = nullptr;
    double *tri_area = nullptr, *tri_bary = nullptr;
    int32_t *tet_tri = nullptr, *nbr = nullptr, *nbr_face = nullptr, *tris = nullptr, *tri_tets = nullptr;

    template <typename T>
    T *alloc(size_t n) {
        T *p = mb_alloc<T>(n);
        owned.push_back(p);
        return p;
    }
    ~stg_meshbuild() {
        for (void *p : owned) (void)hipFree(p);
    }
};

namespace {

thread_local std::string g_mb_error;

template <typename F>
int mb_guard(stg_meshbuild *m, F &&f) {
    try {
        f();
        return STG_OK;
    } catch (const MbError &x) {
        if (m) m->err = x.msg;
        g_mb_error = x.msg;
        return x.code;
    } catch (const std::bad_alloc &) {
        g_mb_error = "host allocation failed";
        return STG_ERR_SYS;
    }
}

void build(stg_meshbuild *m, const double *verts, int32_t nv, const int32_t *tets, int32_t nt, const int32_t *stris,
           int32_t ns) {
    if (!verts || !tets || nv <= 0 || nt <= 0) throw MbError{STG_ERR_ARG, "empty or null mesh"};
    if (ns < 0 || (ns > 0 && !stris)) throw MbError{STG_ERR_ARG, "bad supplied triangles"};
    for (size_t i = 0; i < 4 * (size_t)nt; ++i)
        if (tets[i] < 0 || tets[i] >= nv) throw MbError{STG_ERR_ARG, "tet vertex index out of range"};
    for (size_t i = 0; i < 3 * (size_t)ns; ++i)
        if (stris[i] < 0 || stris[i] >= nv) throw MbError{STG_ERR_ARG, "triangle vertex index out of range"};
    if ((long)ns + 4L * nt > INT32_MAX) throw MbError{STG_ERR_ARG, "mesh too large"};
    m->nt = nt;
    m->nv = nv;
    const long nf = (long)ns + 4L * nt;   // face positions
    const long nq = 4L * nt;              // tet faces
    hipStream_t st = nullptr;  // the null stream: one build at a time, synchronised at the end
    // inputs
    double *d_verts = mb_alloc<double>(3 * (size_t)nv);
    int32_t *d_tets = mb_alloc<int32_t>(4 * (size_t)nt);
    int32_t *d_stris = mb_alloc<int32_t>(3 * (size_t)ns);
    std::vector<void *> tmp = {d_verts, d_tets, d_stris};
    try {
        MB_HIP(hipMemcpy(d_verts, verts, 3 * (size_t)nv * sizeof(double), hipMemcpyHostToDevice));
        MB_HIP(hipMemcpy(d_tets, tets, 4 * (size_t)nt * sizeof(int32_t), hipMemcpyHostToDevice));
        if (ns) MB_HIP(hipMemcpy(d_stris, stris, 3 * (size_t)ns * sizeof(int32_t), hipMemcpyHostToDevice));
        int32_t *d_err = mb_alloc<int32_t>(1);
        tmp.push_back(d_err);
        MB_HIP(hipMemset(d_err, 0, sizeof(int32_t)));
        m->tet_vol = m->alloc<double>(nt);
        m->tet_bary = m->alloc<double>(3 * (size_t)nt);
        hipLaunchKernelGGL(k_tet_geom, grid_for(nt), dim3(kB), 0, st, nt, d_verts, d_tets, m->tet_vol,
                           m->tet_bary, d_err);
        // face keys and the two stable sorts
        uint64_t *hi = mb_alloc<uint64_t>(nf), *hi2 = mb_alloc<uint64_t>(nf);
        uint32_t *lo = mb_alloc<uint32_t>(nf), *lo2 = mb_alloc<uint32_t>(nf);
        int32_t *id = mb_alloc<int32_t>(nf), *id2 = mb_alloc<int32_t>(nf);
        tmp.insert(tmp.end(), {hi, hi2, lo, lo2, id, id2});
        hipLaunchKernelGGL(k_face_keys, grid_for(nf), dim3(kB), 0, st, nf, ns, nv, d_tets, d_stris, hi, lo, id);
        unsigned hi_bits = 1;
        while (hi_bits < 64 && ((uint64_t)1 << hi_bits) < (uint64_t)nv * (uint64_t)nv) ++hi_bits;
        unsigned lo_bits = 1;
        while (lo_bits < 32 && ((uint64_t)1 << lo_bits) < (uint64_t)nv) ++lo_bits;
        int32_t *start = mb_alloc<int32_t>(nf), *mark = mb_alloc<int32_t>(nf), *tri_no = mb_alloc<int32_t>(nf);
        tmp.insert(tmp.end(), {start, mark, tri_no});
        size_t tb1 = 0, tb2 = 0, tb3 = 0;
        MB_HIP(rocprim::radix_sort_pairs(nullptr, tb1, lo, lo2, id, id2, (size_t)nf, 0u, lo_bits, st));
        MB_HIP(rocprim::radix_sort_pairs(nullptr, tb2, hi2, hi, id2, id, (size_t)nf, 0u, hi_bits, st));
        MB_HIP(rocprim::exclusive_scan(nullptr, tb3, mark, tri_no, 0, (size_t)nf, rocprim::plus<int32_t>(), st));
        void *temp = mb_alloc<unsigned char>(std::max(tb1, std::max(tb2, tb3)));
        tmp.push_back(temp);
        // by c (stable: position order within equal c) ...
        MB_HIP(rocprim::radix_sort_pairs(temp, tb1, lo, lo2, id, id2, (size_t)nf, 0u, lo_bits, st));
        // ... then by a * nv + b, stable, so the order is (a, b, c), ties in position order
        hipLaunchKernelGGL(k_gather<uint64_t>, grid_for(nf), dim3(kB), 0, st, nf, hi, id2, hi2);
        MB_HIP(rocprim::radix_sort_pairs(temp, tb2, hi2, hi, id2, id, (size_t)nf, 0u, hi_bits, st));
        // c in the final order (the sorts leave their input keys untouched)
        hipLaunchKernelGGL(k_gather<uint32_t>, grid_for(nf), dim3(kB), 0, st, nf, (const uint32_t *)lo,
                           (const int32_t *)id, lo2);
        hipLaunchKernelGGL(k_fill, grid_for(nf), dim3(kB), 0, st, nf, mark, 0);
        hipLaunchKernelGGL(k_starts, grid_for(nf), dim3(kB), 0, st, nf, hi, lo2, (const int32_t *)id, start, mark);
        // triangle number of each first appearance: the marks before it
        MB_HIP(rocprim::exclusive_scan(temp, tb3, mark, tri_no, 0, (size_t)nf, rocprim::plus<int32_t>(), st));
        int32_t last_no = 0, last_mark = 0;
        MB_HIP(hipMemcpy(&last_no, tri_no + (nf - 1), sizeof(int32_t), hipMemcpyDeviceToHost));
        MB_HIP(hipMemcpy(&last_mark, mark + (nf - 1), sizeof(int32_t), hipMemcpyDeviceToHost));
        const int32_t ntri = last_no + last_mark;
        int32_t herr = 0;
        MB_HIP(hipMemcpy(&herr, d_err, sizeof(int32_t), hipMemcpyDeviceToHost));
        if (herr & 1) throw MbError{STG_ERR_ARG, "degenerate tetrahedron"};
        m->ntri = ntri;
        m->tris = m->alloc<int32_t>(3 * (size_t)ntri);
        m->tri_tets = m->alloc<int32_t>(2 * (size_t)ntri);
        m->tet_tri = m->alloc<int32_t>(nq);
        m->nbr = m->alloc<int32_t>(nq);
        m->nbr_face = m->alloc<int32_t>(nq);
        hipLaunchKernelGGL(k_fill, grid_for(nq), dim3(kB), 0, st, nq, m->nbr, -1);
        hipLaunchKernelGGL(k_fill, grid_for(nq), dim3(kB), 0, st, nq, m->nbr_face, -1);
        hipLaunchKernelGGL(k_tri_build, grid_for(nf), dim3(kB), 0, st, nf, ns, nv, (const uint64_t *)hi,
                           (const uint32_t *)lo2, (const int32_t *)id, (const int32_t *)start,
                           (const int32_t *)tri_no, m->tris, m->tri_tets, m->tet_tri, m->nbr, m->nbr_face, d_err);
        MB_HIP(hipMemcpy(&herr, d_err, sizeof(int32_t), hipMemcpyDeviceToHost));
        if (herr & 2) throw MbError{STG_ERR_ARG, "non-manifold mesh: a triangle is shared by more than two tets"};
        if (herr & 4) throw MbError{STG_ERR_ARG, "a supplied triangle is not a face of any tetrahedron"};
        m->tri_area = m->alloc<double>(ntri);
        m->tri_bary = m->alloc<double>(3 * (size_t)ntri);
        hipLaunchKernelGGL(k_tri_geom, grid_for(ntri), dim3(kB), 0, st, ntri, d_verts, m->tris, m->tri_area,
                           m->tri_bary);
        m->face_area = m->alloc<double>(nq);
        m->face_dist = m->alloc<double>(nq);
        hipLaunchKernelGGL(k_face_measures, grid_for(nq), dim3(kB), 0, st, nq, m->tet_tri, m->tri_area, m->nbr,
                           m->tet_bary, m->face_area, m->face_dist);
        MB_HIP(hipGetLastError());
        MB_HIP(hipDeviceSynchronize());
    } catch (...) {
        for (void *p : tmp) (void)hipFree(p);
        throw;
    }
    for (void *p : tmp) (void)hipFree(p);
}

template <typename T>
void fetch(T *dst, const T *src, size_t n) {
    if (dst && n) MB_HIP(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyDeviceToHost));
}

}  // namespace

extern "C" {

int stg_mesh_build(const double *verts, int32_t n_verts, const int32_t *tets, int32_t n_tets, const int32_t *tris,
                   int32_t n_tris, stg_meshbuild **out) {
    if (!out) return STG_ERR_ARG;
    *out = nullptr;
    stg_meshbuild *m = nullptr;
    try {
        m = new stg_meshbuild();
    } catch (const std::bad_alloc &) {
        g_mb_error = "host allocation failed";
        return STG_ERR_SYS;
    }
    const int rc = mb_guard(m, [&] { build(m, verts, n_verts, tets, n_tets, tris, n_tris); });
    if (rc != STG_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return STG_OK;
}

int stg_mesh_build_sizes(const stg_meshbuild *m, int32_t *n_tets, int32_t *n_tris) {
    if (!m || !n_tets || !n_tris) return STG_ERR_ARG;
    *n_tets = m->nt;
    *n_tris = m->ntri;
    return STG_OK;
}

int stg_mesh_build_fetch(stg_meshbuild *m, double *tet_vol, double *tet_bary, int32_t *tet_tri,
                         int32_t *tet_nbr, int32_t *tet_nbr_face, double *tet_face_area,
                         double *tet_face_dist, int32_t *tris, int32_t *tri_tets, double *tri_area,
                         double *tri_bary) {
    if (!m) return STG_ERR_ARG;
    return mb_guard(m, [&] {
        const size_t nt = (size_t)m->nt, nf = 4 * nt, ntri = (size_t)m->ntri;
        fetch(tet_vol, m->tet_vol, nt);
        fetch(tet_bary, m->tet_bary, 3 * nt);
        fetch(tet_tri, m->tet_tri, nf);
        fetch(tet_nbr, m->nbr, nf);
        fetch(tet_nbr_face, m->nbr_face, nf);
        fetch(tet_face_area, m->face_area, nf);
        fetch(tet_face_dist, m->face_dist, nf);
        fetch(tris, m->tris, 3 * ntri);
        fetch(tri_tets, m->tri_tets, 2 * ntri);
        fetch(tri_area, m->tri_area, ntri);
        fetch(tri_bary, m->tri_bary, 3 * ntri);
    });
}

int stg_mesh_build_destroy(stg_meshbuild *m) {
    delete m;
    return STG_OK;
}

const char *stg_mesh_build_last_error(void) { return g_mb_error.c_str(); }

}  // extern "C"
