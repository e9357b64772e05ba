// Per-element field operations shared by the one-launch-per-op kernels
// (pic_fields.hip) and the batched one-block-per-simulation field step
// (pic_batch.hip).  Each function handles one element id `t` of its index
// space; callers map threads to ids.  The scalar type T is float for the
// FP32 path and double for the FP64 path.
#pragma once
#include "pic_common.h"

namespace pic {

__device__ inline int64_t slab_to_index(const PicDesc& d, int axis, int pos, int a, int b) {
    if (axis == 0) return fidx(d, pos, a, b);
    if (axis == 1) return fidx(d, a, pos, b);
    return fidx(d, a, b, pos);
}

__host__ __device__ inline void other_dims(const PicDesc& d, int axis, int& da, int& db) {
    da = d.dims[axis == 0 ? 1 : 0];
    db = d.dims[axis == 2 ? 1 : 2];
}

// Number of element ids of each operation on one axis.
__host__ __device__ inline int64_t refresh_elems(const PicDesc& d, int axis) {
    int da, db;
    other_dims(d, axis, da, db);
    return (int64_t)2 * d.g * da * db;
}
__host__ __device__ inline int64_t fold_elems(const PicDesc& d, int axis) {
    int da, db;
    other_dims(d, axis, da, db);
    return d.n[axis] == 1 ? (int64_t)da * db : (int64_t)d.g * da * db;
}

__device__ inline void atomic_add_t(float* p, float v) { unsafeAtomicAdd(p, v); }
__device__ inline void atomic_add_t(double* p, double v) { unsafeAtomicAdd(p, v); }

// ghost_cells.py:142-196: ghost element t (< refresh_elems) of one axis.
template <class T>
__device__ inline void refresh_elem(const PicDesc& d, T* f, int axis, int mode, int64_t t) {
    int g = d.g, n = d.n[axis];
    int da, db;
    other_dims(d, axis, da, db);
    int b = (int)(t % db);
    int64_t r = t / db;
    int a = (int)(r % da);
    int l = (int)(r / da);
    int dst = l < g ? l : n + l;          // upper ghosts n+g .. n+2g-1
    T v = (T)0;
    if (mode == MODE_REDUCED_COPY) {
        v = f[slab_to_index(d, axis, g, a, b)];
    } else if (mode == MODE_PERIODIC) {
        int src = l < g ? n + l : l;      // lower l <- n+l ; upper (l-g) <- g+(l-g)
        v = f[slab_to_index(d, axis, src, a, b)];
    }
    f[slab_to_index(d, axis, dst, a, b)] = v;
}

// ghost_cells.py:218-289: fold element t (< fold_elems) of a non-reduced axis.
// bc: 0 periodic self-exchange, 1 conducting reflection, other: dropped.
template <class T>
__device__ inline void fold_elem(const PicDesc& d, T* f, int axis, int bc, int64_t t) {
    int g = d.g, n = d.n[axis];
    int da, db;
    other_dims(d, axis, da, db);
    int b = (int)(t % db);
    int64_t r = t / db;
    int a = (int)(r % da);
    int l = (int)(r / da);
    int64_t lo_i = slab_to_index(d, axis, l, a, b);
    int64_t hi_i = slab_to_index(d, axis, n + g + l, a, b);
    T lo = f[lo_i];
    T hi = f[hi_i];
    f[lo_i] = (T)0;
    f[hi_i] = (T)0;
    if (bc == 0) {
        atomic_add_t(&f[slab_to_index(d, axis, n + l, a, b)], lo);
        atomic_add_t(&f[slab_to_index(d, axis, g + l, a, b)], hi);
    } else if (bc == 1) {
        atomic_add_t(&f[slab_to_index(d, axis, g + l, a, b)], -lo);
        atomic_add_t(&f[slab_to_index(d, axis, n + l, a, b)], -hi);
    }
}

// ghost_cells.py:218-235: fold element t of a reduced (single-cell) axis.
template <class T>
__device__ inline void fold_reduced_elem(const PicDesc& d, T* f, int axis, int bc, int64_t t) {
    int g = d.g;
    int da, db;
    other_dims(d, axis, da, db);
    int b = (int)(t % db);
    int a = (int)(t / db);
    T lower = (T)0, upper = (T)0;
    for (int l = 0; l < g; ++l) {
        int64_t i0 = slab_to_index(d, axis, l, a, b);
        int64_t i1 = slab_to_index(d, axis, 1 + g + l, a, b);
        lower += f[i0];
        upper += f[i1];
        f[i0] = (T)0;
        f[i1] = (T)0;
    }
    T s = lower + upper;
    int64_t c = slab_to_index(d, axis, g, a, b);
    if (bc == 0) f[c] += s;
    else if (bc == 1) f[c] += -s;
}

// Value of a refreshed copy of f at (i,j,k) where exactly one coordinate may
// sit one cell outside the interior along `axis`.
template <class T>
__device__ inline T refreshed_at(const T* __restrict__ f, const PicDesc& d, const int mode[3], int i, int j, int k,
                                 int axis) {
    int c[3] = {i, j, k};
    int src = ghost_source(c[axis], d.n[axis], d.g, mode[axis]);
    if (src < 0) return (T)0;
    c[axis] = src;
    return f[fidx(d, c[0], c[1], c[2])];
}

template <class T>
struct YeeConstT {
    T hdt, dt, c2, inv_eps, id[3];
    int mode[3];
};
using YeeConst = YeeConstT<float>;

template <class T = float>
__host__ __device__ inline YeeConstT<T> yee_const(const PicDesc& d) {
    YeeConstT<T> y;
    y.hdt = (T)(d.dt / 2);
    y.dt = (T)d.dt;
    y.c2 = (T)(d.C * d.C);
    y.inv_eps = (T)(1.0 / d.eps);
    for (int a = 0; a < 3; ++a) {
        y.id[a] = (T)(1.0 / d.d[a]);
        y.mode[a] = refresh_mode(d, a, 0);
    }
    return y;
}

// first_order_yee.py:96-142 at interior array point (i, j, k):
// B_out = B_in - (dt/2) curl E with forward differences.
template <class T>
__device__ inline void yee_b_point(const PicDesc& d, const YeeConstT<T>& y, const T* const E[3], const T* const B[3],
                                   T* const out[3], int i, int j, int k) {
    int64_t c = fidx(d, i, j, k);
    T ex = E[0][c], ey = E[1][c], ez = E[2][c];
    T ez_yp = refreshed_at(E[2], d, y.mode, i, j + 1, k, 1);
    T ey_zp = refreshed_at(E[1], d, y.mode, i, j, k + 1, 2);
    T ex_zp = refreshed_at(E[0], d, y.mode, i, j, k + 1, 2);
    T ez_xp = refreshed_at(E[2], d, y.mode, i + 1, j, k, 0);
    T ey_xp = refreshed_at(E[1], d, y.mode, i + 1, j, k, 0);
    T ex_yp = refreshed_at(E[0], d, y.mode, i, j + 1, k, 1);
    T curl_x = (ez_yp - ez) * y.id[1] - (ey_zp - ey) * y.id[2];
    T curl_y = (ex_zp - ex) * y.id[2] - (ez_xp - ez) * y.id[0];
    T curl_z = (ey_xp - ey) * y.id[0] - (ex_yp - ex) * y.id[1];
    out[0][c] = B[0][c] - y.hdt * curl_x;
    out[1][c] = B[1][c] - y.hdt * curl_y;
    out[2][c] = B[2][c] - y.hdt * curl_z;
}

// first_order_yee.py:12-72 at interior array point (i, j, k):
// E_out = E_in + dt (C^2 curl B - J/eps) with backward differences.
template <class T>
__device__ inline void yee_e_point(const PicDesc& d, const YeeConstT<T>& y, const T* const E[3], const T* const B[3],
                                   const T* const J[3], T* const out[3], int i, int j, int k) {
    int64_t c = fidx(d, i, j, k);
    T bx = B[0][c], by = B[1][c], bz = B[2][c];
    T bz_ym = refreshed_at(B[2], d, y.mode, i, j - 1, k, 1);
    T by_zm = refreshed_at(B[1], d, y.mode, i, j, k - 1, 2);
    T bx_zm = refreshed_at(B[0], d, y.mode, i, j, k - 1, 2);
    T bz_xm = refreshed_at(B[2], d, y.mode, i - 1, j, k, 0);
    T by_xm = refreshed_at(B[1], d, y.mode, i - 1, j, k, 0);
    T bx_ym = refreshed_at(B[0], d, y.mode, i, j - 1, k, 1);
    T curl_x = (bz - bz_ym) * y.id[1] - (by - by_zm) * y.id[2];
    T curl_y = (bx - bx_zm) * y.id[2] - (bz - bz_xm) * y.id[0];
    T curl_z = (by - by_xm) * y.id[0] - (bx - bx_ym) * y.id[1];
    out[0][c] = E[0][c] + (y.c2 * curl_x - J[0][c] * y.inv_eps) * y.dt;
    out[1][c] = E[1][c] + (y.c2 * curl_y - J[1][c] * y.inv_eps) * y.dt;
    out[2][c] = E[2][c] + (y.c2 * curl_z - J[2][c] * y.inv_eps) * y.dt;
}

// PML coordinate stretch of the six curl derivatives (PML.py:251-359): for a
// derivative d along axis a at interior point (i, j, k), with b = exp(-sigma_a
// dt) from the ghosted profile of axis a, the ADE memory becomes
// m' = b m + (b - 1) d and the derivative d + m'.  Derivative order (and memory
// order): d?z/dy, d?y/dz, d?x/dz, d?z/dx, d?y/dx, d?x/dy; curl = (d0 - d1,
// d2 - d3, d4 - d5).  mem[] are interior (Nx, Ny, Nz) arrays updated in place.
// line[a] (optional) holds profile a along its own axis, for profiles that do
// not vary along the other two: one cached value per point instead of three
// reads of field-shaped arrays.
struct PmlArgs {
    float* mem[6];
    const float* sigma[3];
    const float* line[3];
    float dt;
};

__device__ inline void pml_stretch6(const PicDesc& d, const PmlArgs& pml, int i, int j, int k, float dv[6]) {
    const int64_t c = fidx(d, i, j, k);
    const int64_t mi = ((int64_t)(i - d.g) * d.n[1] + (j - d.g)) * d.n[2] + (k - d.g);
    const int idx[3] = {i, j, k};
    float b[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) b[a] = expf(-(pml.line[a] ? pml.line[a][idx[a]] : pml.sigma[a][c]) * pml.dt);
    constexpr int kAxis[6] = {1, 2, 2, 0, 0, 1};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const float bb = b[kAxis[q]];
        const float m = bb * pml.mem[q][mi] + (bb - 1.0f) * dv[q];
        if (bb != 1.0f) pml.mem[q][mi] = m;     // sigma = 0: m' = m, nothing to write
        dv[q] += m;
    }
}

// yee_b_point with the PML stretch of the E derivatives (first_order_yee.py:127-142).
__device__ inline void yee_b_point_pml(const PicDesc& d, const YeeConst& y, const PmlArgs& pml,
                                       const float* const E[3], const float* const B[3], float* const out[3], int i,
                                       int j, int k) {
    int64_t c = fidx(d, i, j, k);
    float ex = E[0][c], ey = E[1][c], ez = E[2][c];
    float dv[6] = {(refreshed_at(E[2], d, y.mode, i, j + 1, k, 1) - ez) * y.id[1],
                   (refreshed_at(E[1], d, y.mode, i, j, k + 1, 2) - ey) * y.id[2],
                   (refreshed_at(E[0], d, y.mode, i, j, k + 1, 2) - ex) * y.id[2],
                   (refreshed_at(E[2], d, y.mode, i + 1, j, k, 0) - ez) * y.id[0],
                   (refreshed_at(E[1], d, y.mode, i + 1, j, k, 0) - ey) * y.id[0],
                   (refreshed_at(E[0], d, y.mode, i, j + 1, k, 1) - ex) * y.id[1]};
    pml_stretch6(d, pml, i, j, k, dv);
    out[0][c] = B[0][c] - y.hdt * (dv[0] - dv[1]);
    out[1][c] = B[1][c] - y.hdt * (dv[2] - dv[3]);
    out[2][c] = B[2][c] - y.hdt * (dv[4] - dv[5]);
}

// yee_e_point with the PML stretch of the B derivatives (first_order_yee.py:49-72).
__device__ inline void yee_e_point_pml(const PicDesc& d, const YeeConst& y, const PmlArgs& pml,
                                       const float* const E[3], const float* const B[3], const float* const J[3],
                                       float* const out[3], int i, int j, int k) {
    int64_t c = fidx(d, i, j, k);
    float bx = B[0][c], by = B[1][c], bz = B[2][c];
    float dv[6] = {(bz - refreshed_at(B[2], d, y.mode, i, j - 1, k, 1)) * y.id[1],
                   (by - refreshed_at(B[1], d, y.mode, i, j, k - 1, 2)) * y.id[2],
                   (bx - refreshed_at(B[0], d, y.mode, i, j, k - 1, 2)) * y.id[2],
                   (bz - refreshed_at(B[2], d, y.mode, i - 1, j, k, 0)) * y.id[0],
                   (by - refreshed_at(B[1], d, y.mode, i - 1, j, k, 0)) * y.id[0],
                   (bx - refreshed_at(B[0], d, y.mode, i, j - 1, k, 1)) * y.id[1]};
    pml_stretch6(d, pml, i, j, k, dv);
    out[0][c] = E[0][c] + (y.c2 * (dv[0] - dv[1]) - J[0][c] * y.inv_eps) * y.dt;
    out[1][c] = E[1][c] + (y.c2 * (dv[2] - dv[3]) - J[1][c] * y.inv_eps) * y.dt;
    out[2][c] = E[2][c] + (y.c2 * (dv[4] - dv[5]) - J[2][c] * y.inv_eps) * y.dt;
}

}  // namespace pic
