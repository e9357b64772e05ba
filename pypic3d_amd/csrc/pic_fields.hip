// Field kernels: fill, ghost refresh/fold, conducting walls, filters, Yee
// curl updates and the electrostatic centred gradient.
//
// The Yee kernels read neighbours through ghost_source(), i.e. they see the
// values a freshly refreshed copy of the input would hold
// (first_order_yee.py:31, :115 refresh before differencing), so the caller's
// input arrays are never modified -- the reference is functional.
#include "pic_field_ops.h"

namespace {

using namespace pic;

constexpr int kBlock = 256;

template <class T>
__global__ void fill_kernel(T* __restrict__ p, int64_t n, T v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// ghost_cells.py:142-196: one axis of the refresh, for up to three fields.
template <class T>
__global__ void refresh_axis_kernel(PicDesc d, T* f0, T* f1, T* f2, int axis, int mode) {
    const int64_t per = refresh_elems(d, axis);
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per * 3) return;
    int which = (int)(t / per);
    T* f = which == 0 ? f0 : (which == 1 ? f1 : f2);
    if (f == nullptr) return;
    refresh_elem(d, f, axis, mode, t - which * per);
}

// ghost_cells.py:218-289: one axis of the fold (non-reduced axis).
template <class T>
__global__ void fold_axis_kernel(PicDesc d, T* f, int axis, int bc) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= fold_elems(d, axis)) return;
    fold_elem(d, f, axis, bc, t);
}

// ghost_cells.py:218-235: fold of a reduced (single-cell) axis.
template <class T>
__global__ void fold_reduced_kernel(PicDesc d, T* f, int axis, int bc) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= fold_elems(d, axis)) return;
    fold_reduced_elem(d, f, axis, bc, t);
}

// Zero the wall planes g and n+g-1 of one axis (ghost_cells.py:344-362).
template <class T>
__global__ void zero_planes_kernel(PicDesc d, T* f, int axis) {
    int da, db;
    other_dims(d, axis, da, db);
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)2 * da * db) return;
    int which = (int)(t / ((int64_t)da * db));
    int64_t r = t - (int64_t)which * da * db;
    int b = (int)(r % db);
    int a = (int)(r / db);
    int pos = which == 0 ? d.g : d.n[axis] + d.g - 1;
    f[slab_to_index(d, axis, pos, a, b)] = (T)0;
}

// Fill the ghosts of one axis from the adjacent interior plane
// (ghost_cells.py:365-386).
template <class T>
__global__ void constant_ghosts_kernel(PicDesc d, T* f, int axis) {
    int g = d.g, n = d.n[axis];
    int da, db;
    other_dims(d, axis, da, db);
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)2 * g * da * db) return;
    int b = (int)(t % db);
    int64_t r = t / db;
    int a = (int)(r % da);
    int l = (int)(r / da);
    int dst = l < g ? l : n + l;
    int src = l < g ? g : n + g - 1;
    f[slab_to_index(d, axis, dst, a, b)] = f[slab_to_index(d, axis, src, a, b)];
}

// filters.py:234-291: interior <- 3x3x3 correlation; ghosts copied.
template <bool kBilinear, class T>
__global__ void filter_kernel(PicDesc d, const T* __restrict__ in, T* __restrict__ out, T alpha) {
    int64_t total = field_size(d);
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int k = (int)(t % d.dims[2]);
    int64_t r = t / d.dims[2];
    int j = (int)(r % d.dims[1]);
    int i = (int)(r / d.dims[1]);
    int g = d.g;
    bool inside = i >= g && i < d.n[0] + g && j >= g && j < d.n[1] + g && k >= g && k < d.n[2] + g;
    if (!inside) {
        out[t] = in[t];
        return;
    }
    const int64_t sx = (int64_t)d.dims[1] * d.dims[2], sy = d.dims[2];
    if (kBilinear) {
        const T w1[3] = {(T)1, (T)2, (T)1};
        T acc = (T)0;
        for (int a = -1; a <= 1; ++a)
            for (int b = -1; b <= 1; ++b)
                for (int c = -1; c <= 1; ++c)
                    acc += (w1[a + 1] * w1[b + 1] * w1[c + 1] / (T)64) * in[t + a * sx + b * sy + c];
        out[t] = acc;
    } else {
        T nw = ((T)1 - alpha) / (T)6;
        T faces = in[t - sx] + in[t + sx] + in[t - sy] + in[t + sy] + in[t - 1] + in[t + 1];
        out[t] = alpha * in[t] + nw * faces;
    }
}

template <class T>
struct F3T {
    const T* p[3];
};
template <class T>
struct W3T {
    T* p[3];
};
using F3 = F3T<float>;
using W3f = W3T<float>;
struct W6f {
    float* p[6];
};

// first_order_yee.py:96-142: B_out = B_in - (dt/2) curl E (forward diffs).
template <class T>
__global__ void yee_b_kernel(PicDesc d, YeeConstT<T> y, F3T<T> E, F3T<T> B, W3T<T> out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.n[2]) return;
    yee_b_point(d, y, E.p, B.p, out.p, blockIdx.z + d.g, blockIdx.y + d.g, k + d.g);
}

// first_order_yee.py:12-72: E_out = E_in + dt (C^2 curl B - J/eps) (backward diffs).
template <class T>
__global__ void yee_e_kernel(PicDesc d, YeeConstT<T> y, F3T<T> E, F3T<T> B, F3T<T> J, W3T<T> out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.n[2]) return;
    yee_e_point(d, y, E.p, B.p, J.p, out.p, blockIdx.z + d.g, blockIdx.y + d.g, k + d.g);
}

__global__ void yee_b_pml_kernel(PicDesc d, YeeConst y, PmlArgs pml, F3 E, F3 B, W3f out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.n[2]) return;
    yee_b_point_pml(d, y, pml, E.p, B.p, out.p, blockIdx.z + d.g, blockIdx.y + d.g, k + d.g);
}

__global__ void yee_e_pml_kernel(PicDesc d, YeeConst y, PmlArgs pml, F3 E, F3 B, F3 J, W3f out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.n[2]) return;
    yee_e_point_pml(d, y, pml, E.p, B.p, J.p, out.p, blockIdx.z + d.g, blockIdx.y + d.g, k + d.g);
}

// stretch_spatial_derivative over six interior derivative arrays in place
// (apply_tiled_pml_to_e_curl / _b_curl without the curl assembly).
__global__ void pml_stretch_kernel(PicDesc d, PmlArgs pml, W6f dv) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.n[2]) return;
    const int i = blockIdx.z + d.g, j = blockIdx.y + d.g;
    const int64_t mi = ((int64_t)blockIdx.z * d.n[1] + blockIdx.y) * d.n[2] + k;
    float v[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) v[q] = dv.p[q][mi];
    pml_stretch6(d, pml, i, j, k + d.g, v);
#pragma unroll
    for (int q = 0; q < 6; ++q) dv.p[q][mi] = v[q];
}

// electrostatic_yee.py:212-246: E interior = -(phi[+1]-phi[-1])/(2d), reading
// phi's own (already boundary-applied) ghosts.
__global__ void efield_from_phi_kernel(PicDesc d, const float* __restrict__ phi, float* __restrict__ ex,
                                       float* __restrict__ ey, float* __restrict__ ez, float h2x, float h2y,
                                       float h2z) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    int i = blockIdx.z;
    if (k >= d.n[2]) return;
    i += d.g;
    j += d.g;
    k += d.g;
    int64_t c = fidx(d, i, j, k);
    const int64_t sx = (int64_t)d.dims[1] * d.dims[2], sy = d.dims[2];
    ex[c] = -1.0f * (phi[c + sx] - phi[c - sx]) * h2x;
    ey[c] = -1.0f * (phi[c + sy] - phi[c - sy]) * h2y;
    ez[c] = -1.0f * (phi[c + 1] - phi[c - 1]) * h2z;
}

bool desc_ok(const PicDesc* d) {
    if (!d) return false;
    if (d->g < 1) return false;
    for (int a = 0; a < 3; ++a) {
        if (d->n[a] < 1 || d->dims[a] != d->n[a] + 2 * d->g) return false;
        // a non-reduced axis narrower than the guard band would alias slabs
        if (d->n[a] > 1 && d->n[a] < d->g) return false;
    }
    return true;
}

}  // namespace

// Host-side launch sequences, for the FP32 (float) and FP64 (double) fields.
template <class T>
static int refresh_t(const PicDesc* d, T* f0, T* f1, T* f2, int bc_type, hipStream_t s) {
    for (int axis = 0; axis < 3; ++axis) {
        int da = d->dims[axis == 0 ? 1 : 0], db = d->dims[axis == 2 ? 1 : 2];
        int64_t work = (int64_t)2 * d->g * da * db * 3;
        hipLaunchKernelGGL(refresh_axis_kernel<T>, dim3(grid_1d(work, kBlock)), dim3(kBlock), 0, s, *d, f0, f1, f2,
                           axis, refresh_mode(*d, axis, bc_type));
        PIC_CHECK_LAUNCH();
    }
    return 0;
}

// Shared with other translation units.
int pic_refresh_fields(const PicDesc* d, float* f0, float* f1, float* f2, int bc_type, hipStream_t s) {
    return refresh_t<float>(d, f0, f1, f2, bc_type, s);
}

static dim3 interior_grid(const PicDesc* d, int block) {
    return dim3((unsigned)((d->n[2] + block - 1) / block), (unsigned)d->n[1], (unsigned)d->n[0]);
}

static int interior_block(const PicDesc* d) { return d->n[2] >= 128 ? 128 : 64; }

template <class T>
static int fill_t(T* ptr, int64_t n, T value, hipStream_t s) {
    PIC_REQUIRE(ptr != nullptr && n >= 0);
    if (n == 0) return 0;
    int blocks = grid_1d(n, kBlock);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(fill_kernel<T>, dim3(blocks), dim3(kBlock), 0, s, ptr, n, value);
    PIC_CHECK_LAUNCH();
    return 0;
}

template <class T>
static int halo_refresh_t(const PicDesc* desc, T* field, int bc_type, hipStream_t s) {
    PIC_REQUIRE(desc_ok(desc) && field && (bc_type == 0 || bc_type == 1));
    return refresh_t<T>(desc, field, nullptr, nullptr, bc_type, s);
}

template <class T>
static int halo_fold_t(const PicDesc* desc, T* field, int bc_type, hipStream_t s) {
    PIC_REQUIRE(desc_ok(desc) && field && (bc_type == 0 || bc_type == 1));
    for (int axis = 0; axis < 3; ++axis) {
        int bc = bc_type == 0 ? desc->bc[axis] : desc->pbc[axis];
        int da = desc->dims[axis == 0 ? 1 : 0], db = desc->dims[axis == 2 ? 1 : 2];
        if (desc->n[axis] == 1) {
            hipLaunchKernelGGL(fold_reduced_kernel<T>, dim3(grid_1d((int64_t)da * db, kBlock)), dim3(kBlock), 0, s,
                               *desc, field, axis, bc);
        } else {
            hipLaunchKernelGGL(fold_axis_kernel<T>, dim3(grid_1d((int64_t)desc->g * da * db, kBlock)), dim3(kBlock),
                               0, s, *desc, field, axis, bc);
        }
        PIC_CHECK_LAUNCH();
    }
    return 0;
}

template <class T>
static int zero_boundary_t(const PicDesc* desc, T* field, int axis, hipStream_t s) {
    PIC_REQUIRE(desc_ok(desc) && field && axis >= 0 && axis < 3);
    if (desc->bc[axis] == 1) {
        int da = desc->dims[axis == 0 ? 1 : 0], db = desc->dims[axis == 2 ? 1 : 2];
        hipLaunchKernelGGL(zero_planes_kernel<T>, dim3(grid_1d((int64_t)2 * da * db, kBlock)), dim3(kBlock), 0, s,
                           *desc, field, axis);
        PIC_CHECK_LAUNCH();
    }
    return refresh_t<T>(desc, field, nullptr, nullptr, 0, s);
}

template <class T>
static int constant_boundary_t(const PicDesc* desc, T* field, int axis, hipStream_t s) {
    PIC_REQUIRE(desc_ok(desc) && field && axis >= 0 && axis < 3);
    int rc = refresh_t<T>(desc, field, nullptr, nullptr, 0, s);
    if (rc) return rc;
    if (desc->bc[axis] == 1) {
        int da = desc->dims[axis == 0 ? 1 : 0], db = desc->dims[axis == 2 ? 1 : 2];
        hipLaunchKernelGGL(constant_ghosts_kernel<T>, dim3(grid_1d((int64_t)2 * desc->g * da * db, kBlock)),
                           dim3(kBlock), 0, s, *desc, field, axis);
        PIC_CHECK_LAUNCH();
    }
    return 0;
}

template <bool kBilinear, class T>
static int filter_t(const PicDesc* desc, const T* in, T* out, double alpha, hipStream_t s) {
    PIC_REQUIRE(desc_ok(desc) && in && out && in != out);
    int64_t n = field_size(*desc);
    hipLaunchKernelGGL((filter_kernel<kBilinear, T>), dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, s, *desc, in, out,
                       (T)alpha);
    PIC_CHECK_LAUNCH();
    return 0;
}

template <class T>
static int yee_b_t(const PicDesc* desc, const T* const E[3], const T* const B_in[3], T* const B_out[3],
                   hipStream_t s) {
    PIC_REQUIRE(desc_ok(desc) && E && B_in && B_out);
    for (int c = 0; c < 3; ++c) PIC_REQUIRE(E[c] && B_in[c] && B_out[c]);
    for (int c = 0; c < 3; ++c)
        for (int e = 0; e < 3; ++e) PIC_REQUIRE((const T*)B_out[c] != E[e]);
    int block = interior_block(desc);
    hipLaunchKernelGGL(yee_b_kernel<T>, interior_grid(desc, block), dim3(block), 0, s, *desc, yee_const<T>(*desc),
                       F3T<T>{{E[0], E[1], E[2]}}, F3T<T>{{B_in[0], B_in[1], B_in[2]}},
                       W3T<T>{{B_out[0], B_out[1], B_out[2]}});
    PIC_CHECK_LAUNCH();
    return refresh_t<T>(desc, B_out[0], B_out[1], B_out[2], 0, s);
}

template <class T>
static int yee_e_t(const PicDesc* desc, const T* const E_in[3], const T* const B[3], const T* const J[3],
                   T* const E_out[3], hipStream_t s) {
    PIC_REQUIRE(desc_ok(desc) && E_in && B && J && E_out);
    for (int c = 0; c < 3; ++c) PIC_REQUIRE(E_in[c] && B[c] && J[c] && E_out[c]);
    for (int c = 0; c < 3; ++c)
        for (int e = 0; e < 3; ++e) PIC_REQUIRE((const T*)E_out[c] != B[e] && (const T*)E_out[c] != J[e]);
    int block = interior_block(desc);
    hipLaunchKernelGGL(yee_e_kernel<T>, interior_grid(desc, block), dim3(block), 0, s, *desc, yee_const<T>(*desc),
                       F3T<T>{{E_in[0], E_in[1], E_in[2]}}, F3T<T>{{B[0], B[1], B[2]}}, F3T<T>{{J[0], J[1], J[2]}},
                       W3T<T>{{E_out[0], E_out[1], E_out[2]}});
    PIC_CHECK_LAUNCH();
    return refresh_t<T>(desc, E_out[0], E_out[1], E_out[2], 0, s);
}

extern "C" {

int pic_abi_version(void) { return PIC_ABI_VERSION; }

int pic_fill_f32(float* ptr, int64_t n, float value, void* stream) { return fill_t(ptr, n, value, as_stream(stream)); }

int pic_halo_refresh(const PicDesc* desc, float* field, int bc_type, void* stream) {
    return halo_refresh_t(desc, field, bc_type, as_stream(stream));
}

int pic_halo_refresh_f64(const PicDesc* desc, double* field, int bc_type, void* stream) {
    return halo_refresh_t(desc, field, bc_type, as_stream(stream));
}

int pic_halo_fold(const PicDesc* desc, float* field, int bc_type, void* stream) {
    return halo_fold_t(desc, field, bc_type, as_stream(stream));
}

int pic_halo_fold_f64(const PicDesc* desc, double* field, int bc_type, void* stream) {
    return halo_fold_t(desc, field, bc_type, as_stream(stream));
}

int pic_zero_boundary(const PicDesc* desc, float* field, int axis, void* stream) {
    return zero_boundary_t(desc, field, axis, as_stream(stream));
}

int pic_zero_boundary_f64(const PicDesc* desc, double* field, int axis, void* stream) {
    return zero_boundary_t(desc, field, axis, as_stream(stream));
}

int pic_constant_boundary(const PicDesc* desc, float* field, int axis, void* stream) {
    return constant_boundary_t(desc, field, axis, as_stream(stream));
}

int pic_constant_boundary_f64(const PicDesc* desc, double* field, int axis, void* stream) {
    return constant_boundary_t(desc, field, axis, as_stream(stream));
}

int pic_filter_digital(const PicDesc* desc, const float* in, float* out, double alpha, void* stream) {
    return filter_t<false>(desc, in, out, alpha, as_stream(stream));
}

int pic_filter_digital_f64(const PicDesc* desc, const double* in, double* out, double alpha, void* stream) {
    return filter_t<false>(desc, in, out, alpha, as_stream(stream));
}

int pic_filter_bilinear(const PicDesc* desc, const float* in, float* out, void* stream) {
    return filter_t<true>(desc, in, out, 1.0, as_stream(stream));
}

int pic_filter_bilinear_f64(const PicDesc* desc, const double* in, double* out, void* stream) {
    return filter_t<true>(desc, in, out, 1.0, as_stream(stream));
}

int pic_yee_update_b(const PicDesc* desc, const float* const E[3], const float* const B_in[3], float* const B_out[3],
                     void* stream) {
    return yee_b_t(desc, E, B_in, B_out, as_stream(stream));
}

int pic_yee_update_b_f64(const PicDesc* desc, const double* const E[3], const double* const B_in[3],
                         double* const B_out[3], void* stream) {
    return yee_b_t(desc, E, B_in, B_out, as_stream(stream));
}

int pic_yee_update_e(const PicDesc* desc, const float* const E_in[3], const float* const B[3], const float* const J[3],
                     float* const E_out[3], void* stream) {
    return yee_e_t(desc, E_in, B, J, E_out, as_stream(stream));
}

int pic_yee_update_e_f64(const PicDesc* desc, const double* const E_in[3], const double* const B[3],
                         const double* const J[3], double* const E_out[3], void* stream) {
    return yee_e_t(desc, E_in, B, J, E_out, as_stream(stream));
}

static bool pml_ok(const PicPml* pml) {
    if (!pml) return false;
    for (int q = 0; q < 6; ++q)
        if (!pml->mem[q]) return false;
    for (int a = 0; a < 3; ++a)
        if (!pml->sigma[a]) return false;
    return true;
}

static PmlArgs pml_args(const PicDesc* desc, const PicPml* pml) {
    PmlArgs a;
    for (int q = 0; q < 6; ++q) a.mem[q] = pml->mem[q];
    for (int c = 0; c < 3; ++c) {
        a.sigma[c] = pml->sigma[c];
        a.line[c] = pml->sigma_line[c];
    }
    a.dt = (float)desc->dt;          // the stretch uses the full step in both updates (PML.py:291, 337)
    return a;
}

int pic_yee_update_b_pml(const PicDesc* desc, const float* const E[3], const float* const B_in[3],
                         float* const B_out[3], const PicPml* pml, void* stream) {
    PIC_REQUIRE(desc_ok(desc) && E && B_in && B_out && pml_ok(pml));
    for (int c = 0; c < 3; ++c) PIC_REQUIRE(E[c] && B_in[c] && B_out[c]);
    for (int c = 0; c < 3; ++c)
        for (int e = 0; e < 3; ++e) PIC_REQUIRE((const float*)B_out[c] != E[e]);
    hipStream_t s = as_stream(stream);
    int block = interior_block(desc);
    hipLaunchKernelGGL(yee_b_pml_kernel, interior_grid(desc, block), dim3(block), 0, s, *desc, yee_const(*desc),
                       pml_args(desc, pml), F3{{E[0], E[1], E[2]}}, F3{{B_in[0], B_in[1], B_in[2]}},
                       W3f{{B_out[0], B_out[1], B_out[2]}});
    PIC_CHECK_LAUNCH();
    return pic_refresh_fields(desc, B_out[0], B_out[1], B_out[2], 0, s);
}

int pic_yee_update_e_pml(const PicDesc* desc, const float* const E_in[3], const float* const B[3],
                         const float* const J[3], float* const E_out[3], const PicPml* pml, void* stream) {
    PIC_REQUIRE(desc_ok(desc) && E_in && B && J && E_out && pml_ok(pml));
    for (int c = 0; c < 3; ++c) PIC_REQUIRE(E_in[c] && B[c] && J[c] && E_out[c]);
    for (int c = 0; c < 3; ++c)
        for (int e = 0; e < 3; ++e) PIC_REQUIRE((const float*)E_out[c] != B[e] && (const float*)E_out[c] != J[e]);
    hipStream_t s = as_stream(stream);
    int block = interior_block(desc);
    hipLaunchKernelGGL(yee_e_pml_kernel, interior_grid(desc, block), dim3(block), 0, s, *desc, yee_const(*desc),
                       pml_args(desc, pml), F3{{E_in[0], E_in[1], E_in[2]}}, F3{{B[0], B[1], B[2]}},
                       F3{{J[0], J[1], J[2]}}, W3f{{E_out[0], E_out[1], E_out[2]}});
    PIC_CHECK_LAUNCH();
    return pic_refresh_fields(desc, E_out[0], E_out[1], E_out[2], 0, s);
}

int pic_pml_stretch(const PicDesc* desc, float* const deriv[6], const PicPml* pml, void* stream) {
    PIC_REQUIRE(desc_ok(desc) && deriv && pml_ok(pml));
    W6f dv;
    for (int q = 0; q < 6; ++q) {
        PIC_REQUIRE(deriv[q]);
        dv.p[q] = deriv[q];
    }
    hipStream_t s = as_stream(stream);
    int block = interior_block(desc);
    hipLaunchKernelGGL(pml_stretch_kernel, interior_grid(desc, block), dim3(block), 0, s, *desc, pml_args(desc, pml),
                       dv);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_efield_from_phi(const PicDesc* desc, const float* phi, float* const E[3], void* stream) {
    PIC_REQUIRE(desc_ok(desc) && phi && E && E[0] && E[1] && E[2]);
    hipStream_t s = as_stream(stream);
    int block = interior_block(desc);
    hipLaunchKernelGGL(efield_from_phi_kernel, interior_grid(desc, block), dim3(block), 0, s, *desc, phi, E[0], E[1],
                       E[2], (float)(1.0 / (2.0 * desc->d[0])), (float)(1.0 / (2.0 * desc->d[1])),
                       (float)(1.0 / (2.0 * desc->d[2])));
    PIC_CHECK_LAUNCH();
    return pic_refresh_fields(desc, E[0], E[1], E[2], 0, s);
}

}  // extern "C"
