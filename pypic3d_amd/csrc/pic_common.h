// Shared device helpers for the gfx950 PIC kernels.
//
// Index convention: a field component is float[(Nx+2g)][(Ny+2g)][(Nz+2g)],
// z fastest; physical cell i sits at array index i+g on every axis.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "../../include/pypic3d_amd.h"

#define PIC_CHECK_LAUNCH()                                   \
    do {                                                     \
        hipError_t e__ = hipGetLastError();                  \
        if (e__ != hipSuccess) return (int)e__;              \
    } while (0)

#define PIC_REQUIRE(cond)                                    \
    do {                                                     \
        if (!(cond)) return -1;                              \
    } while (0)

static inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

__host__ __device__ inline int64_t fidx(const PicDesc& d, int i, int j, int k) {
    return ((int64_t)i * d.dims[1] + j) * d.dims[2] + k;
}

__host__ __device__ inline int64_t field_size(const PicDesc& d) {
    return (int64_t)d.dims[0] * d.dims[1] * d.dims[2];
}

// Neighbour modes an axis presents to stencils once its ghosts are refreshed.
enum AxisMode : int { MODE_PERIODIC = 0, MODE_ZERO = 1, MODE_REDUCED_COPY = 2 };

// Refresh mode of one axis for the field (bc_type 0) or particle (1) codes
// (ghost_cells.py:142-215: reduced periodic copies interior[g], any other
// non-periodic one-device axis receives zeros).
__host__ __device__ inline int refresh_mode(const PicDesc& d, int axis, int bc_type) {
    int bc = bc_type == 0 ? d.bc[axis] : d.pbc[axis];
    bool reduced = d.n[axis] == 1;
    if (bc != 0) return MODE_ZERO;
    return reduced ? MODE_REDUCED_COPY : MODE_PERIODIC;
}

// Value a refreshed field holds at array index `idx` along `axis`, given the
// interior values, when idx is one past the interior (forward) or one before
// it (backward).  Returns the source index or -1 for a zero ghost.
__device__ inline int ghost_source(int idx, int n, int g, int mode) {
    if (idx >= g && idx < n + g) return idx;
    if (mode == MODE_ZERO) return -1;
    if (mode == MODE_REDUCED_COPY) return g;
    // periodic: lower ghost l <- n + l, upper ghost n+g+l <- g + l
    return idx < g ? idx + n : idx - n;
}

// JAX scatter index handling for mode="drop": negative indices are first
// normalised NumPy-style, then out-of-range writes are dropped.
__device__ inline bool drop_index(int& i, int L) {
    if (i < 0) i += L;
    return (unsigned)i < (unsigned)L;
}

// anchor = floor((x - origin)/spacing) (CIC) or round-half-even (TSC)
// (grid_and_stencil.py:92-118), evaluated in FP64 like the FP64 reference so
// cell indices agree bit for bit.  Clamped so the int conversion of a
// garbage position cannot overflow downstream index arithmetic.
__device__ inline int anchor_of(double pos, double org, double sp, int sf) {
    double t = (pos - org) / sp;
    t = sf == 1 ? floor(t) : rint(t);
    t = fmin(fmax(t, -1.0e8), 1.0e8);
    return (int)t;
}

// hipcc fuses a*b + c into one multiply-add by default; the reference rounds
// the product and the sum separately.  These helpers switch contraction off
// for their own expressions so their roundings are the reference's.

// Node coordinate anchor*spacing + origin rounded like the FP64 reference.
__device__ inline double node_coord(int a, double sp, double org) {
#pragma clang fp contract(off)
    return (double)a * sp + org;
}

// x + v*dt rounded like the FP64 reference (update_tiled_particle_positions).
__device__ inline double step_pos(double x, double v, double dt) {
#pragma clang fp contract(off)
    return x + v * dt;
}

// (x - org) - (a + 1/2) d, the face offset of J_from_rhov.py:139-141.
__device__ inline double face_offset(double x, double org, int a, double dd) {
#pragma clang fp contract(off)
    return (x - org) - (a + 0.5) * dd;
}

// shapes.py:5-54 -- weights of stencil points (-1, 0, +1) for r = delta / d.
__device__ inline void shape3(double r, int sf, double w[3]) {
    if (sf == 1) {
        w[0] = 0.0;
        w[1] = 1.0 - r;
        w[2] = r;
    } else {
        w[0] = 0.5 * (0.5 - r) * (0.5 - r);
        w[1] = 0.75 - r * r;
        w[2] = 0.5 * (0.5 + r) * (0.5 + r);
    }
}

__device__ inline void atomic_add_f32(float* p, float v) {
    // hardware global_atomic_add_f32 on coarse-grained device memory
    unsafeAtomicAdd(p, v);
}

// positive modulo (jnp.mod of an int); the in-range case skips the division
__device__ inline int pmod(int a, int L) {
    if ((unsigned)a < (unsigned)L) return a;
    int r = a % L;
    return r < 0 ? r + L : r;
}

// An empty set (no species or no slots) may come with null planes: the
// entry points return before touching them.
inline bool particles_ok(const PicSpecies* sp, const PicParticles* p) {
    if (!sp || !p) return false;
    if (sp->count < 0 || sp->count > PIC_MAX_SPECIES || sp->slots < 0) return false;
    if (sp->count == 0 || sp->slots == 0) return true;
    for (int a = 0; a < 3; ++a)
        if (!p->x[a] || !p->u[a]) return false;
    return p->active != nullptr;
}

// species of flat slot i (slots are species-major, S <= PIC_MAX_SPECIES)
__device__ inline int species_of(int64_t i, int slots, int count) {
    int s = 0;
    int64_t end = slots;
    while (s + 1 < count && i >= end) {
        ++s;
        end += slots;
    }
    return s;
}

static inline int grid_1d(int64_t n, int block) {
    int64_t b = (n + block - 1) / block;
    if (b < 1) b = 1;
    if (b > 0x7fffffff) b = 0x7fffffff;
    return (int)b;
}
