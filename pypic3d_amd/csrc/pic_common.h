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

// Logical block of a launch of `blocks` workgroups such that each of the 8
// XCDs (which receive workgroups round-robin by id) runs one contiguous range
// of logical blocks, in order: neighbouring tiles then share their halo nodes
// through the same L2.  A bijection on [0, blocks).
__device__ inline int xcd_block(int id, int blocks) {
    constexpr int kXcds = 8;
    const int q = blocks / kXcds, rem = blocks % kXcds;
    const int x = id % kXcds, local = id / kXcds;
    return x * q + min(x, rem) + local;
}

// Lane i of the wave receives lane i - kOff's value (kOff < 16) inside its
// aligned row of 16 lanes; lanes whose source leaves the row receive 0.
template <int kOff>
__device__ inline int dpp_row_shr(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x110 + kOff, 0xf, 0xf, true);
}

// Lane i of the wave receives lane i + kOff's value (kOff < 16) inside its
// aligned row of 16 lanes; lanes whose source leaves the row receive 0.
template <int kOff>
__device__ inline int dpp_row_shl(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x100 + kOff, 0xf, 0xf, true);
}

template <int kOff>
__device__ inline void row_count_up(int r, int& seg, int& p) {
    const bool take = seg == 0;
    const int sb = dpp_row_shr<kOff>(seg), pb = dpp_row_shr<kOff>(p);
    p += take ? pb : 0;
    seg |= r >= kOff ? sb : 1;
}

template <int kOff>
__device__ inline void row_count_down(int r, int& seg, int& s) {
    const bool take = seg == 0;
    const int sb = dpp_row_shl<kOff>(seg), cb = dpp_row_shl<kOff>(s);
    s += take ? cb : 0;
    seg |= r + kOff < 16 ? sb : 1;
}

// Position of each lane in its run of equal keys in consecutive lanes of its
// aligned row of 16: p = lanes from the run's first lane to this one, s =
// lanes from this one to the run's last lane (both counting this lane), so the
// run's last lane in the row is lane + s - 1 and holds the run's length in p.
// Every lane of the wave must execute it; DPP moves are unconditional.
__device__ inline void row_run_position(int key, int& p, int& s) {
    const int r = (int)(threadIdx.x & 15u);
    const int kprev = dpp_row_shr<1>(key), knext = dpp_row_shl<1>(key);
    int head = ((r == 0) | (kprev != key)) ? 1 : 0, tail = ((r == 15) | (knext != key)) ? 1 : 0;
    p = 1;
    s = 1;
    row_count_up<1>(r, head, p);
    row_count_up<2>(r, head, p);
    row_count_up<4>(r, head, p);
    row_count_up<8>(r, head, p);
    row_count_down<1>(r, tail, s);
    row_count_down<2>(r, tail, s);
    row_count_down<4>(r, tail, s);
    row_count_down<8>(r, tail, s);
}

// Lane i receives lane i + 1's value inside its row of 16 (0 for the last lane).
__device__ inline int dpp_row_shl1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x101, 0xf, 0xf, true); }

// One step of a segmented inclusive scan inside rows of 16 lanes: `seg` is
// set where a run head lies among the kOff lanes ending at this one; a lane
// without one adds the partial sum of the lane kOff below it.
// The DPP moves are evaluated unconditionally and selected afterwards: a
// move placed under a lane condition would run with those lanes disabled, and
// lanes reading from a disabled lane receive 0.
template <int kOff, int kN>
__device__ inline void row_scan_step(int r, int& seg, float (&v)[kN]) {
    const bool take = seg == 0;
    const int sb = dpp_row_shr<kOff>(seg);
    const int below = r >= kOff ? sb : 1;
#pragma unroll
    for (int m = 0; m < kN; ++m) {
        const float o = __int_as_float(dpp_row_shr<kOff>(__float_as_int(v[m])));
        v[m] += take ? o : 0.0f;
    }
    seg |= below;
}

// Sums kN values over each run of equal keys in consecutive lanes inside
// each aligned row of 16 lanes: afterwards the last lane of a run within its
// row holds the run's row total, and the call returns true on exactly those
// lanes (other lanes hold partial sums, which the caller discards).  Equal
// keys in lanes that are not consecutive form separate runs.  Every lane of
// the wave must execute it.
template <int kN>
__device__ inline bool row_run_sums(int key, float (&v)[kN]) {
    const int r = (int)(threadIdx.x & 15u);
    const int kprev = dpp_row_shr<1>(key), knext = dpp_row_shl1(key);
    int seg = ((r == 0) | (kprev != key)) ? 1 : 0;    // run head
    row_scan_step<1>(r, seg, v);
    row_scan_step<2>(r, seg, v);
    row_scan_step<4>(r, seg, v);
    row_scan_step<8>(r, seg, v);
    return (r == 15) | (knext != key);
}

// kP consecutive floats from 16-byte (kP = 4) or 8-byte (kP = 2) aligned memory
template <int kP>
__device__ inline void store_vec(float* p, const float v[kP]) {
    if constexpr (kP == 4) *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    else *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
}

template <int kP>
__device__ inline void load_vec(const float* p, float v[kP]) {
    if constexpr (kP == 4) {
        const float4 q = *reinterpret_cast<const float4*>(p);
        v[0] = q.x;
        v[1] = q.y;
        v[2] = q.z;
        v[3] = q.w;
    } else {
        const float2 q = *reinterpret_cast<const float2*>(p);
        v[0] = q.x;
        v[1] = q.y;
    }
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
