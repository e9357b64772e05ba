// Fused particle phase of the electrodynamic Esirkepov step
// (evolve.py:36-43 push, :70-79 deposit + position update + refresh).
//
// One block owns kPerBlock consecutive particle slots, processed in rounds of
// one slot per thread.  Per particle: gather E+E_ext / B+B_ext on the Yee
// staggered grids, Boris push, then the Esirkepov deposit.
//
// Deposit organisation.  With cell-sorted particles the 64 lanes of a wave
// cover two or three cells, so per-lane global atomics would hammer the same
// twelve addresses.  The common case -- 3-D, first-order shape, particle
// stays in its cell on every axis -- deposits exactly 12 values at fixed
// offsets from its cell (Jx at (a, b..b+1, c..c+1), Jy at (a..a+1, b,
// c..c+1), Jz at (a..a+1, b..b+1, c)).  Those lanes are combined by a wave
// segmented inclusive scan over runs of equal cell; only the last lane of
// each run issues the 12 atomics.  Every other particle (a face crossing,
// second-order shape, reduced axes) is put on a per-block LDS queue and
// deposited after the rounds by the general 5-slot path, with the whole
// block working on the queue so there is no divergence cost in the rounds.
//
// Anchors use an FP32 estimate and fall back to the exact FP64 division of
// the reference whenever the estimate is within a margin of a cell face, so
// cell indices still agree bit for bit with an FP64 evaluation.
#pragma once
#include "pic_particle_math.h"

namespace {

using namespace pic;

constexpr int kThreads = 256;

// Anchor from an FP32 estimate t of (x - origin)/spacing: taken as is when it
// sits farther than `margin` from a face (floor) or a half-cell (round, TSC);
// otherwise the exact FP64 division of the reference on x_exact decides.  The
// margin covers the FP32 error of the estimate (a few ulp of |t|).
__device__ inline int anchor_est(float t, float mag, double x_exact, double org, double sp, int sf) {
    float tt = sf == 1 ? t : t + 0.5f;
    float f = floorf(tt);
    float fr = tt - f;
    // FP32 error of the estimate <= a few ulp of (|x| + |origin|) / spacing
    float margin = 1.0e-4f + 1.0e-6f * mag;
    if (fr >= margin && fr <= 1.0f - margin) {
        f = fminf(fmaxf(f, -1.0e8f), 1.0e8f);
        return (int)f;
    }
    return anchor_of(x_exact, org, sp, sf);
}

__device__ inline int anchor_fast(float x, float org32, float isp32, double org, double sp, int sf) {
    return anchor_est((x - org32) * isp32, (fabsf(x) + fabsf(org32)) * isp32, (double)x, org, sp, sf);
}

__device__ inline void shape3f(float r, int sf, float w[3]) {
    if (sf == 1) {
        w[0] = 0.0f;
        w[1] = 1.0f - r;
        w[2] = r;
    } else {
        w[0] = 0.5f * (0.5f - r) * (0.5f - r);
        w[1] = 0.75f - r * r;
        w[2] = 0.5f * (0.5f + r) * (0.5f + r);
    }
}

struct AxisConst {
    float org_c, org_v, isp_c, isp_v, inv_d;
};

// Gather stencil on one axis for one grid (center or vertex); also returns
// the anchor and the FP64-exact offset from the anchor node.
__device__ inline void gather_axis_fast(float x, float org32, float isp32, double org, double sp, int sf, int L,
                                        bool active, int g, GatherAxis& s, int& anchor, float& delta) {
    int a = anchor_fast(x, org32, isp32, org, sp, sf);
    double dl = (double)x - node_coord(a, sp, org);
    anchor = a;
    delta = (float)dl;
    float w[3];
    shape3f((float)dl * isp32, sf, w);
    if (active) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            s.p[q] = pmod(a - 1 + q, L);
            s.w[q] = w[q];
        }
        s.k0 = sf == 1 ? 1 : 0;
        s.k1 = 3;
    } else {
        s.p[0] = g;
        s.w[0] = w[0] + w[1] + w[2];
        s.k0 = 0;
        s.k1 = 1;
    }
}

// First-order (CIC) stencil on an active axis: points a, a+1 (the weight of
// a-1 is exactly zero, shapes.py:46-48), wrapped into the ghosted line like
// jnp.mod in prepare_particle_axis_stencil (grid_and_stencil.py:141-166).
struct CicAxis {
    int p0, p1;
    float w0, w1;
};

__device__ inline void cic_axis(float x, float org32, float isp32, double org, double sp, int L, CicAxis& s,
                                int& anchor, float& delta) {
    int a = anchor_fast(x, org32, isp32, org, sp, 1);
    double dl = (double)x - node_coord(a, sp, org);
    anchor = a;
    delta = (float)dl;
    float r = (float)dl * isp32;
    s.p0 = pmod(a, L);
    s.p1 = pmod(a + 1, L);
    s.w0 = 1.0f - r;
    s.w1 = r;
}

__device__ inline float gather_cic(const float* __restrict__ f, const float* __restrict__ fe, const PicDesc& d,
                                   const CicAxis& ax, const CicAxis& ay, const CicAxis& az) {
    // 32-bit unsigned offsets on a uniform base (field_size < 2^31 is checked at launch)
    const uint32_t sy = (uint32_t)d.dims[2], sx = (uint32_t)d.dims[1] * sy;
    const uint32_t r00 = (uint32_t)ax.p0 * sx + (uint32_t)ay.p0 * sy, r01 = (uint32_t)ax.p0 * sx + (uint32_t)ay.p1 * sy;
    const uint32_t r10 = (uint32_t)ax.p1 * sx + (uint32_t)ay.p0 * sy, r11 = (uint32_t)ax.p1 * sx + (uint32_t)ay.p1 * sy;
    const uint32_t z0 = (uint32_t)az.p0, z1 = (uint32_t)az.p1;
    float v000 = f[r00 + z0], v001 = f[r00 + z1], v010 = f[r01 + z0], v011 = f[r01 + z1];
    float v100 = f[r10 + z0], v101 = f[r10 + z1], v110 = f[r11 + z0], v111 = f[r11 + z1];
    if (fe) {
        v000 += fe[r00 + z0];
        v001 += fe[r00 + z1];
        v010 += fe[r01 + z0];
        v011 += fe[r01 + z1];
        v100 += fe[r10 + z0];
        v101 += fe[r10 + z1];
        v110 += fe[r11 + z0];
        v111 += fe[r11 + z1];
    }
    float c00 = v000 * az.w0 + v001 * az.w1, c01 = v010 * az.w0 + v011 * az.w1;
    float c10 = v100 * az.w0 + v101 * az.w1, c11 = v110 * az.w0 + v111 * az.w1;
    return ax.w0 * (ay.w0 * c00 + ay.w1 * c01) + ax.w1 * (ay.w0 * c10 + ay.w1 * c11);
}

// Deposit targets.  GlobalSink adds straight into the J arrays (JAX scatter
// "drop" index semantics, Esirkepov.py:232-240).  BoxSink accumulates into an
// LDS box of B^3 nodes per component whose origin is `o` (array indices) and
// falls back to the global arrays for nodes outside it; the caller flushes
// the box into J with the same drop semantics.
struct GlobalSink {
    float* J[3];
    __device__ inline void add(const PicDesc& d, int c, int i, int j, int k, float v) const {
        deposit_at(J[c], d, i, j, k, v);
    }
};

struct BoxSink {
    float* J[3];
    float* box;
    int o[3];
    int B;
    __device__ inline void add(const PicDesc& d, int c, int i, int j, int k, float v) const {
        if (v == 0.0f) return;
        const int bi = i - o[0], bj = j - o[1], bk = k - o[2];
        if ((unsigned)bi < (unsigned)B && (unsigned)bj < (unsigned)B && (unsigned)bk < (unsigned)B)
            atomicAdd(&box[((c * B + bi) * B + bj) * B + bk], v);
        else
            deposit_at(J[c], d, i, j, k, v);
    }
};

// LDS box with per-axis extents e[] from origin o[] (array indices), three
// components; used when a block's particles cover a small, data-dependent
// range of cells.  Nodes outside (or `on` false) go to the global arrays.
struct RangeBoxSink {
    float* J[3];
    float* box;
    int o[3], e[3];
    bool on;
    __device__ inline void add(const PicDesc& d, int c, int i, int j, int k, float v) const {
        if (v == 0.0f) return;
        const int bi = i - o[0], bj = j - o[1], bk = k - o[2];
        if (on && (unsigned)bi < (unsigned)e[0] && (unsigned)bj < (unsigned)e[1] && (unsigned)bk < (unsigned)e[2])
            atomicAdd(&box[((c * e[0] + bi) * e[1] + bj) * e[2] + bk], v);
        else
            deposit_at(J[c], d, i, j, k, v);
    }
};

// First-order Esirkepov for a particle whose anchor moves by at most one cell
// per axis, in a 3-slot frame starting at b = min(a_old, a_new): S0, D = S1 -
// S0 (FP64) and the running sum C of D at slots 0 and 1 (slot 2's sum is 0).
// The same values the 5-slot new-anchor frame of Esirkepov.py:158-189 holds
// at those points; all other slots of that frame carry zero.
struct CrossAxis {
    int b;
    float S0[3], D[3], C0, C1;
};

__device__ inline bool cross_axis(double xo, double xn, double org, double sp, double inv_d, float org32,
                                  float isp32, CrossAxis& e) {
    float mag = (fabsf((float)xo) + fabsf(org32) + 1.0f) * isp32;
    int ao = anchor_est(((float)xo - org32) * isp32, mag, xo, org, sp, 1);
    int an = anchor_est(((float)xn - org32) * isp32, mag + fabsf(((float)xn - (float)xo) * isp32), xn, org, sp, 1);
    if (an - ao > 1 || ao - an > 1) return false;
    double ro = (xo - node_coord(ao, sp, org)) * inv_d;
    double rn = (xn - node_coord(an, sp, org)) * inv_d;
    int b = min(ao, an);
    double s0[3] = {0.0, 0.0, 0.0}, s1[3] = {0.0, 0.0, 0.0};
    if (ao == b) { s0[0] = 1.0 - ro; s0[1] = ro; } else { s0[1] = 1.0 - ro; s0[2] = ro; }
    if (an == b) { s1[0] = 1.0 - rn; s1[1] = rn; } else { s1[1] = 1.0 - rn; s1[2] = rn; }
    e.b = b;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        e.S0[k] = (float)s0[k];
        e.D[k] = (float)(s1[k] - s0[k]);
    }
    e.C0 = (float)(s1[0] - s0[0]);
    e.C1 = (float)((s1[0] - s0[0]) + (s1[1] - s0[1]));
    return true;
}

__device__ inline float sel3(int k, float a0, float a1, float a2) { return k == 0 ? a0 : (k == 1 ? a1 : a2); }

// All 54 terms (component c, differenced slot i in {0,1}, 3x3 slots (j, k) of
// the other two axes) of one particle, fully unrolled so every per-axis value
// stays in registers.
template <int kC, class Sink>
__device__ inline void crossing_component(const PicDesc& d, double qw, const CrossAxis& A, const CrossAxis& P,
                                          const CrossAxis& Q, const Sink& sink) {
    constexpr int o1 = kC == 0 ? 1 : 0, o2 = kC == 2 ? 1 : 2;
    const float pre = (float)(-(qw / (d.d[o1] * d.d[o2])) / d.dt);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const float ci = i == 0 ? A.C0 : A.C1;
        if (ci == 0.0f) continue;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                float y = P.S0[j] * Q.S0[k] + 0.5f * (P.S0[j] * Q.D[k] + P.D[j] * Q.S0[k]) +
                          (1.0f / 3.0f) * P.D[j] * Q.D[k];
                int ic = A.b + i, i1 = P.b + j, i2 = Q.b + k;
                int ix = kC == 0 ? ic : i1;
                int iy = kC == 0 ? i1 : (kC == 1 ? ic : i2);
                int iz = kC == 2 ? ic : i2;
                sink.add(d, kC, ix, iy, iz, pre * ci * y);
            }
        }
    }
}

template <class Sink>
__device__ inline void crossing_deposit(const PicDesc& d, double qw, const CrossAxis& cx, const CrossAxis& cy,
                                        const CrossAxis& cz, const Sink& sink) {
    crossing_component<0>(d, qw, cx, cy, cz, sink);
    crossing_component<1>(d, qw, cy, cx, cz, sink);
    crossing_component<2>(d, qw, cz, cx, cy, sink);
}

// Face-crossing deposit with reduced axes (Esirkepov.py:241-329 in the
// 3-slot frame): kDims = 2 (x, y active, z reduced) or 1 (x active).  The
// active components sum the running differences C of their own axis times the
// factor of the other active axes (S0 + D/2 in 2-D, 1 in 1-D); a reduced
// component c carries q v_c / dV times the factor of all active axes.
template <int kDims, class Sink>
__device__ inline void crossing_deposit_reduced(const PicDesc& d, const SpeciesK& k, const float v[3],
                                                const CrossAxis& cx, const CrossAxis& cy, const Sink& sink) {
    const int g = d.g;
    if constexpr (kDims == 2) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const float cix = i == 0 ? cx.C0 : cx.C1, ciy = i == 0 ? cy.C0 : cy.C1;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                if (cix != 0.0f) sink.add(d, 0, cx.b + i, cy.b + j, g, k.pre[0] * cix * (cy.S0[j] + 0.5f * cy.D[j]));
                if (ciy != 0.0f) sink.add(d, 1, cx.b + j, cy.b + i, g, k.pre[1] * ciy * (cx.S0[j] + 0.5f * cx.D[j]));
            }
        }
        const float pz = k.q_dv * v[2];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
                float y = cx.S0[j] * cy.S0[kk] + 0.5f * (cx.S0[j] * cy.D[kk] + cx.D[j] * cy.S0[kk]) +
                          (1.0f / 3.0f) * cx.D[j] * cy.D[kk];
                sink.add(d, 2, cx.b + j, cy.b + kk, g, pz * y);
            }
    } else {
        sink.add(d, 0, cx.b, g, g, k.pre[0] * cx.C0);
        sink.add(d, 0, cx.b + 1, g, g, k.pre[0] * cx.C1);
        const float py = k.q_dv * v[1], pz = k.q_dv * v[2];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float f1 = cx.S0[j] + 0.5f * cx.D[j];
            sink.add(d, 1, cx.b + j, g, g, py * f1);
            sink.add(d, 2, cx.b + j, g, g, pz * f1);
        }
    }
}

// LDS copy of E+E_ext and B+B_ext around a tile: component c occupies
// box[c*B^3 ...], node (i, j, k) of the ghosted arrays at
// ((i - o0)*B + (j - o1))*B + (k - o2), filled from index pmod(i, L) like the
// global gather reads it.
struct FieldBox {
    const float* box;
    int o[3];
    int B;
};

__device__ inline float gather_box(const FieldBox& fb, int c, int a0, int a1, int a2, const CicAxis& ax,
                                   const CicAxis& ay, const CicAxis& az) {
    const int B = fb.B;
    const float* p = fb.box + c * B * B * B + ((a0 - fb.o[0]) * B + (a1 - fb.o[1])) * B + (a2 - fb.o[2]);
    const float v000 = p[0], v001 = p[1], v010 = p[B], v011 = p[B + 1];
    const float v100 = p[B * B], v101 = p[B * B + 1], v110 = p[B * B + B], v111 = p[B * B + B + 1];
    float c00 = v000 * az.w0 + v001 * az.w1, c01 = v010 * az.w0 + v011 * az.w1;
    float c10 = v100 * az.w0 + v101 * az.w1, c11 = v110 * az.w0 + v111 * az.w1;
    return ax.w0 * (ay.w0 * c00 + ay.w1 * c01) + ax.w1 * (ay.w0 * c10 + ay.w1 * c11);
}

// Per-block LDS: the face-crossing queue and the species constants.
template <int kRounds>
struct FusedShared {
    int q_idx[kThreads * kRounds];
    float q_v[3][kThreads * kRounds];
    int q_n;
    SpeciesK spk[PIC_MAX_SPECIES];
};

// Regular-case deposit layout of a first-order fast path (kDims active axes
// x..; the remaining axes are reduced): number of values and, for value m,
// its J component and its offset from the particle's cell.
template <int kDims>
struct RegularLayout {
    static constexpr int nv = kDims == 3 ? 12 : (kDims == 2 ? 8 : 5);
    __device__ static inline int comp(int m) {
        if (kDims == 3) return m >> 2;
        if (kDims == 2) return m < 2 ? 0 : (m < 4 ? 1 : 2);
        return m < 1 ? 0 : (m < 3 ? 1 : 2);
    }
    // node offset (dx, dy, dz) of value m from the particle's anchor cell
    __device__ static inline void delta(int m, int& dx, int& dy, int& dz) {
        dx = dy = dz = 0;
        if (kDims == 3) {
            const int q = m & 3, hi = q >> 1, lo = q & 1;
            if (m < 4) { dy = hi; dz = lo; }            // Jx: (a, b..b+1, c..c+1)
            else if (m < 8) { dx = hi; dz = lo; }       // Jy: (a..a+1, b, c..c+1)
            else { dx = hi; dy = lo; }                  // Jz: (a..a+1, b..b+1, c)
        } else if (kDims == 2) {
            if (m < 2) dy = m;                          // Jx: (a, b..b+1)
            else if (m < 4) dx = m - 2;                 // Jy: (a..a+1, b)
            else { dx = (m - 4) >> 1; dy = (m - 4) & 1; }   // Jz: (a..a+1, b..b+1)
        } else if (m >= 1) {
            dx = (m - 1) & 1;                           // Jy, Jz: (a..a+1)
        }
    }
};

// CIC gather with NX x NY x NZ points (2 on an active axis, 1 on a reduced one).
template <int NX, int NY, int NZ>
__device__ inline float gather_n(const float* __restrict__ f, const float* __restrict__ fe, uint32_t sx, uint32_t sy,
                                 const CicAxis& ax, const CicAxis& ay, const CicAxis& az) {
    const uint32_t px[2] = {(uint32_t)ax.p0, (uint32_t)ax.p1}, py[2] = {(uint32_t)ay.p0, (uint32_t)ay.p1};
    const uint32_t pz[2] = {(uint32_t)az.p0, (uint32_t)az.p1};
    const float wx[2] = {ax.w0, ax.w1}, wy[2] = {ay.w0, ay.w1}, wz[2] = {az.w0, az.w1};
    float acc = 0.0f;
#pragma unroll
    for (int i = 0; i < NX; ++i) {
        float ai = 0.0f;
#pragma unroll
        for (int j = 0; j < NY; ++j) {
            const uint32_t row = px[i] * sx + py[j] * sy;
            float aj = 0.0f;
#pragma unroll
            for (int k = 0; k < NZ; ++k) {
                float v = f[row + pz[k]];
                if (fe) v += fe[row + pz[k]];
                aj += v * wz[k];
            }
            ai += wy[j] * aj;
        }
        acc += wx[i] * ai;
    }
    return acc;
}

// A reduced axis collapses the stencil to index g with the summed weight
// (grid_and_stencil.py:168-239), which is one for the first-order shape.
__device__ inline void reduced_cic(int g, CicAxis& s) {
    s.p0 = s.p1 = g;
    s.w0 = 1.0f;
    s.w1 = 0.0f;
}

// The particle phase for the block's kThreads * kRounds consecutive slots
// [base, base + kThreads * kRounds) of one simulation.  kDims = 3, 2, 1:
// first-order fast path with that many leading active axes (x, y, z / x, y /
// x); kDims = 0: any shape and reduced-axis layout through the general 5-slot
// path.  `d` and `sp` may live in kernel arguments or in device memory.
template <int kDims, int kRounds, class Sink, bool kFieldBox = false>
__device__ inline void fused_block(const PicDesc& d, const PicSpecies& sp, const PicParticles& in,
                                   const FieldPtrs& fp, const PicParticles& out, const Sink& sink, int64_t n,
                                   uint32_t base, FusedShared<kRounds>& sh, const FieldBox& fb = FieldBox{}) {
    constexpr bool kFast = kDims > 0;
    using Lay = RegularLayout<kDims == 0 ? 3 : kDims>;
    constexpr int NV = Lay::nv;
    if (threadIdx.x == 0) sh.q_n = 0;
    if (threadIdx.x < sp.count) sh.spk[threadIdx.x] = species_constants(d, sp, threadIdx.x);
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int sf = d.shape_factor;
    AxisConst ac[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        ac[a].org_c = (float)d.org_c[a];
        ac[a].org_v = (float)d.org_v[a];
        ac[a].isp_c = (float)(1.0 / d.sp_c[a]);
        ac[a].isp_v = (float)(1.0 / d.sp_v[a]);
        ac[a].inv_d = (float)(1.0 / d.d[a]);
    }
    const uint32_t sy = (uint32_t)d.dims[2];
    const uint32_t sx = (uint32_t)d.dims[1] * sy;
    const float fdt = (float)d.dt;

    for (int r = 0; r < kRounds; ++r) {
        const int local = r * kThreads + threadIdx.x;
        const uint32_t i = base + (uint32_t)local;
        const bool valid = i < n;
        const int s = valid ? species_of(i, sp.slots, sp.count) : 0;
        bool act = valid && in.active[i] != 0;
        int key = -1 - lane;
        float val[NV];
#pragma unroll
        for (int m = 0; m < NV; ++m) val[m] = 0.0f;
        int cell = 0, a0 = 0, a1 = 0, a2 = 0;

        if (act) {
            float xf[3] = {in.x[0][i], in.x[1][i], in.x[2][i]};
            float v[3] = {in.u[0][i], in.u[1][i], in.u[2][i]};
            // ---- gather (boris.py:302-433) on center and vertex lines
            int anc[3], anv[3];
            float dc[3], dv[3];
            float E[3], B[3];
            if constexpr (kFast) {
                CicAxis gc[3], gv[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    if (a < kDims) {
                        cic_axis(xf[a], ac[a].org_c, ac[a].isp_c, d.org_c[a], d.sp_c[a], d.dims[a], gc[a], anc[a],
                                 dc[a]);
                        cic_axis(xf[a], ac[a].org_v, ac[a].isp_v, d.org_v[a], d.sp_v[a], d.dims[a], gv[a], anv[a],
                                 dv[a]);
                    } else {
                        reduced_cic(d.g, gc[a]);
                        reduced_cic(d.g, gv[a]);
                        anc[a] = anv[a] = 0;
                        dc[a] = dv[a] = 0.0f;
                    }
                }
                bool boxed = false;
                if constexpr (kFieldBox && kDims == 3) {
                    boxed = true;
#pragma unroll
                    for (int a = 0; a < 3; ++a)
                        boxed = boxed && (unsigned)(anc[a] - fb.o[a]) < (unsigned)(fb.B - 1) &&
                                (unsigned)(anv[a] - fb.o[a]) < (unsigned)(fb.B - 1);
                }
                if (boxed) {
                    E[0] = gather_box(fb, 0, anv[0], anc[1], anc[2], gv[0], gc[1], gc[2]);
                    E[1] = gather_box(fb, 1, anc[0], anv[1], anc[2], gc[0], gv[1], gc[2]);
                    E[2] = gather_box(fb, 2, anc[0], anc[1], anv[2], gc[0], gc[1], gv[2]);
                    B[0] = gather_box(fb, 3, anc[0], anv[1], anv[2], gc[0], gv[1], gv[2]);
                    B[1] = gather_box(fb, 4, anv[0], anc[1], anv[2], gv[0], gc[1], gv[2]);
                    B[2] = gather_box(fb, 5, anv[0], anv[1], anc[2], gv[0], gv[1], gc[2]);
                } else if constexpr (kDims == 3) {
                    E[0] = gather_cic(fp.E[0], fp.Ee[0], d, gv[0], gc[1], gc[2]);
                    E[1] = gather_cic(fp.E[1], fp.Ee[1], d, gc[0], gv[1], gc[2]);
                    E[2] = gather_cic(fp.E[2], fp.Ee[2], d, gc[0], gc[1], gv[2]);
                    B[0] = gather_cic(fp.B[0], fp.Be[0], d, gc[0], gv[1], gv[2]);
                    B[1] = gather_cic(fp.B[1], fp.Be[1], d, gv[0], gc[1], gv[2]);
                    B[2] = gather_cic(fp.B[2], fp.Be[2], d, gv[0], gv[1], gc[2]);
                } else {
                    constexpr int NY = kDims >= 2 ? 2 : 1;
                    E[0] = gather_n<2, NY, 1>(fp.E[0], fp.Ee[0], sx, sy, gv[0], gc[1], gc[2]);
                    E[1] = gather_n<2, NY, 1>(fp.E[1], fp.Ee[1], sx, sy, gc[0], gv[1], gc[2]);
                    E[2] = gather_n<2, NY, 1>(fp.E[2], fp.Ee[2], sx, sy, gc[0], gc[1], gv[2]);
                    B[0] = gather_n<2, NY, 1>(fp.B[0], fp.Be[0], sx, sy, gc[0], gv[1], gv[2]);
                    B[1] = gather_n<2, NY, 1>(fp.B[1], fp.Be[1], sx, sy, gv[0], gc[1], gv[2]);
                    B[2] = gather_n<2, NY, 1>(fp.B[2], fp.Be[2], sx, sy, gv[0], gv[1], gc[2]);
                }
            } else {
                GatherAxis gc[3], gv[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    bool axact = d.n[a] > 1;
                    gather_axis_fast(xf[a], ac[a].org_c, ac[a].isp_c, d.org_c[a], d.sp_c[a], sf, d.dims[a], axact,
                                     d.g, gc[a], anc[a], dc[a]);
                    gather_axis_fast(xf[a], ac[a].org_v, ac[a].isp_v, d.org_v[a], d.sp_v[a], sf, d.dims[a], axact,
                                     d.g, gv[a], anv[a], dv[a]);
                }
                E[0] = gather3(fp.E[0], fp.Ee[0], d, gv[0], gc[1], gc[2]);
                E[1] = gather3(fp.E[1], fp.Ee[1], d, gc[0], gv[1], gc[2]);
                E[2] = gather3(fp.E[2], fp.Ee[2], d, gc[0], gc[1], gv[2]);
                B[0] = gather3(fp.B[0], fp.Be[0], d, gc[0], gv[1], gv[2]);
                B[1] = gather3(fp.B[1], fp.Be[1], d, gv[0], gc[1], gv[2]);
                B[2] = gather3(fp.B[2], fp.Be[2], d, gv[0], gv[1], gc[2]);
            }
            (void)anv;
            (void)dv;
            // ---- push (boris.py:14-124), masked by update_u (particle_push.py:134-142)
            float nv[3] = {v[0], v[1], v[2]};
            const SpeciesK k = sh.spk[s];
            push_velocity(nv, E, B, k.h, d);
#pragma unroll
            for (int a = 0; a < 3; ++a)
                if ((k.uu >> a) & 1u) v[a] = nv[a];
            // ---- new position (Esirkepov.py:125-127) and regular-case test
            double xn[3];
            bool regular = kFast;
            float S0[3][2], D[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                bool ux = (k.ux >> a) & 1u;
                xn[a] = ux ? step_pos((double)xf[a], (double)v[a], d.dt) : (double)xf[a];
                if (!kFast || a >= kDims || !regular) continue;
                float dvt = ux ? v[a] * fdt : 0.0f;
                int an = anchor_est((xf[a] + dvt - ac[a].org_c) * ac[a].isp_c,
                                    (fabsf(xf[a]) + fabsf(dvt) + fabsf(ac[a].org_c)) * ac[a].isp_c, xn[a],
                                    d.org_c[a], d.sp_c[a], 1);
                regular = regular && an == anc[a] && anc[a] >= 0 && anc[a] + 1 < d.dims[a];
                float r0 = dc[a] * ac[a].inv_d;
                S0[a][0] = 1.0f - r0;
                S0[a][1] = r0;
                D[a] = dvt * ac[a].inv_d;      // S_new - S_old = (+D at a+1, -D at a)
            }
            if (regular) {
                // Esirkepov weights in difference form with D = (-D, +D) at (a, a+1):
                // W = D_c * (S0 S0 + (S0 D + D S0)/2 + D D/3) over the other active axes,
                // J_c = -(q/(d_o1 d_o2 dt)) * cumsum_c(W) -> value at slot a only; a
                // reduced component c carries q v_c / dV times the factor of the
                // active axes (Esirkepov.py:197-214).
                if constexpr (kDims == 3) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        int o1 = c == 0 ? 1 : 0, o2 = c == 2 ? 1 : 2;
                        float pre = k.pre[c] * (-D[c]);
#pragma unroll
                        for (int j = 0; j < 2; ++j) {
                            float dj = j ? D[o1] : -D[o1];
#pragma unroll
                            for (int kk = 0; kk < 2; ++kk) {
                                float dk = kk ? D[o2] : -D[o2];
                                float y = S0[o1][j] * S0[o2][kk] + 0.5f * (S0[o1][j] * dk + dj * S0[o2][kk]) +
                                          (1.0f / 3.0f) * dj * dk;
                                val[c * 4 + j * 2 + kk] = pre * y;
                            }
                        }
                    }
                    a0 = anc[0];
                    a1 = anc[1];
                    a2 = anc[2];
                } else if constexpr (kDims == 2) {
                    float px = k.pre[0] * (-D[0]), py = k.pre[1] * (-D[1]), pz = k.q_dv * v[2];
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        float dxj = j ? D[0] : -D[0], dyj = j ? D[1] : -D[1];
                        val[j] = px * (S0[1][j] + 0.5f * dyj);
                        val[2 + j] = py * (S0[0][j] + 0.5f * dxj);
#pragma unroll
                        for (int kk = 0; kk < 2; ++kk) {
                            float dyk = kk ? D[1] : -D[1];
                            val[4 + j * 2 + kk] = pz * (S0[0][j] * S0[1][kk] + 0.5f * (S0[0][j] * dyk + dxj * S0[1][kk]) +
                                                        (1.0f / 3.0f) * dxj * dyk);
                        }
                    }
                    a0 = anc[0];
                    a1 = anc[1];
                    a2 = d.g;
                } else {
                    float py = k.q_dv * v[1], pz = k.q_dv * v[2];
                    val[0] = k.pre[0] * (-D[0]);
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        float f1 = S0[0][j] + 0.5f * (j ? D[0] : -D[0]);
                        val[1 + j] = py * f1;
                        val[3 + j] = pz * f1;
                    }
                    a0 = anc[0];
                    a1 = a2 = d.g;
                }
                cell = (int)fidx(d, a0, a1, a2);
                key = cell;
            } else {
                int slot = atomicAdd(&sh.q_n, 1);
                sh.q_idx[slot] = local;
                sh.q_v[0][slot] = v[0];
                sh.q_v[1][slot] = v[1];
                sh.q_v[2][slot] = v[2];
            }
            // ---- position update + particle boundaries (particle_tile_communication.py)
#pragma unroll
            for (int a = 0; a < 3; ++a) apply_axis_boundary(d.pbc[a], d.wind[a], xn[a], v[a], act);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                out.x[a][i] = act ? (float)xn[a] : 0.0f;
                out.u[a][i] = act ? v[a] : 0.0f;
            }
            out.active[i] = act ? 1 : 0;
        } else if (valid) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                out.x[a][i] = 0.0f;
                out.u[a][i] = 0.0f;
            }
            out.active[i] = 0;
        }

        // ---- wave segmented inclusive scan over runs of equal cell
        if constexpr (kFast) {
            unsigned long long any_regular = __ballot(key >= 0);
            if (any_regular) {
                int prev = __shfl_up(key, 1);
                bool head = lane == 0 || prev != key;
                unsigned long long heads = __ballot(head);
                unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
                int seg = __popcll(heads & upto);
                for (int dd = 1; dd < 64; dd <<= 1) {
                    int so = __shfl_up(seg, dd);
                    bool same = lane >= dd && so == seg;
                    if (__ballot(same) == 0ull) break;
#pragma unroll
                    for (int m = 0; m < NV; ++m) {
                        float o = __shfl_up(val[m], dd);
                        if (same) val[m] += o;
                    }
                }
                bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
                if (tail && key >= 0) {
#pragma unroll
                    for (int m = 0; m < NV; ++m) {
                        int dx, dy, dz;
                        Lay::delta(m, dx, dy, dz);
                        sink.add(d, Lay::comp(m), a0 + dx, a1 + dy, a2 + dz, val[m]);
                    }
                }
            }
        }
    }
    __syncthreads();

    // ---- queued particles (Esirkepov.py:105-331)
    const int nq = sh.q_n;
    float* J[3] = {sink.J[0], sink.J[1], sink.J[2]};
    if constexpr (kDims == 3) {
        // one lane per queued particle, all terms unrolled (no scratch arrays)
        for (int q = threadIdx.x; q < nq; q += kThreads) {
            uint32_t i = base + (uint32_t)sh.q_idx[q];
            const SpeciesK& k = sh.spk[species_of(i, sp.slots, sp.count)];
            float v[3] = {sh.q_v[0][q], sh.q_v[1][q], sh.q_v[2][q]};
            double xo[3], xn[3];
            load_pos(in, i, xo);
#pragma unroll
            for (int a = 0; a < 3; ++a) xn[a] = ((k.ux >> a) & 1u) ? step_pos(xo[a], (double)v[a], d.dt) : xo[a];
            CrossAxis cx, cy, cz;
            bool ok = cross_axis(xo[0], xn[0], d.org_c[0], d.sp_c[0], 1.0 / d.d[0], ac[0].org_c, ac[0].isp_c, cx);
            ok = cross_axis(xo[1], xn[1], d.org_c[1], d.sp_c[1], 1.0 / d.d[1], ac[1].org_c, ac[1].isp_c, cy) && ok;
            ok = cross_axis(xo[2], xn[2], d.org_c[2], d.sp_c[2], 1.0 / d.d[2], ac[2].org_c, ac[2].isp_c, cz) && ok;
            if (ok) crossing_deposit(d, k.qw, cx, cy, cz, sink);
            else esirkepov_particle(d, k.qw, xo, xn, v, J);   // moved two or more cells
        }
    } else if constexpr (kDims == 2 || kDims == 1) {
        for (int q = threadIdx.x; q < nq; q += kThreads) {
            uint32_t i = base + (uint32_t)sh.q_idx[q];
            const SpeciesK& k = sh.spk[species_of(i, sp.slots, sp.count)];
            float v[3] = {sh.q_v[0][q], sh.q_v[1][q], sh.q_v[2][q]};
            double xo[3], xn[3];
            load_pos(in, i, xo);
#pragma unroll
            for (int a = 0; a < 3; ++a) xn[a] = ((k.ux >> a) & 1u) ? step_pos(xo[a], (double)v[a], d.dt) : xo[a];
            CrossAxis cx, cy;
            bool ok = cross_axis(xo[0], xn[0], d.org_c[0], d.sp_c[0], 1.0 / d.d[0], ac[0].org_c, ac[0].isp_c, cx);
            if (kDims == 2)
                ok = cross_axis(xo[1], xn[1], d.org_c[1], d.sp_c[1], 1.0 / d.d[1], ac[1].org_c, ac[1].isp_c, cy) && ok;
            if (ok) crossing_deposit_reduced<kDims>(d, k, v, cx, cy, sink);
            else esirkepov_particle(d, k.qw, xo, xn, v, J);   // moved two or more cells
        }
    } else {
        for (int q = threadIdx.x; q < nq; q += kThreads) {
            uint32_t i = base + (uint32_t)sh.q_idx[q];
            const SpeciesK& k = sh.spk[species_of(i, sp.slots, sp.count)];
            float v[3] = {sh.q_v[0][q], sh.q_v[1][q], sh.q_v[2][q]};
            double xo[3], xn[3];
            load_pos(in, i, xo);
#pragma unroll
            for (int a = 0; a < 3; ++a) xn[a] = ((k.ux >> a) & 1u) ? step_pos(xo[a], (double)v[a], d.dt) : xo[a];
            esirkepov_particle(d, k.qw, xo, xn, v, J);
        }
    }
}

// First-order fast-path dimensionality of a descriptor (0: general path).
__host__ __device__ inline int fused_kind(const PicDesc& d) {
    if (d.shape_factor != 1) return 0;
    const bool a0 = d.n[0] > 1, a1 = d.n[1] > 1, a2 = d.n[2] > 1;
    if (a0 && a1 && a2) return 3;
    if (a0 && a1 && !a2) return 2;
    if (a0 && !a1 && !a2) return 1;
    return 0;
}


// The fused block with its J accumulated in an LDS box sized to the block's
// particles: the block first bounds the center-grid anchors of its slots
// [base, base + kThreads*kRounds) (FP32 estimate, padded by two nodes below
// and three above for the rounding, the stencil and face crossings), then
// runs fused_block into the box and adds the box into J once.  A block whose
// box does not fit deposits straight into J.  `box` holds kRangeBoxCap floats.
constexpr int kRangeBoxCap = 8192;

template <int kRounds>
__device__ inline void fused_block_boxed(const PicDesc& d, const PicSpecies& sp, const PicParticles& in,
                                         const FieldPtrs& fp, const PicParticles& out, float* const J[3], int64_t n,
                                         uint32_t base, FusedShared<kRounds>& sh, float* box, int* bmin, int* bmax) {
    if (threadIdx.x < 3) {
        bmin[threadIdx.x] = INT_MAX;
        bmax[threadIdx.x] = INT_MIN;
    }
    __syncthreads();
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (int r = 0; r < kRounds; ++r) {
        const uint32_t i = base + (uint32_t)(r * kThreads + threadIdx.x);
        if (i < n && in.active[i]) {
            for (int a = 0; a < 3; ++a) {
                const float t = (in.x[a][i] - (float)d.org_c[a]) * (float)(1.0 / d.sp_c[a]);
                const int c = (int)fminf(fmaxf(floorf(t), -1.0e6f), 1.0e6f);
                lo[a] = min(lo[a], c);
                hi[a] = max(hi[a], c);
            }
        }
    }
    for (int a = 0; a < 3; ++a) {
        if (lo[a] <= hi[a]) {
            atomicMin(&bmin[a], lo[a]);
            atomicMax(&bmax[a], hi[a]);
        }
    }
    __syncthreads();
    RangeBoxSink sink;
    for (int c = 0; c < 3; ++c) sink.J[c] = J[c];
    sink.box = box;
    sink.on = bmax[0] >= bmin[0];
    int64_t vol = 1;
    for (int a = 0; a < 3; ++a) {
        if (d.n[a] == 1) {
            sink.o[a] = d.g;           // reduced axis: every node sits at index g
            sink.e[a] = 1;
        } else {
            sink.o[a] = bmin[a] - 2;
            sink.e[a] = (int)min((int64_t)bmax[a] - bmin[a] + 6, (int64_t)kRangeBoxCap);
        }
        vol *= sink.e[a];
    }
    sink.on = sink.on && 3 * vol <= kRangeBoxCap;
    const int nbox = sink.on ? (int)(3 * vol) : 0;
    for (int e = threadIdx.x; e < nbox; e += kThreads) box[e] = 0.0f;
    __syncthreads();
    switch (fused_kind(d)) {
        case 3: fused_block<3, kRounds>(d, sp, in, fp, out, sink, n, base, sh); break;
        case 2: fused_block<2, kRounds>(d, sp, in, fp, out, sink, n, base, sh); break;
        case 1: fused_block<1, kRounds>(d, sp, in, fp, out, sink, n, base, sh); break;
        default: fused_block<0, kRounds>(d, sp, in, fp, out, sink, n, base, sh); break;
    }
    __syncthreads();
    const int v1 = sink.e[1] * sink.e[2], v3 = (int)vol;
    for (int e = threadIdx.x; e < nbox; e += kThreads) {
        const float val = box[e];
        if (val == 0.0f) continue;
        const int c = e / v3, r = e % v3;
        deposit_at(J[c], d, sink.o[0] + r / v1, sink.o[1] + (r / sink.e[2]) % sink.e[1], sink.o[2] + r % sink.e[2],
                   val);
    }
}

// The six field components (plus the external fields) of the kFB^3 nodes from
// array index fo into the LDS field box (periodic wrap of the indices).  A
// thread's node offsets are the same for every component, so they are
// computed once.
template <int kFB>
__device__ inline void fill_field_box(const PicDesc& d, const FieldPtrs& fp, const int fo[3], float* fbox) {
    constexpr int kFV = kFB * kFB * kFB;
    constexpr int kQ = (kFV + kThreads - 1) / kThreads;
    int off[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const int r = (int)threadIdx.x + q * kThreads;
        const int i = pmod(fo[0] + r / (kFB * kFB), d.dims[0]), j = pmod(fo[1] + (r / kFB) % kFB, d.dims[1]);
        const int k = pmod(fo[2] + r % kFB, d.dims[2]);
        off[q] = (int)fidx(d, i, j, k);
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        const float* F = c < 3 ? fp.E[c] : fp.B[c - 3];
        const float* Fe = c < 3 ? fp.Ee[c] : fp.Be[c - 3];
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const int r = (int)threadIdx.x + q * kThreads;
            if (kFV % kThreads == 0 || r < kFV) {
                float v = F[off[q]];
                if (Fe) v += Fe[off[q]];
                fbox[c * kFV + r] = v;
            }
        }
    }
}

// Adds the non-zero nodes of the kC-component LDS current box (kB^3 nodes
// per component from array index o) into J, dropping indices past the array
// end and wrapping negative ones once (deposit_at); node offsets computed
// once per thread.
template <int kB, int kC = 3>
__device__ inline void flush_current_box(const PicDesc& d, const int o[3], const float* box, float* Jx, float* Jy,
                                         float* Jz) {
    constexpr int kB3 = kB * kB * kB;
    constexpr int kQ = (kB3 + kThreads - 1) / kThreads;
    float* const J[3] = {Jx, Jy, Jz};
    int off[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const int r = (int)threadIdx.x + q * kThreads;
        int i = o[0] + r / (kB * kB), j = o[1] + (r / kB) % kB, k = o[2] + r % kB;
        const bool in = (kB3 % kThreads == 0 || r < kB3) && drop_index(i, d.dims[0]) && drop_index(j, d.dims[1]) &&
                        drop_index(k, d.dims[2]);
        off[q] = in ? (int)fidx(d, i, j, k) : -1;
    }
#pragma unroll
    for (int c = 0; c < kC; ++c) {
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const int r = (int)threadIdx.x + q * kThreads;
            if (off[q] < 0) continue;
            const float v = box[c * kB3 + r];
            if (v != 0.0f) atomic_add_f32(J[c] + off[q], v);
        }
    }
}

}  // namespace
