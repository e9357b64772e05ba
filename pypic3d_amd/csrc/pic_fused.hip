// Single-simulation launch of the fused particle phase (the block function
// and its design notes live in pic_fused_block.h).
#include <algorithm>
#include <cmath>

#include "pic_exact_grid.h"
#include "pic_fused_block.h"

namespace {

// kRounds: particle slots per thread (block = kThreads * kRounds slots);
// kMinBlocks: resident-block target handed to the register allocator.
template <int kDims, int kRounds, int kMinBlocks>
__global__ __launch_bounds__(kThreads, kMinBlocks) void fused_push_esirkepov_v2(PicDesc d, PicSpecies sp,
                                                                                 PicParticles in, FieldPtrs fp,
                                                                                 PicParticles out, float* Jx,
                                                                                 float* Jy, float* Jz, int64_t n) {
    __shared__ FusedShared<kRounds> sh;
    const uint32_t base = blockIdx.x * (uint32_t)(kThreads * kRounds);   // n < 2^31 is checked at launch
    const GlobalSink sink{{Jx, Jy, Jz}};
    fused_block<kDims, kRounds>(d, sp, in, fp, out, sink, n, base, sh);
}

// Tile-ordered launch (particles sorted by pic_tile_keys): block b owns the
// slot range [starts[b], starts[b+1]) of one (species, tile) pair, i.e. the
// particles that were in that T^3-cell tile at the last sort.  Run sums and
// face-crossing deposits accumulate in an LDS box of (T+4)^3 nodes per
// component (the tile, one node below and three above for drift and the
// crossing frame); deposits that fall outside it go straight to J.  The box is
// added into J once per block, so J atomics scale with the tile surface
// instead of with the number of runs.
struct TileGeom {
    int T;
    int tiles;          // tiles per species
    int nt[3];          // tiles per axis
};

template <int kRounds, int kT, int kMinBlocks>
__global__ __launch_bounds__(kThreads, kMinBlocks) void fused_tiled_kernel(PicDesc d, PicSpecies sp, PicParticles in,
                                                                  FieldPtrs fp, PicParticles out, float* Jx,
                                                                  float* Jy, float* Jz,
                                                                  const int32_t* __restrict__ starts, TileGeom tg) {
    constexpr int kB = kT + 4;
    constexpr int kBox = 3 * kB * kB * kB;
    constexpr int kFB = kT + 4;
    constexpr int kFV = kFB * kFB * kFB;
    __shared__ FusedShared<kRounds> sh;
    __shared__ float box[kBox];
    __shared__ float fbox[6 * kFV];
    const int b = blockIdx.x;
    const int64_t begin = starts[b], end = starts[b + 1];
    if (begin >= end) return;
    const int t = b % (tg.tiles + 1);
    int o[3] = {0, 0, 0}, fo[3] = {0, 0, 0};
    if (t < tg.tiles) {
        const int tx = t / (tg.nt[1] * tg.nt[2]), ty = (t / tg.nt[2]) % tg.nt[1], tz = t % tg.nt[2];
        const int c0[3] = {tx * kT, ty * kT, tz * kT};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            o[a] = c0[a] + d.g - 1;
            fo[a] = c0[a] + d.g - 2;
        }
    }
    for (int e = threadIdx.x; e < kBox; e += kThreads) box[e] = 0.0f;
    // E+E_ext, B+B_ext on the (T+4)^3 nodes from T+g-2 (gather anchors of the
    // tile's particles, drifted by up to a cell, and their +1 neighbours)
    for (int e = threadIdx.x; e < 6 * kFV; e += kThreads) {
        const int c = e / kFV, r = e % kFV;
        const int i = pmod(fo[0] + r / (kFB * kFB), d.dims[0]), j = pmod(fo[1] + (r / kFB) % kFB, d.dims[1]);
        const int k = pmod(fo[2] + r % kFB, d.dims[2]);
        const int64_t id = fidx(d, i, j, k);
        const float* F = c < 3 ? fp.E[c] : fp.B[c - 3];
        const float* Fe = c < 3 ? fp.Ee[c] : fp.Be[c - 3];
        float v = F[id];
        if (Fe) v += Fe[id];
        fbox[e] = v;
    }
    __syncthreads();
    const BoxSink sink{{Jx, Jy, Jz}, box, {o[0], o[1], o[2]}, kB};
    const FieldBox fb{fbox, {fo[0], fo[1], fo[2]}, kFB};
    for (int64_t base = begin; base < end; base += kThreads * kRounds) {
        fused_block<3, kRounds, BoxSink, true>(d, sp, in, fp, out, sink, end, (uint32_t)base, sh, fb);
        __syncthreads();
    }
    float* const J[3] = {Jx, Jy, Jz};
    for (int e = threadIdx.x; e < kBox; e += kThreads) {
        const float v = box[e];
        if (v == 0.0f) continue;
        const int c = e / (kB * kB * kB), r = e % (kB * kB * kB);
        deposit_at(J[c], d, o[0] + r / (kB * kB), o[1] + (r / kB) % kB, o[2] + r % kB, v);
    }
}

// Grouped variant of the tile-ordered launch.  Thread t of a chunk owns kP
// consecutive sorted slots [base + kP t, base + kP t + kP), loaded and stored
// as one kP-wide vector per plane when the group lies inside the block's
// range and the planes are aligned.  Sorted slots of one cell are contiguous,
// so a thread's particles mostly share a cell: their twelve regular-case
// values are summed in registers and added into the LDS box when the cell
// changes and once at the end of the group, which replaces the wave
// segmented scan.  Face crossings go through the block's queue as in
// fused_block.
// Per-block LDS queue of face-crossing particles (slot offset from the chunk
// base and the pushed velocity); a full queue sends further particles to the
// rare list.
constexpr int kCrossCap = 512;
struct CrossQueue {
    int q_idx[kCrossCap];
    float q_f[6][kCrossCap];
    int q_n;
};

// Add the twelve regular-case values of a particle anchored at cell
// (a0, a1, a2) (RegularLayout<3> order) into the box, or into J when the
// 2x2x2 node block around the cell leaves the box.
template <int kB>
__device__ inline void flush_cell(const PicDesc& d, const BoxSink& sink, int a0, int a1, int a2, const float acc[12]) {
    const int bi = a0 - sink.o[0], bj = a1 - sink.o[1], bk = a2 - sink.o[2];
    if ((unsigned)bi < (unsigned)(kB - 1) && (unsigned)bj < (unsigned)(kB - 1) && (unsigned)bk < (unsigned)(kB - 1)) {
        float* p = sink.box + (bi * kB + bj) * kB + bk;
#pragma unroll
        for (int m = 0; m < 12; ++m) {
            int dx, dy, dz;
            RegularLayout<3>::delta(m, dx, dy, dz);
            atomicAdd(p + (m >> 2) * kB * kB * kB + (dx * kB + dy) * kB + dz, acc[m]);
        }
    } else {
#pragma unroll
        for (int m = 0; m < 12; ++m) {
            int dx, dy, dz;
            RegularLayout<3>::delta(m, dx, dy, dz);
            deposit_at(sink.J[m >> 2], d, a0 + dx, a1 + dy, a2 + dz, acc[m]);
        }
    }
}

// Periodic wrap of wrap_periodic_position (grid_and_stencil.py:14-35) with
// the one-period cases done without fmod (fmod(s, W) = s - W exactly for
// W <= s < 2W, and = s for -W < s < 0); anything else takes the general form.
__device__ inline void periodic_wrap(double wind, double& x) {
    const double half = 0.5 * wind;
    if (x >= half || x < -half) {
        const double s = x + half;
        double m;
        if (s >= wind && s < 2.0 * wind) {
            m = s - wind;
        } else if (s < 0.0 && s > -wind) {
            m = s + wind;
        } else {
            m = fmod(s, wind);
            if (m != 0.0 && ((m < 0.0) != (wind < 0.0))) m += wind;
        }
        double w = m - half;
        if (w == -half && x >= half) w = half;
        x = w;
    }
}

// One particle of the tile kernel through the general route: global gather
// with FP64 anchors, push, 3-slot (or 5-slot) Esirkepov deposit, position
// update, boundaries and store.  Run by fused_tile_rare_kernel for the rare
// particles the tile kernel's fast loop does not cover.
template <class Sink>
__device__ inline void tile_particle_general(const PicDesc& d, const SpeciesK& k, const PicParticles& in,
                                             const FieldPtrs& fp, const PicParticles& out, const Sink& sink,
                                             int64_t i, int64_t src) {
    double xo[3];
    load_pos(in, src, xo);
    float v[3] = {in.u[0][src], in.u[1][src], in.u[2][src]};
    CicAxis gc[3], gv[3];
    int anc[3], anv[3];
    float dc[3], dv[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float x = (float)xo[a];
        cic_axis(x, (float)d.org_c[a], (float)(1.0 / d.sp_c[a]), d.org_c[a], d.sp_c[a], d.dims[a], gc[a], anc[a],
                 dc[a]);
        cic_axis(x, (float)d.org_v[a], (float)(1.0 / d.sp_v[a]), d.org_v[a], d.sp_v[a], d.dims[a], gv[a], anv[a],
                 dv[a]);
    }
    float E[3], B[3];
    E[0] = gather_cic(fp.E[0], fp.Ee[0], d, gv[0], gc[1], gc[2]);
    E[1] = gather_cic(fp.E[1], fp.Ee[1], d, gc[0], gv[1], gc[2]);
    E[2] = gather_cic(fp.E[2], fp.Ee[2], d, gc[0], gc[1], gv[2]);
    B[0] = gather_cic(fp.B[0], fp.Be[0], d, gc[0], gv[1], gv[2]);
    B[1] = gather_cic(fp.B[1], fp.Be[1], d, gv[0], gc[1], gv[2]);
    B[2] = gather_cic(fp.B[2], fp.Be[2], d, gv[0], gv[1], gc[2]);
    float nv[3] = {v[0], v[1], v[2]};
    push_velocity(nv, E, B, k.h, d);
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if ((k.uu >> a) & 1u) v[a] = nv[a];
    double xn[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) xn[a] = ((k.ux >> a) & 1u) ? step_pos(xo[a], (double)v[a], d.dt) : xo[a];
    CrossAxis cx, cy, cz;
    bool ok = cross_axis(xo[0], xn[0], d.org_c[0], d.sp_c[0], 1.0 / d.d[0], (float)d.org_c[0],
                         (float)(1.0 / d.sp_c[0]), cx);
    ok = cross_axis(xo[1], xn[1], d.org_c[1], d.sp_c[1], 1.0 / d.d[1], (float)d.org_c[1], (float)(1.0 / d.sp_c[1]),
                    cy) && ok;
    ok = cross_axis(xo[2], xn[2], d.org_c[2], d.sp_c[2], 1.0 / d.d[2], (float)d.org_c[2], (float)(1.0 / d.sp_c[2]),
                    cz) && ok;
    if (ok) {
        crossing_deposit(d, k.qw, cx, cy, cz, sink);
    } else {
        float* const J[3] = {sink.J[0], sink.J[1], sink.J[2]};
        esirkepov_particle(d, k.qw, xo, xn, v, J);
    }
    bool act = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) apply_axis_boundary(d.pbc[a], d.wind[a], xn[a], v[a], act);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        out.x[a][i] = act ? (float)xn[a] : 0.0f;
        out.u[a][i] = act ? v[a] : 0.0f;
    }
    out.active[i] = act ? 1 : 0;
}

// Tile-ordered launch on an exact grid.  Block b owns the sorted slot range
// [starts[b], starts[b+1]) of one (species, tile) pair.  Thread t of a chunk
// owns kP consecutive slots; sorted slots of one cell are contiguous, so a
// thread's particles mostly share a cell and their twelve regular-case values
// are summed in registers and added into the LDS current box when the cell
// changes; at the end of the group the lanes of each 16-lane row first sum
// their values over equal cells (row_run_sums), so one lane per cell and row
// adds into the box.  Fields are
// gathered from the LDS field box.  Particles that cross a face go to the
// block's crossing queue (3-slot deposit into the box after the chunk).
// Anything else -- a particle outside the tile's field box or the exact
// anchor range, a step of a cell or more on some axis, a crossing that finds
// the queue full -- is appended to the rare list (rare[0] = count, slots from
// rare[1]) without being stored; fused_tile_rare_kernel then runs it whole.
// Keeping those paths out of this kernel keeps its register count low.


__device__ inline void rare_push(int32_t* rare, int64_t i) {
    const int p = atomicAdd(rare, 1);
    rare[1 + p] = (int32_t)i;
}

// 3-slot crossing frame of one axis (the values cross_axis computes in FP64)
// from the quantities the fast loop queues: the anchor step st = a_new -
// a_old in {-1, 0, +1}, ro = (x_old - node_old)/d and f = (x_new -
// node_new)/d, or the step itself v dt/d when st = 0 (no cancellation then).
// rel is the old anchor relative to the box origin; b becomes the frame's
// first node relative to it.
__device__ inline void cross_frame(int rel, int st, float ro, float f, CrossAxis& e) {
    if (st == 0) {                    // frame (a, a+1, -): S0 = (1-ro, ro), D = (-dl, dl)
        e.b = rel;
        e.S0[0] = 1.0f - ro;
        e.S0[1] = ro;
        e.S0[2] = 0.0f;
        e.D[0] = -f;
        e.D[1] = f;
        e.D[2] = 0.0f;
        e.C0 = -f;
        e.C1 = 0.0f;
    } else if (st > 0) {              // frame (a, a+1, a+2): S0 = (1-ro, ro, 0), S1 = (0, 1-rn, rn)
        e.b = rel;
        e.S0[0] = 1.0f - ro;
        e.S0[1] = ro;
        e.S0[2] = 0.0f;
        e.D[0] = ro - 1.0f;
        e.D[2] = f;
        e.D[1] = -(e.D[0] + f);
        e.C0 = e.D[0];
        e.C1 = -f;
    } else {                          // frame (a-1, a, a+1): S0 = (0, 1-ro, ro), S1 = (1-rn, rn, 0)
        e.b = rel - 1;
        e.S0[0] = 0.0f;
        e.S0[1] = 1.0f - ro;
        e.S0[2] = ro;
        e.D[0] = 1.0f - f;
        e.D[2] = -ro;
        e.D[1] = -(e.D[0] - ro);
        e.C0 = e.D[0];
        e.C1 = ro;
    }
}

// Face-crossing deposit (the 54 terms of crossing_deposit) straight into the
// LDS box when the particle's 3x3x3 frame lies inside it: one base address,
// the term offsets are compile-time constants, no per-term bounds checks.
// pre[c] = -(q w / (d_p d_q dt)) as in SpeciesK.
template <int kC, int kB>
__device__ inline void crossing_component_box(const float pre, const CrossAxis& A, const CrossAxis& P,
                                              const CrossAxis& Q, float* p) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const float ci = i == 0 ? A.C0 : A.C1;
        if (ci == 0.0f) continue;
        const float w = pre * ci;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float y = P.S0[j] * Q.S0[k] + 0.5f * (P.S0[j] * Q.D[k] + P.D[j] * Q.S0[k]) +
                                (1.0f / 3.0f) * P.D[j] * Q.D[k];
                const int dx = kC == 0 ? i : j;
                const int dy = kC == 0 ? j : (kC == 1 ? i : k);
                const int dz = kC == 2 ? i : k;
                atomicAdd(p + kC * kB * kB * kB + (dx * kB + dy) * kB + dz, w * y);
            }
        }
        __builtin_amdgcn_sched_barrier(0);   // bound the live terms to one row of nine
    }
}

// Frames given relative to the box origin; the caller has checked that
// every frame base lies in [0, kB - 3].
template <int kB>
__device__ inline void crossing_deposit_box(const SpeciesK& k, const CrossAxis& cx, const CrossAxis& cy,
                                            const CrossAxis& cz, float* box) {
    float* p = box + (cx.b * kB + cy.b) * kB + cz.b;
    crossing_component_box<0, kB>(k.pre[0], cx, cy, cz, p);
    crossing_component_box<1, kB>(k.pre[1], cy, cx, cz, p);
    crossing_component_box<2, kB>(k.pre[2], cz, cx, cy, p);
}

template <int kT, int kP, int kMinBlocks, int kG, bool kGather>
__global__ __launch_bounds__(kThreads, kMinBlocks) void fused_tile_exact_kernel(
    PicDesc d, PicSpecies sp, PicParticles in, FieldPtrs fp, PicParticles out, float* Jx, float* Jy, float* Jz,
    const int32_t* __restrict__ starts, TileGeom tg, ExactGrid eg, int32_t* __restrict__ rare, int vec_ok, SpeciesTable st,
    const int32_t* __restrict__ gather) {
    constexpr int kB = kT + 4;
    constexpr int kB3 = kB * kB * kB;
    constexpr int kFB = kT + 4;
    constexpr int kFV = kFB * kFB * kFB;
    constexpr int kChunk = kThreads * kP * kG;     // a thread owns kG * kP consecutive slots
    __shared__ CrossQueue sh;
    __shared__ float box[3 * kB3];
    __shared__ float fbox[6 * kFV];
    const int b = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int64_t begin = starts[b], end = starts[b + 1];
    if (begin >= end) return;
    const int t = b % (tg.tiles + 1);
    const int s = b / (tg.tiles + 1);          // blocks are grouped per species
    int o[3] = {0, 0, 0}, fo[3] = {0, 0, 0};
    // interior: every particle the fast path handles (anchors inside the field
    // box, cells c0 - 2 .. c0 + T + 1, a step of under one cell) stays inside
    // [-W/2, W/2) on every axis, so its particle boundaries do nothing and are
    // skipped (a block-uniform branch)
    bool interior = t < tg.tiles;
    if (t < tg.tiles) {
        const int tx = t / (tg.nt[1] * tg.nt[2]), ty = (t / tg.nt[2]) % tg.nt[1], tz = t % tg.nt[2];
        const int c0[3] = {tx * kT, ty * kT, tz * kT};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            o[a] = c0[a] + d.g - 1;
            fo[a] = c0[a] + d.g - 2;
            interior = interior && c0[a] >= 3 && c0[a] + kT + 3 <= d.n[a];
        }
    }
    for (int e = threadIdx.x; e < 3 * kB3; e += kThreads) box[e] = 0.0f;
    fill_field_box<kFB>(d, fp, fo, fbox);
    const BoxSink sink{{Jx, Jy, Jz}, box, {o[0], o[1], o[2]}, kB};
    const FieldBox fb{fbox, {fo[0], fo[1], fo[2]}, kFB};
    const SpeciesK k = st.k[s];
    const float fdt = (float)d.dt;

    for (int64_t base = begin & ~(int64_t)(kP - 1); base < end; base += kChunk) {
        if (threadIdx.x == 0) sh.q_n = 0;
        __syncthreads();    // also orders the box fills before the first deposits
        float acc[12];
#pragma unroll
        for (int m = 0; m < 12; ++m) acc[m] = 0.0f;
        int h0 = 0, h1 = 0, h2 = 0;
        bool have = false;
#pragma unroll 1
        for (int g = 0; g < kG; ++g) {
        const int64_t s0 = base + (int64_t)kP * (kG * (int)threadIdx.x + g);
        const bool full = vec_ok && s0 >= begin && s0 + kP <= end;
        // The thread's kP slots are loaded at once, with one kP-wide vector load
        // per plane when the group lies inside the range (the lanes of a wave
        // then read contiguous memory, and all loads share one round trip).
        // One particle per iteration (the loop stays rolled to bound register
        // pressure): slot j's inputs are taken from the front of PX/PU/pact,
        // which shift by one, and its outputs enter at the back, so after kP
        // iterations PX/PU/oact hold the outputs in slot order for the vector
        // stores.  Slots sent to the rare list are stored as zeros here and
        // rewritten by the rare kernel.
        float PX[3][kP], PU[3][kP];
        uint32_t pact = 0u, oact = 0u;
        if (!kGather && full) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                load_vec<kP>(in.x[a] + s0, PX[a]);
                load_vec<kP>(in.u[a] + s0, PU[a]);
            }
            if constexpr (kP == 4) pact = *reinterpret_cast<const uint32_t*>(in.active + s0);
            else pact = *reinterpret_cast<const uint16_t*>(in.active + s0);
        } else if (kGather && full) {
            // slot i's input is in.*[gather[i]]: the permutation of a re-bin,
            // applied by this launch while it reads.  The flags are loaded
            // first and combined only after every load has been issued (the
            // scheduling barrier keeps the combination from being hoisted
            // between the loads, where each use of a flag would wait for all
            // loads issued before it)
            float gf[kP];
            load_vec<kP>(reinterpret_cast<const float*>(gather + s0), gf);
            uint32_t ab[kP];
#pragma unroll
            for (int j = 0; j < kP; ++j) ab[j] = in.active[__float_as_int(gf[j])];
#pragma unroll
            for (int j = 0; j < kP; ++j) {
                const int64_t src = __float_as_int(gf[j]);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    PX[a][j] = in.x[a][src];
                    PU[a][j] = in.u[a][src];
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < kP; ++j) pact |= (ab[j] != 0u ? 1u : 0u) << (8 * j);
        } else {
#pragma unroll
            for (int j = 0; j < kP; ++j) {
                const int64_t i = s0 + j;
                const bool inr = i >= begin && i < end;
                const int64_t src = kGather ? (inr ? (int64_t)gather[i] : 0) : i;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    PX[a][j] = inr ? in.x[a][src] : 0.0f;
                    PU[a][j] = inr ? in.u[a][src] : 0.0f;
                }
                pact |= (inr && in.active[src] != 0 ? 1u : 0u) << (8 * j);
            }
        }
#pragma unroll 1
        for (int j = 0; j < kP; ++j) {
            const float xf[3] = {PX[0][0], PX[1][0], PX[2][0]};
            float v[3] = {PU[0][0], PU[1][0], PU[2][0]};
            const bool act0 = (pact & 0xffu) != 0u;
            pact >>= 8;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
#pragma unroll
                for (int q = 0; q + 1 < kP; ++q) {
                    PX[a][q] = PX[a][q + 1];
                    PU[a][q] = PU[a][q + 1];
                }
            }
            const int64_t i = s0 + j;
            float ox[3] = {0.0f, 0.0f, 0.0f}, ou[3] = {0.0f, 0.0f, 0.0f};
            uint32_t oa = 0u;
            do {
            if (!act0) break;          // inactive, or outside [begin, end)
            // ---- anchors on both node lines (grid_and_stencil.py:92-166), exact in FP32
            int anc[3], anv[3];
            float dc[3], dv[3], nodec[3];
            bool fast = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                fast = fast && xf[a] >= eg.xlo[a] && xf[a] < eg.xhi[a];
                exact_anchors(xf[a], eg.org_c[a], eg.sp_c[a], eg.isp_c[a], eg.half_sp[a], anc[a], dc[a], nodec[a],
                              anv[a], dv[a]);
                fast = fast && (unsigned)(anc[a] - fo[a]) < (unsigned)(kFB - 1) &&
                       (unsigned)(anv[a] - fo[a]) < (unsigned)(kFB - 1);
            }
            if (!fast) {
                rare_push(rare, i);
                break;
            }
            // ---- gather (boris.py:302-433) from the tile's field box
            CicAxis gc[3], gv[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float rc = dc[a] * eg.isp_c[a], rv = dv[a] * eg.isp_c[a];
                gc[a].p0 = gc[a].p1 = gv[a].p0 = gv[a].p1 = 0;
                gc[a].w0 = 1.0f - rc;
                gc[a].w1 = rc;
                gv[a].w0 = 1.0f - rv;
                gv[a].w1 = rv;
            }
            float E[3], Bf[3];
            E[0] = gather_box(fb, 0, anv[0], anc[1], anc[2], gv[0], gc[1], gc[2]);
            E[1] = gather_box(fb, 1, anc[0], anv[1], anc[2], gc[0], gv[1], gc[2]);
            E[2] = gather_box(fb, 2, anc[0], anc[1], anv[2], gc[0], gc[1], gv[2]);
            Bf[0] = gather_box(fb, 3, anc[0], anv[1], anv[2], gc[0], gv[1], gv[2]);
            Bf[1] = gather_box(fb, 4, anv[0], anc[1], anv[2], gv[0], gc[1], gv[2]);
            Bf[2] = gather_box(fb, 5, anv[0], anv[1], anc[2], gv[0], gv[1], gc[2]);
            // ---- push (boris.py:14-124), masked by update_u (particle_push.py:134-142)
            float nv[3] = {v[0], v[1], v[2]};
            push_velocity(nv, E, Bf, k.h, d);
#pragma unroll
            for (int a = 0; a < 3; ++a)
                if ((k.uu >> a) & 1u) v[a] = nv[a];
            // ---- new position (Esirkepov.py:125-127, FP64) and regular-case test: the
            // new position stays in the old cell, away from its faces by more than the
            // FP64 rounding of floor((x - origin)/spacing).  The test runs on the FP32
            // offset (x - node) + v dt with a band twice as wide (ExactGrid::band_*),
            // which implies the FP64 one; a particle inside the extra band width takes
            // the crossing queue, which treats a stay in the cell like any other step
            double xn[3];
            bool regular = true, short_step = true;
            float S0[3][2], D[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const bool ux = (k.ux >> a) & 1u;
                xn[a] = ux ? step_pos((double)xf[a], (double)v[a], d.dt) : (double)xf[a];
                const float dvt = ux ? v[a] * fdt : 0.0f;
                const float e = dc[a] + dvt;      // a frozen axis tests its offset the same way
                regular = regular && e >= eg.band_lo[a] && e < eg.band_hi[a] && anc[a] >= 0 && anc[a] + 1 < d.dims[a];
                short_step = short_step && fabsf(dvt) < 0.99f * eg.sp_c[a];   // anchor moves by one node at most
                const float r0 = dc[a] * eg.inv_d[a];
                S0[a][0] = 1.0f - r0;
                S0[a][1] = r0;
                D[a] = dvt * eg.inv_d[a];
            }
            if (!regular) {
                // face crossing: the new anchor by bracketing the FP64 new position
                // between the exact nodes next to the old cell (outside the face band)
                // and a frame inside the box, else the rare list; the crossing pass
                // rebuilds the frame from (step, ro, f, old anchor)
                bool ok = short_step;
                int packed = 0;
                float fro[3], ff[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float hi32 = nodec[a] + eg.sp_c[a];
                    const int st = xn[a] < (double)nodec[a] ? -1 : (xn[a] >= (double)hi32 ? 1 : 0);
                    const float noden = st == 0 ? nodec[a] : (st < 0 ? nodec[a] - eg.sp_c[a] : hi32);
                    ok = ok && xn[a] >= (double)noden + eg.tol[a] &&
                         xn[a] < (double)(noden + eg.sp_c[a]) - eg.tol[a];
                    const int rel = anc[a] - o[a];
                    ok = ok && (unsigned)(rel + (st < 0 ? -1 : 0)) <= (unsigned)(kB - 3);
                    fro[a] = dc[a] * eg.inv_d[a];
                    ff[a] = st == 0 ? D[a] : (float)(xn[a] - (double)noden) * eg.inv_d[a];
                    packed |= ((st + 1) << (10 + 2 * a)) | ((rel & 31) << (16 + 5 * a));
                }
                const int slot = ok ? atomicAdd(&sh.q_n, 1) : kCrossCap;
                if (slot >= kCrossCap) {
                    rare_push(rare, i);
                    break;
                }
                sh.q_idx[slot] = packed;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    sh.q_f[a][slot] = fro[a];
                    sh.q_f[3 + a][slot] = ff[a];
                }
            }
            if (regular) {
                // difference-form Esirkepov weights of a particle that stays in its
                // cell (Esirkepov.py:158-214): for component c over the other axes p, q,
                // W = D_c (S0p S0q + (S0p dq + dp S0q)/2 + dp dq/3) with d = -D, +D at
                // slots 0, 1, i.e. W = D_c (Ap Aq + dp dq/12) with A = S0 + d/2;
                // J_c = -(q/(d_p d_q dt)) W at slot a only
                float A[3][2];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    A[a][1] = S0[a][1] + 0.5f * D[a];
                    A[a][0] = S0[a][0] - 0.5f * D[a];
                }
                float val[12];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int o1 = c == 0 ? 1 : 0, o2 = c == 2 ? 1 : 2;
                    const float pre = k.pre[c] * (-D[c]);
                    const float e = D[o1] * D[o2] * (1.0f / 12.0f);
                    val[c * 4 + 0] = pre * fmaf(A[o1][0], A[o2][0], e);
                    val[c * 4 + 1] = pre * fmaf(A[o1][0], A[o2][1], -e);
                    val[c * 4 + 2] = pre * fmaf(A[o1][1], A[o2][0], -e);
                    val[c * 4 + 3] = pre * fmaf(A[o1][1], A[o2][1], e);
                }
                const bool same = have && anc[0] == h0 && anc[1] == h1 && anc[2] == h2;
                if (have && !same) flush_cell<kB>(d, sink, h0, h1, h2, acc);
#pragma unroll
                for (int m = 0; m < 12; ++m) acc[m] = (same ? acc[m] : 0.0f) + val[m];
                h0 = anc[0];
                h1 = anc[1];
                h2 = anc[2];
                have = true;
            }
            // ---- position update + particle boundaries (particle_tile_communication.py)
            bool act = true;
            if (!interior) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    if (d.pbc[a] == 0) periodic_wrap(d.wind[a], xn[a]);
                    else apply_axis_boundary(d.pbc[a], d.wind[a], xn[a], v[a], act);
                }
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                ox[a] = act ? (float)xn[a] : 0.0f;
                ou[a] = act ? v[a] : 0.0f;
            }
            oa = act ? 1u : 0u;
            } while (false);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                PX[a][kP - 1] = ox[a];
                PU[a][kP - 1] = ou[a];
            }
            oact = (oact >> 8) | (oa << (8 * (kP - 1)));
        }
        if (full) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                store_vec<kP>(out.x[a] + s0, PX[a]);
                store_vec<kP>(out.u[a] + s0, PU[a]);
            }
            if constexpr (kP == 4) *reinterpret_cast<uint32_t*>(out.active + s0) = oact;
            else *reinterpret_cast<uint16_t*>(out.active + s0) = (uint16_t)oact;
        } else {
#pragma unroll
            for (int j = 0; j < kP; ++j) {
                const int64_t i = s0 + j;
                if (i < begin || i >= end) continue;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    out.x[a][i] = PX[a][j];
                    out.u[a][i] = PU[a][j];
                }
                out.active[i] = (uint8_t)((oact >> (8 * j)) & 0xffu);
            }
        }
        }   // slot groups
        {
            // the thread's last cell: equal cells of neighbouring lanes are summed
            // first, so one lane per cell and 16-lane row adds into the box (a
            // cell outside the box keeps a key of its own)
            const int bi = h0 - o[0], bj = h1 - o[1], bk = h2 - o[2];
            const bool inbox = (unsigned)bi < (unsigned)(kB - 1) && (unsigned)bj < (unsigned)(kB - 1) &&
                               (unsigned)bk < (unsigned)(kB - 1);
            const int key = !have ? -1 : (inbox ? (bi * kB + bj) * kB + bk : -2 - (int)(threadIdx.x & 63u));
            if (row_run_sums(key, acc) && have) flush_cell<kB>(d, sink, h0, h1, h2, acc);
        }
        __syncthreads();
        // ---- queued face crossings (3-slot frame, Esirkepov.py:105-331)
        const int nq = min(sh.q_n, kCrossCap);
        for (int q = threadIdx.x; q < nq; q += kThreads) {
            const int e = sh.q_idx[q];
            CrossAxis ca[3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
                cross_frame((e >> (16 + 5 * a)) & 31, ((e >> (10 + 2 * a)) & 3) - 1, sh.q_f[a][q], sh.q_f[3 + a][q],
                            ca[a]);
            crossing_deposit_box<kB>(k, ca[0], ca[1], ca[2], box);
        }
        __syncthreads();
    }
    flush_current_box<kB>(d, o, box, Jx, Jy, Jz);
}

// The rare list of fused_tile_exact_kernel: each listed particle runs the
// general route with its deposit added straight into J.
__global__ __launch_bounds__(kThreads) void fused_tile_rare_kernel(PicDesc d, PicSpecies sp, PicParticles in,
                                                                   FieldPtrs fp, PicParticles out, float* Jx,
                                                                   float* Jy, float* Jz,
                                                                   const int32_t* __restrict__ rare,
                                                                   const int32_t* __restrict__ gather) {
    const int n = rare[0];
    const GlobalSink sink{{Jx, Jy, Jz}};
    for (int q = blockIdx.x * kThreads + threadIdx.x; q < n; q += gridDim.x * kThreads) {
        const int64_t i = rare[1 + q];
        const SpeciesK k = species_constants(d, sp, species_of(i, sp.slots, sp.count));
        tile_particle_general(d, k, in, fp, out, sink, i, gather ? (int64_t)gather[i] : i);
    }
}

__global__ void zero_count_kernel(int32_t* c) { *c = 0; }

// Direct-current branch of time_loop_electrodynamic (evolve.py:44-67) on
// tile-sorted particles of an exact 3-D first-order grid, fused: gather from
// the tile's LDS field box, push, x += v dt/2 (FP64, rounded to FP32 like
// pic_advance_positions), particle boundaries, the J_from_rhov deposit at the
// half position (J_from_rhov.py:88-200: node weights 1 - dn/d, dn/d and face
// weights 1 - df/d, df/d with df = dn - d/2 on the 2x2x2 nodes of the centre
// anchor, summed in registers over a thread's equal-anchor particles and
// added into an LDS J box), x += v dt/2, boundaries, store.  Particles the
// fast loop does not cover are listed for fused_direct_rare_kernel.  kVec:
// a thread's four slots are stored with one 16-byte store per plane at the
// end (24 more registers); otherwise each slot is stored as it finishes,
// which keeps the kernel at 128 VGPRs for 4 waves per SIMD.
template <int kT, int kMinBlocks, bool kVec, int kP = 4, bool kGather = false>
__global__ __launch_bounds__(kThreads, kMinBlocks) void fused_tile_direct_kernel(
    PicDesc d, PicSpecies sp, PicParticles in, FieldPtrs fp, PicParticles out, float* Jx, float* Jy, float* Jz,
    const int32_t* __restrict__ starts, TileGeom tg, ExactGrid eg, int32_t* __restrict__ rare, int vec_ok, SpeciesTable st,
    const int32_t* __restrict__ gather) {
    constexpr int kB = kT + 5;
    constexpr int kB3 = kB * kB * kB;
    constexpr int kFB = kT + 4;
    constexpr int kFV = kFB * kFB * kFB;
    constexpr int kChunk = kThreads * kP;
    __shared__ float box[3 * kB3];
    __shared__ float fbox[6 * kFV];
    const int b = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int64_t begin = starts[b], end = starts[b + 1];
    if (begin >= end) return;
    const int t = b % (tg.tiles + 1);
    const int s = b / (tg.tiles + 1);
    int o[3] = {0, 0, 0}, fo[3] = {0, 0, 0};
    // interior: a fast-path particle (anchors inside the field box, cells
    // c0 - 2 .. c0 + T + 1) that moves less than a cell per step stays inside
    // [-W/2, W/2) on every axis, so its particle boundaries do nothing
    bool interior = t < tg.tiles;
    if (t < tg.tiles) {
        const int tx = t / (tg.nt[1] * tg.nt[2]), ty = (t / tg.nt[2]) % tg.nt[1], tz = t % tg.nt[2];
        const int c0[3] = {tx * kT, ty * kT, tz * kT};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            o[a] = fo[a] = c0[a] + d.g - 2;
            interior = interior && c0[a] >= 3 && c0[a] + kT + 3 <= d.n[a];
        }
    }
    const float fdt = (float)d.dt;
    for (int e = threadIdx.x; e < 3 * kB3; e += kThreads) box[e] = 0.0f;
    fill_field_box<kFB>(d, fp, fo, fbox);
    __syncthreads();
    const FieldBox fb{fbox, {fo[0], fo[1], fo[2]}, kFB};
    const SpeciesK k = st.k[s];
    const double hdt = 0.5 * d.dt;
    auto flush = [&](int h0, int h1, int h2, const float* acc) {
        float* p = box + ((h0 - o[0]) * kB + (h1 - o[1])) * kB + (h2 - o[2]);
#pragma unroll
        for (int m = 0; m < 24; ++m) {
            const int c = m >> 3, n = m & 7;
            atomicAdd(p + c * kB3 + ((n >> 2) * kB + ((n >> 1) & 1)) * kB + (n & 1), acc[m]);
        }
    };

    for (int64_t base = begin & ~(int64_t)(kP - 1); base < end; base += kChunk) {
        const int64_t s0 = base + (int64_t)kP * threadIdx.x;
        float acc[24];
#pragma unroll
        for (int m = 0; m < 24; ++m) acc[m] = 0.0f;
        int h0 = 0, h1 = 0, h2 = 0;
        bool have = false;
        // kVec: OX/OU first hold the group's inputs (one kP-wide load per plane
        // when the group lies inside the range) and take its outputs slot by
        // slot, as in fused_tile_exact_kernel
        float OX[3][kP], OU[3][kP];
        uint32_t pact = 0u, oact = 0u;
        const bool full = vec_ok && s0 >= begin && s0 + kP <= end;
        if constexpr (kVec) {
            if (!kGather && full) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    load_vec<kP>(in.x[a] + s0, OX[a]);
                    load_vec<kP>(in.u[a] + s0, OU[a]);
                }
                if constexpr (kP == 4) pact = *reinterpret_cast<const uint32_t*>(in.active + s0);
                else pact = *reinterpret_cast<const uint16_t*>(in.active + s0);
            } else if (kGather && full) {
                // slot i's input is in.*[gather[i]] (a re-bin's permutation)
                float gf[kP];
                load_vec<kP>(reinterpret_cast<const float*>(gather + s0), gf);
#pragma unroll
                for (int j = 0; j < kP; ++j) {
                    const int64_t src = __float_as_int(gf[j]);
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        OX[a][j] = in.x[a][src];
                        OU[a][j] = in.u[a][src];
                    }
                    pact |= (in.active[src] != 0 ? 1u : 0u) << (8 * j);
                }
            } else {
#pragma unroll
                for (int j = 0; j < kP; ++j) {
                    const int64_t i = s0 + j;
                    const bool inr = i >= begin && i < end;
                    const int64_t src = kGather ? (inr ? (int64_t)gather[i] : 0) : i;
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        OX[a][j] = inr ? in.x[a][src] : 0.0f;
                        OU[a][j] = inr ? in.u[a][src] : 0.0f;
                    }
                    pact |= (inr && in.active[src] != 0 ? 1u : 0u) << (8 * j);
                }
            }
        }
#pragma unroll 1
        for (int j = 0; j < kP; ++j) {
            const int64_t i = s0 + j;
            float ox[3] = {0.0f, 0.0f, 0.0f}, ou[3] = {0.0f, 0.0f, 0.0f};
            uint32_t oa = 0u;
            float xf[3], v[3];
            bool act0;
            if constexpr (kVec) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    xf[a] = OX[a][0];
                    v[a] = OU[a][0];
#pragma unroll
                    for (int q = 0; q + 1 < kP; ++q) {
                        OX[a][q] = OX[a][q + 1];
                        OU[a][q] = OU[a][q + 1];
                    }
                }
                act0 = (pact & 0xffu) != 0u;
                pact >>= 8;
            }
            do {
            if constexpr (!kVec) {
                if (i < begin || i >= end) break;
                const int64_t src = kGather ? (int64_t)gather[i] : i;
                act0 = in.active[src] != 0;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    xf[a] = in.x[a][src];
                    v[a] = in.u[a][src];
                }
                loads_together(xf, v);
            }
            if (!act0) break;
            int anc[3], anv[3];
            float dc[3], dv[3], nodec[3];
            bool fast = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                fast = fast && xf[a] >= eg.xlo[a] && xf[a] < eg.xhi[a];
                exact_anchors(xf[a], eg.org_c[a], eg.sp_c[a], eg.isp_c[a], eg.half_sp[a], anc[a], dc[a], nodec[a],
                              anv[a], dv[a]);
                fast = fast && (unsigned)(anc[a] - fo[a]) < (unsigned)(kFB - 1) &&
                       (unsigned)(anv[a] - fo[a]) < (unsigned)(kFB - 1);
            }
            if (!fast) {
                rare_push(rare, i);
                break;
            }
            CicAxis gc[3], gv[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float rc = dc[a] * eg.isp_c[a], rv = dv[a] * eg.isp_c[a];
                gc[a].p0 = gc[a].p1 = gv[a].p0 = gv[a].p1 = 0;
                gc[a].w0 = 1.0f - rc;
                gc[a].w1 = rc;
                gv[a].w0 = 1.0f - rv;
                gv[a].w1 = rv;
            }
            float E[3], Bf[3];
            E[0] = gather_box(fb, 0, anv[0], anc[1], anc[2], gv[0], gc[1], gc[2]);
            E[1] = gather_box(fb, 1, anc[0], anv[1], anc[2], gc[0], gv[1], gc[2]);
            E[2] = gather_box(fb, 2, anc[0], anc[1], anv[2], gc[0], gc[1], gv[2]);
            Bf[0] = gather_box(fb, 3, anc[0], anv[1], anv[2], gc[0], gv[1], gv[2]);
            Bf[1] = gather_box(fb, 4, anv[0], anc[1], anv[2], gv[0], gc[1], gv[2]);
            Bf[2] = gather_box(fb, 5, anv[0], anv[1], anc[2], gv[0], gv[1], gc[2]);
            float nv[3] = {v[0], v[1], v[2]};
            push_velocity(nv, E, Bf, k.h, d);
#pragma unroll
            for (int a = 0; a < 3; ++a)
                if ((k.uu >> a) & 1u) v[a] = nv[a];
            // ---- x += v dt/2, boundaries (evolve.py:48-52)
            bool act = true;
            bool skip_bc = interior;
#pragma unroll
            for (int a = 0; a < 3; ++a) skip_bc = skip_bc && fabsf(v[a]) * fdt < 0.99f * eg.sp_c[a];
            float xh[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                double t = ((k.ux >> a) & 1u) ? (double)(float)step_pos((double)xf[a], (double)v[a], hdt)
                                              : (double)xf[a];
                if (!skip_bc) {
                    if (d.pbc[a] == 0) periodic_wrap(d.wind[a], t);
                    else apply_axis_boundary(d.pbc[a], d.wind[a], t, v[a], act);
                }
                xh[a] = (float)t;
            }
            if (act) {
                // ---- J_from_rhov at the half position (anchors exact in FP32)
                int an[3];
                float dn[3];
                bool inbox = true;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    int av2;
                    float nd, dv2;
                    inbox = inbox && xh[a] >= eg.xlo[a] && xh[a] < eg.xhi[a];
                    exact_anchors(xh[a], eg.org_c[a], eg.sp_c[a], eg.isp_c[a], eg.half_sp[a], an[a], dn[a], nd, av2,
                                  dv2);
                    inbox = inbox && (unsigned)(an[a] - o[a]) <= (unsigned)(kB - 2);
                }
                if (!inbox) {
                    rare_push(rare, i);
                    break;
                }
                float wn[3][2], wf[3][2];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float r = dn[a] * eg.inv_d[a];
                    wn[a][0] = 1.0f - r;
                    wn[a][1] = r;
                    const float rf = (dn[a] - eg.half_sp[a]) * eg.inv_d[a];
                    wf[a][0] = 1.0f - rf;
                    wf[a][1] = rf;
                }
                const float qv[3] = {k.q_dv * v[0], k.q_dv * v[1], k.q_dv * v[2]};
                float val[24];
#pragma unroll
                for (int n = 0; n < 8; ++n) {
                    const int xi = n >> 2, yi = (n >> 1) & 1, zi = n & 1;
                    val[n] = qv[0] * wf[0][xi] * wn[1][yi] * wn[2][zi];
                    val[8 + n] = qv[1] * wn[0][xi] * wf[1][yi] * wn[2][zi];
                    val[16 + n] = qv[2] * wn[0][xi] * wn[1][yi] * wf[2][zi];
                }
                const bool same = have && an[0] == h0 && an[1] == h1 && an[2] == h2;
                if (have && !same) flush(h0, h1, h2, acc);
#pragma unroll
                for (int m = 0; m < 24; ++m) acc[m] = (same ? acc[m] : 0.0f) + val[m];
                h0 = an[0];
                h1 = an[1];
                h2 = an[2];
                have = true;
                // ---- x += v dt/2, boundaries (evolve.py:62-66)
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    double t2 = ((k.ux >> a) & 1u) ? (double)(float)step_pos((double)xh[a], (double)v[a], hdt)
                                                   : (double)xh[a];
                    if (!skip_bc) {
                        if (d.pbc[a] == 0) periodic_wrap(d.wind[a], t2);
                        else apply_axis_boundary(d.pbc[a], d.wind[a], t2, v[a], act);
                    }
                    ox[a] = (float)t2;
                    ou[a] = v[a];
                }
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    ox[a] = act ? ox[a] : 0.0f;
                    ou[a] = act ? ou[a] : 0.0f;
                }
                oa = act ? 1u : 0u;
            }
            } while (false);
            if constexpr (kVec) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    OX[a][kP - 1] = ox[a];
                    OU[a][kP - 1] = ou[a];
                }
                oact = (oact >> 8) | (oa << (8 * (kP - 1)));
            } else if (i >= begin && i < end) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    out.x[a][i] = ox[a];
                    out.u[a][i] = ou[a];
                }
                out.active[i] = (uint8_t)oa;
            }
        }
        {
            // the group's last cell, summed over equal cells of neighbouring lanes first
            const int key = have ? ((h0 - o[0]) * kB + (h1 - o[1])) * kB + (h2 - o[2]) : -1;
            if (row_run_sums(key, acc) && have) flush(h0, h1, h2, acc);
        }
        if constexpr (!kVec) continue;
        if (full) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                store_vec<kP>(out.x[a] + s0, OX[a]);
                store_vec<kP>(out.u[a] + s0, OU[a]);
            }
            if constexpr (kP == 4) *reinterpret_cast<uint32_t*>(out.active + s0) = oact;
            else *reinterpret_cast<uint16_t*>(out.active + s0) = (uint16_t)oact;
        } else {
#pragma unroll
            for (int j = 0; j < kP; ++j) {
                const int64_t i = s0 + j;
                if (i < begin || i >= end) continue;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    out.x[a][i] = OX[a][j];
                    out.u[a][i] = OU[a][j];
                }
                out.active[i] = (uint8_t)((oact >> (8 * j)) & 0xffu);
            }
        }
    }
    __syncthreads();
    flush_current_box<kB>(d, o, box, Jx, Jy, Jz);
}

// The rare list of fused_tile_direct_kernel: each listed particle runs the
// direct branch through the general route (global gather, FP64 anchors, the
// 3^3 deposit into J).
__global__ __launch_bounds__(kThreads) void fused_direct_rare_kernel(PicDesc d, PicSpecies sp, PicParticles in,
                                                                     FieldPtrs fp, PicParticles out, float* Jx,
                                                                     float* Jy, float* Jz,
                                                                     const int32_t* __restrict__ rare,
                                                                     const int32_t* __restrict__ gather) {
    const int n = rare[0];
    const GlobalSink sink{{Jx, Jy, Jz}};
    for (int q = blockIdx.x * kThreads + threadIdx.x; q < n; q += gridDim.x * kThreads) {
        const int64_t i = rare[1 + q];
        const int64_t src = gather ? (int64_t)gather[i] : i;      // input slot (a re-bin's permutation)
        const SpeciesK k = species_constants(d, sp, species_of(i, sp.slots, sp.count));
        double pos[3];
        load_pos(in, src, pos);
        float v[3] = {in.u[0][src], in.u[1][src], in.u[2][src]};
        float E[3], B[3];
        CicAxis gc[3], gv[3];
        for (int a = 0; a < 3; ++a) {
            int an;
            float dl;
            cic_axis((float)pos[a], (float)d.org_c[a], (float)(1.0 / d.sp_c[a]), d.org_c[a], d.sp_c[a], d.dims[a],
                     gc[a], an, dl);
            cic_axis((float)pos[a], (float)d.org_v[a], (float)(1.0 / d.sp_v[a]), d.org_v[a], d.sp_v[a], d.dims[a],
                     gv[a], an, dl);
        }
        E[0] = gather_cic(fp.E[0], fp.Ee[0], d, gv[0], gc[1], gc[2]);
        E[1] = gather_cic(fp.E[1], fp.Ee[1], d, gc[0], gv[1], gc[2]);
        E[2] = gather_cic(fp.E[2], fp.Ee[2], d, gc[0], gc[1], gv[2]);
        B[0] = gather_cic(fp.B[0], fp.Be[0], d, gc[0], gv[1], gv[2]);
        B[1] = gather_cic(fp.B[1], fp.Be[1], d, gv[0], gc[1], gv[2]);
        B[2] = gather_cic(fp.B[2], fp.Be[2], d, gv[0], gv[1], gc[2]);
        float nv[3] = {v[0], v[1], v[2]};
        push_velocity(nv, E, B, k.h, d);
        for (int a = 0; a < 3; ++a)
            if ((k.uu >> a) & 1u) v[a] = nv[a];
        bool act = true;
        double xh[3];
        for (int a = 0; a < 3; ++a) {
            xh[a] = ((k.ux >> a) & 1u) ? (double)(float)step_pos(pos[a], (double)v[a], 0.5 * d.dt) : pos[a];
            apply_axis_boundary(d.pbc[a], d.wind[a], xh[a], v[a], act);
            xh[a] = (double)(float)xh[a];
        }
        if (act) {
            direct_deposit(d, k.q_dv, xh, v, sink);
            for (int a = 0; a < 3; ++a) {
                double t2 = ((k.ux >> a) & 1u) ? (double)(float)step_pos(xh[a], (double)v[a], 0.5 * d.dt) : xh[a];
                apply_axis_boundary(d.pbc[a], d.wind[a], t2, v[a], act);
                xh[a] = t2;
            }
        }
        for (int a = 0; a < 3; ++a) {
            out.x[a][i] = act ? (float)xh[a] : 0.0f;
            out.u[a][i] = act ? v[a] : 0.0f;
        }
        out.active[i] = act ? 1 : 0;
    }
}

// Slot-ordered launch for 1-D and 2-D first-order runs (and the general
// path): each block accumulates J in an LDS box sized to its particles.
template <int kRounds>
__global__ __launch_bounds__(kThreads, 1) void fused_boxed_kernel(PicDesc d, PicSpecies sp, PicParticles in,
                                                                  FieldPtrs fp, PicParticles out, float* Jx,
                                                                  float* Jy, float* Jz, int64_t n) {
    __shared__ FusedShared<kRounds> sh;
    __shared__ float box[kRangeBoxCap];
    __shared__ int bmin[3], bmax[3];
    float* const J[3] = {Jx, Jy, Jz};
    fused_block_boxed<kRounds>(d, sp, in, fp, out, J, n, blockIdx.x * (uint32_t)(kThreads * kRounds), sh, box, bmin,
                               bmax);
}

}  // namespace

extern "C" {

int pic_fused_push_esirkepov(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                             const float* const E[3], const float* const B[3], const float* const E_ext[3],
                             const float* const B_ext[3], PicParticles* out, float* const J[3], void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok(sp, in) && particles_ok(sp, out) && E && B && J);
    PIC_REQUIRE(desc->n[0] > 1 || desc->n[1] > 1 || desc->n[2] > 1);
    const bool empty = (int64_t)sp->count * sp->slots == 0;
    for (int a = 0; a < 3; ++a)
        PIC_REQUIRE(E[a] && B[a] && J[a] && (empty || (in->x[a] != out->x[a] && in->u[a] != out->u[a])));
    PIC_REQUIRE(empty || in->active != out->active);
    PIC_REQUIRE(field_size(*desc) < ((int64_t)1 << 31));
    int64_t n = (int64_t)sp->count * sp->slots;
    PIC_REQUIRE(n < ((int64_t)1 << 31) - 8 * kThreads);
    if (n == 0) return 0;
    const int kind = fused_kind(*desc);
    hipStream_t st = as_stream(stream);
    FieldPtrs fp = make_fields(E, B, E_ext, B_ext);
#define PIC_FUSED_LAUNCH(DIMS, R, MB)                                                                           \
    hipLaunchKernelGGL((fused_push_esirkepov_v2<DIMS, R, MB>), dim3(grid_1d(n, kThreads * R)), dim3(kThreads), 0, \
                       st, *desc, *sp, *in, fp, *out, J[0], J[1], J[2], n)
    if (kind == 3) {
        switch (desc->variant) {
            case 1: PIC_FUSED_LAUNCH(3, 4, 3); break;
            case 2: PIC_FUSED_LAUNCH(3, 4, 4); break;
            case 3: PIC_FUSED_LAUNCH(3, 8, 1); break;
            case 4: PIC_FUSED_LAUNCH(3, 2, 1); break;
            case 5: PIC_FUSED_LAUNCH(3, 8, 3); break;
            default: PIC_FUSED_LAUNCH(3, 4, 1); break;
        }
    } else {
        hipLaunchKernelGGL(fused_boxed_kernel<4>, dim3(grid_1d(n, kThreads * 4)), dim3(kThreads), 0, st, *desc, *sp,
                           *in, fp, *out, J[0], J[1], J[2], n);
    }
#undef PIC_FUSED_LAUNCH
    PIC_CHECK_LAUNCH();
    return 0;
}

static int fused_direct_tiled(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                              const float* const E[3], const float* const B[3], const float* const E_ext[3],
                              const float* const B_ext[3], PicParticles* out, float* const J[3], int32_t tile,
                              const int32_t* starts, int32_t min_blocks, int32_t* workspace, const int32_t* gather,
                              void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok(sp, in) && particles_ok(sp, out) && E && B && J && starts && workspace);
    PIC_REQUIRE(fused_kind(*desc) == 3 && (tile == 4 || tile == 8) && (min_blocks == 1 || min_blocks == 4));
    ExactGrid eg;
    PIC_REQUIRE(make_exact_grid(*desc, eg));
    for (int a = 0; a < 3; ++a) PIC_REQUIRE(desc->d[a] == desc->sp_c[a]);
    const bool empty = (int64_t)sp->count * sp->slots == 0;
    for (int a = 0; a < 3; ++a)
        PIC_REQUIRE(E[a] && B[a] && J[a] && (empty || (in->x[a] != out->x[a] && in->u[a] != out->u[a])));
    PIC_REQUIRE(empty || in->active != out->active);
    PIC_REQUIRE(field_size(*desc) < ((int64_t)1 << 31));
    int64_t n = (int64_t)sp->count * sp->slots;
    PIC_REQUIRE(n < ((int64_t)1 << 31) - 8 * kThreads);
    if (n == 0) return 0;
    TileGeom tg;
    tg.T = tile;
    for (int a = 0; a < 3; ++a) tg.nt[a] = (desc->n[a] + tile - 1) / tile;
    tg.tiles = tg.nt[0] * tg.nt[1] * tg.nt[2];
    const int64_t blocks = (int64_t)sp->count * (tg.tiles + 1);
    PIC_REQUIRE(blocks < ((int64_t)1 << 31));
    hipStream_t st = as_stream(stream);
    FieldPtrs fp = make_fields(E, B, E_ext, B_ext);
    int vec_ok = 1;
    for (int a = 0; a < 3; ++a)
        vec_ok &= ((((uintptr_t)out->x[a]) | ((uintptr_t)out->u[a]) | ((uintptr_t)in->x[a]) |
                    ((uintptr_t)in->u[a])) & 15u) == 0;
    vec_ok &= ((((uintptr_t)out->active) | ((uintptr_t)in->active)) & 3u) == 0;
    vec_ok &= (((uintptr_t)gather) & 15u) == 0;
    hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(1), 0, st, workspace);
    const SpeciesTable stab = species_table(*desc, *sp);
#define PIC_DIRECT_LAUNCH2(T, MB, V, P)                                                                              \
    do {                                                                                                             \
        if (gather)                                                                                                  \
            hipLaunchKernelGGL((fused_tile_direct_kernel<T, MB, V, P, true>), dim3((unsigned)blocks), dim3(kThreads), \
                               0, st, *desc, *sp, *in, fp, *out, J[0], J[1], J[2], starts, tg, eg, workspace, vec_ok, \
                               stab, gather);                                                                        \
        else                                                                                                         \
            hipLaunchKernelGGL((fused_tile_direct_kernel<T, MB, V, P, false>), dim3((unsigned)blocks),               \
                               dim3(kThreads), 0, st, *desc, *sp, *in, fp, *out, J[0], J[1], J[2], starts, tg, eg,   \
                               workspace, vec_ok, stab, gather);                                                     \
    } while (0)
#define PIC_DIRECT_LAUNCH(T, MB, V) PIC_DIRECT_LAUNCH2(T, MB, V, 4)
    if (tile == 8) {
        if (min_blocks == 4) PIC_DIRECT_LAUNCH(8, 4, false);
        else PIC_DIRECT_LAUNCH(8, 1, true);
    } else {
        if (min_blocks == 4) PIC_DIRECT_LAUNCH2(4, 4, true, 2);
        else PIC_DIRECT_LAUNCH(4, 1, true);
    }
#undef PIC_DIRECT_LAUNCH
#undef PIC_DIRECT_LAUNCH2
    const int64_t rare_grid = std::min<int64_t>(std::max<int64_t>(grid_1d(n, kThreads), 1), 1024);
    hipLaunchKernelGGL(fused_direct_rare_kernel, dim3((unsigned)rare_grid), dim3(kThreads), 0, st, *desc, *sp, *in,
                       fp, *out, J[0], J[1], J[2], workspace, gather);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_fused_push_direct_tiled(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                                const float* const E[3], const float* const B[3], const float* const E_ext[3],
                                const float* const B_ext[3], PicParticles* out, float* const J[3], int32_t tile,
                                const int32_t* starts, int32_t min_blocks, int32_t* workspace, void* stream) {
    return fused_direct_tiled(desc, sp, in, E, B, E_ext, B_ext, out, J, tile, starts, min_blocks, workspace, nullptr,
                              stream);
}

int pic_fused_push_direct_tiled_gather(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                                       const int32_t* gather, const float* const E[3], const float* const B[3],
                                       const float* const E_ext[3], const float* const B_ext[3], PicParticles* out,
                                       float* const J[3], int32_t tile, const int32_t* starts, int32_t min_blocks,
                                       int32_t* workspace, void* stream) {
    PIC_REQUIRE(sp && (gather || (int64_t)sp->count * sp->slots == 0));
    return fused_direct_tiled(desc, sp, in, E, B, E_ext, B_ext, out, J, tile, starts, min_blocks, workspace, gather,
                              stream);
}

int pic_tile_exact_grid(const PicDesc* desc) {
    ExactGrid eg;
    return desc && desc_ok(desc) && make_exact_grid(*desc, eg) ? 1 : 0;
}

static int fused_esirkepov_tiled(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                                 const float* const E[3], const float* const B[3], const float* const E_ext[3],
                                 const float* const B_ext[3], PicParticles* out, float* const J[3], int32_t tile,
                                 const int32_t* starts, int32_t rounds, int32_t min_blocks, int32_t* workspace,
                                 const int32_t* gather, void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok(sp, in) && particles_ok(sp, out) && E && B && J && starts);
    PIC_REQUIRE(fused_kind(*desc) == 3 && (tile == 4 || tile == 8) &&
                (rounds == 2 || rounds == 4 || rounds == 8));
    PIC_REQUIRE(min_blocks == 1 || min_blocks == 3 || min_blocks == 4);
    const bool empty = (int64_t)sp->count * sp->slots == 0;
    for (int a = 0; a < 3; ++a)
        PIC_REQUIRE(E[a] && B[a] && J[a] && (empty || (in->x[a] != out->x[a] && in->u[a] != out->u[a])));
    PIC_REQUIRE(empty || in->active != out->active);
    PIC_REQUIRE(field_size(*desc) < ((int64_t)1 << 31));
    int64_t n = (int64_t)sp->count * sp->slots;
    PIC_REQUIRE(n < ((int64_t)1 << 31) - 8 * kThreads);
    if (n == 0) return 0;
    TileGeom tg;
    tg.T = tile;
    for (int a = 0; a < 3; ++a) tg.nt[a] = (desc->n[a] + tile - 1) / tile;
    tg.tiles = tg.nt[0] * tg.nt[1] * tg.nt[2];
    const int64_t blocks = (int64_t)sp->count * (tg.tiles + 1);
    PIC_REQUIRE(blocks < ((int64_t)1 << 31));
    hipStream_t st = as_stream(stream);
    FieldPtrs fp = make_fields(E, B, E_ext, B_ext);
#define PIC_TILED_LAUNCH(R, T, MB)                                                                              \
    hipLaunchKernelGGL((fused_tiled_kernel<R, T, MB>), dim3((unsigned)blocks), dim3(kThreads), 0, st, *desc, *sp, \
                       *in, fp, *out, J[0], J[1], J[2], starts, tg)
#define PIC_TILED_MB(R, T)                              \
    do {                                                \
        if (min_blocks == 3) PIC_TILED_LAUNCH(R, T, 3); \
        else if (min_blocks == 4) PIC_TILED_LAUNCH(R, T, 4); \
        else PIC_TILED_LAUNCH(R, T, 1);                 \
    } while (0)
    ExactGrid eg;
    if (desc->variant != 1 && make_exact_grid(*desc, eg)) {
        PIC_REQUIRE(workspace);
        // vector loads and stores of a thread's slot group need 16-byte aligned planes
        int vec_ok = 1;
        for (int a = 0; a < 3; ++a)
            vec_ok &= ((((uintptr_t)out->x[a]) | ((uintptr_t)out->u[a]) | ((uintptr_t)in->x[a]) |
                        ((uintptr_t)in->u[a])) & 15u) == 0;
        vec_ok &= ((((uintptr_t)out->active) | ((uintptr_t)in->active)) & 3u) == 0;
        vec_ok &= (((uintptr_t)gather) & 15u) == 0;
        hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(1), 0, st, workspace);
        const SpeciesTable stab = species_table(*desc, *sp);
#define PIC_EXACT_LAUNCH(T, P, MB, G)                                                                               \
        do {                                                                                                        \
            if (gather)                                                                                             \
                hipLaunchKernelGGL((fused_tile_exact_kernel<T, P, MB, G, true>), dim3((unsigned)blocks),            \
                                   dim3(kThreads), 0, st, *desc, *sp, *in, fp, *out, J[0], J[1], J[2], starts, tg,  \
                                   eg, workspace, vec_ok, stab, gather);                                            \
            else                                                                                                    \
                hipLaunchKernelGGL((fused_tile_exact_kernel<T, P, MB, G, false>), dim3((unsigned)blocks),           \
                                   dim3(kThreads), 0, st, *desc, *sp, *in, fp, *out, J[0], J[1], J[2], starts, tg,  \
                                   eg, workspace, vec_ok, stab, gather);                                            \
        } while (0)
#define PIC_EXACT_MB(T, P, G)                                  \
        do {                                                   \
            if (min_blocks == 3) PIC_EXACT_LAUNCH(T, P, 3, G); \
            else if (min_blocks == 4) PIC_EXACT_LAUNCH(T, P, 4, G); \
            else PIC_EXACT_LAUNCH(T, P, 1, G);                 \
        } while (0)
        // rounds: slots per thread (8: two groups of four)
        if (tile == 8) {
            if (rounds == 2) PIC_EXACT_MB(8, 2, 1);
            else if (rounds == 4) PIC_EXACT_MB(8, 4, 1);
            else PIC_EXACT_MB(8, 4, 2);
        } else {
            if (rounds == 2) PIC_EXACT_MB(4, 2, 1);
            else if (rounds == 4) PIC_EXACT_MB(4, 4, 1);
            else PIC_EXACT_MB(4, 4, 2);
        }
#undef PIC_EXACT_MB
#undef PIC_EXACT_LAUNCH
        const int64_t rare_grid = std::min<int64_t>(std::max<int64_t>(grid_1d(n, kThreads), 1), 1024);
        hipLaunchKernelGGL(fused_tile_rare_kernel, dim3((unsigned)rare_grid), dim3(kThreads), 0, st, *desc, *sp, *in,
                           fp, *out, J[0], J[1], J[2], workspace, gather);
    } else if (tile == 8) {
        if (rounds == 2) PIC_TILED_MB(2, 8);
        else PIC_TILED_MB(4, 8);
    } else {
        if (rounds == 2) PIC_TILED_MB(2, 4);
        else PIC_TILED_MB(4, 4);
    }
#undef PIC_TILED_MB
#undef PIC_TILED_LAUNCH
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_fused_push_esirkepov_tiled(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                                   const float* const E[3], const float* const B[3], const float* const E_ext[3],
                                   const float* const B_ext[3], PicParticles* out, float* const J[3], int32_t tile,
                                   const int32_t* starts, int32_t rounds, int32_t min_blocks, int32_t* workspace,
                                   void* stream) {
    return fused_esirkepov_tiled(desc, sp, in, E, B, E_ext, B_ext, out, J, tile, starts, rounds, min_blocks,
                                 workspace, nullptr, stream);
}

int pic_fused_push_esirkepov_tiled_gather(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                                          const int32_t* gather, const float* const E[3], const float* const B[3],
                                          const float* const E_ext[3], const float* const B_ext[3],
                                          PicParticles* out, float* const J[3], int32_t tile, const int32_t* starts,
                                          int32_t rounds, int32_t min_blocks, int32_t* workspace, void* stream) {
    ExactGrid eg;
    PIC_REQUIRE(desc && desc_ok(desc) && sp && desc->variant != 1 && make_exact_grid(*desc, eg));
    PIC_REQUIRE(gather || (int64_t)sp->count * sp->slots == 0);
    return fused_esirkepov_tiled(desc, sp, in, E, B, E_ext, B_ext, out, J, tile, starts, rounds, min_blocks,
                                 workspace, gather, stream);
}

}  // extern "C"
