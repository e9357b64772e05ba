// Particle kernels: field gather + Boris push, position advance, particle
// boundaries, Esirkepov / direct / rho deposition, and the fused particle phase
// of the electrodynamic Esirkepov step.
//
// Precision plan.  Particle state and fields are FP32.  Every per-axis scalar
// that decides an index or feeds a weight difference is evaluated in FP64 from
// the FP32 inputs: anchors use the reference's exact division (so cell indices
// agree bit for bit with an FP64 evaluation), and the Esirkepov weight
// differences D = S(new) - S(old) are formed in FP64 so the ~v*dt/dx sized
// differences do not lose digits to cancellation.  The 3-D weight products and
// all accumulation are FP32.
#include "pic_exact_grid.h"
#include "pic_fused_block.h"

namespace {

using namespace pic;

constexpr int kBlock = 256;

// ------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------

// kCic: first-order shape, two-point stencils per active axis (the weight of
// point a-1 is exactly zero, shapes.py:46-48) and index g on reduced axes,
// 8 loads per component; otherwise the general 3^3 gather.
template <bool kCic>
__global__ void push_kernel(PicDesc d, PicSpecies sp, PicParticles in, FieldPtrs fp, float* ux_o, float* uy_o,
                            float* uz_o, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = species_of(i, sp.slots, sp.count);
    float v[3] = {in.u[0][i], in.u[1][i], in.u[2][i]};
    float nv[3] = {v[0], v[1], v[2]};
    if (in.active[i]) {
        float E[3], B[3];
        if constexpr (kCic) {
            CicAxis gc[3], gv[3];
            for (int a = 0; a < 3; ++a) {
                if (d.n[a] > 1) {
                    const float x = in.x[a][i];
                    int an;
                    float dl;
                    cic_axis(x, (float)d.org_c[a], (float)(1.0 / d.sp_c[a]), d.org_c[a], d.sp_c[a], d.dims[a], gc[a],
                             an, dl);
                    cic_axis(x, (float)d.org_v[a], (float)(1.0 / d.sp_v[a]), d.org_v[a], d.sp_v[a], d.dims[a], gv[a],
                             an, dl);
                } else {
                    reduced_cic(d.g, gc[a]);
                    reduced_cic(d.g, gv[a]);
                }
            }
            E[0] = gather_cic(fp.E[0], fp.Ee[0], d, gv[0], gc[1], gc[2]);
            E[1] = gather_cic(fp.E[1], fp.Ee[1], d, gc[0], gv[1], gc[2]);
            E[2] = gather_cic(fp.E[2], fp.Ee[2], d, gc[0], gc[1], gv[2]);
            B[0] = gather_cic(fp.B[0], fp.Be[0], d, gc[0], gv[1], gv[2]);
            B[1] = gather_cic(fp.B[1], fp.Be[1], d, gv[0], gc[1], gv[2]);
            B[2] = gather_cic(fp.B[2], fp.Be[2], d, gv[0], gv[1], gc[2]);
        } else {
            double pos[3];
            load_pos(in, i, pos);
            gather_EB(d, fp, pos, E, B);
        }
        float h = (float)(sp.charge[s] * d.dt / (2.0 * sp.mass[s]));
        push_velocity(nv, E, B, h, d);
        for (int a = 0; a < 3; ++a)
            if (!sp.update_u[s][a]) nv[a] = v[a];
    }
    ux_o[i] = nv[0];
    uy_o[i] = nv[1];
    uz_o[i] = nv[2];
}

__global__ void advance_kernel(PicDesc d, PicSpecies sp, PicParticles in, double dt, float* xo, float* yo, float* zo,
                               int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = species_of(i, sp.slots, sp.count);
    bool act = in.active[i] != 0;
    float* out[3] = {xo, yo, zo};
    for (int a = 0; a < 3; ++a) {
        float x = in.x[a][i];
        if (act && sp.update_x[s][a]) x = (float)step_pos((double)x, (double)in.u[a][i], dt);
        out[a][i] = x;
    }
}

__global__ void boundaries_kernel(PicDesc d, PicParticles io, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool act = io.active[i] != 0;
    double x[3];
    float u[3];
    for (int a = 0; a < 3; ++a) {
        x[a] = (double)io.x[a][i];
        u[a] = io.u[a][i];
        apply_axis_boundary(d.pbc[a], d.wind[a], x[a], u[a], act);
    }
    for (int a = 0; a < 3; ++a) {
        io.x[a][i] = act ? (float)x[a] : 0.0f;
        io.u[a][i] = act ? u[a] : 0.0f;
    }
    io.active[i] = act ? 1 : 0;
}

__global__ void esirkepov_kernel(PicDesc d, PicSpecies sp, PicParticles in, float* Jx, float* Jy, float* Jz,
                                 int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!in.active[i]) return;
    int s = species_of(i, sp.slots, sp.count);
    double xo[3], xn[3];
    float v[3] = {in.u[0][i], in.u[1][i], in.u[2][i]};
    load_pos(in, i, xo);
    for (int a = 0; a < 3; ++a) xn[a] = sp.update_x[s][a] ? step_pos(xo[a], (double)v[a], d.dt) : xo[a];
    float* J[3] = {Jx, Jy, Jz};
    esirkepov_particle(d, sp.charge[s] * sp.weight[s], xo, xn, v, J);
}

// J_from_rhov deposit of particle i (J_from_rhov.py:88-200) into a sink.
template <class Sink>
__device__ inline void direct_particle(const PicDesc& d, float dqv, const PicParticles& in, int64_t i,
                                       const Sink& sink) {
    double pos[3];
    load_pos(in, i, pos);
    const float u[3] = {in.u[0][i], in.u[1][i], in.u[2][i]};
    direct_deposit(d, dqv, pos, u, sink);
}

// compute_rho deposit of particle i (rho.py:72-153) into component 0 of a sink.
template <class Sink>
__device__ inline void rho_particle(const PicDesc& d, float qd, const PicParticles& in, int64_t i, const Sink& sink) {
    double pos[3];
    load_pos(in, i, pos);
    NodeAxis ax[3];
    for (int a = 0; a < 3; ++a)
        node_axis(pos[a], d.org_c[a], d.sp_c[a], d.d[a], d.shape_factor, d.n[a] == 1, d.g, ax[a]);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (a < ax[0].k0 || a >= ax[0].k1 || b < ax[1].k0 || b >= ax[1].k1 || c < ax[2].k0 || c >= ax[2].k1)
                    continue;
                sink.add(d, 0, ax[0].p[a], ax[1].p[b], ax[2].p[c], qd * ax[0].wn[a] * ax[1].wn[b] * ax[2].wn[c]);
            }
}

__device__ inline float charge_density(const PicDesc& d, const PicSpecies& sp, int s) {
    return (float)(sp.charge[s] * sp.weight[s] / (d.d[0] * d.d[1] * d.d[2]));
}

// Deposit sinks with plain global atomics (slot-ordered launches).
struct GlobalDeposit {
    float* J[3];
    __device__ inline void add(const PicDesc& d, int c, int i, int j, int k, float v) const {
        deposit_at(J[c], d, i, j, k, v);
    }
};

__global__ void direct_kernel(PicDesc d, PicSpecies sp, PicParticles in, float* Jx, float* Jy, float* Jz, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!in.active[i]) return;
    const GlobalDeposit sink{{Jx, Jy, Jz}};
    direct_particle(d, charge_density(d, sp, species_of(i, sp.slots, sp.count)), in, i, sink);
}

__global__ void rho_kernel(PicDesc d, PicSpecies sp, PicParticles in, float* rho, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!in.active[i]) return;
    const GlobalDeposit sink{{rho, rho, rho}};
    rho_particle(d, charge_density(d, sp, species_of(i, sp.slots, sp.count)), in, i, sink);
}

// LDS box sink for the tile-ordered deposits: kC components of B^3 nodes from
// origin o (array indices); nodes outside go straight to the global arrays.
template <int kC, int kB>
struct DepositBox {
    float* J[3];
    float* box;
    int o[3];
    __device__ inline void add(const PicDesc& d, int c, int i, int j, int k, float v) const {
        if (v == 0.0f) return;
        const int bi = i - o[0], bj = j - o[1], bk = k - o[2];
        if ((unsigned)bi < (unsigned)kB && (unsigned)bj < (unsigned)kB && (unsigned)bk < (unsigned)kB)
            atomicAdd(&box[((c * kB + bi) * kB + bj) * kB + bk], v);
        else
            deposit_at(J[c], d, i, j, k, v);
    }
};

// Tile-ordered deposits (particles sorted by pic_tile_keys): block b owns the
// slots [starts[b], starts[b+1]) of one (species, tile); deposits accumulate in
// an LDS box of (T+5)^3 nodes from T*tile + g - 2 and are added into the
// arrays once per block.
template <int kC, int kT, bool kRho>
__global__ __launch_bounds__(kBlock) void tiled_deposit_kernel(PicDesc d, PicSpecies sp, PicParticles in, float* J0,
                                                               float* J1, float* J2,
                                                               const int32_t* __restrict__ starts, int tiles,
                                                               int nt1, int nt2) {
    constexpr int kB = kT + 5;
    constexpr int kV = kB * kB * kB;
    __shared__ float box[kC * kV];
    const int b = blockIdx.x;
    const int64_t begin = starts[b], end = starts[b + 1];
    if (begin >= end) return;
    const int t = b % (tiles + 1);
    int o[3] = {0, 0, 0};
    if (t < tiles) {
        o[0] = (t / (nt1 * nt2)) * kT + d.g - 2;
        o[1] = ((t / nt2) % nt1) * kT + d.g - 2;
        o[2] = (t % nt2) * kT + d.g - 2;
    }
    for (int e = threadIdx.x; e < kC * kV; e += kBlock) box[e] = 0.0f;
    __syncthreads();
    const DepositBox<kC, kB> sink{{J0, J1, J2}, box, {o[0], o[1], o[2]}};
    for (int64_t i = begin + threadIdx.x; i < end; i += kBlock) {
        if (!in.active[i]) continue;
        const float q = charge_density(d, sp, species_of(i, sp.slots, sp.count));
        if constexpr (kRho) rho_particle(d, q, in, i, sink);
        else direct_particle(d, q, in, i, sink);
    }
    __syncthreads();
    float* const J[3] = {J0, J1, J2};
    for (int e = threadIdx.x; e < kC * kV; e += kBlock) {
        const float v = box[e];
        if (v == 0.0f) continue;
        const int c = e / kV, r = e % kV;
        deposit_at(J[c], d, o[0] + r / (kB * kB), o[1] + (r / kB) % kB, o[2] + r % kB, v);
    }
}

// Tile-ordered rho / J_from_rhov deposit for 3-D first-order runs on an exact
// grid whose spacing equals the cell size.  Thread t of a chunk owns four
// consecutive sorted slots; a CIC particle deposits on the 2x2x2 nodes
// (a..a+1) of its centre anchor a (node weights 1 - dn/d, dn/d; J face weights
// 1 - df/d, df/d with df = dn - d/2, J_from_rhov.py:110-198, rho.py:72-150),
// so particles of one cell share all 8*kC targets: their values are summed in
// registers while the anchor stays and added into the LDS box when it
// changes.  Anchors come from exact_anchors (FP32, equal to the FP64
// evaluation); particles outside the exact range or the box take the general
// per-particle deposit after the group.
//
// A thread's four slots are read with one 16-byte load per plane (4 bytes of
// activity flags) when the group lies inside the block's range and the planes
// are aligned (vec_ok), so a wave reads 1 KiB of contiguous memory per plane;
// the next chunk's group is loaded before the current one is deposited, so
// the memory latency overlaps the arithmetic.
template <int kC, int kT, bool kRho>
__global__ __launch_bounds__(kBlock) void tiled_deposit_group_kernel(PicDesc d, PicSpecies sp, PicParticles in,
                                                                     float* J0, float* J1, float* J2,
                                                                     const int32_t* __restrict__ starts, int tiles,
                                                                     int nt1, int nt2, ExactGrid eg, int vec_ok) {
    constexpr int kB = kT + 5;
    constexpr int kV = kB * kB * kB;
    constexpr int kP = 4;
    constexpr int kN = 8 * kC;
    constexpr int kU = kRho ? 1 : 3;       // velocity planes held per group (none read for rho)
    constexpr int kChunk = kBlock * kP;
    __shared__ float box[kC * kV];
    const int b = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int64_t begin = starts[b], end = starts[b + 1];
    if (begin >= end) return;
    const int t = b % (tiles + 1);
    const int s = b / (tiles + 1);
    int o[3] = {0, 0, 0};
    if (t < tiles) {
        o[0] = (t / (nt1 * nt2)) * kT + d.g - 2;
        o[1] = ((t / nt2) % nt1) * kT + d.g - 2;
        o[2] = (t % nt2) * kT + d.g - 2;
    }
    for (int e = threadIdx.x; e < kC * kV; e += kBlock) box[e] = 0.0f;
    __syncthreads();
    const DepositBox<kC, kB> sink{{J0, J1, J2}, box, {o[0], o[1], o[2]}};
    const float q = charge_density(d, sp, s);
    auto flush = [&](int h0, int h1, int h2, const float* acc) {
        float* p = box + ((h0 - o[0]) * kB + (h1 - o[1])) * kB + (h2 - o[2]);
#pragma unroll
        for (int m = 0; m < kN; ++m) {
            const int c = m >> 3, n = m & 7;
            atomicAdd(p + c * kV + ((n >> 2) * kB + ((n >> 1) & 1)) * kB + (n & 1), acc[m]);
        }
    };
    struct Group {
        float x[3][kP], u[kU][kP];
        uint32_t act;      // byte j: slot j is inside [begin, end) and active
    };
    auto load = [&](int64_t s0, Group& g) {
        if (vec_ok && s0 >= begin && s0 + kP <= end) {
#pragma unroll
            for (int a = 0; a < 3; ++a) load_vec<kP>(in.x[a] + s0, g.x[a]);
            if constexpr (!kRho) {
#pragma unroll
                for (int a = 0; a < 3; ++a) load_vec<kP>(in.u[a] + s0, g.u[a]);
            }
            g.act = *reinterpret_cast<const uint32_t*>(in.active + s0);
        } else {
            g.act = 0u;
#pragma unroll
            for (int j = 0; j < kP; ++j) {
                const int64_t i = s0 + j;
                const bool inr = i >= begin && i < end;
#pragma unroll
                for (int a = 0; a < 3; ++a) g.x[a][j] = inr ? in.x[a][i] : 0.0f;
                if constexpr (!kRho) {
#pragma unroll
                    for (int a = 0; a < 3; ++a) g.u[a][j] = inr ? in.u[a][i] : 0.0f;
                }
                g.act |= (inr && in.active[i] != 0 ? 1u : 0u) << (8 * j);
            }
        }
    };
    int64_t base = begin & ~(int64_t)(kP - 1);
    Group cur;
    load(base + (int64_t)kP * threadIdx.x, cur);
    for (; base < end; base += kChunk) {
        const int64_t s0 = base + (int64_t)kP * threadIdx.x;
        Group nxt;
        nxt.act = 0u;
        if (base + kChunk < end) load(s0 + kChunk, nxt);
        float acc[kN];
#pragma unroll
        for (int m = 0; m < kN; ++m) acc[m] = 0.0f;
        int h0 = 0, h1 = 0, h2 = 0;
        bool have = false;
        uint32_t general = 0u;
#pragma unroll
        for (int j = 0; j < kP; ++j) {
            if (((cur.act >> (8 * j)) & 0xffu) == 0u) continue;
            int an[3];
            float dn[3];
            bool fast = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float x = cur.x[a][j];
                int av;
                float nd, dv;
                fast = fast && x >= eg.xlo[a] && x < eg.xhi[a];
                exact_anchors(x, eg.org_c[a], eg.sp_c[a], eg.isp_c[a], eg.half_sp[a], an[a], dn[a], nd, av, dv);
                fast = fast && (unsigned)(an[a] - o[a]) <= (unsigned)(kB - 2);
            }
            if (!fast) {
                general |= 1u << j;
                continue;
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
            float val[kN];
            if constexpr (kRho) {
#pragma unroll
                for (int n = 0; n < 8; ++n) val[n] = q * wn[0][n >> 2] * wn[1][(n >> 1) & 1] * wn[2][n & 1];
            } else {
                const float qv[3] = {q * cur.u[0][j], q * cur.u[1][j], q * cur.u[2][j]};
#pragma unroll
                for (int n = 0; n < 8; ++n) {
                    const int xi = n >> 2, yi = (n >> 1) & 1, zi = n & 1;
                    val[n] = qv[0] * wf[0][xi] * wn[1][yi] * wn[2][zi];
                    val[8 + n] = qv[1] * wn[0][xi] * wf[1][yi] * wn[2][zi];
                    val[16 + n] = qv[2] * wn[0][xi] * wn[1][yi] * wf[2][zi];
                }
            }
            const bool same = have && an[0] == h0 && an[1] == h1 && an[2] == h2;
            if (have && !same) flush(h0, h1, h2, acc);
#pragma unroll
            for (int m = 0; m < kN; ++m) acc[m] = (same ? acc[m] : 0.0f) + val[m];
            h0 = an[0];
            h1 = an[1];
            h2 = an[2];
            have = true;
        }
        // the group's last cell: equal cells of neighbouring lanes are summed
        // first, so one lane per cell and 16-lane row adds into the box
        const int key = have ? ((h0 - o[0]) * kB + (h1 - o[1])) * kB + (h2 - o[2]) : -1;
        if (row_run_sums(key, acc) && have) flush(h0, h1, h2, acc);
        while (general) {
            const int j = __builtin_ctz(general);
            general &= general - 1u;
            if constexpr (kRho) rho_particle(d, q, in, s0 + j, sink);
            else direct_particle(d, q, in, s0 + j, sink);
        }
        cur = nxt;
    }
    __syncthreads();
    flush_current_box<kB, kC>(d, o, box, J0, J1, J2);
}

}  // namespace

extern "C" {

int pic_push(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, const float* const E[3],
             const float* const B[3], const float* const E_ext[3], const float* const B_ext[3], float* const u_out[3],
             void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok(sp, in) && E && B && u_out);
    int64_t n = (int64_t)sp->count * sp->slots;
    for (int a = 0; a < 3; ++a) PIC_REQUIRE(E[a] && B[a] && (u_out[a] || n == 0));
    if (n == 0) return 0;
    if (desc->shape_factor == 1)
        hipLaunchKernelGGL(push_kernel<true>, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *desc, *sp,
                           *in, make_fields(E, B, E_ext, B_ext), u_out[0], u_out[1], u_out[2], n);
    else
        hipLaunchKernelGGL(push_kernel<false>, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *desc,
                           *sp, *in, make_fields(E, B, E_ext, B_ext), u_out[0], u_out[1], u_out[2], n);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_advance_positions(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, double dt_scale,
                          float* const x_out[3], void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok(sp, in) && x_out);
    int64_t n = (int64_t)sp->count * sp->slots;
    if (n == 0) return 0;
    PIC_REQUIRE(x_out[0] && x_out[1] && x_out[2]);
    hipLaunchKernelGGL(advance_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *desc, *sp, *in,
                       desc->dt * dt_scale, x_out[0], x_out[1], x_out[2], n);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_particle_boundaries(const PicDesc* desc, const PicSpecies* sp, PicParticles* io, void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok(sp, io));
    int64_t n = (int64_t)sp->count * sp->slots;
    if (n == 0) return 0;
    hipLaunchKernelGGL(boundaries_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *desc, *io, n);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_deposit_esirkepov(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, float* const J[3],
                          void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok(sp, in) && J && J[0] && J[1] && J[2]);
    PIC_REQUIRE(desc->n[0] > 1 || desc->n[1] > 1 || desc->n[2] > 1);
    int64_t n = (int64_t)sp->count * sp->slots;
    if (n == 0) return 0;
    hipLaunchKernelGGL(esirkepov_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *desc, *sp, *in,
                       J[0], J[1], J[2], n);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_deposit_direct(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, float* const J[3],
                       void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok(sp, in) && J && J[0] && J[1] && J[2]);
    int64_t n = (int64_t)sp->count * sp->slots;
    if (n == 0) return 0;
    hipLaunchKernelGGL(direct_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *desc, *sp, *in,
                       J[0], J[1], J[2], n);
    PIC_CHECK_LAUNCH();
    return 0;
}

static int tiled_deposit(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, float* const J[3],
                         int32_t tile, const int32_t* starts, bool rho, void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok(sp, in) && J && J[0] && starts && (tile == 4 || tile == 8));
    if (!rho) PIC_REQUIRE(J[1] && J[2]);
    int nt[3];
    for (int a = 0; a < 3; ++a) nt[a] = (desc->n[a] + tile - 1) / tile;
    const int64_t tiles = (int64_t)nt[0] * nt[1] * nt[2];
    const int64_t blocks = (int64_t)sp->count * (tiles + 1);
    PIC_REQUIRE(blocks < ((int64_t)1 << 31));
    if ((int64_t)sp->count * sp->slots == 0) return 0;
    hipStream_t st = as_stream(stream);
    float* j1 = rho ? J[0] : J[1];
    float* j2 = rho ? J[0] : J[2];
#define PIC_TD(C, T, R)                                                                                           \
    hipLaunchKernelGGL((tiled_deposit_kernel<C, T, R>), dim3((unsigned)blocks), dim3(kBlock), 0, st, *desc, *sp, *in, \
                       J[0], j1, j2, starts, (int)tiles, nt[1], nt[2])
    ExactGrid eg;
    bool grouped = desc->variant != 1 && desc->shape_factor == 1 && desc->n[0] > 1 && desc->n[1] > 1 &&
                   desc->n[2] > 1 && make_exact_grid(*desc, eg);
    for (int a = 0; a < 3 && grouped; ++a) grouped = desc->d[a] == desc->sp_c[a];
#define PIC_TDG(C, T, R)                                                                                            \
    hipLaunchKernelGGL((tiled_deposit_group_kernel<C, T, R>), dim3((unsigned)blocks), dim3(kBlock), 0, st, *desc, \
                       *sp, *in, J[0], j1, j2, starts, (int)tiles, nt[1], nt[2], eg, vec_ok)
    int vec_ok = 1;
    for (int a = 0; a < 3; ++a) {
        vec_ok &= (((uintptr_t)in->x[a]) & 15u) == 0;
        if (!rho) vec_ok &= (((uintptr_t)in->u[a]) & 15u) == 0;
    }
    vec_ok &= (((uintptr_t)in->active) & 3u) == 0;
    if (grouped) {
        if (rho) {
            if (tile == 8) PIC_TDG(1, 8, true);
            else PIC_TDG(1, 4, true);
        } else {
            if (tile == 8) PIC_TDG(3, 8, false);
            else PIC_TDG(3, 4, false);
        }
    } else if (rho) {
        if (tile == 8) PIC_TD(1, 8, true);
        else PIC_TD(1, 4, true);
    } else {
        if (tile == 8) PIC_TD(3, 8, false);
        else PIC_TD(3, 4, false);
    }
#undef PIC_TDG
#undef PIC_TD
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_deposit_direct_tiled(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, float* const J[3],
                             int32_t tile, const int32_t* starts, void* stream) {
    return tiled_deposit(desc, sp, in, J, tile, starts, false, stream);
}

int pic_deposit_rho_tiled(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, float* rho, int32_t tile,
                          const int32_t* starts, void* stream) {
    float* const J[3] = {rho, rho, rho};
    return tiled_deposit(desc, sp, in, J, tile, starts, true, stream);
}

int pic_deposit_rho(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, float* rho, void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok(sp, in) && rho);
    int64_t n = (int64_t)sp->count * sp->slots;
    if (n == 0) return 0;
    hipLaunchKernelGGL(rho_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *desc, *sp, *in, rho,
                       n);
    PIC_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
