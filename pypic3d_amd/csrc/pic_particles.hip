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
#include "pic_particle_math.h"

namespace {

using namespace pic;

constexpr int kBlock = 256;

// ------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------

__global__ void push_kernel(PicDesc d, PicSpecies sp, PicParticles in, FieldPtrs fp, float* ux_o, float* uy_o,
                            float* uz_o, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = species_of(i, sp.slots, sp.count);
    float v[3] = {in.u[0][i], in.u[1][i], in.u[2][i]};
    float nv[3] = {v[0], v[1], v[2]};
    if (in.active[i]) {
        double pos[3];
        load_pos(in, i, pos);
        float E[3], B[3];
        gather_EB(d, fp, pos, E, B);
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

__global__ void direct_kernel(PicDesc d, PicSpecies sp, PicParticles in, float* Jx, float* Jy, float* Jz, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!in.active[i]) return;
    int s = species_of(i, sp.slots, sp.count);
    double pos[3];
    load_pos(in, i, pos);
    NodeAxis ax[3];
    for (int a = 0; a < 3; ++a)
        node_axis(pos[a], d.org_c[a], d.sp_c[a], d.d[a], d.shape_factor, d.n[a] == 1, d.g, ax[a]);
    float dq = (float)(sp.charge[s] * sp.weight[s] / (d.d[0] * d.d[1] * d.d[2]));
    float q[3] = {dq * in.u[0][i], dq * in.u[1][i], dq * in.u[2][i]};
    for (int a = ax[0].k0; a < ax[0].k1; ++a)
        for (int b = ax[1].k0; b < ax[1].k1; ++b)
            for (int c = ax[2].k0; c < ax[2].k1; ++c) {
                int ix = ax[0].p[a], iy = ax[1].p[b], iz = ax[2].p[c];
                deposit_at(Jx, d, ix, iy, iz, q[0] * ax[0].wf[a] * ax[1].wn[b] * ax[2].wn[c]);
                deposit_at(Jy, d, ix, iy, iz, q[1] * ax[0].wn[a] * ax[1].wf[b] * ax[2].wn[c]);
                deposit_at(Jz, d, ix, iy, iz, q[2] * ax[0].wn[a] * ax[1].wn[b] * ax[2].wf[c]);
            }
}

__global__ void rho_kernel(PicDesc d, PicSpecies sp, PicParticles in, float* rho, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!in.active[i]) return;
    int s = species_of(i, sp.slots, sp.count);
    double pos[3];
    load_pos(in, i, pos);
    NodeAxis ax[3];
    for (int a = 0; a < 3; ++a)
        node_axis(pos[a], d.org_c[a], d.sp_c[a], d.d[a], d.shape_factor, d.n[a] == 1, d.g, ax[a]);
    float qd = (float)(sp.charge[s] * sp.weight[s] / (d.d[0] * d.d[1] * d.d[2]));
    for (int a = ax[0].k0; a < ax[0].k1; ++a)
        for (int b = ax[1].k0; b < ax[1].k1; ++b)
            for (int c = ax[2].k0; c < ax[2].k1; ++c)
                deposit_at(rho, d, ax[0].p[a], ax[1].p[b], ax[2].p[c], qd * ax[0].wn[a] * ax[1].wn[b] * ax[2].wn[c]);
}

}  // namespace

extern "C" {

int pic_push(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, const float* const E[3],
             const float* const B[3], const float* const E_ext[3], const float* const B_ext[3], float* const u_out[3],
             void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok(sp, in) && E && B && u_out);
    for (int a = 0; a < 3; ++a) PIC_REQUIRE(E[a] && B[a] && u_out[a]);
    int64_t n = (int64_t)sp->count * sp->slots;
    if (n == 0) return 0;
    hipLaunchKernelGGL(push_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *desc, *sp, *in,
                       make_fields(E, B, E_ext, B_ext), u_out[0], u_out[1], u_out[2], n);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_advance_positions(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, double dt_scale,
                          float* const x_out[3], void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok(sp, in) && x_out && x_out[0] && x_out[1] && x_out[2]);
    int64_t n = (int64_t)sp->count * sp->slots;
    if (n == 0) return 0;
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
