// FP64 particle path: gather + push, position advance, particle boundaries,
// Esirkepov / direct / rho deposits and the energy diagnostic on double
// particle planes and double fields -- the precision the reference runs in
// (its CLI enables jax_enable_x64, __main__.py:186-187).  One thread per slot,
// any dimensionality and shape order, global FP64 atomics for the deposits;
// meant for small FP64 runs such as the 64-cell two-stream case, not for
// throughput.  The arithmetic follows the reference's own expressions in
// FP64 (the FP32 path's hardware reciprocals and factorized weights are
// FP32 tunings and are not used here).  The field operations of the FP64 path
// (refresh, fold, walls, filters, Yee) are the templated field kernels of
// pic_fields.hip.
#include "pic_particle_math.h"

namespace {

using namespace pic;

constexpr int kBlock = 256;
constexpr int kMaxPartials = 1024;

struct PartsD {
    double* x[3];
    double* u[3];
    uint8_t* active;
};

struct FieldsD {
    const double* E[3];
    const double* B[3];
    const double* Ee[3];
    const double* Be[3];
};

inline PartsD parts_d(const PicParticlesF64* p) {
    PartsD q;
    for (int a = 0; a < 3; ++a) {
        q.x[a] = p->x[a];
        q.u[a] = p->u[a];
    }
    q.active = p->active;
    return q;
}

inline bool particles_ok_d(const PicSpecies* sp, const PicParticlesF64* p) {
    if (!sp || !p) return false;
    if (sp->count < 0 || sp->count > PIC_MAX_SPECIES || sp->slots < 0) return false;
    if (sp->count == 0 || sp->slots == 0) return true;
    for (int a = 0; a < 3; ++a)
        if (!p->x[a] || !p->u[a]) return false;
    return p->active != nullptr;
}

// ---- gather (boris.py:302-433, particle_push.py:56-94) -------------------

struct AxisD {
    int p[3];
    double w[3];
    int k0, k1;
};

// prepare_particle_axis_stencil with periodic wrap of the points and the
// collapse of an inactive axis to index g with the summed weight
__device__ inline void gather_axis_d(double pos, double org, double sp, int sf, int L, bool active, int g, AxisD& s) {
    const int a = anchor_of(pos, org, sp, sf);
    double w[3];
    shape3((pos - node_coord(a, sp, org)) / sp, sf, w);
    if (active) {
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

__device__ inline double gather3_d(const double* __restrict__ f, const double* __restrict__ fe, const PicDesc& d,
                                   const AxisD& ax, const AxisD& ay, const AxisD& az) {
    double acc = 0.0;
    for (int i = ax.k0; i < ax.k1; ++i)
        for (int j = ay.k0; j < ay.k1; ++j) {
            const int64_t row = ((int64_t)ax.p[i] * d.dims[1] + ay.p[j]) * d.dims[2];
            for (int k = az.k0; k < az.k1; ++k) {
                double v = f[row + az.p[k]];
                if (fe) v += fe[row + az.p[k]];
                acc += v * ax.w[i] * ay.w[j] * az.w[k];
            }
        }
    return acc;
}

// Six staggered components (particle_push.py:63-68)
__device__ inline void gather_eb_d(const PicDesc& d, const FieldsD& fp, const double pos[3], double E[3],
                                   double B[3]) {
    AxisD c[3], v[3];
    for (int a = 0; a < 3; ++a) {
        const bool act = d.n[a] > 1;
        gather_axis_d(pos[a], d.org_c[a], d.sp_c[a], d.shape_factor, d.dims[a], act, d.g, c[a]);
        gather_axis_d(pos[a], d.org_v[a], d.sp_v[a], d.shape_factor, d.dims[a], act, d.g, v[a]);
    }
    E[0] = gather3_d(fp.E[0], fp.Ee[0], d, v[0], c[1], c[2]);
    E[1] = gather3_d(fp.E[1], fp.Ee[1], d, c[0], v[1], c[2]);
    E[2] = gather3_d(fp.E[2], fp.Ee[2], d, c[0], c[1], v[2]);
    B[0] = gather3_d(fp.B[0], fp.Be[0], d, c[0], v[1], v[2]);
    B[1] = gather3_d(fp.B[1], fp.Be[1], d, v[0], c[1], v[2]);
    B[2] = gather3_d(fp.B[2], fp.Be[2], d, v[0], v[1], c[2]);
}

// ---- pushers (boris.py:14-124, higuera_cary.py:5-113) ----------------------

__device__ inline void cross_d(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ inline double dot_d(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// boris_single_particle (boris.py:14-56) and relativistic_boris_single_particle (:60-124)
__device__ inline void boris_d(double v[3], const double E[3], const double B[3], double h, bool rel, double C) {
    double um[3], t[3], up[3], s[3], c1[3];
    const double C2 = C * C;
    if (rel) {
        const double gamma = 1.0 / sqrt(1.0 - dot_d(v, v) / C2);
        for (int a = 0; a < 3; ++a) um[a] = v[a] * gamma + h * E[a];
        const double gm = sqrt(1.0 + dot_d(um, um) / C2);
        for (int a = 0; a < 3; ++a) t[a] = h * B[a] / gm;
    } else {
        for (int a = 0; a < 3; ++a) {
            um[a] = v[a] + h * E[a];
            t[a] = h * B[a];
        }
    }
    cross_d(um, t, c1);
    for (int a = 0; a < 3; ++a) up[a] = um[a] + c1[a];
    const double t2 = dot_d(t, t);
    for (int a = 0; a < 3; ++a) s[a] = 2.0 * t[a] / (1.0 + t2);
    cross_d(up, s, c1);
    double nu[3];
    for (int a = 0; a < 3; ++a) nu[a] = um[a] + c1[a] + h * E[a];
    if (rel) {
        const double ng = sqrt(1.0 + dot_d(nu, nu) / C2);
        for (int a = 0; a < 3; ++a) v[a] = nu[a] / ng;
    } else {
        for (int a = 0; a < 3; ++a) v[a] = nu[a];
    }
}

__device__ inline void higuera_cary_d(double v[3], const double E[3], const double B[3], double h, double C) {
    const double C2 = C * C;
    const double g0 = 1.0 / sqrt(1.0 - dot_d(v, v) / C2);
    double ue[3], beta[3], t[3], um[3], c1[3];
    for (int a = 0; a < 3; ++a) {
        ue[a] = v[a] * g0 + h * E[a];
        beta[a] = h * B[a];
    }
    const double b2 = dot_d(beta, beta);
    const double ustar = dot_d(ue, beta) / C;
    const double sigma = (1.0 + dot_d(ue, ue) / C2) - b2;
    const double gam = sqrt((sigma + sqrt(sigma * sigma + 4.0 * (b2 + ustar * ustar))) / 2.0);
    for (int a = 0; a < 3; ++a) t[a] = beta[a] / gam;
    const double s = 1.0 / (1.0 + dot_d(t, t));
    const double uet = dot_d(ue, t);
    cross_d(ue, t, c1);
    for (int a = 0; a < 3; ++a) um[a] = s * (ue[a] + uet * t[a] + c1[a]);
    cross_d(um, t, c1);
    double nu[3];
    for (int a = 0; a < 3; ++a) nu[a] = um[a] + h * E[a] + c1[a];
    const double ng = sqrt(1.0 + dot_d(nu, nu) / C2);
    for (int a = 0; a < 3; ++a) v[a] = nu[a] / ng;
}

// ---- Esirkepov (Esirkepov.py:105-331, 365-504) -----------------------------

// Per-axis 5-slot frame of the new anchor: S0 = old weights rolled by the
// anchor shift, S1 = new weights; inactive axes collapse to one point at g.
struct EsirD {
    int base;
    bool active;
    double S0[5], S1[5];
};

__device__ inline void esir_axis_d(double xo, double xn, double org, double sp, double dd, int sf, bool active, int g,
                                   EsirD& e) {
    const int ao = anchor_of(xo, org, sp, sf);
    const int an = anchor_of(xn, org, sp, sf);
    double wo[3], wn[3];
    shape3((xo - node_coord(ao, sp, org)) / dd, sf, wo);   // deposits divide by the dynamic d (:164-175)
    shape3((xn - node_coord(an, sp, org)) / dd, sf, wn);
    const double Sn[5] = {0.0, wn[0], wn[1], wn[2], 0.0};
    const double So[5] = {0.0, wo[0], wo[1], wo[2], 0.0};
    const int shift = an - ao;
    e.active = active;
    if (!active) {
        double tn = 0.0, to = 0.0;
        for (int i = 0; i < 5; ++i) {
            tn += Sn[i];
            to += So[pmod(i + shift, 5)];
        }
        for (int i = 0; i < 5; ++i) e.S0[i] = e.S1[i] = 0.0;
        e.S1[0] = tn;
        e.S0[0] = to;
        e.base = g;
        return;
    }
    for (int i = 0; i < 5; ++i) {
        e.S1[i] = Sn[i];
        e.S0[i] = So[pmod(i + shift, 5)];    // jnp.roll(w5, -shift)
    }
    e.base = an - 2;
}

__device__ inline void add_drop_d(double* J, const PicDesc& d, int ix, int iy, int iz, double v) {
    if (v == 0.0) return;
    if (!drop_index(ix, d.dims[0]) || !drop_index(iy, d.dims[1]) || !drop_index(iz, d.dims[2])) return;
    unsafeAtomicAdd(&J[fidx(d, ix, iy, iz)], v);
}

// One particle: W per component in the reference's 1/3-1/6 form
// (_3D/_2d/_1d_esirkepov_weights), the running sum along active current
// axes (:228-230), the factor -q w/(d_o1 d_o2 dt) of an active axis or
// q w v/(dx dy dz) of a reduced one (:197-214).
__device__ inline void esirkepov_particle_d(const PicDesc& d, double qw, const double xo[3], const double xn[3],
                                            const double v[3], double* const J[3]) {
    EsirD ax[3];
    bool act[3];
    for (int a = 0; a < 3; ++a) {
        act[a] = d.n[a] > 1;
        esir_axis_d(xo[a], xn[a], d.org_c[a], d.sp_c[a], d.d[a], d.shape_factor, act[a], d.g, ax[a]);
    }
    const double dV = d.d[0] * d.d[1] * d.d[2];
    const int ni[3] = {act[0] ? 5 : 1, act[1] ? 5 : 1, act[2] ? 5 : 1};
    for (int c = 0; c < 3; ++c) {
        const int p = c == 0 ? 1 : 0, q = c == 2 ? 1 : 2;      // the other two axes
        const double pre = act[c] ? -(qw / (d.d[p] * d.d[q])) / d.dt : qw * v[c] / dV;
        for (int j = 0; j < ni[p]; ++j) {
            for (int k = 0; k < ni[q]; ++k) {
                // factor of the other axes (active ones vary with j, k; reduced ones are one point)
                double f;
                if (act[p] && act[q]) {
                    f = (1.0 / 3.0) * (ax[p].S1[j] * ax[q].S1[k] + ax[p].S0[j] * ax[q].S0[k]) +
                        (1.0 / 6.0) * (ax[p].S1[j] * ax[q].S0[k] + ax[p].S0[j] * ax[q].S1[k]);
                } else if (act[p]) {
                    f = 0.5 * (ax[p].S1[j] + ax[p].S0[j]);
                } else if (act[q]) {
                    f = 0.5 * (ax[q].S1[k] + ax[q].S0[k]);
                } else {
                    f = 1.0;
                }
                if (act[c]) {
                    // reduced other axes: the 2-D / 1-D weight of the in-plane current
                    // carries the collapsed axis through f (its totals are 1 for S0 and S1)
                    double run = 0.0;
                    for (int i = 0; i < 5; ++i) {
                        run += (ax[c].S1[i] - ax[c].S0[i]) * f;
                        int idx[3] = {ax[0].base, ax[1].base, ax[2].base};
                        idx[c] += i;
                        if (act[p]) idx[p] += j;
                        if (act[q]) idx[q] += k;
                        add_drop_d(J[c], d, idx[0], idx[1], idx[2], pre * run);
                    }
                } else {
                    int idx[3] = {ax[0].base, ax[1].base, ax[2].base};
                    if (act[p]) idx[p] += j;
                    if (act[q]) idx[q] += k;
                    add_drop_d(J[c], d, idx[0], idx[1], idx[2], pre * f);
                }
            }
        }
    }
}

// ---- direct current and charge (J_from_rhov.py:88-200, rho.py:72-153) -----

struct NodeD {
    int p[3];
    double wn[3], wf[3];
    int k0, k1;
};

__device__ inline void node_axis_d(double pos, double org, double sp, double dd, int sf, bool reduced, int g,
                                   NodeD& s) {
    const int a = anchor_of(pos, org, sp, sf);
    double wn[3], wf[3];
    shape3((pos - node_coord(a, sp, org)) / dd, sf, wn);
    shape3(face_offset(pos, org, a, dd) / dd, sf, wf);   // face offset (J_from_rhov.py:139-141)
    if (reduced) {
        s.p[0] = g;
        s.wn[0] = wn[0] + wn[1] + wn[2];
        s.wf[0] = wf[0] + wf[1] + wf[2];
        s.k0 = 0;
        s.k1 = 1;
        return;
    }
    for (int q = 0; q < 3; ++q) {
        s.p[q] = a - 1 + q;                     // local_bc=2: no wrap, drop out of range
        s.wn[q] = wn[q];
        s.wf[q] = wf[q];
    }
    s.k0 = sf == 1 ? 1 : 0;
    s.k1 = 3;
}

// particle_tile_communication.py:41-59 + wrap_periodic_position
// (grid_and_stencil.py:14-35).  The reference maps every periodic position
// through mod(x + W/2, W) - W/2, in range or not, so its result carries the
// rounding of that round trip; the FP64 path reproduces it bit for bit
// (mod(s, W) = s for 0 <= s < W needs no division).
__device__ inline void axis_boundary_d(int bc, double wind, double& x, double& u, bool& active) {
    const double half = 0.5 * wind;
    if (bc == 0) {
        const double s = x + half;
        double m = s;
        if (!(s >= 0.0 && s < wind)) {
            m = fmod(s, wind);
            if (m != 0.0 && ((m < 0.0) != (wind < 0.0))) m += wind;
        }
        double w = m - half;
        if (w == -half && x >= half) w = half;
        x = w;
    } else if (bc == 1) {
        const bool flip = (x >= half) || (x <= -half);
        if (x > half) x = 2.0 * half - x;
        else if (x < -half) x = -2.0 * half - x;
        if (flip) u = -u;
    } else if (bc == 2) {
        active = active && (x <= half) && (x >= -half);
    }
}

__device__ inline double block_sum(double v) {
    __shared__ double sh[kBlock];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double out = sh[0];
    __syncthreads();
    return out;
}

// ---- kernels --------------------------------------------------------------

__global__ void push_d_kernel(PicDesc d, PicSpecies sp, PartsD in, FieldsD fp, double* ux, double* uy, double* uz,
                              int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = species_of(i, sp.slots, sp.count);
    const double v[3] = {in.u[0][i], in.u[1][i], in.u[2][i]};
    double nv[3] = {v[0], v[1], v[2]};
    if (in.active[i]) {
        const double pos[3] = {in.x[0][i], in.x[1][i], in.x[2][i]};
        double E[3], B[3];
        gather_eb_d(d, fp, pos, E, B);
        const double h = sp.charge[s] * d.dt / (2.0 * sp.mass[s]);
        if (d.pusher == 1) higuera_cary_d(nv, E, B, h, d.C);
        else boris_d(nv, E, B, h, d.relativistic != 0, d.C);
        for (int a = 0; a < 3; ++a)
            if (!sp.update_u[s][a]) nv[a] = v[a];
    }
    ux[i] = nv[0];
    uy[i] = nv[1];
    uz[i] = nv[2];
}

__global__ void advance_d_kernel(PicSpecies sp, PartsD in, double dt, double* xo, double* yo, double* zo, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = species_of(i, sp.slots, sp.count);
    const bool act = in.active[i] != 0;
    double* out[3] = {xo, yo, zo};
    for (int a = 0; a < 3; ++a) {
        double x = in.x[a][i];
        if (act && sp.update_x[s][a]) x = step_pos(x, in.u[a][i], dt);
        out[a][i] = x;
    }
}

__global__ void boundaries_d_kernel(PicDesc d, PartsD io, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool act = io.active[i] != 0;
    double x[3], u[3];
    for (int a = 0; a < 3; ++a) {
        x[a] = io.x[a][i];
        u[a] = io.u[a][i];
        axis_boundary_d(d.pbc[a], d.wind[a], x[a], u[a], act);
    }
    for (int a = 0; a < 3; ++a) {
        io.x[a][i] = act ? x[a] : 0.0;
        io.u[a][i] = act ? u[a] : 0.0;
    }
    io.active[i] = act ? 1 : 0;
}

__global__ void esirkepov_d_kernel(PicDesc d, PicSpecies sp, PartsD in, double* Jx, double* Jy, double* Jz,
                                   int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !in.active[i]) return;
    const int s = species_of(i, sp.slots, sp.count);
    const double xo[3] = {in.x[0][i], in.x[1][i], in.x[2][i]};
    const double v[3] = {in.u[0][i], in.u[1][i], in.u[2][i]};
    double xn[3];
    for (int a = 0; a < 3; ++a) xn[a] = sp.update_x[s][a] ? step_pos(xo[a], v[a], d.dt) : xo[a];   // (:125-127)
    double* const J[3] = {Jx, Jy, Jz};
    esirkepov_particle_d(d, sp.charge[s] * sp.weight[s], xo, xn, v, J);
}

// kRho: compute_rho's node-weight deposit into J[0]; otherwise J_from_rhov
template <bool kRho>
__global__ void node_deposit_d_kernel(PicDesc d, PicSpecies sp, PartsD in, double* Jx, double* Jy, double* Jz,
                                      int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !in.active[i]) return;
    const int s = species_of(i, sp.slots, sp.count);
    const double dq = sp.charge[s] * sp.weight[s] / (d.d[0] * d.d[1] * d.d[2]);
    NodeD ax[3];
    for (int a = 0; a < 3; ++a)
        node_axis_d(in.x[a][i], d.org_c[a], d.sp_c[a], d.d[a], d.shape_factor, d.n[a] == 1, d.g, ax[a]);
    const double u[3] = {in.u[0][i], in.u[1][i], in.u[2][i]};
    for (int a = ax[0].k0; a < ax[0].k1; ++a)
        for (int b = ax[1].k0; b < ax[1].k1; ++b)
            for (int c = ax[2].k0; c < ax[2].k1; ++c) {
                const int ix = ax[0].p[a], iy = ax[1].p[b], iz = ax[2].p[c];
                if (kRho) {
                    add_drop_d(Jx, d, ix, iy, iz, dq * ax[0].wn[a] * ax[1].wn[b] * ax[2].wn[c]);
                } else {
                    add_drop_d(Jx, d, ix, iy, iz, dq * u[0] * ax[0].wf[a] * ax[1].wn[b] * ax[2].wn[c]);
                    add_drop_d(Jy, d, ix, iy, iz, dq * u[1] * ax[0].wn[a] * ax[1].wf[b] * ax[2].wn[c]);
                    add_drop_d(Jz, d, ix, iy, iz, dq * u[2] * ax[0].wn[a] * ax[1].wn[b] * ax[2].wf[c]);
                }
            }
}

__global__ void energy_fields_d_kernel(PicDesc d, FieldsD f, double* partials) {
    const int64_t total = (int64_t)d.n[0] * d.n[1] * d.n[2];
    double e2 = 0.0, b2 = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(t % d.n[2]) + d.g;
        const int64_t r = t / d.n[2];
        const int j = (int)(r % d.n[1]) + d.g, i = (int)(r / d.n[1]) + d.g;
        const int64_t c = fidx(d, i, j, k);
        for (int a = 0; a < 3; ++a) {
            e2 += f.E[a][c] * f.E[a][c];
            b2 += f.B[a][c] * f.B[a][c];
        }
    }
    const double se = block_sum(e2);
    const double sb = block_sum(b2);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = se;
        partials[2 * blockIdx.x + 1] = sb;
    }
}

__global__ void energy_kinetic_d_kernel(PicSpecies sp, PartsD in, double C, double* partials, int64_t n) {
    double ke = 0.0;
    const double C2 = C * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!in.active[i]) continue;
        const int s = species_of(i, sp.slots, sp.count);
        const double m = sp.mass[s] * sp.weight[s];
        const double v2 = in.u[0][i] * in.u[0][i] + in.u[1][i] * in.u[1][i] + in.u[2][i] * in.u[2][i];
        const double gamma = 1.0 / sqrt(1.0 - v2 / C2);
        const double mom2 = (m * gamma) * (m * gamma) * v2;
        ke += sqrt(mom2 * C2 + m * m * C2 * C2) - m * C2;
    }
    const double s = block_sum(ke);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void energy_finalize_d(const double* fpart, int nf, const double* kpart, int nk, double ce, double cb,
                                  double* out) {
    double e = 0.0, b = 0.0, k = 0.0;
    for (int i = threadIdx.x; i < nf; i += blockDim.x) {
        e += fpart[2 * i];
        b += fpart[2 * i + 1];
    }
    for (int i = threadIdx.x; i < nk; i += blockDim.x) k += kpart[i];
    e = block_sum(e);
    b = block_sum(b);
    k = block_sum(k);
    if (threadIdx.x == 0) {
        out[0] = ce * e;
        out[1] = cb * b;
        out[2] = k;
    }
}

int partial_blocks(int64_t work) {
    int64_t b = (work + kBlock - 1) / kBlock;
    return (int)(b < 1 ? 1 : (b > kMaxPartials ? kMaxPartials : b));
}

FieldsD fields_d(const double* const E[3], const double* const B[3], const double* const Ee[3],
                 const double* const Be[3]) {
    FieldsD f;
    for (int a = 0; a < 3; ++a) {
        f.E[a] = E[a];
        f.B[a] = B[a];
        f.Ee[a] = Ee ? Ee[a] : nullptr;
        f.Be[a] = Be ? Be[a] : nullptr;
    }
    return f;
}

}  // namespace

extern "C" {

int pic_push_f64(const PicDesc* desc, const PicSpecies* sp, const PicParticlesF64* in, const double* const E[3],
                 const double* const B[3], const double* const E_ext[3], const double* const B_ext[3],
                 double* const u_out[3], void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok_d(sp, in) && E && B && u_out);
    for (int a = 0; a < 3; ++a) PIC_REQUIRE(E[a] && B[a] && u_out[a]);
    const int64_t n = (int64_t)sp->count * sp->slots;
    if (n == 0) return 0;
    hipLaunchKernelGGL(push_d_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *desc, *sp,
                       parts_d(in), fields_d(E, B, E_ext, B_ext), u_out[0], u_out[1], u_out[2], n);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_advance_positions_f64(const PicDesc* desc, const PicSpecies* sp, const PicParticlesF64* in, double dt_scale,
                              double* const x_out[3], void* stream) {
    PIC_REQUIRE(desc && particles_ok_d(sp, in) && x_out);
    for (int a = 0; a < 3; ++a) PIC_REQUIRE(x_out[a]);
    const int64_t n = (int64_t)sp->count * sp->slots;
    if (n == 0) return 0;
    hipLaunchKernelGGL(advance_d_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *sp,
                       parts_d(in), desc->dt * dt_scale, x_out[0], x_out[1], x_out[2], n);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_particle_boundaries_f64(const PicDesc* desc, const PicSpecies* sp, PicParticlesF64* io, void* stream) {
    PIC_REQUIRE(desc && particles_ok_d(sp, io));
    const int64_t n = (int64_t)sp->count * sp->slots;
    if (n == 0) return 0;
    hipLaunchKernelGGL(boundaries_d_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *desc,
                       parts_d(io), n);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_deposit_esirkepov_f64(const PicDesc* desc, const PicSpecies* sp, const PicParticlesF64* in,
                              double* const J[3], void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok_d(sp, in) && J && J[0] && J[1] && J[2]);
    PIC_REQUIRE(desc->n[0] > 1 || desc->n[1] > 1 || desc->n[2] > 1);
    const int64_t n = (int64_t)sp->count * sp->slots;
    if (n == 0) return 0;
    hipLaunchKernelGGL(esirkepov_d_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *desc, *sp,
                       parts_d(in), J[0], J[1], J[2], n);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_deposit_direct_f64(const PicDesc* desc, const PicSpecies* sp, const PicParticlesF64* in, double* const J[3],
                           void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok_d(sp, in) && J && J[0] && J[1] && J[2]);
    const int64_t n = (int64_t)sp->count * sp->slots;
    if (n == 0) return 0;
    hipLaunchKernelGGL(node_deposit_d_kernel<false>, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream),
                       *desc, *sp, parts_d(in), J[0], J[1], J[2], n);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_deposit_rho_f64(const PicDesc* desc, const PicSpecies* sp, const PicParticlesF64* in, double* rho,
                        void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok_d(sp, in) && rho);
    const int64_t n = (int64_t)sp->count * sp->slots;
    if (n == 0) return 0;
    hipLaunchKernelGGL(node_deposit_d_kernel<true>, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream),
                       *desc, *sp, parts_d(in), rho, rho, rho, n);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_energy_f64(const PicDesc* desc, const PicSpecies* sp, const PicParticlesF64* in, const double* const E[3],
                   const double* const B[3], double* out, void* workspace, int64_t workspace_bytes, void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok_d(sp, in) && E && B && out && workspace);
    for (int a = 0; a < 3; ++a) PIC_REQUIRE(E[a] && B[a]);
    PIC_REQUIRE(workspace_bytes >= (int64_t)3 * kMaxPartials * 8);
    hipStream_t s = as_stream(stream);
    double* fpart = static_cast<double*>(workspace);
    double* kpart = fpart + 2 * kMaxPartials;
    const int nf = partial_blocks((int64_t)desc->n[0] * desc->n[1] * desc->n[2]);
    const int64_t n = (int64_t)sp->count * sp->slots;
    const int nk = partial_blocks(n);
    hipLaunchKernelGGL(energy_fields_d_kernel, dim3(nf), dim3(kBlock), 0, s, *desc,
                       fields_d(E, B, nullptr, nullptr), fpart);
    PIC_CHECK_LAUNCH();
    hipLaunchKernelGGL(energy_kinetic_d_kernel, dim3(nk), dim3(kBlock), 0, s, *sp, parts_d(in), desc->C, kpart, n);
    PIC_CHECK_LAUNCH();
    const double dV = desc->d[0] * desc->d[1] * desc->d[2];
    hipLaunchKernelGGL(energy_finalize_d, dim3(1), dim3(kBlock), 0, s, fpart, nf, kpart, nk, 0.5 * desc->eps * dV,
                       0.5 / desc->mu * dV, out);
    PIC_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"

namespace {

// wrap_periodic_position (grid_and_stencil.py:14-35) on an array of
// positions: mod(x + W/2, W) - W/2 for every element, +W/2 kept at +W/2.
template <class T>
__global__ void wrap_periodic_kernel(T* __restrict__ x, int64_t n, double wind) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = (double)x[i], u = 0.0;
    bool act = true;
    axis_boundary_d(0, wind, v, u, act);
    x[i] = (T)v;
}

template <class T>
int wrap_periodic_t(T* x, int64_t n, double wind, hipStream_t s) {
    PIC_REQUIRE((x || n == 0) && n >= 0 && wind > 0.0);
    if (n == 0) return 0;
    hipLaunchKernelGGL(wrap_periodic_kernel<T>, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, s, x, n, wind);
    PIC_CHECK_LAUNCH();
    return 0;
}

}  // namespace

extern "C" {

int pic_wrap_periodic(float* x, int64_t n, double wind, void* stream) {
    return wrap_periodic_t(x, n, wind, as_stream(stream));
}

int pic_wrap_periodic_f64(double* x, int64_t n, double wind, void* stream) {
    return wrap_periodic_t(x, n, wind, as_stream(stream));
}

}  // extern "C"
