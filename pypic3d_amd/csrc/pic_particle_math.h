// Per-particle device math shared by the particle kernels: field gather,
// Boris pushers, Esirkepov / direct / charge stencils and particle
// boundaries.  See pic_particles.hip for the precision plan.
#pragma once

#include "pic_common.h"

namespace pic {

// ------------------------------------------------------------------------
// gather (boris.py:302-433, particle_push.py:56-94)
// ------------------------------------------------------------------------

struct GatherAxis {
    int p[3];
    float w[3];
    int k0, k1;     // live stencil slots [k0, k1)
};

__device__ inline void gather_axis(double pos, double org, double sp, int sf, int L, bool active, int g,
                                   GatherAxis& s) {
    int a = anchor_of(pos, org, sp, sf);
    double delta = pos - node_coord(a, sp, org);
    double w[3];
    shape3(delta / sp, sf, w);
    if (active) {
        for (int q = 0; q < 3; ++q) {
            s.p[q] = pmod(a - 1 + q, L);
            s.w[q] = (float)w[q];
        }
        s.k0 = sf == 1 ? 1 : 0;   // CIC weight of point -1 is exactly 0 (shapes.py:46-48)
        s.k1 = 3;
    } else {
        s.p[0] = g;
        s.w[0] = (float)(w[0] + w[1] + w[2]);
        s.k0 = 0;
        s.k1 = 1;
    }
}

__device__ inline float gather3(const float* __restrict__ f, const float* __restrict__ fext, const PicDesc& d,
                                const GatherAxis& ax, const GatherAxis& ay, const GatherAxis& az) {
    float acc = 0.0f;
    for (int i = ax.k0; i < ax.k1; ++i) {
        for (int j = ay.k0; j < ay.k1; ++j) {
            float wij = ax.w[i] * ay.w[j];
            int64_t row = ((int64_t)ax.p[i] * d.dims[1] + ay.p[j]) * d.dims[2];
            for (int k = az.k0; k < az.k1; ++k) {
                float v = f[row + az.p[k]];
                if (fext) v += fext[row + az.p[k]];
                acc += v * (wij * az.w[k]);
            }
        }
    }
    return acc;
}

struct FieldPtrs {
    const float* E[3];
    const float* B[3];
    const float* Ee[3];
    const float* Be[3];
};

// Per-species constants of the particle kernels, evaluated once per block in
// FP64 with the same expressions the per-particle code used, so the FP32
// values handed to the particle math are unchanged:
//   qw  = charge * weight                              (Esirkepov.py:102)
//   h   = charge * dt / (2 mass)                       (boris.py:14-124)
//   pre = -(qw / (d_o1 d_o2)) / dt per current axis    (Esirkepov.py:197-214)
//   q_dv = qw / (dx dy dz), the factor of a reduced component (:197-214)
//   ux, uu = update_x / update_u as bit masks          (particle_class.py:5-14)
struct SpeciesK {
    double qw;
    float h;
    float pre[3];
    float q_dv;
    uint32_t ux, uu;
};

__host__ __device__ inline SpeciesK species_constants(const PicDesc& d, const PicSpecies& sp, int s) {
    SpeciesK k;
    k.qw = sp.charge[s] * sp.weight[s];
    k.h = (float)(sp.charge[s] * d.dt / (2.0 * sp.mass[s]));
    k.pre[0] = (float)(-(k.qw / (d.d[1] * d.d[2])) / d.dt);
    k.pre[1] = (float)(-(k.qw / (d.d[0] * d.d[2])) / d.dt);
    k.pre[2] = (float)(-(k.qw / (d.d[0] * d.d[1])) / d.dt);
    k.q_dv = (float)(k.qw / (d.d[0] * d.d[1] * d.d[2]));
    k.ux = k.uu = 0u;
    for (int a = 0; a < 3; ++a) {
        k.ux |= (sp.update_x[s][a] ? 1u : 0u) << a;
        k.uu |= (sp.update_u[s][a] ? 1u : 0u) << a;
    }
    return k;
}

// Species constants of every species, computed once on the host and passed by
// value, so a block reads its species' constants with scalar loads instead of
// evaluating the FP64 divisions in every lane.
struct SpeciesTable {
    SpeciesK k[PIC_MAX_SPECIES];
};

inline SpeciesTable species_table(const PicDesc& d, const PicSpecies& sp) {
    SpeciesTable t{};
    for (int s = 0; s < sp.count && s < PIC_MAX_SPECIES; ++s) t.k[s] = species_constants(d, sp, s);
    return t;
}

// Six staggered components (particle_push.py:63-68): Ex (v,c,c), Ey (c,v,c),
// Ez (c,c,v), Bx (c,v,v), By (v,c,v), Bz (v,v,c).
__device__ inline void gather_EB(const PicDesc& d, const FieldPtrs& fp, const double pos[3], float E[3],
                                 float B[3]) {
    GatherAxis c[3], v[3];
    for (int a = 0; a < 3; ++a) {
        bool act = d.n[a] > 1;
        gather_axis(pos[a], d.org_c[a], d.sp_c[a], d.shape_factor, d.dims[a], act, d.g, c[a]);
        gather_axis(pos[a], d.org_v[a], d.sp_v[a], d.shape_factor, d.dims[a], act, d.g, v[a]);
    }
    E[0] = gather3(fp.E[0], fp.Ee[0], d, v[0], c[1], c[2]);
    E[1] = gather3(fp.E[1], fp.Ee[1], d, c[0], v[1], c[2]);
    E[2] = gather3(fp.E[2], fp.Ee[2], d, c[0], c[1], v[2]);
    B[0] = gather3(fp.B[0], fp.Be[0], d, c[0], v[1], v[2]);
    B[1] = gather3(fp.B[1], fp.Be[1], d, v[0], c[1], v[2]);
    B[2] = gather3(fp.B[2], fp.Be[2], d, v[0], v[1], c[2]);
}

// ------------------------------------------------------------------------
// pushers (boris.py:14-124)
// ------------------------------------------------------------------------

__device__ inline void cross(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// Boris rotation with hardware reciprocal / reciprocal-square-root (1 ulp),
// the same algebra as boris.py:14-124: gamma = 1/sqrt(1 - v^2/C^2),
// u- = gamma v + hE, t = hB / sqrt(1 + u-^2/C^2), u' = u- + u- x t,
// s = 2t / (1 + t^2), u+ = u- + u' x s, u = u+ + hE, v = u / sqrt(1 + u^2/C^2).
__device__ inline void boris_push(float v[3], const float E[3], const float B[3], float h, int relativistic,
                                  float C) {
    float um[3], t[3], up[3], s[3], c1[3];
    const float invC2 = __builtin_amdgcn_rcpf(C * C);
    if (relativistic) {
        float v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        float gamma = __builtin_amdgcn_rsqf(1.0f - v2 * invC2);
        for (int a = 0; a < 3; ++a) um[a] = v[a] * gamma + h * E[a];
        float igm = __builtin_amdgcn_rsqf(1.0f + (um[0] * um[0] + um[1] * um[1] + um[2] * um[2]) * invC2);
        for (int a = 0; a < 3; ++a) t[a] = h * B[a] * igm;
    } else {
        for (int a = 0; a < 3; ++a) {
            um[a] = v[a] + h * E[a];
            t[a] = h * B[a];
        }
    }
    cross(um, t, c1);
    for (int a = 0; a < 3; ++a) up[a] = um[a] + c1[a];
    float f = 2.0f * __builtin_amdgcn_rcpf(1.0f + t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    for (int a = 0; a < 3; ++a) s[a] = t[a] * f;
    cross(up, s, c1);
    float nu[3];
    for (int a = 0; a < 3; ++a) nu[a] = um[a] + c1[a] + h * E[a];
    if (relativistic) {
        float ing = __builtin_amdgcn_rsqf(1.0f + (nu[0] * nu[0] + nu[1] * nu[1] + nu[2] * nu[2]) * invC2);
        for (int a = 0; a < 3; ++a) v[a] = nu[a] * ing;
    } else {
        for (int a = 0; a < 3; ++a) v[a] = nu[a];
    }
}

// Higuera-Cary (higuera_cary.py:57-113): u = gamma v, u_e = u + hE,
// gamma_hc = sqrt((sigma + sqrt(sigma^2 + 4 (beta^2 + u*^2))) / 2) with
// sigma = (1 + u_e^2/C^2) - beta^2, u* = u_e.beta / C, beta = hB, t = beta / gamma_hc,
// u_m = (u_e + (u_e.t) t + u_e x t) / (1 + t^2), u = u_m + hE + u_m x t.
__device__ inline void higuera_cary_push(float v[3], const float E[3], const float B[3], float h, float C) {
    const float invC2 = __builtin_amdgcn_rcpf(C * C);
    float v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    float gamma0 = __builtin_amdgcn_rsqf(1.0f - v2 * invC2);
    float ue[3], beta[3], t[3], um[3], c1[3];
    for (int a = 0; a < 3; ++a) {
        ue[a] = v[a] * gamma0 + h * E[a];
        beta[a] = h * B[a];
    }
    float b2 = beta[0] * beta[0] + beta[1] * beta[1] + beta[2] * beta[2];
    float ustar = (ue[0] * beta[0] + ue[1] * beta[1] + ue[2] * beta[2]) / C;
    float sigma = (1.0f + (ue[0] * ue[0] + ue[1] * ue[1] + ue[2] * ue[2]) * invC2) - b2;
    float gam = sqrtf(0.5f * (sigma + sqrtf(sigma * sigma + 4.0f * (b2 + ustar * ustar))));
    float ig = __builtin_amdgcn_rcpf(gam);
    for (int a = 0; a < 3; ++a) t[a] = beta[a] * ig;
    float s = __builtin_amdgcn_rcpf(1.0f + t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    float uet = ue[0] * t[0] + ue[1] * t[1] + ue[2] * t[2];
    cross(ue, t, c1);
    for (int a = 0; a < 3; ++a) um[a] = s * (ue[a] + uet * t[a] + c1[a]);
    cross(um, t, c1);
    float nu[3];
    for (int a = 0; a < 3; ++a) nu[a] = um[a] + h * E[a] + c1[a];
    float ing = __builtin_amdgcn_rsqf(1.0f + (nu[0] * nu[0] + nu[1] * nu[1] + nu[2] * nu[2]) * invC2);
    for (int a = 0; a < 3; ++a) v[a] = nu[a] * ing;
}

// particle_push.py:109-132 dispatch
__device__ inline void push_velocity(float v[3], const float E[3], const float B[3], float h, const PicDesc& d) {
    if (d.pusher == 1) higuera_cary_push(v, E, B, h, (float)d.C);
    else boris_push(v, E, B, h, d.relativistic, (float)d.C);
}

// ------------------------------------------------------------------------
// Esirkepov (Esirkepov.py:105-331, 365-504)
// ------------------------------------------------------------------------

// Per-axis 5-slot stencil in the frame of the NEW anchor (points a_new-2 ..
// a_new+2, Esirkepov.py:158-189): S0 = old weights rolled by the anchor shift,
// D = S_new - S0.  Inactive axes collapse to one point at g (:28-45).
struct EsirAxis {
    int base;          // array index of slot 0 (active) or of the single point
    int live_lo, live_hi;    // slots where S0 or S_new is nonzero
    int dlo, dhi;      // first / last slot with D != 0
    float S0[5];
    float D[5];
    float C[5];        // running sum of D (the cumsum of Esirkepov.py:228)
};

__device__ inline void esir_axis(double xo, double xn, double org, double sp, double dd, int sf, bool active, int g,
                                 EsirAxis& e) {
    int ao = anchor_of(xo, org, sp, sf);
    int an = anchor_of(xn, org, sp, sf);
    double wo[3], wn[3];
    shape3((xo - node_coord(ao, sp, org)) / dd, sf, wo);   // deposits divide by the dynamic d (:164-175)
    shape3((xn - node_coord(an, sp, org)) / dd, sf, wn);
    double S1[5] = {0.0, wn[0], wn[1], wn[2], 0.0};
    double Sp[5] = {0.0, wo[0], wo[1], wo[2], 0.0};
    int shift = an - ao;
    double S0[5], D[5];
    for (int i = 0; i < 5; ++i) {
        S0[i] = Sp[pmod(i + shift, 5)];            // jnp.roll(w5, -shift)
        D[i] = S1[i] - S0[i];
    }
    if (!active) {
        double tn = S1[0] + S1[1] + S1[2] + S1[3] + S1[4];
        double to = S0[0] + S0[1] + S0[2] + S0[3] + S0[4];
        e.base = g;
        e.live_lo = e.live_hi = 0;
        e.dlo = 0;
        e.dhi = 0;
        e.S0[0] = (float)to;
        e.D[0] = (float)(tn - to);
        e.C[0] = 0.0f;
        return;
    }
    e.base = an - 2;
    int lo = 5, hi = -1, dlo = 5, dhi = -1;
    double run = 0.0;
    for (int i = 0; i < 5; ++i) {
        if (S0[i] != 0.0 || S1[i] != 0.0) {
            lo = min(lo, i);
            hi = max(hi, i);
        }
        if (D[i] != 0.0) {
            dlo = min(dlo, i);
            dhi = max(dhi, i);
        }
        run += D[i];
        e.S0[i] = (float)S0[i];
        e.D[i] = (float)D[i];
        e.C[i] = (float)run;
    }
    e.live_lo = lo;
    e.live_hi = hi;
    e.dlo = dlo;
    e.dhi = dhi;
}

// Factor of the non-differenced axes: none -> 1; one -> S0 + D/2; two ->
// S0 S0 + (S0 D + D S0)/2 + D D/3 (Esirkepov.py:365-504 in difference form).
__device__ inline float factor1(const EsirAxis& a, int j) { return a.S0[j] + 0.5f * a.D[j]; }

__device__ inline float factor2(const EsirAxis& a, int j, const EsirAxis& b, int k) {
    return a.S0[j] * b.S0[k] + 0.5f * (a.S0[j] * b.D[k] + a.D[j] * b.S0[k]) + (1.0f / 3.0f) * a.D[j] * b.D[k];
}

__device__ inline void deposit_at(float* J, const PicDesc& d, int ix, int iy, int iz, float v) {
    if (v == 0.0f) return;
    if (!drop_index(ix, d.dims[0]) || !drop_index(iy, d.dims[1]) || !drop_index(iz, d.dims[2])) return;
    atomic_add_f32(&J[fidx(d, ix, iy, iz)], v);
}

__device__ inline void esirkepov_particle(const PicDesc& d, double qw, const double xo[3], const double xn[3],
                                          const float v[3], float* const J[3]) {
    EsirAxis ax[3];
    bool act[3];
    for (int a = 0; a < 3; ++a) {
        act[a] = d.n[a] > 1;
        esir_axis(xo[a], xn[a], d.org_c[a], d.sp_c[a], d.d[a], d.shape_factor, act[a], d.g, ax[a]);
    }
    const double dV = d.d[0] * d.d[1] * d.d[2];
    for (int c = 0; c < 3; ++c) {
        int o1 = c == 0 ? 1 : 0, o2 = c == 2 ? 1 : 2;
        // factor axes: active axes other than c
        int fa[2], nf = 0;
        for (int a = 0; a < 3; ++a)
            if (a != c && act[a]) fa[nf++] = a;
        float pre;
        int i_lo, i_hi;
        if (act[c]) {
            pre = (float)(-(qw / (d.d[o1] * d.d[o2])) / d.dt);
            i_lo = ax[c].dlo;
            i_hi = ax[c].dhi;        // C[dhi] == 0 in exact arithmetic
        } else {
            pre = (float)(qw * (double)v[c] / dV);
            i_lo = 0;
            i_hi = 1;
        }
        int j_lo = 0, j_hi = 0, k_lo = 0, k_hi = 0;
        if (nf >= 1) {
            j_lo = ax[fa[0]].live_lo;
            j_hi = ax[fa[0]].live_hi;
        }
        if (nf == 2) {
            k_lo = ax[fa[1]].live_lo;
            k_hi = ax[fa[1]].live_hi;
        }
        for (int i = i_lo; i < i_hi; ++i) {
            float ci = act[c] ? ax[c].C[i] : 1.0f;
            if (ci == 0.0f) continue;
            for (int j = j_lo; j <= j_hi; ++j) {
                for (int k = k_lo; k <= k_hi; ++k) {
                    float y = nf == 0 ? 1.0f : (nf == 1 ? factor1(ax[fa[0]], j) : factor2(ax[fa[0]], j, ax[fa[1]], k));
                    int idx[3];
                    idx[0] = ax[0].base;
                    idx[1] = ax[1].base;
                    idx[2] = ax[2].base;
                    if (act[c]) idx[c] += i;
                    if (nf >= 1) idx[fa[0]] += j;
                    if (nf == 2) idx[fa[1]] += k;
                    deposit_at(J[c], d, idx[0], idx[1], idx[2], pre * ci * y);
                }
            }
        }
    }
}

// ------------------------------------------------------------------------
// direct current and charge (J_from_rhov.py:88-200, rho.py:72-153)
// ------------------------------------------------------------------------

struct NodeAxis {
    int p[3];
    float wn[3], wf[3];
    int k0, k1;
};

__device__ inline void node_axis(double pos, double org, double sp, double dd, int sf, bool reduced, int g,
                                 NodeAxis& s) {
    int a = anchor_of(pos, org, sp, sf);
    double dn = pos - node_coord(a, sp, org);
    double df = face_offset(pos, org, a, dd);   // J_from_rhov.py:139-141
    double wn[3], wf[3];
    shape3(dn / dd, sf, wn);
    shape3(df / dd, sf, wf);
    if (reduced) {
        s.p[0] = g;
        s.wn[0] = (float)(wn[0] + wn[1] + wn[2]);
        s.wf[0] = (float)(wf[0] + wf[1] + wf[2]);
        s.k0 = 0;
        s.k1 = 1;
        return;
    }
    for (int q = 0; q < 3; ++q) {
        s.p[q] = a - 1 + q;                         // local_bc=2: no wrap, drop out of range
        s.wn[q] = (float)wn[q];
        s.wf[q] = (float)wf[q];
    }
    s.k0 = sf == 1 ? 1 : 0;
    s.k1 = 3;
}

// J_from_rhov deposit (J_from_rhov.py:88-200) of a particle at pos with
// velocity u into a sink (add(d, c, i, j, k, v)): node weights along the other
// axes, face weights along the current's axis, dqv = q w / (dx dy dz).
template <class Sink>
__device__ inline void direct_deposit(const PicDesc& d, float dqv, const double pos[3], const float u[3],
                                      const Sink& sink) {
    NodeAxis ax[3];
    for (int a = 0; a < 3; ++a)
        node_axis(pos[a], d.org_c[a], d.sp_c[a], d.d[a], d.shape_factor, d.n[a] == 1, d.g, ax[a]);
    const float q[3] = {dqv * u[0], dqv * u[1], dqv * u[2]};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (a < ax[0].k0 || a >= ax[0].k1 || b < ax[1].k0 || b >= ax[1].k1 || c < ax[2].k0 || c >= ax[2].k1)
                    continue;
                const int ix = ax[0].p[a], iy = ax[1].p[b], iz = ax[2].p[c];
                sink.add(d, 0, ix, iy, iz, q[0] * ax[0].wf[a] * ax[1].wn[b] * ax[2].wn[c]);
                sink.add(d, 1, ix, iy, iz, q[1] * ax[0].wn[a] * ax[1].wf[b] * ax[2].wn[c]);
                sink.add(d, 2, ix, iy, iz, q[2] * ax[0].wn[a] * ax[1].wn[b] * ax[2].wf[c]);
            }
}

__device__ inline void load_pos(const PicParticles& P, int64_t i, double x[3]) {
    x[0] = (double)P.x[0][i];
    x[1] = (double)P.x[1][i];
    x[2] = (double)P.x[2][i];
}

// particle_tile_communication.py:41-59 + wrap_periodic_position
// (grid_and_stencil.py:14-35), in FP64.
__device__ inline void apply_axis_boundary(int bc, double wind, double& x, float& u, bool& active) {
    double half = 0.5 * wind;
    if (bc == 0) {
        if (x >= half || x < -half) {
            double m = fmod(x + half, wind);
            if (m != 0.0 && ((m < 0.0) != (wind < 0.0))) m += wind;
            double w = m - half;
            if (w == -half && x >= half) w = half;
            x = w;
        }
    } else if (bc == 1) {
        bool flip = (x >= half) || (x <= -half);
        if (x > half) x = 2.0 * half - x;
        else if (x < -half) x = -2.0 * half - x;
        if (flip) u = -u;
    } else if (bc == 2) {
        active = active && (x <= half) && (x >= -half);
    }
}


inline bool desc_ok(const PicDesc* d) {
    if (!d) return false;
    if (d->g < 1 || (d->shape_factor != 1 && d->shape_factor != 2)) return false;
    for (int a = 0; a < 3; ++a)
        if (d->n[a] < 1 || d->dims[a] != d->n[a] + 2 * d->g) return false;
    return true;
}

inline FieldPtrs make_fields(const float* const E[3], const float* const B[3], const float* const Ee[3],
                      const float* const Be[3]) {
    FieldPtrs f;
    for (int a = 0; a < 3; ++a) {
        f.E[a] = E[a];
        f.B[a] = B[a];
        f.Ee[a] = Ee ? Ee[a] : nullptr;
        f.Be[a] = Be ? Be[a] : nullptr;
    }
    return f;
}

}  // namespace pic
