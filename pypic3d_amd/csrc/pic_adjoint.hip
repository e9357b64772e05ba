// Adjoint (reverse-mode) kernels of the electrodynamic Esirkepov step
// (evolve.py:16-103) for periodic fields and particles.
//
// The forward step, per particle and then on the grid:
//   E_p, B_p = gather(E + E_ext, B + B_ext at x)            (particle_push.py)
//   v'       = push(v, E_p, B_p)        on update_u axes    (boris.py / higuera_cary.py)
//   J       += Esirkepov(x, x + v' dt, v'), fold + refresh  (Esirkepov.py)
//   x'       = x + v' dt                on update_x axes, periodic wrap
//   B_h = B - (dt/2) curl_f E;  E' = F(E + dt (C^2 curl_b B_h - J/eps));
//   B' = F(B_h - (dt/2) curl_f E')                          (first_order_yee.py)
// with curl_f / curl_b the forward / backward difference curls and F the
// digital filter (identity at alpha = 1).
//
// Reverse mode, adjoints in FP64:
//  * Yee: on interior arrays with periodic wrap the transpose of curl_f is
//    curl_b and vice versa, and the symmetric digital filter is self-adjoint.
//  * fold + refresh of J transposes to "read the adjoint at the raw node's
//    periodic partner": the deposit stencil reads the interior adjoint of J
//    at the partner of each (ghosted) node.
//  * deposit -> gather: each particle reads the adjoint of J on its Esirkepov
//    stencil and contracts it with dJ/d(x, v'), computed by forward-mode dual
//    numbers through the same 5-slot Esirkepov algebra.
//  * gather -> scatter: the adjoint of E_p, B_p is scattered (atomics) onto
//    the interior partners of the gather nodes, the transpose of the ghost
//    refresh the gather reads; dE_p/dx comes from the weight derivatives.
//  * push: the 3x9 Jacobian of (v, E_p, B_p) -> v' by dual numbers, applied
//    transposed.
#include "pic_particle_math.h"

namespace {

using namespace pic;

constexpr int kThreads = 256;

// periodic partner (interior index) of a ghosted index along one axis
__device__ inline int partner(int r, int g, int n) {
    int i = r - g;
    if ((unsigned)i < (unsigned)n) return i;   // interior: no wrap
    i %= n;
    return i < 0 ? i + n : i;
}

__device__ inline int64_t partner_index(const PicDesc& d, int i, int j, int k) {
    return ((int64_t)partner(i, d.g, d.n[0]) * d.n[1] + partner(j, d.g, d.n[1])) * d.n[2] + partner(k, d.g, d.n[2]);
}


// ------------------------------------------------------------------ duals
template <int N>
struct Dual {
    double v;
    double t[N];
};

template <int N>
__device__ inline Dual<N> dconst(double x) {
    Dual<N> r;
    r.v = x;
#pragma unroll
    for (int i = 0; i < N; ++i) r.t[i] = 0.0;
    return r;
}

template <int N>
__device__ inline Dual<N> dvar(double x, int k) {
    Dual<N> r = dconst<N>(x);
    r.t[k] = 1.0;
    return r;
}

template <int N>
__device__ inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.t[i] = a.t[i] + b.t[i];
    return r;
}

template <int N>
__device__ inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.t[i] = a.t[i] - b.t[i];
    return r;
}

template <int N>
__device__ inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.t[i] = a.t[i] * b.v + a.v * b.t[i];
    return r;
}

template <int N>
__device__ inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.t[i] = (a.t[i] - r.v * b.t[i]) * ib;
    return r;
}

template <int N>
__device__ inline Dual<N> operator+(const Dual<N>& a, double b) {
    Dual<N> r = a;
    r.v += b;
    return r;
}

template <int N>
__device__ inline Dual<N> operator+(double b, const Dual<N>& a) {
    return a + b;
}

template <int N>
__device__ inline Dual<N> operator-(const Dual<N>& a, double b) {
    Dual<N> r = a;
    r.v -= b;
    return r;
}

template <int N>
__device__ inline Dual<N> operator-(double b, const Dual<N>& a) {
    Dual<N> r;
    r.v = b - a.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.t[i] = -a.t[i];
    return r;
}

template <int N>
__device__ inline Dual<N> operator*(const Dual<N>& a, double b) {
    Dual<N> r;
    r.v = a.v * b;
#pragma unroll
    for (int i = 0; i < N; ++i) r.t[i] = a.t[i] * b;
    return r;
}

template <int N>
__device__ inline Dual<N> operator*(double b, const Dual<N>& a) {
    return a * b;
}

template <int N>
__device__ inline Dual<N> operator/(const Dual<N>& a, double b) {
    return a * (1.0 / b);
}

template <int N>
__device__ inline Dual<N> operator/(double a, const Dual<N>& b) {
    return dconst<N>(a) / b;
}

template <int N>
__device__ inline Dual<N> dsqrt(const Dual<N>& a) {
    Dual<N> r;
    r.v = sqrt(a.v);
    const double k = 0.5 / r.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.t[i] = a.t[i] * k;
    return r;
}

__device__ inline double dsqrt(double a) { return sqrt(a); }

template <class T>
__device__ inline void tcross(const T a[3], const T b[3], T c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

template <class T>
__device__ inline T tdot(const T a[3], const T b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// ---------------------------------------------------------------- pushers
// boris_single_particle (boris.py:14-56), relativistic_boris_single_particle
// (boris.py:60-124) and higuera_cary_single_particle (higuera_cary.py:57-113)
// in FP64 over a generic scalar (the dual numbers give their Jacobians).
template <class T>
__device__ inline void push_t(const T v[3], const T E[3], const T B[3], double h, int pusher, int relativistic,
                              double C, T out[3]) {
    const double C2 = C * C;
    if (pusher == 1) {
        T u[3], ue[3], beta[3], t[3], um[3], c1[3], nu[3];
        T g0 = 1.0 / dsqrt(1.0 - tdot(v, v) / C2);
        for (int a = 0; a < 3; ++a) {
            u[a] = v[a] * g0;
            ue[a] = u[a] + E[a] * h;
            beta[a] = B[a] * h;
        }
        T b2 = tdot(beta, beta);
        T ustar = tdot(ue, beta) / C;
        T sigma = (1.0 + tdot(ue, ue) / C2) - b2;
        T gam = dsqrt((sigma + dsqrt(sigma * sigma + 4.0 * (b2 + ustar * ustar))) / 2.0);
        for (int a = 0; a < 3; ++a) t[a] = beta[a] / gam;
        T s = 1.0 / (1.0 + tdot(t, t));
        T uet = tdot(ue, t);
        tcross(ue, t, c1);
        for (int a = 0; a < 3; ++a) um[a] = s * (ue[a] + uet * t[a] + c1[a]);
        tcross(um, t, c1);
        for (int a = 0; a < 3; ++a) nu[a] = um[a] + E[a] * h + c1[a];
        T ng = dsqrt(1.0 + tdot(nu, nu) / C2);
        for (int a = 0; a < 3; ++a) out[a] = nu[a] / ng;
        return;
    }
    T um[3], t[3], up[3], s[3], c1[3];
    T gm;
    if (relativistic) {
        T g = 1.0 / dsqrt(1.0 - tdot(v, v) / C2);
        for (int a = 0; a < 3; ++a) um[a] = v[a] * g + E[a] * h;
        gm = dsqrt(1.0 + tdot(um, um) / C2);
        for (int a = 0; a < 3; ++a) t[a] = B[a] * h / gm;
    } else {
        for (int a = 0; a < 3; ++a) {
            um[a] = v[a] + E[a] * h;
            t[a] = B[a] * h;
        }
    }
    tcross(um, t, c1);
    for (int a = 0; a < 3; ++a) up[a] = um[a] + c1[a];
    T tt = 1.0 + tdot(t, t);
    for (int a = 0; a < 3; ++a) s[a] = 2.0 * t[a] / tt;
    tcross(up, s, c1);
    T nu[3];
    for (int a = 0; a < 3; ++a) nu[a] = um[a] + c1[a] + E[a] * h;
    if (relativistic) {
        T ng = dsqrt(1.0 + tdot(nu, nu) / C2);
        for (int a = 0; a < 3; ++a) out[a] = nu[a] / ng;
    } else {
        for (int a = 0; a < 3; ++a) out[a] = nu[a];
    }
}

// Reverse mode of the (relativistic) Boris push of push_t, in FP64: given
// lo = dL/d(v'), returns lin = dL/d(v, E_p, B_p) in one backward sweep over
// the forward values (the tangent passes of push_t take nine sweeps).  Each
// line below is the transpose of one forward statement, in reverse order;
// for c = a x b the transposes are la += b x lc, lb += lc x a.
__device__ inline void boris_vjp(const double v[3], const double E[3], const double B[3], double h,
                                 int relativistic, double C, const double lo[3], double lin[9]) {
    const double C2 = C * C;
    double um[3], t[3], up[3], s[3], c1[3], nu[3];
    double g = 1.0, gm = 1.0;
    if (relativistic) {
        g = 1.0 / sqrt(1.0 - tdot(v, v) / C2);
        for (int a = 0; a < 3; ++a) um[a] = v[a] * g + E[a] * h;
        gm = sqrt(1.0 + tdot(um, um) / C2);
        for (int a = 0; a < 3; ++a) t[a] = B[a] * h / gm;
    } else {
        for (int a = 0; a < 3; ++a) {
            um[a] = v[a] + E[a] * h;
            t[a] = B[a] * h;
        }
    }
    tcross(um, t, c1);
    for (int a = 0; a < 3; ++a) up[a] = um[a] + c1[a];
    const double tt = 1.0 + tdot(t, t);
    for (int a = 0; a < 3; ++a) s[a] = 2.0 * t[a] / tt;
    tcross(up, s, c1);
    for (int a = 0; a < 3; ++a) nu[a] = um[a] + c1[a] + E[a] * h;
    // out = nu / ng
    double lnu[3];
    if (relativistic) {
        const double ng = sqrt(1.0 + tdot(nu, nu) / C2);
        const double k = tdot(lo, nu) / (C2 * ng * ng * ng);
        for (int a = 0; a < 3; ++a) lnu[a] = lo[a] / ng - k * nu[a];
    } else {
        for (int a = 0; a < 3; ++a) lnu[a] = lo[a];
    }
    // nu = um + c2 + E h,  c2 = up x s
    double lum[3], lE[3], lup[3], ls[3], lt[3];
    for (int a = 0; a < 3; ++a) {
        lum[a] = lnu[a];
        lE[a] = h * lnu[a];
    }
    tcross(s, lnu, lup);
    tcross(lnu, up, ls);
    // s = 2 t / tt
    const double st = tdot(ls, t);
    for (int a = 0; a < 3; ++a) lt[a] = 2.0 * ls[a] / tt - 4.0 * t[a] * st / (tt * tt);
    // up = um + c1,  c1 = um x t
    double w[3];
    tcross(t, lup, w);
    for (int a = 0; a < 3; ++a) lum[a] += lup[a] + w[a];
    tcross(lup, um, w);
    for (int a = 0; a < 3; ++a) lt[a] += w[a];
    double lv[3];
    if (relativistic) {
        // t = B h / gm,  gm = sqrt(1 + um.um / C2)
        const double lgm = -tdot(t, lt) / gm;
        for (int a = 0; a < 3; ++a) {
            lin[6 + a] = h * lt[a] / gm;
            lum[a] += lgm * um[a] / (C2 * gm);
        }
        // um = v g + E h,  g = (1 - v.v / C2)^(-1/2)
        const double lg = tdot(v, lum);
        for (int a = 0; a < 3; ++a) lv[a] = g * lum[a] + lg * g * g * g * v[a] / C2;
    } else {
        for (int a = 0; a < 3; ++a) {
            lin[6 + a] = h * lt[a];
            lv[a] = lum[a];
        }
    }
    for (int a = 0; a < 3; ++a) {
        lin[a] = lv[a];
        lin[3 + a] = lE[a] + h * lum[a];
    }
}

// ----------------------------------------------------------------- gather
// One axis of interpolate_field_to_particles (boris.py:127-258): stencil
// points, weights and their x-derivatives (weights divide by the grid spacing).
// kN = 3 holds the reference's 3-point stencil of either shape; kN = 2 holds
// only the two points a, a+1 that carry the first-order weights (the
// reference's third point has weight 0 for CIC), for first-order callers.
template <int kN>
struct AdjAxisN {
    int p[kN];
    int r0;                 // unwrapped array index of point 0 (the scatter box coordinate)
    double w[kN], dw[kN];
    int n;
};
using AdjAxis = AdjAxisN<3>;

// First-order two-point axis: the nonzero points of the 3-point stencil in
// the same order, so every sum over points adds the same terms in the same
// order (the zero-weight terms of the 3-point form add exact zeros).
__device__ inline void adj_gather_axis(double x, double org, double sp, int sf, int L, bool active, int g,
                                       AdjAxisN<2>& s) {
    (void)sf;               // first order only
    const int a = anchor_of(x, org, sp, 1);
    const double r = (x - node_coord(a, sp, org)) / sp;
    if (active) {
        s.p[0] = pmod(a, L);
        s.p[1] = pmod(a + 1, L);
        s.w[0] = 1.0 - r;
        s.w[1] = r;
        s.dw[0] = -1.0 / sp;
        s.dw[1] = 1.0 / sp;
        s.r0 = a;
        s.n = 2;
    } else {
        s.p[0] = g;
        s.r0 = g;
        s.w[0] = (1.0 - r) + r;
        s.dw[0] = 0.0;
        s.n = 1;
    }
}

__device__ inline void adj_gather_axis(double x, double org, double sp, int sf, int L, bool active, int g,
                                       AdjAxis& s) {
    int a = anchor_of(x, org, sp, sf);
    double r = (x - node_coord(a, sp, org)) / sp;
    double w[3], dw[3];
    if (sf == 1) {
        w[0] = 0.0; w[1] = 1.0 - r; w[2] = r;
        dw[0] = 0.0; dw[1] = -1.0 / sp; dw[2] = 1.0 / sp;
    } else {
        w[0] = 0.5 * (0.5 - r) * (0.5 - r);
        w[1] = 0.75 - r * r;
        w[2] = 0.5 * (0.5 + r) * (0.5 + r);
        dw[0] = -(0.5 - r) / sp;
        dw[1] = -2.0 * r / sp;
        dw[2] = (0.5 + r) / sp;
    }
    if (active) {
        for (int q = 0; q < 3; ++q) {
            s.p[q] = pmod(a - 1 + q, L);
            s.w[q] = w[q];
            s.dw[q] = dw[q];
        }
        s.r0 = a - 1;
        s.n = 3;
    } else {
        s.p[0] = g;
        s.r0 = g;
        s.w[0] = w[0] + w[1] + w[2];
        s.dw[0] = 0.0;      // the collapsed weight is the partition of unity
        s.n = 1;
    }
}

template <int kN>
struct AdjGridN {
    AdjAxisN<kN> c[3], v[3];
};
using AdjGrid = AdjGridN<3>;

// value and d/dx of one staggered component; `stag` bit a set = vertex line on axis a
template <int kN>
__device__ inline void gather_grad(const PicDesc& d, const AdjGridN<kN>& gr, int stag, const float* F,
                                   const float* Fe, double& val, double grad[3]) {
    const AdjAxisN<kN>& ax = (stag & 1) ? gr.v[0] : gr.c[0];
    const AdjAxisN<kN>& ay = (stag & 2) ? gr.v[1] : gr.c[1];
    const AdjAxisN<kN>& az = (stag & 4) ? gr.v[2] : gr.c[2];
    val = 0.0;
    grad[0] = grad[1] = grad[2] = 0.0;
#pragma unroll
    for (int i = 0; i < kN; ++i)
#pragma unroll
        for (int j = 0; j < kN; ++j)
#pragma unroll
            for (int k = 0; k < kN; ++k) {
                if (i >= ax.n || j >= ay.n || k >= az.n) continue;
                int64_t id = fidx(d, ax.p[i], ay.p[j], az.p[k]);
                double f = (double)F[id];
                if (Fe) f += (double)Fe[id];
                val += f * ax.w[i] * ay.w[j] * az.w[k];
                grad[0] += f * ax.dw[i] * ay.w[j] * az.w[k];
                grad[1] += f * ax.w[i] * ay.dw[j] * az.w[k];
                grad[2] += f * ax.w[i] * ay.w[j] * az.dw[k];
            }
}

// Scatter box: the block's gather nodes (unwrapped indices) in LDS for the
// six components; nodes are added to the interior adjoint arrays at their
// periodic partners once per block.
struct ScatterBox {
    double* box;
    int o[3], e[3];
    bool on;
};

template <int kN>
__device__ inline void scatter_adj(const PicDesc& d, const AdjGridN<kN>& gr, int stag, double lam, double* G,
                                   const ScatterBox& sb, int comp) {
    if (lam == 0.0) return;
    const AdjAxisN<kN>& ax = (stag & 1) ? gr.v[0] : gr.c[0];
    const AdjAxisN<kN>& ay = (stag & 2) ? gr.v[1] : gr.c[1];
    const AdjAxisN<kN>& az = (stag & 4) ? gr.v[2] : gr.c[2];
#pragma unroll
    for (int i = 0; i < kN; ++i)
#pragma unroll
        for (int j = 0; j < kN; ++j)
#pragma unroll
            for (int k = 0; k < kN; ++k) {
                if (i >= ax.n || j >= ay.n || k >= az.n) continue;
                double w = ax.w[i] * ay.w[j] * az.w[k];
                if (w == 0.0) continue;
                if (sb.on) {
                    const int bi = ax.r0 + i - sb.o[0], bj = ay.r0 + j - sb.o[1], bk = az.r0 + k - sb.o[2];
                    atomicAdd(&sb.box[((comp * sb.e[0] + bi) * sb.e[1] + bj) * sb.e[2] + bk], lam * w);
                } else {
                    unsafeAtomicAdd(&G[partner_index(d, ax.p[i], ay.p[j], az.p[k])], lam * w);
                }
            }
}

// Staggering (particle_push.py:63-68): Ex (v,c,c) Ey (c,v,c) Ez (c,c,v)
// Bx (c,v,v) By (v,c,v) Bz (v,v,c); bit a = vertex on axis a.
__device__ constexpr int kStagE[3] = {1, 2, 4};
__device__ constexpr int kStagB[3] = {6, 5, 3};

// -------------------------------------------------------------- Esirkepov
// Esirkepov.py:105-331 over dual numbers with tangents (x_old[3], v'[3]).
using D6 = Dual<6>;

struct DualAxis {
    int base;
    int lo, hi;        // structural support of S0 and S_new in the 5-slot frame
    D6 S0[5], D[5], C[5];
};

__device__ inline void shape3_dual(const D6& r, int sf, D6 w[3]) {
    if (sf == 1) {
        w[0] = dconst<6>(0.0);
        w[1] = 1.0 - r;
        w[2] = r;
    } else {
        D6 a = 0.5 - r, b = 0.5 + r;
        w[0] = 0.5 * (a * a);
        w[1] = 0.75 - r * r;
        w[2] = 0.5 * (b * b);
    }
}

__device__ inline void dual_axis(const D6& xo, const D6& xn, double org, double sp, double dd, int sf, bool active,
                                 int g, DualAxis& e) {
    int ao = anchor_of(xo.v, org, sp, sf);
    int an = anchor_of(xn.v, org, sp, sf);
    D6 wo[3], wn[3];
    shape3_dual((xo - node_coord(ao, sp, org)) / dd, sf, wo);
    shape3_dual((xn - node_coord(an, sp, org)) / dd, sf, wn);
    D6 zero = dconst<6>(0.0);
    D6 S1[5] = {zero, wn[0], wn[1], wn[2], zero};
    D6 Sp[5] = {zero, wo[0], wo[1], wo[2], zero};
    int shift = an - ao;
    D6 run = zero;
    for (int i = 0; i < 5; ++i) {
        e.S0[i] = Sp[pmod(i + shift, 5)];          // jnp.roll(w5, -shift)
        e.D[i] = S1[i] - e.S0[i];
    }
    if (!active) {
        D6 tn = S1[0] + S1[1] + S1[2] + S1[3] + S1[4];
        D6 to = e.S0[0] + e.S0[1] + e.S0[2] + e.S0[3] + e.S0[4];
        e.base = g;
        e.lo = e.hi = 0;
        e.S0[0] = to;
        e.D[0] = tn - to;
        e.C[0] = zero;
        return;
    }
    e.base = an - 2;
    // S_new lives on slots {2,3} (CIC) / {1,2,3} (TSC); S0 on the same slots
    // moved by -shift; a roll by two or more wraps, so keep the whole frame then
    const int s_lo = sf == 1 ? 2 : 1;
    if (shift > 1 || shift < -1) {
        e.lo = 0;
        e.hi = 4;
    } else {
        e.lo = min(s_lo, s_lo - shift);
        e.hi = max(3, 3 - shift);
    }
    for (int i = 0; i < 5; ++i) {
        run = run + e.D[i];
        e.C[i] = run;
    }
}

// Contracts the adjoint of J (interior array; a raw deposit node reads its
// periodic partner, the transpose of fold + refresh) with dJ/d(x_old, v').
// Loops run over the structural support of the stencil: terms outside it are
// constant zero in the forward deposit.
__device__ inline void esirkepov_adjoint(const PicDesc& d, double qw, const D6 xo[3], const D6 xn[3],
                                               const D6 v[3], const double* const lamJ[3], double g6[6]) {
    DualAxis ax[3];
    bool act[3];
    for (int a = 0; a < 3; ++a) {
        act[a] = d.n[a] > 1;
        dual_axis(xo[a], xn[a], d.org_c[a], d.sp_c[a], d.d[a], d.shape_factor, act[a], d.g, ax[a]);
    }
    const double dV = d.d[0] * d.d[1] * d.d[2];
    for (int c = 0; c < 3; ++c) {
        int o1 = c == 0 ? 1 : 0, o2 = c == 2 ? 1 : 2;
        int fa[2], nf = 0;
        for (int a = 0; a < 3; ++a)
            if (a != c && act[a]) fa[nf++] = a;
        D6 pre;
        int i0, i1;
        if (act[c]) {
            pre = dconst<6>(-(qw / (d.d[o1] * d.d[o2])) / d.dt);
            i0 = ax[c].lo;
            i1 = ax[c].hi;          // the running sum vanishes at the top of the support
        } else {
            pre = v[c] * (qw / dV);
            i0 = 0;
            i1 = 1;
        }
        const int j0 = nf >= 1 ? ax[fa[0]].lo : 0, j1 = nf >= 1 ? ax[fa[0]].hi : 0;
        const int k0 = nf == 2 ? ax[fa[1]].lo : 0, k1 = nf == 2 ? ax[fa[1]].hi : 0;
        const double* L = lamJ[c];
        for (int i = i0; i < i1; ++i) {
            D6 ci = act[c] ? ax[c].C[i] : dconst<6>(1.0);
            for (int j = j0; j <= j1; ++j)
                for (int k = k0; k <= k1; ++k) {
                    int idx[3] = {ax[0].base, ax[1].base, ax[2].base};
                    if (act[c]) idx[c] += i;
                    if (nf >= 1) idx[fa[0]] += j;
                    if (nf == 2) idx[fa[1]] += k;
                    if (!drop_index(idx[0], d.dims[0]) || !drop_index(idx[1], d.dims[1]) ||
                        !drop_index(idx[2], d.dims[2]))
                        continue;
                    double lam = L[partner_index(d, idx[0], idx[1], idx[2])];
                    if (lam == 0.0) continue;
                    D6 y;
                    if (nf == 0) {
                        y = dconst<6>(1.0);
                    } else if (nf == 1) {
                        const DualAxis& A = ax[fa[0]];
                        y = A.S0[j] + 0.5 * A.D[j];
                    } else {
                        const DualAxis& A = ax[fa[0]];
                        const DualAxis& B = ax[fa[1]];
                        y = A.S0[j] * B.S0[k] + 0.5 * (A.S0[j] * B.D[k] + A.D[j] * B.S0[k]) +
                            (1.0 / 3.0) * (A.D[j] * B.D[k]);
                    }
                    D6 val = pre * ci * y;
                    for (int m = 0; m < 6; ++m) g6[m] += lam * val.t[m];
                }
        }
    }
}

// First-order Esirkepov adjoint in the 3-slot frame of a particle whose
// anchor moves by at most one cell per active axis (slots b = min(a_old,
// a_new) .. b+2; all other slots of the 5-slot frame hold zero).  Weights are
// linear in the positions, so the derivatives are exact: per axis, S0 depends
// on x_old only and S_new on x_new only (dS/dx = +-1/d).  Returns false when
// an axis moves further (the caller uses the general dual-number path).
struct Cic3 {
    int b;
    float S0[3], D[3], dS0[3], dS1[3];    // dS0 = dS0/dx_old, dS1 = dS_new/dx_new
};

__device__ inline bool cic3_axis(double xo, double xn, double org, double sp, double dd, Cic3& e) {
    const int ao = anchor_of(xo, org, sp, 1), an = anchor_of(xn, org, sp, 1);
    if (an - ao > 1 || ao - an > 1) return false;
    const double ro = (xo - node_coord(ao, sp, org)) / dd, rn = (xn - node_coord(an, sp, org)) / dd;
    const int b = min(ao, an);
    double s0[3] = {0.0, 0.0, 0.0}, s1[3] = {0.0, 0.0, 0.0};
    const float id = (float)(1.0 / dd);
#pragma unroll
    for (int q = 0; q < 3; ++q) e.dS0[q] = e.dS1[q] = 0.0f;
    if (ao == b) { s0[0] = 1.0 - ro; s0[1] = ro; e.dS0[0] = -id; e.dS0[1] = id; }
    else { s0[1] = 1.0 - ro; s0[2] = ro; e.dS0[1] = -id; e.dS0[2] = id; }
    if (an == b) { s1[0] = 1.0 - rn; s1[1] = rn; e.dS1[0] = -id; e.dS1[1] = id; }
    else { s1[1] = 1.0 - rn; s1[2] = rn; e.dS1[1] = -id; e.dS1[2] = id; }
    e.b = b;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        e.S0[q] = (float)s0[q];
        e.D[q] = (float)(s1[q] - s0[q]);
    }
    return true;
}

// Adds dL/dx_old (go[3]), dL/dx_new (gn[3]) and the direct dL/dv' of reduced
// components (gv[3]) for lambda = the interior adjoint of J.
__device__ inline bool cic_esirkepov_adjoint(const PicDesc& d, double qw, const double xo[3], const double xn[3],
                                             const double v[3], const double* const lamJ[3], double go[3],
                                             double gn[3], double gv[3]) {
    Cic3 ax[3];
    bool act[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        act[a] = d.n[a] > 1;
        if (act[a] && !cic3_axis(xo[a], xn[a], d.org_c[a], d.sp_c[a], d.d[a], ax[a])) return false;
        if (!act[a]) ax[a].b = d.g;
    }
    const double dV = d.d[0] * d.d[1] * d.d[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int o1 = c == 0 ? 1 : 0, o2 = c == 2 ? 1 : 2;
        // factor axes: the active axes among o1, o2 (in that order)
        const bool f1 = act[o1], f2 = act[o2];
        const int A = f1 ? o1 : o2, Bx = o2;
        const int nf = (f1 ? 1 : 0) + (f2 ? 1 : 0);
        const bool ac = act[c];
        const double pre = ac ? -(qw / (d.d[o1] * d.d[o2])) / d.dt : qw * v[c] / dV;
        const double* L = lamJ[c];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            if (!ac && i > 0) continue;
            // running sum of D along c (slots 0, 1) and its derivatives
            float C = 1.0f, dCo = 0.0f, dCn = 0.0f;
            if (ac) {
                C = ax[c].D[0] + (i ? ax[c].D[1] : 0.0f);
                dCo = -(ax[c].dS0[0] + (i ? ax[c].dS0[1] : 0.0f));
                dCn = ax[c].dS1[0] + (i ? ax[c].dS1[1] : 0.0f);
            }
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                if (nf == 0 && j > 0) continue;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if (nf < 2 && k > 0) continue;
                    int idx[3] = {ax[0].b, ax[1].b, ax[2].b};
                    if (ac) idx[c] += i;
                    if (nf >= 1) idx[A] += j;
                    if (nf == 2) idx[Bx] += k;
                    if (!drop_index(idx[0], d.dims[0]) || !drop_index(idx[1], d.dims[1]) ||
                        !drop_index(idx[2], d.dims[2]))
                        continue;
                    const double lam = L[partner_index(d, idx[0], idx[1], idx[2])];
                    if (lam == 0.0) continue;
                    float F = 1.0f, dFAo = 0.0f, dFAn = 0.0f, dFBo = 0.0f, dFBn = 0.0f;
                    if (nf == 1) {
                        const Cic3& P = ax[A];
                        F = P.S0[j] + 0.5f * P.D[j];
                        dFAo = 0.5f * P.dS0[j];
                        dFAn = 0.5f * P.dS1[j];
                    } else if (nf == 2) {
                        const Cic3& P = ax[A];
                        const Cic3& Q = ax[Bx];
                        F = P.S0[j] * Q.S0[k] + 0.5f * (P.S0[j] * Q.D[k] + P.D[j] * Q.S0[k]) +
                            (1.0f / 3.0f) * P.D[j] * Q.D[k];
                        dFAo = P.dS0[j] * (0.5f * Q.S0[k] + Q.D[k] * (1.0f / 6.0f));
                        dFAn = P.dS1[j] * (0.5f * Q.S0[k] + Q.D[k] * (1.0f / 3.0f));
                        dFBo = Q.dS0[k] * (0.5f * P.S0[j] + P.D[j] * (1.0f / 6.0f));
                        dFBn = Q.dS1[k] * (0.5f * P.S0[j] + P.D[j] * (1.0f / 3.0f));
                    }
                    const double lp = lam * pre;
                    if (ac) {
                        go[c] += lp * dCo * F;
                        gn[c] += lp * dCn * F;
                    } else {
                        gv[c] += lam * (qw / dV) * F;
                    }
                    if (nf >= 1) {
                        go[A] += lp * C * dFAo;
                        gn[A] += lp * C * dFAn;
                    }
                    if (nf == 2) {
                        go[Bx] += lp * C * dFBo;
                        gn[Bx] += lp * C * dFBn;
                    }
                }
            }
        }
    }
    return true;
}

// ----------------------------------------------------------- particle kernel
struct LamPtrs {
    const double* J[3];
    double* x[3];
    double* v[3];
    double* E[3];
    double* B[3];
};

constexpr int kAdjThreads = 128;

// Gather values (no derivatives) of the six components at x.
template <int kN>
__device__ inline void gather_values(const PicDesc& d, const FieldPtrs& fp, const double x[3], double Ep[3],
                                     double Bp[3]) {
    AdjGridN<kN> gr;
    for (int a = 0; a < 3; ++a) {
        bool act = d.n[a] > 1;
        adj_gather_axis(x[a], d.org_c[a], d.sp_c[a], d.shape_factor, d.dims[a], act, d.g, gr.c[a]);
        adj_gather_axis(x[a], d.org_v[a], d.sp_v[a], d.shape_factor, d.dims[a], act, d.g, gr.v[a]);
    }
    double g3[3];
    for (int c = 0; c < 3; ++c) {
        gather_grad(d, gr, kStagE[c], fp.E[c], fp.Ee[c], Ep[c], g3);
        gather_grad(d, gr, kStagB[c], fp.B[c], fp.Be[c], Bp[c], g3);
    }
}

// The phases keep few values live at a time (register pressure sets the
// occupancy of this FP64 kernel): (1) gather values, push values, position
// and Esirkepov adjoints -> lambda(x), lambda(v'); (2) push Jacobian, one
// tangent direction per pass, contracted with lambda(v') -> lambda(v, E_p,
// B_p); (3) gather stencils again for dE_p/dx and the scatter of lambda(E_p,
// B_p) through the block's LDS box.
// kAnalytic: the first-order 3-slot path; a particle that moves two or more
// cells on an axis is appended to `defer` (untouched here) for the general
// pass.  !kAnalytic: the general dual-number path, over all slots (only ==
// nullptr) or over the `*count` slots listed in `only`.
// kPush: 1 = Boris reverse sweep only, 2 = tangent passes only, 0 = chosen
// per launch from the descriptor (both compiled in).
// kPart: 0 = all three phases in one kernel; 1 = phases 1-2, writing lambda(x)
// so far and lambda(v) into lp and lambda(E_p, B_p) with a done flag into
// `scratch` (7 doubles per slot, plane-major); 2 = phase 3 for the flagged
// slots.  The split keeps each kernel at two waves per SIMD (together they
// need more than 256 registers).
template <bool kAnalytic, int kMinBlocks, int kPush = 0, int kPart = 0>
__global__ __launch_bounds__(kAdjThreads, kMinBlocks) void adjoint_particles_kernel(
    PicDesc d, PicSpecies sp, PicParticles in, FieldPtrs fp, LamPtrs lp, int64_t n, int32_t* defer,
    int32_t* defer_count, const int32_t* only, const int32_t* only_count, double* __restrict__ scratch) {
    constexpr int kCap = 3072;   // doubles of LDS for the scatter box (6 components)
    __shared__ double sbox[kCap];
    __shared__ int bmin[3], bmax[3];
    const int64_t m_all = only ? (int64_t)*only_count : n;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < m_all; base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = base + threadIdx.x;
    const int64_t i = t < m_all ? (only ? (int64_t)only[t] : t) : n;
    if (threadIdx.x < 3) {
        bmin[threadIdx.x] = INT_MAX;
        bmax[threadIdx.x] = INT_MIN;
    }
    __syncthreads();
    bool live = i < n && in.active[i];
    if constexpr (kPart == 2) live = i < n && scratch[6 * n + i] == 1.0;
    if (kPart != 2 && i < n && !live) {
        for (int a = 0; a < 3; ++a) {
            lp.x[a][i] = 0.0;
            lp.v[a][i] = 0.0;
        }
    }
    const int64_t ii = live ? i : 0;
    const int s = species_of(ii, sp.slots, sp.count);
    const SpeciesK k = species_constants(d, sp, s);
    const double h = sp.charge[s] * d.dt / (2.0 * sp.mass[s]);
    double x[3], v[3], lx[3], lv[3];
    for (int a = 0; a < 3; ++a) {
        x[a] = live ? (double)in.x[a][ii] : 0.0;
        v[a] = live ? (double)in.u[a][ii] : 0.0;
        lx[a] = live ? lp.x[a][ii] : 0.0;   // adjoint of x' (the wrap is the identity map's derivative)
        lv[a] = live ? lp.v[a][ii] : 0.0;   // adjoint of v'
    }
    constexpr int kPts = kAnalytic ? 2 : 3;   // the analytic path runs first-order stencils only
    double lin[9];
    if constexpr (kPart != 2) {
    // ---- phase 1
    double Ep[3], Bp[3];
    gather_values<kPts>(d, fp, x, Ep, Bp);
    double vp[3];
    {
        double vn[3];
        push_t(v, Ep, Bp, h, d.pusher, d.relativistic, d.C, vn);
        for (int a = 0; a < 3; ++a) vp[a] = ((k.uu >> a) & 1u) ? vn[a] : v[a];
    }
    for (int a = 0; a < 3; ++a)
        if ((k.ux >> a) & 1u) lv[a] += d.dt * lx[a];   // x' = x + v' dt
    {
        double xnp[3];
        for (int a = 0; a < 3; ++a) xnp[a] = ((k.ux >> a) & 1u) ? step_pos(x[a], vp[a], d.dt) : x[a];
        double go[3] = {0, 0, 0}, gn[3] = {0, 0, 0}, gvd[3] = {0, 0, 0};
        if constexpr (kAnalytic) {
            if (live && !cic_esirkepov_adjoint(d, k.qw, x, xnp, vp, lp.J, go, gn, gvd)) {
                defer[atomicAdd(defer_count, 1)] = (int32_t)i;
                live = false;           // finished by the general pass
            }
            for (int a = 0; a < 3; ++a) {
                lx[a] += go[a] + gn[a];
                if ((k.ux >> a) & 1u) lv[a] += d.dt * gn[a];
                lv[a] += gvd[a];
            }
        } else if (live) {
            D6 xo[3], xnew[3], vd[3];
            for (int a = 0; a < 3; ++a) {
                xo[a] = dvar<6>(x[a], a);
                vd[a] = dvar<6>(vp[a], 3 + a);
                if ((k.ux >> a) & 1u) {
                    xnew[a] = xo[a] + vd[a] * d.dt;
                    xnew[a].v = xnp[a];    // same rounding as the forward anchor
                } else {
                    xnew[a] = xo[a];
                }
            }
            double g6[6] = {0, 0, 0, 0, 0, 0};
            esirkepov_adjoint(d, k.qw, xo, xnew, vd, lp.J, g6);
            for (int a = 0; a < 3; ++a) {
                lx[a] += g6[a];
                lv[a] += g6[3 + a];
            }
        }
    }
    // ---- phase 2: lambda(v, E_p, B_p) = J_push^T lambda(v') on update_u axes
    const bool vjp = kPush == 1 || (kPush == 0 && d.pusher == 0 && d.variant != 1);
    if (kPush != 2 && vjp) {
        double lo[3];
        for (int a = 0; a < 3; ++a) lo[a] = ((k.uu >> a) & 1u) ? lv[a] : 0.0;
        boris_vjp(v, Ep, Bp, h, d.relativistic, d.C, lo, lin);
        for (int a = 0; a < 3; ++a)
            if (!((k.uu >> a) & 1u)) lin[a] += lv[a];        // frozen component: v' = v
    } else
#pragma unroll 1
    for (int m = 0; m < 9; ++m) {
        Dual<1> vv[3], EE[3], BB[3], out[3];
        for (int a = 0; a < 3; ++a) {
            vv[a] = dconst<1>(v[a]);
            EE[a] = dconst<1>(Ep[a]);
            BB[a] = dconst<1>(Bp[a]);
        }
        if (m < 3) vv[m].t[0] = 1.0;
        else if (m < 6) EE[m - 3].t[0] = 1.0;
        else BB[m - 6].t[0] = 1.0;
        push_t(vv, EE, BB, h, d.pusher, d.relativistic, d.C, out);
        double acc = 0.0;
        for (int a = 0; a < 3; ++a) acc += ((k.uu >> a) & 1u) ? lv[a] * out[a].t[0] : 0.0;
        if (m < 3 && !((k.uu >> m) & 1u)) acc += lv[m];   // frozen component: v' = v
        lin[m] = acc;
    }
    }   // phases 1-2
    if constexpr (kPart == 1) {
        if (live) {
            for (int a = 0; a < 3; ++a) {
                lp.x[a][i] = lx[a];
                lp.v[a][i] = lin[a];
            }
            for (int m = 0; m < 6; ++m) scratch[m * n + i] = lin[3 + m];
            scratch[6 * n + i] = 1.0;
        } else if (i < n) {
            scratch[6 * n + i] = 0.0;
        }
        continue;
    }
    if constexpr (kPart == 2) {
        for (int m = 0; m < 6; ++m) lin[3 + m] = live ? scratch[m * n + i] : 0.0;
    }
    // ---- phase 3: gather stencils, dE_p/dx, scatter box
    AdjGridN<kPts> gr;
    for (int a = 0; a < 3; ++a) {
        bool act = d.n[a] > 1;
        adj_gather_axis(x[a], d.org_c[a], d.sp_c[a], d.shape_factor, d.dims[a], act, d.g, gr.c[a]);
        adj_gather_axis(x[a], d.org_v[a], d.sp_v[a], d.shape_factor, d.dims[a], act, d.g, gr.v[a]);
    }
    if (live) {
        for (int a = 0; a < 3; ++a) {
            const int lo = min(gr.c[a].r0, gr.v[a].r0);
            const int hi = max(gr.c[a].r0 + gr.c[a].n, gr.v[a].r0 + gr.v[a].n) - 1;
            atomicMin(&bmin[a], lo);
            atomicMax(&bmax[a], hi);
        }
    }
    __syncthreads();
    ScatterBox sb;
    sb.box = sbox;
    sb.on = bmax[0] >= bmin[0];
    int64_t vol64 = 1;
    for (int a = 0; a < 3; ++a) {
        sb.o[a] = bmin[a];
        sb.e[a] = sb.on ? (int)min((int64_t)bmax[a] - bmin[a] + 1, (int64_t)kCap + 1) : 1;
        vol64 *= sb.e[a];
    }
    sb.on = sb.on && 6 * vol64 <= kCap;
    const int vol = sb.on ? (int)vol64 : 1;
    if (sb.on)
        for (int e = threadIdx.x; e < 6 * vol; e += blockDim.x) sbox[e] = 0.0;
    __syncthreads();
    if (live) {
        for (int c = 0; c < 3; ++c) {
            double val, g3[3];
            gather_grad(d, gr, kStagE[c], fp.E[c], fp.Ee[c], val, g3);
            for (int a = 0; a < 3; ++a) lx[a] += lin[3 + c] * g3[a];
            gather_grad(d, gr, kStagB[c], fp.B[c], fp.Be[c], val, g3);
            for (int a = 0; a < 3; ++a) lx[a] += lin[6 + c] * g3[a];
            scatter_adj(d, gr, kStagE[c], lin[3 + c], lp.E[c], sb, c);
            scatter_adj(d, gr, kStagB[c], lin[6 + c], lp.B[c], sb, 3 + c);
        }
        for (int a = 0; a < 3; ++a) {
            lp.x[a][i] = lx[a];
            if constexpr (kPart != 2) lp.v[a][i] = lin[a];
        }
    }
    __syncthreads();
    if (sb.on) {
        for (int e = threadIdx.x; e < 6 * vol; e += blockDim.x) {
            const double val = sbox[e];
            if (val == 0.0) continue;
            const int c = e / vol, r = e % vol;
            const int bi = r / (sb.e[1] * sb.e[2]), bj = (r / sb.e[2]) % sb.e[1], bk = r % sb.e[2];
            double* G = c < 3 ? lp.E[c] : lp.B[c - 3];
            unsafeAtomicAdd(&G[partner_index(d, sb.o[0] + bi, sb.o[1] + bj, sb.o[2] + bk)], val);
        }
    }
    __syncthreads();   // the next pass reuses the box
    }
}

__global__ void zero_count_kernel(int32_t* c) { *c = 0; }

// ------------------------------------------------------------ field kernels
// Interior arrays (Nx, Ny, Nz), z fastest, periodic wrap.
__device__ inline int64_t iidx(const PicDesc& d, int i, int j, int k) {
    return ((int64_t)i * d.n[1] + j) * d.n[2] + k;
}

__device__ inline int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

struct D3 {
    const double* p[3];
};
struct W3 {
    double* p[3];
};

// out = a + coef * curl(b); backward = 0 forward differences (update_B's curl E),
// 1 backward differences (update_E's curl B).
__global__ void curl_axpy_kernel(PicDesc d, W3 out, D3 a, D3 b, double coef, int backward) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t tot = (int64_t)d.n[0] * d.n[1] * d.n[2];
    if (t >= tot) return;
    int k = (int)(t % d.n[2]);
    int j = (int)((t / d.n[2]) % d.n[1]);
    int i = (int)(t / ((int64_t)d.n[2] * d.n[1]));
    int s = backward ? -1 : 1;
    // difference along axis ax of component comp: forward (f[+1]-f)/d, backward (f - f[-1])/d
    auto diff = [&](int comp, int ax) -> double {
        int ii = i, jj = j, kk = k;
        if (ax == 0) ii = wrap(i + s, d.n[0]);
        if (ax == 1) jj = wrap(j + s, d.n[1]);
        if (ax == 2) kk = wrap(k + s, d.n[2]);
        double here = b.p[comp][t], there = b.p[comp][iidx(d, ii, jj, kk)];
        return (backward ? here - there : there - here) / d.d[ax];
    };
    double cx = diff(2, 1) - diff(1, 2);
    double cy = diff(0, 2) - diff(2, 0);
    double cz = diff(1, 0) - diff(0, 1);
    double ax0 = a.p[0] ? a.p[0][t] : 0.0, ay0 = a.p[1] ? a.p[1][t] : 0.0, az0 = a.p[2] ? a.p[2][t] : 0.0;
    out.p[0][t] = ax0 + coef * cx;
    out.p[1][t] = ay0 + coef * cy;
    out.p[2][t] = az0 + coef * cz;
}

// digital_filter (filters.py:234-291) on periodic interior data: centre alpha,
// six faces (1 - alpha)/6.  Symmetric, hence its own transpose.
__global__ void filter_f64_kernel(PicDesc d, W3 out, D3 in, double alpha) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t tot = (int64_t)d.n[0] * d.n[1] * d.n[2];
    if (t >= tot) return;
    int k = (int)(t % d.n[2]);
    int j = (int)((t / d.n[2]) % d.n[1]);
    int i = (int)(t / ((int64_t)d.n[2] * d.n[1]));
    double f = (1.0 - alpha) / 6.0;
    for (int c = 0; c < 3; ++c) {
        const double* p = in.p[c];
        double s = p[iidx(d, wrap(i - 1, d.n[0]), j, k)] + p[iidx(d, wrap(i + 1, d.n[0]), j, k)] +
                   p[iidx(d, i, wrap(j - 1, d.n[1]), k)] + p[iidx(d, i, wrap(j + 1, d.n[1]), k)] +
                   p[iidx(d, i, j, wrap(k - 1, d.n[2]))] + p[iidx(d, i, j, wrap(k + 1, d.n[2]))];
        out.p[c][t] = alpha * p[t] + f * s;
    }
}

__global__ void scale_kernel(W3 out, D3 in, double coef, int64_t n) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    for (int c = 0; c < 3; ++c) out.p[c][t] = coef * in.p[c][t];
}

__global__ void expand_kernel(PicDesc d, const double* in, double* out, int64_t tot) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tot) return;
    int k = (int)(t % d.dims[2]);
    int j = (int)((t / d.dims[2]) % d.dims[1]);
    int i = (int)(t / ((int64_t)d.dims[2] * d.dims[1]));
    out[t] = in[iidx(d, partner(i, d.g, d.n[0]), partner(j, d.g, d.n[1]), partner(k, d.g, d.n[2]))];
}

__global__ void fold_add_kernel(PicDesc d, const double* in, double* out, int64_t tot) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tot) return;
    double v = in[t];
    if (v == 0.0) return;
    int k = (int)(t % d.dims[2]);
    int j = (int)((t / d.dims[2]) % d.dims[1]);
    int i = (int)(t / ((int64_t)d.dims[2] * d.dims[1]));
    unsafeAtomicAdd(&out[iidx(d, partner(i, d.g, d.n[0]), partner(j, d.g, d.n[1]), partner(k, d.g, d.n[2]))], v);
}

__global__ void fill_f64_kernel(double* p, int64_t n, double v) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

inline bool periodic_all(const PicDesc* d) {
    for (int a = 0; a < 3; ++a)
        if (d->bc[a] != 0 || d->pbc[a] != 0) return false;
    return true;
}

inline int64_t interior_size(const PicDesc& d) { return (int64_t)d.n[0] * d.n[1] * d.n[2]; }

}  // namespace

extern "C" {

int pic_fill_f64(double* ptr, int64_t n, double value, void* stream) {
    PIC_REQUIRE(ptr || n == 0);
    if (n == 0) return 0;
    hipLaunchKernelGGL(fill_f64_kernel, dim3(grid_1d(n, kThreads)), dim3(kThreads), 0, as_stream(stream), ptr, n,
                       value);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_periodic_expand_f64(const PicDesc* desc, const double* interior, double* ghosted, void* stream) {
    PIC_REQUIRE(desc_ok(desc) && periodic_all(desc) && interior && ghosted && (const double*)ghosted != interior);
    int64_t tot = field_size(*desc);
    hipLaunchKernelGGL(expand_kernel, dim3(grid_1d(tot, kThreads)), dim3(kThreads), 0, as_stream(stream), *desc,
                       interior, ghosted, tot);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_periodic_fold_add_f64(const PicDesc* desc, const double* ghosted, double* interior, void* stream) {
    PIC_REQUIRE(desc_ok(desc) && periodic_all(desc) && interior && ghosted && ghosted != (const double*)interior);
    int64_t tot = field_size(*desc);
    hipLaunchKernelGGL(fold_add_kernel, dim3(grid_1d(tot, kThreads)), dim3(kThreads), 0, as_stream(stream), *desc,
                       ghosted, interior, tot);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_adjoint_yee(const PicDesc* desc, double* const lamE[3], double* const lamB[3], double* const lamJ[3],
                    double* const work[6], void* stream) {
    PIC_REQUIRE(desc_ok(desc) && periodic_all(desc) && lamE && lamB && lamJ && work);
    for (int a = 0; a < 3; ++a) PIC_REQUIRE(lamE[a] && lamB[a] && lamJ[a]);
    for (int a = 0; a < 6; ++a) PIC_REQUIRE(work[a]);
    const PicDesc& d = *desc;
    hipStream_t st = as_stream(stream);
    const int64_t tot = interior_size(d);
    const dim3 grid(grid_1d(tot, kThreads)), blk(kThreads);
    const double half = 0.5 * d.dt;
    W3 LE{{lamE[0], lamE[1], lamE[2]}}, LB{{lamB[0], lamB[1], lamB[2]}}, LJ{{lamJ[0], lamJ[1], lamJ[2]}};
    W3 MU{{work[0], work[1], work[2]}}, NU{{work[3], work[4], work[5]}};
    auto cd = [](const W3& w) { return D3{{w.p[0], w.p[1], w.p[2]}}; };
    const bool filt = d.alpha != 1.0;
    // mu = F^T lamB'  (last half step's filter)
    if (filt) {
        hipLaunchKernelGGL(filter_f64_kernel, grid, blk, 0, st, d, MU, cd(LB), d.alpha);
    } else {
        hipLaunchKernelGGL(scale_kernel, grid, blk, 0, st, MU, cd(LB), 1.0, tot);
    }
    // lamE'' = lamE' - (dt/2) curl_f^T mu = lamE' - (dt/2) curl_b mu
    hipLaunchKernelGGL(curl_axpy_kernel, grid, blk, 0, st, d, LE, cd(LE), cd(MU), -half, 1);
    // nu = F^T lamE''
    if (filt) {
        hipLaunchKernelGGL(filter_f64_kernel, grid, blk, 0, st, d, NU, cd(LE), d.alpha);
    } else {
        hipLaunchKernelGGL(scale_kernel, grid, blk, 0, st, NU, cd(LE), 1.0, tot);
    }
    // lamJ = -(dt/eps) nu
    hipLaunchKernelGGL(scale_kernel, grid, blk, 0, st, LJ, cd(NU), -d.dt / d.eps, tot);
    // lamB_h = mu + dt C^2 curl_b^T nu = mu + dt C^2 curl_f nu
    hipLaunchKernelGGL(curl_axpy_kernel, grid, blk, 0, st, d, LB, cd(MU), cd(NU), d.dt * d.C * d.C, 0);
    // lamE = nu - (dt/2) curl_b lamB_h;  lamB = lamB_h
    hipLaunchKernelGGL(curl_axpy_kernel, grid, blk, 0, st, d, LE, cd(NU), cd(LB), -half, 1);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_adjoint_particles(const PicDesc* desc, const PicSpecies* sp, const PicParticles* state,
                          const float* const E[3], const float* const B[3], const float* const E_ext[3],
                          const float* const B_ext[3], const double* const lamJ[3], double* const lam_x[3],
                          double* const lam_v[3], double* const lamE[3], double* const lamB[3], int32_t min_blocks,
                          int32_t* workspace, double* scratch, void* stream) {
    PIC_REQUIRE(desc_ok(desc) && periodic_all(desc) && particles_ok(sp, state) && E && B && lamJ && lam_x &&
                lam_v && lamE && lamB && workspace);
    int64_t n = (int64_t)sp->count * sp->slots;
    for (int a = 0; a < 3; ++a)
        PIC_REQUIRE(E[a] && B[a] && lamJ[a] && ((lam_x[a] && lam_v[a]) || n == 0) && lamE[a] && lamB[a]);
    PIC_REQUIRE(n < ((int64_t)1 << 31));
    if (n == 0) return 0;
    FieldPtrs fp = make_fields(E, B, E_ext, B_ext);
    LamPtrs lp;
    for (int a = 0; a < 3; ++a) {
        lp.J[a] = lamJ[a];
        lp.x[a] = lam_x[a];
        lp.v[a] = lam_v[a];
        lp.E[a] = lamE[a];
        lp.B[a] = lamB[a];
    }
    int32_t* count = workspace;
    int32_t* list = workspace + 1;
    const dim3 grid(grid_1d(n, kAdjThreads)), blk(kAdjThreads);
    const dim3 grid2(grid_1d(n, kAdjThreads) < 512 ? grid_1d(n, kAdjThreads) : 512);
    hipStream_t st = as_stream(stream);
    if (desc->shape_factor != 1) {
        hipLaunchKernelGGL((adjoint_particles_kernel<false, 1>), grid, blk, 0, st, *desc, *sp, *state, fp, lp, n,
                           nullptr, nullptr, nullptr, nullptr, nullptr);
        PIC_CHECK_LAUNCH();
        return 0;
    }
    PIC_REQUIRE(scratch);
    hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(1), 0, st, count);
    const bool vjp = desc->pusher == 0 && desc->variant != 1;
    // first-order path: phases 1-2, then phase 3; unforced, each lands at two
    // waves per SIMD (about 185 VGPRs), where a register target would spill
    (void)min_blocks;
#define PIC_ADJ_LAUNCH(PUSH, PART)                                                                                \
    hipLaunchKernelGGL((adjoint_particles_kernel<true, 1, PUSH, PART>), grid, blk, 0, st, *desc, *sp, *state, fp,  \
                       lp, n, list, count, nullptr, nullptr, scratch)
    if (vjp) PIC_ADJ_LAUNCH(1, 1);
    else PIC_ADJ_LAUNCH(2, 1);
    PIC_ADJ_LAUNCH(1, 2);
#undef PIC_ADJ_LAUNCH
    hipLaunchKernelGGL((adjoint_particles_kernel<false, 1>), grid2, blk, 0, st, *desc, *sp, *state, fp, lp, n, nullptr,
                       nullptr, list, count, nullptr);
    PIC_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
