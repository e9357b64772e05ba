// Electrostatic Poisson CG (FP32 iterations inside an FP64 refinement loop)
// and the FP64 energy diagnostics.
//
// CG follows solve_poisson_with_conjugate_gradient (electrostatic_yee.py:71-156):
// r0 = rho/eps + lapl(phi), p0 = r0, alpha = r.r / p.(-lapl p), phi += alpha p,
// r -= alpha (-lapl p), beta = r'.r'/r.r, p = r' + beta p.  The reference runs
// it in FP64 to an absolute tolerance of 1e-12 on ||r||.  Here the potential
// is held in FP64 and the residual r = rho/eps + lapl(phi) is recomputed from
// it in FP64 (mixed-precision iterative refinement): each outer pass solves
// -lapl e = r with FP32 CG vectors and FP64 dot products to a relative
// residual of kInnerRtol, then phi += e in FP64.  The outer loop stops at the
// reference's test ||r|| <= max(tol, rtol ||r0||) on the FP64 residual, when a
// pass no longer halves it (the FP64 rounding floor), or after max_iter inner
// iterations in total.
//
// Ghost handling: phi and p are refreshed by _apply_tiled_phi_constant_
// boundaries (:26-43) before every Laplacian in the reference.  Only the face
// neighbours of interior cells enter the 7-point Laplacian, and after that
// sequence a face neighbour across axis a is: the periodic partner (periodic
// axis; a reduced periodic axis is its own partner), the cell itself (the
// last conducting axis, whose ghosts copy the adjacent plane) or zero (an
// earlier conducting axis, whose ghosts the later calls' refresh zeroes).
// The kernels read neighbours through that map on interior-only arrays.
// Without a zero-neighbour axis the operator is singular (constants); the
// mean of the FP64 residual is then removed before each pass (the k = 0
// mode, a neutralising background), so a deposit whose net charge is zero
// only up to FP32 rounding still converges; the reference's FP64 CG relies
// on the net charge being zero to its own rounding.
// The scalars live on the device; every iteration kernel returns at once when
// the device `done` flag is set, so the host queues iterations in batches and
// reads the flag once per batch.
#include "pic_common.h"

#include <cstdio>
#include <mutex>

int pic_refresh_fields(const PicDesc* d, float* f0, float* f1, float* f2, int bc_type, hipStream_t s);
extern "C" int pic_constant_boundary(const PicDesc* desc, float* field, int axis, void* stream);

// Scratch of the allocation-free energy entry point (pic_energy_noalloc):
// the FP64 partial sums of its three reductions, 3 x 1024 doubles = 24 KiB,
// held by every GPU the library runs on until the process ends.  Every
// partial a call reads is written earlier in the same call, so nothing needs
// clearing; calls that overlap on one GPU share the array and race.
static __device__ __align__(256) double scratch[3 * 1024];

namespace {

constexpr int kBlock = 256;
constexpr int kMaxPartials = 1024;
static_assert(sizeof(scratch) == 3 * kMaxPartials * sizeof(double), "energy scratch holds its three partial arrays");
constexpr int kBatch = 32;
constexpr int kMaxOuter = 30;
constexpr double kInnerRtol = 1.0e-5;

enum NeighbourMode : int { NB_PERIODIC = 0, NB_SELF = 1, NB_ZERO = 2 };

struct CGScalars {
    double rr, pAp, alpha, beta, tol2, rr_true, mean;
    int32_t done, iter, max_iter, pad;
};

// interior-array geometry and the face-neighbour map
struct LapMap {
    int n[3];
    int mode[3];
    int singular;   // no zero-neighbour axis: constants span the null space
    int64_t sx, sy;
    double i2[3];   // 1/d^2
};

LapMap lap_map(const PicDesc* d) {
    LapMap M;
    int last = -1;
    for (int a = 0; a < 3; ++a)
        if (d->bc[a] == 1) last = a;
    for (int a = 0; a < 3; ++a) {
        M.n[a] = d->n[a];
        M.mode[a] = d->bc[a] != 1 ? NB_PERIODIC : (a == last ? NB_SELF : NB_ZERO);
        M.i2[a] = 1.0 / (d->d[a] * d->d[a]);
    }
    M.singular = M.mode[0] != NB_ZERO && M.mode[1] != NB_ZERO && M.mode[2] != NB_ZERO;
    M.sy = d->n[2];
    M.sx = (int64_t)d->n[1] * d->n[2];
    return M;
}

__device__ inline double block_sum(double v) {
    __shared__ double sh[kBlock];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    double out = sh[0];
    __syncthreads();
    return out;
}

// value of the face neighbour at offset dir (+-1) along an axis of the
// interior array: index q, extent n, stride st
template <class T>
__device__ inline T face_nb(const T* __restrict__ f, int64_t c, int q, int n, int64_t st, int mode, int dir, T fc) {
    int qq = q + dir;
    if (qq >= 0 && qq < n) return f[c + dir * st];
    if (mode == NB_PERIODIC) return f[c + (dir > 0 ? -(int64_t)(n - 1) : (int64_t)(n - 1)) * st];
    if (mode == NB_SELF) return fc;
    return (T)0;
}

// lapl f at interior cell (i, j, k) (electrostatic_yee.py:98-103) in T
template <class T>
__device__ inline T lapl_at(const T* __restrict__ f, const LapMap& M, int64_t c, int i, int j, int k) {
    const T fc = f[c];
    const T two = (T)2;
    T lx = face_nb(f, c, i, M.n[0], M.sx, M.mode[0], 1, fc) + face_nb(f, c, i, M.n[0], M.sx, M.mode[0], -1, fc) -
           two * fc;
    T ly = face_nb(f, c, j, M.n[1], M.sy, M.mode[1], 1, fc) + face_nb(f, c, j, M.n[1], M.sy, M.mode[1], -1, fc) -
           two * fc;
    T lz = face_nb(f, c, k, M.n[2], (int64_t)1, M.mode[2], 1, fc) +
           face_nb(f, c, k, M.n[2], (int64_t)1, M.mode[2], -1, fc) - two * fc;
    return lx * (T)M.i2[0] + ly * (T)M.i2[1] + lz * (T)M.i2[2];
}

__device__ inline void split(const LapMap& M, int64_t t, int& i, int& j, int& k) {
    k = (int)(t % M.n[2]);
    int64_t r = t / M.n[2];
    j = (int)(r % M.n[1]);
    i = (int)(r / M.n[1]);
}

__device__ inline void interior_coords(const PicDesc& d, int64_t t, int& i, int& j, int& k) {
    k = (int)(t % d.n[2]);
    int64_t r = t / d.n[2];
    j = (int)(r % d.n[1]);
    i = (int)(r / d.n[1]);
    i += d.g;
    j += d.g;
    k += d.g;
}

// phi_d <- interior of phi (ghosted FP32)
__global__ void cg_load_kernel(PicDesc d, const float* __restrict__ phi, double* __restrict__ phid) {
    int64_t total = (int64_t)d.n[0] * d.n[1] * d.n[2];
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_coords(d, t, i, j, k);
        phid[t] = (double)phi[fidx(d, i, j, k)];
    }
}

// sum of the FP64 residual rho/eps + lapl(phi_d) (its mean is removed when
// the operator is singular)
__global__ void cg_residual_sum_kernel(PicDesc d, LapMap M, const float* __restrict__ rho,
                                       const double* __restrict__ phid, double* partials, double inv_eps) {
    int64_t total = (int64_t)M.n[0] * M.n[1] * M.n[2];
    double acc = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        split(M, t, i, j, k);
        acc += (double)rho[fidx(d, i + d.g, j + d.g, k + d.g)] * inv_eps + lapl_at(phid, M, t, i, j, k);
    }
    double s = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void cg_mean_kernel(const double* partials, int nparts, CGScalars* sc, double inv_n) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) acc += partials[b];
    acc = block_sum(acc);
    if (threadIdx.x == 0) sc->mean = acc * inv_n;
}

// FP64 residual r = rho/eps + lapl(phi_d) - mean; starts an inner solve: r
// (FP32), p = r, e = 0; partials of r.r in FP64 (of the FP64 residual)
__global__ void cg_residual_kernel(PicDesc d, LapMap M, const float* __restrict__ rho, const double* __restrict__ phid,
                                   float* __restrict__ r, float* __restrict__ p, float* __restrict__ e,
                                   double* partials, double inv_eps, const CGScalars* sc) {
    int64_t total = (int64_t)M.n[0] * M.n[1] * M.n[2];
    const double mean = M.singular ? sc->mean : 0.0;
    double acc = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        split(M, t, i, j, k);
        double rv = (double)rho[fidx(d, i + d.g, j + d.g, k + d.g)] * inv_eps + lapl_at(phid, M, t, i, j, k) - mean;
        float rf = (float)rv;
        r[t] = rf;
        p[t] = rf;
        e[t] = 0.0f;
        acc += rv * rv;
    }
    double s = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// r.r of the FP64 residual, and the inner solve's start: the inner residual
// is r cast to FP32, its stop threshold max(target, kInnerRtol ||r||)
__global__ void cg_start_kernel(const double* partials, int nparts, CGScalars* sc, double target, int max_iter) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) acc += partials[b];
    double rr = block_sum(acc);
    if (threadIdx.x == 0) {
        double t = fmax(target, kInnerRtol * sqrt(rr));
        sc->rr_true = rr;
        sc->rr = rr;
        sc->tol2 = t * t;
        sc->iter = 0;
        sc->max_iter = max_iter;
        sc->done = (max_iter <= 0 || !(rr > sc->tol2)) ? 1 : 0;
    }
}

__global__ void cg_ap_kernel(LapMap M, const float* __restrict__ p, float* __restrict__ Ap, double* partials,
                             const CGScalars* sc) {
    if (sc->done) return;
    int64_t total = (int64_t)M.n[0] * M.n[1] * M.n[2];
    double acc = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        split(M, t, i, j, k);
        float ap = -lapl_at(p, M, t, i, j, k);
        Ap[t] = ap;
        acc += (double)p[t] * (double)ap;
    }
    double s = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void cg_alpha_kernel(const double* partials, int nparts, CGScalars* sc) {
    if (sc->done) return;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) acc += partials[b];
    double pAp = block_sum(acc);
    if (threadIdx.x == 0) {
        sc->pAp = pAp;
        sc->alpha = sc->rr / pAp;
    }
}

__global__ void cg_update_kernel(int64_t total, float* __restrict__ e, float* __restrict__ r,
                                 const float* __restrict__ p, const float* __restrict__ Ap, double* partials,
                                 const CGScalars* sc) {
    if (sc->done) return;
    float alpha = (float)sc->alpha;
    double acc = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        e[t] += alpha * p[t];
        float rn = r[t] - alpha * Ap[t];
        r[t] = rn;
        acc += (double)rn * (double)rn;
    }
    double s = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void cg_beta_kernel(const double* partials, int nparts, CGScalars* sc) {
    if (sc->done) return;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) acc += partials[b];
    double rr_new = block_sum(acc);
    if (threadIdx.x == 0) {
        sc->beta = rr_new / sc->rr;
        sc->rr = rr_new;
        sc->iter += 1;
        if (sc->iter >= sc->max_iter || !(rr_new > sc->tol2)) sc->done = 1;
    }
}

__global__ void cg_p_kernel(int64_t total, const float* __restrict__ r, float* __restrict__ p, const CGScalars* sc) {
    if (sc->done) return;
    float beta = (float)sc->beta;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x)
        p[t] = r[t] + beta * p[t];
}

// phi_d += e (FP64 accumulation of the FP32 correction)
__global__ void cg_accumulate_kernel(int64_t total, double* __restrict__ phid, const float* __restrict__ e) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x)
        phid[t] += (double)e[t];
}

// phi interior <- phi_d rounded to FP32
__global__ void cg_store_kernel(PicDesc d, const double* __restrict__ phid, float* __restrict__ phi) {
    int64_t total = (int64_t)d.n[0] * d.n[1] * d.n[2];
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_coords(d, t, i, j, k);
        phi[fidx(d, i, j, k)] = (float)phid[t];
    }
}

// ---- energy (utils.py:108-187) ----

__global__ void energy_fields_kernel(PicDesc d, const float* ex, const float* ey, const float* ez, const float* bx,
                                     const float* by, const float* bz, double* partials) {
    int64_t total = (int64_t)d.n[0] * d.n[1] * d.n[2];
    double e2 = 0.0, b2 = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_coords(d, t, i, j, k);
        int64_t c = fidx(d, i, j, k);
        double a = ex[c], b = ey[c], cc = ez[c];
        e2 += a * a + b * b + cc * cc;
        a = bx[c];
        b = by[c];
        cc = bz[c];
        b2 += a * a + b * b + cc * cc;
    }
    double se = block_sum(e2);
    double sb = block_sum(b2);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = se;
        partials[2 * blockIdx.x + 1] = sb;
    }
}

__global__ void energy_kinetic_kernel(PicSpecies sp, PicParticles in, double C, double* partials, int64_t n) {
    double ke = 0.0;
    double C2 = C * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!in.active[i]) continue;
        int s = species_of(i, sp.slots, sp.count);
        double m = sp.mass[s] * sp.weight[s];
        double vx = in.u[0][i], vy = in.u[1][i], vz = in.u[2][i];
        double v2 = vx * vx + vy * vy + vz * vz;
        double gamma = 1.0 / sqrt(1.0 - v2 / C2);
        double mom2 = (m * gamma) * (m * gamma) * v2;
        ke += sqrt(mom2 * C2 + m * m * C2 * C2) - m * C2;
    }
    double s = block_sum(ke);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void energy_finalize(const double* fpart, int nf, const double* kpart, int nk, double ce, double cb,
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

// compute_total_momentum (utils.py:190-202): sum active |v| mass weight
__global__ void momentum_kernel(PicSpecies sp, PicParticles in, double* partials, int64_t n) {
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!in.active[i]) continue;
        int s = species_of(i, sp.slots, sp.count);
        double vx = in.u[0][i], vy = in.u[1][i], vz = in.u[2][i];
        acc += sqrt(vx * vx + vy * vy + vz * vz) * (sp.mass[s] * sp.weight[s]);
    }
    double s = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// sum of rho over the interior (the total charge is this times dV)
__global__ void charge_kernel(PicDesc d, const float* __restrict__ rho, double* partials) {
    int64_t total = (int64_t)d.n[0] * d.n[1] * d.n[2];
    double acc = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_coords(d, t, i, j, k);
        acc += (double)rho[fidx(d, i, j, k)];
    }
    double s = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void sum_finalize(const double* partials, int n, double scale, double* out) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partials[i];
    acc = block_sum(acc);
    if (threadIdx.x == 0) out[0] = scale * acc;
}

int partial_blocks(int64_t work) {
    int64_t b = (work + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > kMaxPartials) b = kMaxPartials;
    return (int)b;
}

bool desc_ok(const PicDesc* d) {
    if (!d || d->g < 1) return false;
    for (int a = 0; a < 3; ++a)
        if (d->n[a] < 1 || d->dims[a] != d->n[a] + 2 * d->g) return false;
    return true;
}

// _apply_tiled_phi_constant_boundaries (electrostatic_yee.py:26-43)
int phi_bc(const PicDesc* d, float* f, hipStream_t s) {
    bool any = false;
    for (int a = 0; a < 3; ++a) {
        if (d->bc[a] == 1) {
            int rc = pic_constant_boundary(d, f, a, (void*)s);
            if (rc) return rc;
            any = true;
        }
    }
    if (!any) return pic_refresh_fields(d, f, nullptr, nullptr, 0, s);
    return 0;
}

struct CGLayout {
    int64_t phid_off, p_off, r_off, ap_off, e_off, part_off, sc_off, total;
};

// workspace: phi_d (FP64) and p, r, Ap, e (FP32) over the interior, FP64
// partial sums, the scalar block
CGLayout cg_layout(const PicDesc* d) {
    auto up = [](int64_t b) { return ((b + 255) / 256) * 256; };
    const int64_t n = (int64_t)d->n[0] * d->n[1] * d->n[2];
    CGLayout L;
    L.phid_off = 0;
    L.p_off = up(n * 8);
    L.r_off = L.p_off + up(n * 4);
    L.ap_off = L.r_off + up(n * 4);
    L.e_off = L.ap_off + up(n * 4);
    L.part_off = L.e_off + up(n * 4);
    L.sc_off = L.part_off + kMaxPartials * 8;
    L.total = L.sc_off + 256;
    return L;
}

// compute_energy's three FP64 reductions with their partial sums in ws
// (2 * kMaxPartials field partials, then kMaxPartials kinetic partials)
int energy_launch(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, const float* const E[3],
                  const float* const B[3], double* out, double* ws, hipStream_t s) {
    double* fpart = ws;
    double* kpart = fpart + 2 * kMaxPartials;
    int64_t interior = (int64_t)desc->n[0] * desc->n[1] * desc->n[2];
    int nf = partial_blocks(interior);
    int64_t n = (int64_t)sp->count * sp->slots;
    int nk = partial_blocks(n);
    hipLaunchKernelGGL(energy_fields_kernel, dim3(nf), dim3(kBlock), 0, s, *desc, E[0], E[1], E[2], B[0], B[1], B[2],
                       fpart);
    PIC_CHECK_LAUNCH();
    hipLaunchKernelGGL(energy_kinetic_kernel, dim3(nk), dim3(kBlock), 0, s, *sp, *in, desc->C, kpart, n);
    PIC_CHECK_LAUNCH();
    double dV = desc->d[0] * desc->d[1] * desc->d[2];
    hipLaunchKernelGGL(energy_finalize, dim3(1), dim3(kBlock), 0, s, fpart, nf, kpart, nk, 0.5 * desc->eps * dV,
                       0.5 / desc->mu * dV, out);
    PIC_CHECK_LAUNCH();
    return 0;
}

// Device address of `scratch` on the current device, looked up once per
// device with that device current (each GPU holds its own copy of the array
// at its own address; the array's name is not a device address in host code).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_address(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

int read_scalars(CGScalars* host, const CGScalars* dev, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(host, dev, sizeof(CGScalars), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return (int)e;
    e = hipStreamSynchronize(s);
    return e == hipSuccess ? 0 : (int)e;
}

}  // namespace

extern "C" {

int64_t pic_cg_workspace_bytes(const PicDesc* desc) {
    if (!desc_ok(desc)) return -1;
    return cg_layout(desc).total;
}

int pic_poisson_cg(const PicDesc* desc, const float* rho, float* phi, double tol, double rtol, int32_t max_iter,
                   void* workspace, int64_t workspace_bytes, int32_t* iterations, double* residual_norm,
                   void* stream) {
    PIC_REQUIRE(desc_ok(desc) && rho && phi && workspace && rho != phi && max_iter >= 0);
    PIC_REQUIRE(tol >= 0.0 && rtol >= 0.0);
    CGLayout Lw = cg_layout(desc);
    PIC_REQUIRE(workspace_bytes >= Lw.total);
    hipStream_t s = as_stream(stream);
    char* base = static_cast<char*>(workspace);
    double* phid = reinterpret_cast<double*>(base + Lw.phid_off);
    float* p = reinterpret_cast<float*>(base + Lw.p_off);
    float* r = reinterpret_cast<float*>(base + Lw.r_off);
    float* Ap = reinterpret_cast<float*>(base + Lw.ap_off);
    float* e = reinterpret_cast<float*>(base + Lw.e_off);
    double* partials = reinterpret_cast<double*>(base + Lw.part_off);
    CGScalars* sc = reinterpret_cast<CGScalars*>(base + Lw.sc_off);

    const LapMap M = lap_map(desc);
    const int64_t interior = (int64_t)desc->n[0] * desc->n[1] * desc->n[2];
    const int nb = partial_blocks(interior);
    const double inv_eps = 1.0 / desc->eps;

    int rc = phi_bc(desc, phi, s);          // phi = apply_bc(phi) (:139)
    if (rc) return rc;
    hipLaunchKernelGGL(cg_load_kernel, dim3(nb), dim3(kBlock), 0, s, *desc, phi, phid);
    PIC_CHECK_LAUNCH();

    CGScalars host;
    int total_iter = 0;
    double target = tol, prev = 0.0, last = 0.0;
    for (int outer = 0; outer < kMaxOuter; ++outer) {
        if (M.singular) {
            hipLaunchKernelGGL(cg_residual_sum_kernel, dim3(nb), dim3(kBlock), 0, s, *desc, M, rho, phid, partials,
                               inv_eps);
            hipLaunchKernelGGL(cg_mean_kernel, dim3(1), dim3(kBlock), 0, s, partials, nb, sc, 1.0 / (double)interior);
        }
        hipLaunchKernelGGL(cg_residual_kernel, dim3(nb), dim3(kBlock), 0, s, *desc, M, rho, phid, r, p, e, partials,
                           inv_eps, sc);
        hipLaunchKernelGGL(cg_start_kernel, dim3(1), dim3(kBlock), 0, s, partials, nb, sc, target,
                           (int)(max_iter - total_iter));
        PIC_CHECK_LAUNCH();
        rc = read_scalars(&host, sc, s);
        if (rc) return rc;
        last = sqrt(host.rr_true);
        if (outer == 0) target = fmax(tol, rtol * last);
        if (!(last > target) || total_iter >= max_iter) break;
        if (outer > 0 && !(last < 0.5 * prev)) break;    // at the FP64 rounding floor
        prev = last;
        // inner FP32 CG on -lapl e = r, e0 = 0, to max(target, kInnerRtol ||r||)
        hipLaunchKernelGGL(cg_start_kernel, dim3(1), dim3(kBlock), 0, s, partials, nb, sc, target,
                           (int)(max_iter - total_iter));
        PIC_CHECK_LAUNCH();
        for (;;) {
            for (int b = 0; b < kBatch; ++b) {
                hipLaunchKernelGGL(cg_ap_kernel, dim3(nb), dim3(kBlock), 0, s, M, p, Ap, partials, sc);
                hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(kBlock), 0, s, partials, nb, sc);
                hipLaunchKernelGGL(cg_update_kernel, dim3(nb), dim3(kBlock), 0, s, interior, e, r, p, Ap, partials,
                                   sc);
                hipLaunchKernelGGL(cg_beta_kernel, dim3(1), dim3(kBlock), 0, s, partials, nb, sc);
                hipLaunchKernelGGL(cg_p_kernel, dim3(nb), dim3(kBlock), 0, s, interior, r, p, sc);
                PIC_CHECK_LAUNCH();
            }
            rc = read_scalars(&host, sc, s);
            if (rc) return rc;
            if (host.done) break;
        }
        total_iter += host.iter;
        hipLaunchKernelGGL(cg_accumulate_kernel, dim3(nb), dim3(kBlock), 0, s, interior, phid, e);
        PIC_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(cg_store_kernel, dim3(nb), dim3(kBlock), 0, s, *desc, phid, phi);
    PIC_CHECK_LAUNCH();
    rc = phi_bc(desc, phi, s);
    if (rc) return rc;
    if (iterations) *iterations = total_iter;
    if (residual_norm) *residual_norm = last;
    return 0;
}

int64_t pic_energy_workspace_bytes(const PicDesc* desc, const PicSpecies* sp) {
    if (!desc_ok(desc) || !sp) return -1;
    return (int64_t)(2 * kMaxPartials + kMaxPartials) * 8;
}

int pic_energy(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, const float* const E[3],
               const float* const B[3], double* out, void* workspace, int64_t workspace_bytes, void* stream) {
    PIC_REQUIRE(desc_ok(desc) && sp && in && E && B && out && workspace);
    for (int a = 0; a < 3; ++a) PIC_REQUIRE(E[a] && B[a]);
    PIC_REQUIRE(particles_ok(sp, in));
    PIC_REQUIRE(workspace_bytes >= pic_energy_workspace_bytes(desc, sp));
    return energy_launch(desc, sp, in, E, B, out, static_cast<double*>(workspace), as_stream(stream));
}

int pic_energy_noalloc(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, const float* const E[3],
                       const float* const B[3], double* out, void* stream) {
    PIC_REQUIRE(desc_ok(desc) && sp && in && E && B && out);
    for (int a = 0; a < 3; ++a) PIC_REQUIRE(E[a] && B[a]);
    PIC_REQUIRE(particles_ok(sp, in));
    // the partial sums this call needs; never more than the scratch holds
    // (the partial count is capped at kMaxPartials per reduction)
    const int64_t need = pic_energy_workspace_bytes(desc, sp);
    if (need < 0 || need > (int64_t)sizeof(scratch)) {
        fprintf(stderr, "pic_energy_noalloc: needs %lld scratch bytes, the static scratch holds %zu\n",
                (long long)need, sizeof(scratch));
        return -2;
    }
    double* ws = nullptr;
    int rc = scratch_address(&ws);
    if (rc) return rc;
    return energy_launch(desc, sp, in, E, B, out, ws, as_stream(stream));
}

/* compute_total_momentum (utils.py:190-202), FP64 accumulation. */
int pic_total_momentum(const PicSpecies* sp, const PicParticles* in, double* out, void* workspace,
                       int64_t workspace_bytes, void* stream) {
    PIC_REQUIRE(sp && in && out && workspace);
    PIC_REQUIRE(particles_ok(sp, in));
    PIC_REQUIRE(workspace_bytes >= (int64_t)kMaxPartials * 8);
    hipStream_t s = as_stream(stream);
    double* part = static_cast<double*>(workspace);
    int64_t n = (int64_t)sp->count * sp->slots;
    int nk = partial_blocks(n);
    hipLaunchKernelGGL(momentum_kernel, dim3(nk), dim3(kBlock), 0, s, *sp, *in, part, n);
    PIC_CHECK_LAUNCH();
    hipLaunchKernelGGL(sum_finalize, dim3(1), dim3(kBlock), 0, s, part, nk, 1.0, out);
    PIC_CHECK_LAUNCH();
    return 0;
}

/* Total charge sum_interior rho dV, FP64 accumulation of the FP32 density. */
int pic_total_charge(const PicDesc* desc, const float* rho, double* out, void* workspace, int64_t workspace_bytes,
                     void* stream) {
    PIC_REQUIRE(desc_ok(desc) && rho && out && workspace);
    PIC_REQUIRE(workspace_bytes >= (int64_t)kMaxPartials * 8);
    hipStream_t s = as_stream(stream);
    double* part = static_cast<double*>(workspace);
    int nf = partial_blocks((int64_t)desc->n[0] * desc->n[1] * desc->n[2]);
    hipLaunchKernelGGL(charge_kernel, dim3(nf), dim3(kBlock), 0, s, *desc, rho, part);
    PIC_CHECK_LAUNCH();
    hipLaunchKernelGGL(sum_finalize, dim3(1), dim3(kBlock), 0, s, part, nf, desc->d[0] * desc->d[1] * desc->d[2],
                       out);
    PIC_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
