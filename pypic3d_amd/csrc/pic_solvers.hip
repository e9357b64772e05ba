// Electrostatic Poisson CG with FP32 vectors and FP64 dot products, and the
// FP64 energy diagnostics.
//
// CG follows solve_poisson_with_conjugate_gradient (electrostatic_yee.py:71-156):
// r0 = rho/eps + lapl(phi), p0 = r0, alpha = r.r / p.(-lapl p), phi += alpha p,
// r -= alpha (-lapl p), beta = r'.r'/r.r, p = r' + beta p, with the phi/p
// boundary handling of _apply_tiled_phi_constant_boundaries (:26-43).  The
// scalars live on the device; every kernel returns at once when the device
// `done` flag is set, so the host can queue iterations in batches and read the
// flag once per batch.  Only phi's interior feeds later iterations, so its
// ghosts are applied once at the end (the reference re-applies them every
// iteration with no effect on the interior).
#include "pic_common.h"

int pic_refresh_fields(const PicDesc* d, float* f0, float* f1, float* f2, int bc_type, hipStream_t s);
extern "C" int pic_constant_boundary(const PicDesc* desc, float* field, int axis, void* stream);

namespace {

constexpr int kBlock = 256;
constexpr int kMaxPartials = 1024;
constexpr int kBatch = 32;

struct CGScalars {
    double rr, pAp, alpha, beta, tol2;
    int32_t done, iter, max_iter, pad;
};

__device__ inline void interior_coords(const PicDesc& d, int64_t t, int& i, int& j, int& k) {
    k = (int)(t % d.n[2]);
    int64_t r = t / d.n[2];
    j = (int)(r % d.n[1]);
    i = (int)(r / d.n[1]);
    i += d.g;
    j += d.g;
    k += d.g;
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

// -lapl f at interior cell c (electrostatic_yee.py:98-103), reading f's ghosts.
__device__ inline float neg_lapl(const float* __restrict__ f, int64_t c, int64_t sx, int64_t sy, float ix2, float iy2,
                                 float iz2) {
    float fc = f[c];
    float l = (f[c + sx] + f[c - sx] - 2.0f * fc) * ix2 + (f[c + sy] + f[c - sy] - 2.0f * fc) * iy2 +
              (f[c + 1] + f[c - 1] - 2.0f * fc) * iz2;
    return -l;
}

struct LapConst {
    int64_t sx, sy;
    float ix2, iy2, iz2;
};

__global__ void cg_init_kernel(PicDesc d, LapConst L, const float* __restrict__ rho, const float* __restrict__ phi,
                               float* __restrict__ r, float* __restrict__ p, double* partials, float inv_eps) {
    int64_t total = (int64_t)d.n[0] * d.n[1] * d.n[2];
    double acc = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_coords(d, t, i, j, k);
        int64_t c = fidx(d, i, j, k);
        float rv = rho[c] * inv_eps - neg_lapl(phi, c, L.sx, L.sy, L.ix2, L.iy2, L.iz2);
        r[c] = rv;
        p[c] = rv;
        acc += (double)rv * (double)rv;
    }
    double s = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void cg_init_finalize(const double* partials, int nparts, CGScalars* sc, double tol, double rtol,
                                 int max_iter) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) acc += partials[b];
    double rr = block_sum(acc);
    if (threadIdx.x == 0) {
        double t = fmax(tol, rtol * sqrt(rr));
        sc->rr = rr;
        sc->tol2 = t * t;
        sc->iter = 0;
        sc->max_iter = max_iter;
        sc->done = (max_iter <= 0 || !(rr > sc->tol2)) ? 1 : 0;
    }
}

__global__ void cg_ap_kernel(PicDesc d, LapConst L, const float* __restrict__ p, float* __restrict__ Ap,
                             double* partials, const CGScalars* sc) {
    if (sc->done) return;
    int64_t total = (int64_t)d.n[0] * d.n[1] * d.n[2];
    double acc = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_coords(d, t, i, j, k);
        int64_t c = fidx(d, i, j, k);
        float ap = neg_lapl(p, c, L.sx, L.sy, L.ix2, L.iy2, L.iz2);
        Ap[c] = ap;
        acc += (double)p[c] * (double)ap;
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

__global__ void cg_update_kernel(PicDesc d, float* __restrict__ phi, float* __restrict__ r, const float* __restrict__ p,
                                 const float* __restrict__ Ap, double* partials, const CGScalars* sc) {
    if (sc->done) return;
    float alpha = (float)sc->alpha;
    int64_t total = (int64_t)d.n[0] * d.n[1] * d.n[2];
    double acc = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_coords(d, t, i, j, k);
        int64_t c = fidx(d, i, j, k);
        phi[c] += alpha * p[c];
        float rn = r[c] - alpha * Ap[c];
        r[c] = rn;
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

__global__ void cg_p_kernel(PicDesc d, const float* __restrict__ r, float* __restrict__ p, const CGScalars* sc) {
    if (sc->done) return;
    float beta = (float)sc->beta;
    int64_t total = (int64_t)d.n[0] * d.n[1] * d.n[2];
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_coords(d, t, i, j, k);
        int64_t c = fidx(d, i, j, k);
        p[c] = r[c] + beta * p[c];
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
    int64_t field_bytes, p_off, r_off, ap_off, part_off, sc_off, total;
};

CGLayout cg_layout(const PicDesc* d) {
    CGLayout L;
    L.field_bytes = ((field_size(*d) * 4 + 255) / 256) * 256;
    L.p_off = 0;
    L.r_off = L.field_bytes;
    L.ap_off = 2 * L.field_bytes;
    L.part_off = 3 * L.field_bytes;
    L.sc_off = L.part_off + kMaxPartials * 8;
    L.total = L.sc_off + 256;
    return L;
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
    CGLayout Lw = cg_layout(desc);
    PIC_REQUIRE(workspace_bytes >= Lw.total);
    hipStream_t s = as_stream(stream);
    char* base = static_cast<char*>(workspace);
    float* p = reinterpret_cast<float*>(base + Lw.p_off);
    float* r = reinterpret_cast<float*>(base + Lw.r_off);
    float* Ap = reinterpret_cast<float*>(base + Lw.ap_off);
    double* partials = reinterpret_cast<double*>(base + Lw.part_off);
    CGScalars* sc = reinterpret_cast<CGScalars*>(base + Lw.sc_off);

    LapConst L;
    L.sx = (int64_t)desc->dims[1] * desc->dims[2];
    L.sy = desc->dims[2];
    L.ix2 = (float)(1.0 / (desc->d[0] * desc->d[0]));
    L.iy2 = (float)(1.0 / (desc->d[1] * desc->d[1]));
    L.iz2 = (float)(1.0 / (desc->d[2] * desc->d[2]));
    int64_t interior = (int64_t)desc->n[0] * desc->n[1] * desc->n[2];
    int nb = partial_blocks(interior);

    // r, p ghosts must hold defined values (p's are rewritten by phi_bc)
    int rc = pic_fill_f32(p, field_size(*desc), 0.0f, stream);
    if (rc) return rc;
    rc = phi_bc(desc, phi, s);
    if (rc) return rc;
    hipLaunchKernelGGL(cg_init_kernel, dim3(nb), dim3(kBlock), 0, s, *desc, L, rho, phi, r, p, partials,
                       (float)(1.0 / desc->eps));
    PIC_CHECK_LAUNCH();
    hipLaunchKernelGGL(cg_init_finalize, dim3(1), dim3(kBlock), 0, s, partials, nb, sc, tol, rtol, (int)max_iter);
    PIC_CHECK_LAUNCH();
    rc = phi_bc(desc, p, s);
    if (rc) return rc;

    CGScalars host;
    for (int it = 0; it < max_iter + kBatch; it += kBatch) {
        hipError_t e = hipMemcpyAsync(&host, sc, sizeof(CGScalars), hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return (int)e;
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return (int)e;
        if (host.done) break;
        for (int b = 0; b < kBatch; ++b) {
            hipLaunchKernelGGL(cg_ap_kernel, dim3(nb), dim3(kBlock), 0, s, *desc, L, p, Ap, partials, sc);
            hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(kBlock), 0, s, partials, nb, sc);
            hipLaunchKernelGGL(cg_update_kernel, dim3(nb), dim3(kBlock), 0, s, *desc, phi, r, p, Ap, partials, sc);
            hipLaunchKernelGGL(cg_beta_kernel, dim3(1), dim3(kBlock), 0, s, partials, nb, sc);
            hipLaunchKernelGGL(cg_p_kernel, dim3(nb), dim3(kBlock), 0, s, *desc, r, p, sc);
            PIC_CHECK_LAUNCH();
            rc = phi_bc(desc, p, s);
            if (rc) return rc;
        }
    }
    rc = phi_bc(desc, phi, s);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(&host, sc, sizeof(CGScalars), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return (int)e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return (int)e;
    if (iterations) *iterations = host.iter;
    if (residual_norm) *residual_norm = sqrt(host.rr);
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
    hipStream_t s = as_stream(stream);
    double* fpart = static_cast<double*>(workspace);
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

}  // extern "C"
