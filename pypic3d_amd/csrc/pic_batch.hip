// Batched ensembles: many independent simulations of uneven grid size and
// particle count advance together, one persistent launch per phase.
//
// Storage: every simulation's particle planes are a slice of one set of batch
// planes (PicBatchSim::particle_offset) and every field component a slice of
// one batch array (PicBatchSim::field_offset), each in the single-simulation
// layout.  The descriptor table lives in device memory.
//
// Particle phase: a work list of (simulation, block) items covers every
// simulation's slots in blocks of kThreads * rounds; a fixed grid of
// persistent blocks walks the list, and each item runs the same block
// function as the single-simulation kernel (fast paths chosen per item).
//
// Field phase: one block per simulation performs the whole grid part of
// time_loop_electrodynamic's Esirkepov branch (evolve.py:70-97): J fold +
// refresh with the particle codes, B half step, E step, B half step, each
// followed by its ghost refresh, with block barriers between the passes.
#include "pic_field_ops.h"
#include "pic_fused_block.h"

namespace {

constexpr int kFieldThreads = 1024;

struct BatchFields {
    const float* E[3];
    float* B[3];
    float* Bh[3];
    float* J[3];
    float* Eo[3];
};

__device__ inline FieldPtrs batch_field_ptrs(const float* const E[3], const float* const B[3], int64_t off) {
    FieldPtrs fp;
    for (int c = 0; c < 3; ++c) {
        fp.E[c] = E[c] + off;
        fp.B[c] = B[c] + off;
        fp.Ee[c] = nullptr;
        fp.Be[c] = nullptr;
    }
    return fp;
}

__device__ inline PicParticles offset_particles(const PicParticles& p, int64_t off) {
    PicParticles q;
    for (int a = 0; a < 3; ++a) {
        q.x[a] = p.x[a] + off;
        q.u[a] = p.u[a] + off;
    }
    q.active = p.active + off;
    return q;
}

struct BatchParticleArgs {
    PicParticles in, out;
    const float* E[3];
    const float* B[3];
    float* J[3];
};

// Each work item runs fused_block_boxed (J accumulated in an LDS box sized to
// the item's particles).
template <int kRounds>
__global__ __launch_bounds__(kThreads, 1) void batch_fused_kernel(const PicBatchSim* __restrict__ sims,
                                                                  const int32_t* __restrict__ work, int32_t n_work,
                                                                  BatchParticleArgs args) {
    __shared__ FusedShared<kRounds> sh;
    __shared__ float box[kRangeBoxCap];
    __shared__ int bmin[3], bmax[3];
    for (int32_t w = blockIdx.x; w < n_work; w += gridDim.x) {
        const int sim = work[2 * w];
        const uint32_t blk = (uint32_t)work[2 * w + 1];
        const PicBatchSim& bs = sims[sim];
        const PicDesc& d = bs.desc;
        const int64_t poff = bs.particle_offset, foff = bs.field_offset;
        const int64_t n = (int64_t)bs.species.count * bs.species.slots;
        const PicParticles in = offset_particles(args.in, poff);
        const PicParticles out = offset_particles(args.out, poff);
        const FieldPtrs fp = batch_field_ptrs(args.E, args.B, foff);
        float* const J[3] = {args.J[0] + foff, args.J[1] + foff, args.J[2] + foff};
        const uint32_t base = blk * (uint32_t)(kThreads * kRounds);
        fused_block_boxed<kRounds>(d, bs.species, in, fp, out, J, n, base, sh, box, bmin, bmax);
        __syncthreads();   // the next item reuses the shared queue and box
    }
}

__device__ inline void block_refresh3(const PicDesc& d, float* const f[3], int bc_type) {
    for (int axis = 0; axis < 3; ++axis) {
        const int mode = refresh_mode(d, axis, bc_type);
        const int64_t m = refresh_elems(d, axis);
        for (int64_t t = threadIdx.x; t < 3 * m; t += blockDim.x) refresh_elem(d, f[t / m], axis, mode, t % m);
        __syncthreads();
    }
}

__device__ inline void interior_ijk(const PicDesc& d, int64_t t, int& i, int& j, int& k) {
    k = (int)(t % d.n[2]) + d.g;
    int64_t r = t / d.n[2];
    j = (int)(r % d.n[1]) + d.g;
    i = (int)(r / d.n[1]) + d.g;
}

__global__ __launch_bounds__(kFieldThreads) void batch_field_kernel(const PicBatchSim* __restrict__ sims,
                                                                     int32_t n_sims, BatchFields bf) {
    for (int32_t sim = blockIdx.x; sim < n_sims; sim += gridDim.x) {
        const PicDesc& d = sims[sim].desc;
        const int64_t off = sims[sim].field_offset;
        const float* E[3];
        float *B[3], *Bh[3], *J[3], *Eo[3];
        for (int c = 0; c < 3; ++c) {
            E[c] = bf.E[c] + off;
            B[c] = bf.B[c] + off;
            Bh[c] = bf.Bh[c] + off;
            J[c] = bf.J[c] + off;
            Eo[c] = bf.Eo[c] + off;
        }
        // J: fold + refresh with the particle codes (Esirkepov.py:357-359)
        for (int axis = 0; axis < 3; ++axis) {
            const int bc = d.pbc[axis];
            const int64_t m = fold_elems(d, axis);
            for (int64_t t = threadIdx.x; t < 3 * m; t += blockDim.x) {
                if (d.n[axis] == 1) fold_reduced_elem(d, J[t / m], axis, bc, t % m);
                else fold_elem(d, J[t / m], axis, bc, t % m);
            }
            __syncthreads();
        }
        block_refresh3(d, J, 1);
        const YeeConst y = yee_const(d);
        const int64_t ni = (int64_t)d.n[0] * d.n[1] * d.n[2];
        int i, j, k;
        // B_h = B - (dt/2) curl E (evolve.py:88)
        for (int64_t t = threadIdx.x; t < ni; t += blockDim.x) {
            interior_ijk(d, t, i, j, k);
            yee_b_point(d, y, E, B, Bh, i, j, k);
        }
        __syncthreads();
        block_refresh3(d, Bh, 0);
        // E' = E + dt (C^2 curl B_h - J/eps) (evolve.py:92)
        for (int64_t t = threadIdx.x; t < ni; t += blockDim.x) {
            interior_ijk(d, t, i, j, k);
            yee_e_point(d, y, E, Bh, J, Eo, i, j, k);
        }
        __syncthreads();
        block_refresh3(d, Eo, 0);
        // B = B_h - (dt/2) curl E' (evolve.py:96)
        for (int64_t t = threadIdx.x; t < ni; t += blockDim.x) {
            interior_ijk(d, t, i, j, k);
            yee_b_point(d, y, Eo, Bh, B, i, j, k);
        }
        __syncthreads();
        block_refresh3(d, B, 0);
    }
}

bool batch_desc_ok(const PicDesc& d) {
    if (d.g < 1 || d.shape_factor < 1 || d.shape_factor > 2) return false;
    bool any = false;
    for (int a = 0; a < 3; ++a) {
        if (d.n[a] < 1 || d.dims[a] != d.n[a] + 2 * d.g) return false;
        if (d.n[a] > 1 && d.n[a] < d.g) return false;
        if (d.bc[a] != 0 || d.pbc[a] != 0) return false;   // periodic fields and particles
        any = any || d.n[a] > 1;
    }
    return any && d.alpha == 1.0;                           // digital filter is the identity
}

}  // namespace

extern "C" {

int pic_batch_check(const PicBatchSim* sims, int32_t n_sims, int64_t particle_slots, int64_t field_elems) {
    PIC_REQUIRE(sims && n_sims > 0);
    for (int32_t s = 0; s < n_sims; ++s) {
        const PicBatchSim& b = sims[s];
        PIC_REQUIRE(batch_desc_ok(b.desc));
        PIC_REQUIRE(b.species.count >= 1 && b.species.count <= PIC_MAX_SPECIES && b.species.slots >= 0);
        const int64_t n = (int64_t)b.species.count * b.species.slots;
        PIC_REQUIRE(b.particle_offset >= 0 && b.particle_offset + n <= particle_slots);
        PIC_REQUIRE(b.field_offset >= 0 && b.field_offset + field_size(b.desc) <= field_elems);
        PIC_REQUIRE(n < ((int64_t)1 << 31) - 8 * kThreads && field_size(b.desc) < ((int64_t)1 << 31));
    }
    return 0;
}

int pic_batch_fused_push_esirkepov(const PicBatchSim* sims, const int32_t* work, int32_t n_work, int32_t rounds,
                                   int32_t grid, const PicParticles* in, const float* const E[3],
                                   const float* const B[3], PicParticles* out, float* const J[3], void* stream) {
    PIC_REQUIRE(sims && work && n_work >= 0 && grid > 0 && in && out && E && B && J);
    PIC_REQUIRE(rounds == 2 || rounds == 4 || rounds == 8);
    for (int a = 0; a < 3; ++a) PIC_REQUIRE(E[a] && B[a] && J[a] && in->x[a] != out->x[a] && in->u[a] != out->u[a]);
    PIC_REQUIRE(in->active != out->active);
    if (n_work == 0) return 0;
    BatchParticleArgs args;
    args.in = *in;
    args.out = *out;
    for (int c = 0; c < 3; ++c) {
        args.E[c] = E[c];
        args.B[c] = B[c];
        args.J[c] = J[c];
    }
    hipStream_t st = as_stream(stream);
    const int g = grid < n_work ? grid : n_work;
    if (rounds == 2) {
        hipLaunchKernelGGL(batch_fused_kernel<2>, dim3(g), dim3(kThreads), 0, st, sims, work, n_work, args);
    } else if (rounds == 4) {
        hipLaunchKernelGGL(batch_fused_kernel<4>, dim3(g), dim3(kThreads), 0, st, sims, work, n_work, args);
    } else {
        hipLaunchKernelGGL(batch_fused_kernel<8>, dim3(g), dim3(kThreads), 0, st, sims, work, n_work, args);
    }
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_batch_field_step(const PicBatchSim* sims, int32_t n_sims, const float* const E_in[3], float* const B[3],
                         float* const B_half[3], float* const J[3], float* const E_out[3], void* stream) {
    PIC_REQUIRE(sims && n_sims > 0 && E_in && B && B_half && J && E_out);
    BatchFields bf;
    for (int c = 0; c < 3; ++c) {
        PIC_REQUIRE(E_in[c] && B[c] && B_half[c] && J[c] && E_out[c]);
        bf.E[c] = E_in[c];
        bf.B[c] = B[c];
        bf.Bh[c] = B_half[c];
        bf.J[c] = J[c];
        bf.Eo[c] = E_out[c];
    }
    for (int c = 0; c < 3; ++c)
        for (int e = 0; e < 3; ++e) PIC_REQUIRE((const float*)E_out[e] != E_in[c] && B_half[e] != B[c]);
    hipLaunchKernelGGL(batch_field_kernel, dim3(n_sims), dim3(kFieldThreads), 0, as_stream(stream), sims, n_sims,
                       bf);
    PIC_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
