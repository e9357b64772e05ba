// Cell keys, stable cell sort and particle permutation.
//
// The key is the reference's flat index (ix*Ny + iy)*Nz + iz with
// ix = clip(floor((x + W/2)/dx), 0, Nx-1) (particle_initialization.py:89-102,
// particle_tile_communication.py:62-79), evaluated in FP64 from the FP32
// position so it matches an FP64 evaluation bit for bit.  Species are kept in
// their own segments and inactive slots sort to the end of their segment.  The
// permutation is rocPRIM's LSD radix sort of (key, slot) pairs, which is
// stable, so it equals np.argsort(key, kind="stable") (:159).
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "pic_common.h"

namespace {

constexpr int kBlock = 256;

__global__ void cell_keys_kernel(PicDesc d, PicSpecies sp, PicParticles in, uint32_t* keys, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = species_of(i, sp.slots, sp.count);
    int64_t ncell = (int64_t)d.n[0] * d.n[1] * d.n[2];
    int64_t key = ncell;
    if (in.active[i]) {
        int c[3];
        for (int a = 0; a < 3; ++a) {
            double t = floor(((double)in.x[a][i] + 0.5 * d.wind[a]) / d.d[a]);
            t = fmin(fmax(t, 0.0), (double)(d.n[a] - 1));      // NaN -> 0 via fmax
            c[a] = (int)t;
        }
        key = ((int64_t)c[0] * d.n[1] + c[1]) * d.n[2] + c[2];
    }
    keys[i] = (uint32_t)(s * (ncell + 1) + key);
}

__global__ void permute_kernel(PicParticles in, PicParticles out, const int32_t* __restrict__ perm, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t src = perm[i];
    if (src < 0 || src >= n) src = i;   // defensive: never read outside the planes
    for (int a = 0; a < 3; ++a) {
        out.x[a][i] = in.x[a][src];
        out.u[a][i] = in.u[a][src];
    }
    out.active[i] = in.active[src];
}

size_t sort_storage(int64_t n, int bits) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                              rocprim::counting_iterator<int32_t>(0), (int32_t*)nullptr, (size_t)n, 0u,
                              (unsigned)bits, (hipStream_t)0);
    return bytes;
}

}  // namespace

extern "C" {

int pic_cell_keys(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, int32_t* keys, void* stream) {
    PIC_REQUIRE(desc && sp && in && keys && in->active && in->x[0] && in->x[1] && in->x[2]);
    PIC_REQUIRE(sp->count >= 1 && sp->count <= PIC_MAX_SPECIES);
    int64_t ncell = (int64_t)desc->n[0] * desc->n[1] * desc->n[2];
    PIC_REQUIRE((int64_t)sp->count * (ncell + 1) < ((int64_t)1 << 31));
    int64_t n = (int64_t)sp->count * sp->slots;
    if (n == 0) return 0;
    hipLaunchKernelGGL(cell_keys_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *desc, *sp, *in,
                       reinterpret_cast<uint32_t*>(keys), n);
    PIC_CHECK_LAUNCH();
    return 0;
}

int64_t pic_sort_workspace_bytes(int64_t n) {
    if (n <= 0) return 0;
    return (int64_t)sort_storage(n, 32) + 256;
}

int pic_sort_permutation(const int32_t* keys, int32_t* keys_sorted, int32_t* perm, int64_t n, int32_t key_bits,
                         void* workspace, int64_t workspace_bytes, void* stream) {
    PIC_REQUIRE(keys && keys_sorted && perm && n >= 0 && key_bits >= 1 && key_bits <= 32);
    PIC_REQUIRE(n < ((int64_t)1 << 31));
    if (n == 0) return 0;
    size_t need = sort_storage(n, key_bits);
    PIC_REQUIRE(workspace != nullptr && (size_t)workspace_bytes >= need);
    size_t bytes = need;
    hipError_t e = rocprim::radix_sort_pairs(workspace, bytes, reinterpret_cast<const uint32_t*>(keys),
                                             reinterpret_cast<uint32_t*>(keys_sorted),
                                             rocprim::counting_iterator<int32_t>(0), perm, (size_t)n, 0u,
                                             (unsigned)key_bits, as_stream(stream));
    if (e != hipSuccess) return (int)e;
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_permute_particles(const PicParticles* in, PicParticles* out, const int32_t* perm, int64_t n, void* stream) {
    PIC_REQUIRE(in && out && perm && n >= 0);
    for (int a = 0; a < 3; ++a) PIC_REQUIRE(in->x[a] && in->u[a] && out->x[a] && out->u[a] && in->x[a] != out->x[a]);
    PIC_REQUIRE(in->active && out->active && in->active != out->active);
    if (n == 0) return 0;
    hipLaunchKernelGGL(permute_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *in, *out, perm,
                       n);
    PIC_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
