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

// Tile-major cell key: cubic tiles of T cells per axis, tiles in x-major
// order, cells inside a tile in x-major order; inactive slots get the key
// after the last tile.  Same cell formula as cell_keys_kernel.
__host__ __device__ inline void tile_counts(const PicDesc& d, int T, int nt[3]) {
    for (int a = 0; a < 3; ++a) nt[a] = (d.n[a] + T - 1) / T;
}

// Per-axis constants of the tile key: W/2 and 1/d in FP64, the tile edge as a
// shift when it is a power of two (else -1).
struct TileKeyConst {
    double half_w[3], inv_d[3];
    int nt[3], shift;
};

__global__ void tile_keys_kernel(PicDesc d, PicSpecies sp, PicParticles in, int T, TileKeyConst kc, uint32_t* keys,
                                 int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t u = (uint32_t)i;     // n < 2^31 is checked at launch
    int s = species_of(i, sp.slots, sp.count);
    const int64_t vol = (int64_t)T * T * T;
    const int64_t span = (int64_t)kc.nt[0] * kc.nt[1] * kc.nt[2] * vol;
    int64_t key = span;
    const bool act = in.active[u] != 0;
    const float x[3] = {in.x[0][u], in.x[1][u], in.x[2][u]};
    if (act) {
        int c[3], tq[3], tr[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double t = floor(((double)x[a] + kc.half_w[a]) * kc.inv_d[a]);
            t = fmin(fmax(t, 0.0), (double)(d.n[a] - 1));
            c[a] = (int)t;
            tq[a] = kc.shift >= 0 ? c[a] >> kc.shift : c[a] / T;
            tr[a] = kc.shift >= 0 ? c[a] & (T - 1) : c[a] % T;
        }
        const int64_t tile = ((int64_t)tq[0] * kc.nt[1] + tq[1]) * kc.nt[2] + tq[2];
        const int64_t local = ((int64_t)tr[0] * T + tr[1]) * T + tr[2];
        key = tile * vol + local;
    }
    keys[u] = (uint32_t)(s * (span + 1) + key);
}

// starts[b] = first sorted position whose key >= the first key of tile b,
// b = s * (tiles + 1) + t (t = tiles: the inactive tail); starts[last] = n.
__global__ void tile_ranges_kernel(const uint32_t* __restrict__ keys, int64_t n, int64_t span, int64_t vol,
                                   int32_t tiles, int32_t species, int32_t* starts) {
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nb = (int64_t)species * (tiles + 1);
    if (b > nb) return;
    if (b == nb) {
        starts[b] = (int32_t)n;
        return;
    }
    const int64_t s = b / (tiles + 1), t = b % (tiles + 1);
    const uint64_t target = (uint64_t)(s * (span + 1) + t * vol);
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if ((uint64_t)keys[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    starts[b] = (int32_t)lo;
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

int64_t pic_tile_count(const PicDesc* desc, int32_t tile) {
    if (!desc || tile < 1) return -1;
    int nt[3];
    tile_counts(*desc, tile, nt);
    return (int64_t)nt[0] * nt[1] * nt[2];
}

int pic_tile_keys(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, int32_t tile, int32_t* keys,
                  void* stream) {
    PIC_REQUIRE(desc && keys && tile >= 1 && particles_ok(sp, in));
    const int64_t span = pic_tile_count(desc, tile) * tile * tile * tile;
    PIC_REQUIRE((int64_t)sp->count * (span + 1) < ((int64_t)1 << 31));
    int64_t n = (int64_t)sp->count * sp->slots;
    PIC_REQUIRE(n < ((int64_t)1 << 31));
    if (n == 0) return 0;
    TileKeyConst kc;
    tile_counts(*desc, tile, kc.nt);
    kc.shift = -1;
    for (int b = 0; b < 31; ++b)
        if (tile == (1 << b)) kc.shift = b;
    for (int a = 0; a < 3; ++a) {
        kc.half_w[a] = 0.5 * desc->wind[a];
        kc.inv_d[a] = 1.0 / desc->d[a];
    }
    hipLaunchKernelGGL(tile_keys_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *desc, *sp, *in,
                       (int)tile, kc, reinterpret_cast<uint32_t*>(keys), n);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_tile_ranges(const PicDesc* desc, const PicSpecies* sp, int32_t tile, const int32_t* keys_sorted,
                    int32_t* starts, void* stream) {
    PIC_REQUIRE(desc && sp && starts && tile >= 1 && sp->count >= 0 && sp->count <= PIC_MAX_SPECIES);
    PIC_REQUIRE(sp->slots >= 0 && (keys_sorted || (int64_t)sp->count * sp->slots == 0));
    const int64_t tiles = pic_tile_count(desc, tile);
    const int64_t vol = (int64_t)tile * tile * tile;
    const int64_t n = (int64_t)sp->count * sp->slots;
    PIC_REQUIRE(n < ((int64_t)1 << 31) && (int64_t)sp->count * (tiles * vol + 1) < ((int64_t)1 << 31));
    const int64_t nb = (int64_t)sp->count * (tiles + 1) + 1;
    hipLaunchKernelGGL(tile_ranges_kernel, dim3(grid_1d(nb, kBlock)), dim3(kBlock), 0, as_stream(stream),
                       reinterpret_cast<const uint32_t*>(keys_sorted), n, tiles * vol, vol, (int32_t)tiles,
                       (int32_t)sp->count, starts);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_cell_keys(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, int32_t* keys, void* stream) {
    PIC_REQUIRE(desc && keys && particles_ok(sp, in));
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
    PIC_REQUIRE(n >= 0 && key_bits >= 1 && key_bits <= 32);
    PIC_REQUIRE(n < ((int64_t)1 << 31));
    if (n == 0) return 0;
    PIC_REQUIRE(keys && keys_sorted && perm);
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
    PIC_REQUIRE(in && out && n >= 0);
    if (n == 0) return 0;
    PIC_REQUIRE(perm);
    for (int a = 0; a < 3; ++a) PIC_REQUIRE(in->x[a] && in->u[a] && out->x[a] && out->u[a] && in->x[a] != out->x[a]);
    PIC_REQUIRE(in->active && out->active && in->active != out->active);
    PIC_REQUIRE(n < ((int64_t)1 << 31));
    hipLaunchKernelGGL(permute_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *in, *out, perm,
                       n);
    PIC_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
