// Single-simulation launch of the fused particle phase (the block function
// and its design notes live in pic_fused_block.h).
#include "pic_fused_block.h"

namespace {

// kRounds: particle slots per thread (block = kThreads * kRounds slots);
// kMinBlocks: resident-block target handed to the register allocator.
template <int kDims, int kRounds, int kMinBlocks>
__global__ __launch_bounds__(kThreads, kMinBlocks) void fused_push_esirkepov_v2(PicDesc d, PicSpecies sp,
                                                                                 PicParticles in, FieldPtrs fp,
                                                                                 PicParticles out, float* Jx,
                                                                                 float* Jy, float* Jz, int64_t n) {
    __shared__ FusedShared<kRounds> sh;
    const uint32_t base = blockIdx.x * (uint32_t)(kThreads * kRounds);   // n < 2^31 is checked at launch
    const GlobalSink sink{{Jx, Jy, Jz}};
    fused_block<kDims, kRounds>(d, sp, in, fp, out, sink, n, base, sh);
}

// Tile-ordered launch (particles sorted by pic_tile_keys): block b owns the
// slot range [starts[b], starts[b+1]) of one (species, tile) pair, i.e. the
// particles that were in that T^3-cell tile at the last sort.  Run sums and
// face-crossing deposits accumulate in an LDS box of (T+4)^3 nodes per
// component (the tile, one node below and three above for drift and the
// crossing frame); deposits that fall outside it go straight to J.  The box is
// added into J once per block, so J atomics scale with the tile surface
// instead of with the number of runs.
struct TileGeom {
    int T;
    int tiles;          // tiles per species
    int nt[3];          // tiles per axis
};

template <int kRounds, int kT, int kMinBlocks>
__global__ __launch_bounds__(kThreads, kMinBlocks) void fused_tiled_kernel(PicDesc d, PicSpecies sp, PicParticles in,
                                                                  FieldPtrs fp, PicParticles out, float* Jx,
                                                                  float* Jy, float* Jz,
                                                                  const int32_t* __restrict__ starts, TileGeom tg) {
    constexpr int kB = kT + 4;
    constexpr int kBox = 3 * kB * kB * kB;
    constexpr int kFB = kT + 4;
    constexpr int kFV = kFB * kFB * kFB;
    __shared__ FusedShared<kRounds> sh;
    __shared__ float box[kBox];
    __shared__ float fbox[6 * kFV];
    const int b = blockIdx.x;
    const int64_t begin = starts[b], end = starts[b + 1];
    if (begin >= end) return;
    const int t = b % (tg.tiles + 1);
    int o[3] = {0, 0, 0}, fo[3] = {0, 0, 0};
    if (t < tg.tiles) {
        const int tx = t / (tg.nt[1] * tg.nt[2]), ty = (t / tg.nt[2]) % tg.nt[1], tz = t % tg.nt[2];
        const int c0[3] = {tx * kT, ty * kT, tz * kT};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            o[a] = c0[a] + d.g - 1;
            fo[a] = c0[a] + d.g - 2;
        }
    }
    for (int e = threadIdx.x; e < kBox; e += kThreads) box[e] = 0.0f;
    // E+E_ext, B+B_ext on the (T+4)^3 nodes from T+g-2 (gather anchors of the
    // tile's particles, drifted by up to a cell, and their +1 neighbours)
    for (int e = threadIdx.x; e < 6 * kFV; e += kThreads) {
        const int c = e / kFV, r = e % kFV;
        const int i = pmod(fo[0] + r / (kFB * kFB), d.dims[0]), j = pmod(fo[1] + (r / kFB) % kFB, d.dims[1]);
        const int k = pmod(fo[2] + r % kFB, d.dims[2]);
        const int64_t id = fidx(d, i, j, k);
        const float* F = c < 3 ? fp.E[c] : fp.B[c - 3];
        const float* Fe = c < 3 ? fp.Ee[c] : fp.Be[c - 3];
        float v = F[id];
        if (Fe) v += Fe[id];
        fbox[e] = v;
    }
    __syncthreads();
    const BoxSink sink{{Jx, Jy, Jz}, box, {o[0], o[1], o[2]}, kB};
    const FieldBox fb{fbox, {fo[0], fo[1], fo[2]}, kFB};
    for (int64_t base = begin; base < end; base += kThreads * kRounds) {
        fused_block<3, kRounds, BoxSink, true>(d, sp, in, fp, out, sink, end, (uint32_t)base, sh, fb);
        __syncthreads();
    }
    float* const J[3] = {Jx, Jy, Jz};
    for (int e = threadIdx.x; e < kBox; e += kThreads) {
        const float v = box[e];
        if (v == 0.0f) continue;
        const int c = e / (kB * kB * kB), r = e % (kB * kB * kB);
        deposit_at(J[c], d, o[0] + r / (kB * kB), o[1] + (r / kB) % kB, o[2] + r % kB, v);
    }
}

// Slot-ordered launch for 1-D and 2-D first-order runs (and the general
// path): each block accumulates J in an LDS box sized to its particles.
template <int kRounds>
__global__ __launch_bounds__(kThreads, 1) void fused_boxed_kernel(PicDesc d, PicSpecies sp, PicParticles in,
                                                                  FieldPtrs fp, PicParticles out, float* Jx,
                                                                  float* Jy, float* Jz, int64_t n) {
    __shared__ FusedShared<kRounds> sh;
    __shared__ float box[kRangeBoxCap];
    __shared__ int bmin[3], bmax[3];
    float* const J[3] = {Jx, Jy, Jz};
    fused_block_boxed<kRounds>(d, sp, in, fp, out, J, n, blockIdx.x * (uint32_t)(kThreads * kRounds), sh, box, bmin,
                               bmax);
}

}  // namespace

extern "C" {

int pic_fused_push_esirkepov(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                             const float* const E[3], const float* const B[3], const float* const E_ext[3],
                             const float* const B_ext[3], PicParticles* out, float* const J[3], void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok(sp, in) && particles_ok(sp, out) && E && B && J);
    PIC_REQUIRE(desc->n[0] > 1 || desc->n[1] > 1 || desc->n[2] > 1);
#pragma unroll
    for (int a = 0; a < 3; ++a) PIC_REQUIRE(E[a] && B[a] && J[a] && in->x[a] != out->x[a] && in->u[a] != out->u[a]);
    PIC_REQUIRE(in->active != out->active);
    PIC_REQUIRE(field_size(*desc) < ((int64_t)1 << 31));
    int64_t n = (int64_t)sp->count * sp->slots;
    PIC_REQUIRE(n < ((int64_t)1 << 31) - 8 * kThreads);
    if (n == 0) return 0;
    const int kind = fused_kind(*desc);
    hipStream_t st = as_stream(stream);
    FieldPtrs fp = make_fields(E, B, E_ext, B_ext);
#define PIC_FUSED_LAUNCH(DIMS, R, MB)                                                                           \
    hipLaunchKernelGGL((fused_push_esirkepov_v2<DIMS, R, MB>), dim3(grid_1d(n, kThreads * R)), dim3(kThreads), 0, \
                       st, *desc, *sp, *in, fp, *out, J[0], J[1], J[2], n)
    if (kind == 3) {
        switch (desc->variant) {
            case 1: PIC_FUSED_LAUNCH(3, 4, 3); break;
            case 2: PIC_FUSED_LAUNCH(3, 4, 4); break;
            case 3: PIC_FUSED_LAUNCH(3, 8, 1); break;
            case 4: PIC_FUSED_LAUNCH(3, 2, 1); break;
            case 5: PIC_FUSED_LAUNCH(3, 8, 3); break;
            default: PIC_FUSED_LAUNCH(3, 4, 1); break;
        }
    } else {
        hipLaunchKernelGGL(fused_boxed_kernel<4>, dim3(grid_1d(n, kThreads * 4)), dim3(kThreads), 0, st, *desc, *sp,
                           *in, fp, *out, J[0], J[1], J[2], n);
    }
#undef PIC_FUSED_LAUNCH
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_fused_push_esirkepov_tiled(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in,
                                   const float* const E[3], const float* const B[3], const float* const E_ext[3],
                                   const float* const B_ext[3], PicParticles* out, float* const J[3], int32_t tile,
                                   const int32_t* starts, int32_t rounds, int32_t min_blocks, void* stream) {
    PIC_REQUIRE(desc_ok(desc) && particles_ok(sp, in) && particles_ok(sp, out) && E && B && J && starts);
    PIC_REQUIRE(fused_kind(*desc) == 3 && (tile == 4 || tile == 8) && (rounds == 2 || rounds == 4));
    PIC_REQUIRE(min_blocks == 1 || min_blocks == 3 || min_blocks == 4);
#pragma unroll
    for (int a = 0; a < 3; ++a) PIC_REQUIRE(E[a] && B[a] && J[a] && in->x[a] != out->x[a] && in->u[a] != out->u[a]);
    PIC_REQUIRE(in->active != out->active);
    PIC_REQUIRE(field_size(*desc) < ((int64_t)1 << 31));
    int64_t n = (int64_t)sp->count * sp->slots;
    PIC_REQUIRE(n < ((int64_t)1 << 31) - 8 * kThreads);
    if (n == 0) return 0;
    TileGeom tg;
    tg.T = tile;
    for (int a = 0; a < 3; ++a) tg.nt[a] = (desc->n[a] + tile - 1) / tile;
    tg.tiles = tg.nt[0] * tg.nt[1] * tg.nt[2];
    const int64_t blocks = (int64_t)sp->count * (tg.tiles + 1);
    PIC_REQUIRE(blocks < ((int64_t)1 << 31));
    hipStream_t st = as_stream(stream);
    FieldPtrs fp = make_fields(E, B, E_ext, B_ext);
#define PIC_TILED_LAUNCH(R, T, MB)                                                                              \
    hipLaunchKernelGGL((fused_tiled_kernel<R, T, MB>), dim3((unsigned)blocks), dim3(kThreads), 0, st, *desc, *sp, \
                       *in, fp, *out, J[0], J[1], J[2], starts, tg)
#define PIC_TILED_MB(R, T)                              \
    do {                                                \
        if (min_blocks == 3) PIC_TILED_LAUNCH(R, T, 3); \
        else if (min_blocks == 4) PIC_TILED_LAUNCH(R, T, 4); \
        else PIC_TILED_LAUNCH(R, T, 1);                 \
    } while (0)
    if (tile == 8) {
        if (rounds == 2) PIC_TILED_MB(2, 8);
        else PIC_TILED_MB(4, 8);
    } else {
        if (rounds == 2) PIC_TILED_MB(2, 4);
        else PIC_TILED_MB(4, 4);
    }
#undef PIC_TILED_MB
#undef PIC_TILED_LAUNCH
    PIC_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
