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
    fused_block<kDims, kRounds>(d, sp, in, fp, out, Jx, Jy, Jz, n, base, sh);
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
    } else if (kind == 2) {
        PIC_FUSED_LAUNCH(2, 4, 1);
    } else if (kind == 1) {
        PIC_FUSED_LAUNCH(1, 4, 1);
    } else {
        PIC_FUSED_LAUNCH(0, 4, 1);
    }
#undef PIC_FUSED_LAUNCH
    PIC_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
