// Exact grids: origins and spacings that make every node coordinate an FP32
// number, so the anchor of an FP32 position can be found in FP32 and agrees
// with the FP64 evaluation floor((x - origin)/spacing) bit for bit.
#pragma once
#include <cmath>

#include "pic_common.h"

namespace {

// Grid constants of an exact grid: on every axis the center line has an FP32
// origin and spacing, the vertex line is the center line shifted by half a
// spacing, and every node org + a*sp with a in [lo, hi + 1] (both lines) is an
// FP32 number (checked on the host).  For a position inside [xlo, xhi) the
// anchor of an FP32 position is then found in FP32 by bracketing it between
// exact node coordinates, and x - node is the correctly rounded difference --
// the value the FP64 evaluation rounds to -- so no FP64 arithmetic is needed
// for the old position.
struct ExactGrid {
    float org_c[3], sp_c[3], isp_c[3], half_sp[3];
    float inv_d[3];
    float xlo[3], xhi[3];
    double tol[3];      // band around a face inside which a new FP64 position takes the queue
    // FP32 form of the same test: an FP32 offset (x - node) + v dt in
    // [band_lo, band_hi) puts the FP64 new position in [node + tol, node + sp - tol)
    // (the FP32 offset is within 2^-23 sp of the FP64 one; the band is 2 tol wide)
    float band_lo[3], band_hi[3];
};

// Center-line anchor and offset of x (xlo <= x < xhi): the FP32 estimate is
// within one node of the anchor, one bracketing step against the exact nodes
// corrects it.  The vertex line follows from the half-cell test.
__device__ inline void exact_anchors(float x, float org, float sp, float isp, float half_sp, int& ac, float& dc,
                                     float& nodec, int& av, float& dv) {
    float f = floorf((x - org) * isp);
    float nd = fmaf(f, sp, org);
    const bool below = x < nd, above = x >= nd + sp;
    f = below ? f - 1.0f : (above ? f + 1.0f : f);
    nd = below ? nd - sp : (above ? nd + sp : nd);
    ac = (int)f;
    dc = x - nd;
    nodec = nd;
    const float mid = nd + half_sp;          // vertex node of index ac
    const bool upper = x >= mid;
    av = upper ? ac : ac - 1;
    dv = upper ? x - mid : x - (mid - sp);
}

// An empty asm that reads a slot's position and velocity: placed before the
// slot's activity test it keeps the compiler from moving these six loads below
// the test, so they are issued with the activity load and share its memory
// round trip instead of waiting for it.
__device__ inline void loads_together(const float (&x)[3], const float (&u)[3]) {
    asm volatile("" ::"v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(u[0]), "v"(u[1]), "v"(u[2]));
}

// Host check for ExactGrid (see its comment); false when any axis fails.
bool make_exact_grid(const PicDesc& d, ExactGrid& eg) {
    for (int a = 0; a < 3; ++a) {
        const double org = d.org_c[a], sp = d.sp_c[a];
        const float o32 = (float)org, s32 = (float)sp, h32 = 0.5f * s32;
        if ((double)o32 != org || (double)s32 != sp || !(sp > 0.0)) return false;
        if (d.sp_v[a] != sp || d.org_v[a] != org + 0.5 * sp || (double)(o32 + h32) != d.org_v[a]) return false;
        if (d.dims[a] > (1 << 20)) return false;
        const int lo = -d.dims[a] - 4, hi = 2 * d.dims[a] + 4;
        for (int n = lo - 1; n <= hi + 1; ++n) {
            const double node = (double)n * sp + org;                   // exact: 24-bit operands
            const float n32 = std::fma((float)n, s32, o32);
            if ((double)n32 != node || (double)(n32 + h32) != node + 0.5 * sp) return false;
        }
        eg.org_c[a] = o32;
        eg.sp_c[a] = s32;
        eg.isp_c[a] = (float)(1.0 / sp);
        eg.half_sp[a] = h32;
        eg.inv_d[a] = (float)(1.0 / d.d[a]);
        eg.xlo[a] = std::fma((float)lo, s32, o32);
        eg.xhi[a] = std::fma((float)hi, s32, o32);
        // floor((x - org)/sp) in FP64 can differ from the exact bracket only within a
        // few FP64 ulps of the quotient (|quotient| < 2^22 here): a band of 2^-20 sp is ample
        eg.tol[a] = std::ldexp(sp, -20);
        eg.band_lo[a] = (float)std::ldexp(sp, -19);                      // exact
        const double hd = sp - std::ldexp(sp, -19);
        float hf = (float)hd;
        if ((double)hf > hd) hf = std::nextafter(hf, 0.0f);                // rounded down
        eg.band_hi[a] = hf;
    }
    return true;
}

}  // namespace
