"""Launch-shape autotuner for the fused particle kernel.

The fused push + Esirkepov kernel is compiled in several launch shapes
(particle slots per thread x resident-block target for the register
allocator, selected by ``PicDesc.variant``).  ``tune_fused`` runs each shape
on the caller's actual particles and fields, times it with HIP events on the
stream the kernel runs on, and returns the fastest.  Outputs go to buffers the
caller is about to overwrite anyway (the next step's destination set and J).
"""

from __future__ import annotations

import torch

from . import _native as N

VARIANTS = {
    0: "4 slots/thread, compiler-chosen registers",
    1: "4 slots/thread, >=3 blocks/CU",
    2: "4 slots/thread, >=4 blocks/CU",
    3: "8 slots/thread, compiler-chosen registers",
    4: "2 slots/thread, compiler-chosen registers",
    5: "8 slots/thread, >=3 blocks/CU",
}


def tune_fused(desc, sp, pp_in, E3, B3, Ee, Be, pp_out, J3, reps=3, variants=None):
    """Returns (best_variant, {variant: mean ms})."""
    s = N.stream_handle()
    variants = list(VARIANTS) if variants is None else list(variants)
    saved = desc.variant
    times = {}
    for v in variants:
        desc.variant = v
        N.call("pic_fused_push_esirkepov", desc, sp, pp_in, N.p3(E3), N.p3(B3), Ee, Be, pp_out, N.p3(J3), s)
        evs = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            N.call("pic_fused_push_esirkepov", desc, sp, pp_in, N.p3(E3), N.p3(B3), Ee, Be, pp_out, N.p3(J3), s)
            b.record()
            evs.append((a, b))
        torch.cuda.synchronize()
        times[v] = sum(a.elapsed_time(b) for a, b in evs) / reps
    desc.variant = saved
    best = min(times, key=times.get)
    return best, times


def tune_tiled(desc, sp, pp_in, E3, B3, Ee, Be, pp_out, J3, tile, starts, workspace, reps=3, rounds=(2, 4, 8),
               min_blocks=(1, 3, 4)):
    """Launch shape of the tile-ordered kernel: ((rounds, min_blocks), {"rounds/min_blocks": mean ms})."""
    s = N.stream_handle()
    times = {}
    for r in rounds:
        for mb in min_blocks:
            args = (desc, sp, pp_in, N.p3(E3), N.p3(B3), Ee, Be, pp_out, N.p3(J3), tile, N.ptr(starts), r, mb,
                    N.ptr(workspace), s)
            N.call("pic_fused_push_esirkepov_tiled", *args)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(reps):
                N.call("pic_fused_push_esirkepov_tiled", *args)
            b.record()
            torch.cuda.synchronize()
            times[(r, mb)] = a.elapsed_time(b) / reps
    best = min(times, key=times.get)
    return best, {f"{r}/{mb}": t for (r, mb), t in times.items()}
