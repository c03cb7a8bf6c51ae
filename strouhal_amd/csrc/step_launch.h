// Launch of one AA pass of a single lattice: picks the kernel specialisation
// (parity, walls, partial cells, LES, forcing, porosity, row staging) from the
// lattice's contents. Each storage format is instantiated in its own
// translation unit (step_fp32.hip, step_fp16.hip, step_bf16.hip).
#pragma once

#include "device.h"

namespace strh {

template <class Conv, int PARITY, bool WALLS, bool NT, bool LES, bool FORCE, bool PLANAR>
static hipError_t launch_step_t(const StepLaunch& L, hipStream_t s) {
    using T = typename Conv::T;
    if (PARITY == 1 && !WALLS && L.rows_bs != 0 && !L.alpha) {
        using St = Store<Conv, kPolAligned>;
        StepParams p = L.p;
#define STRH_ROWS(BSV, LG) hipLaunchKernelGGL((step_odd_rows<St, BSV, NT, LES, FORCE, LG>), dim3(p.blocks), \
                                           dim3(BSV), 0, s, (T*)L.f, (float4*)L.force, L.eps, p)
        if (L.rows_bs == 512) {
            p.blocks = p.N / 512u;
            if (L.large) STRH_ROWS(512, true); else STRH_ROWS(512, false);
        } else {
            p.blocks = p.N / 256u;
            if (L.large) STRH_ROWS(256, true); else STRH_ROWS(256, false);
        }
#undef STRH_ROWS
        return hipGetLastError();
    }
    using St = Store<Conv, PARITY == 0 ? kPolAligned : kPolShifted>;
    StepParams p = L.p;
    p.blocks = (p.N + kStepBS - 1) / kStepBS;
    if (L.alpha) {   // Brinkman lattices: forcing path with per-cell force, no NT / LES
        if (NT || LES) return hipErrorNotSupported;
        hipLaunchKernelGGL((step_kernel<St, PARITY, WALLS, false, false, true, true, true, PLANAR>), dim3(p.blocks),
                           dim3(kStepBS), 0, s, (T*)L.f, L.flags, L.link, L.lidm,
                           (float4*)L.force, L.eps, p, L.alpha, (const uint32_t*)nullptr);
        return hipGetLastError();
    }
    // Steps that write no force run the kernel with the force output compiled
    // out (only walls and partial cells have one).
    const bool out = !((WALLS || NT) && p.writeForce == 0);
#define STRH_STEP(NTV, OUTV, LISTV, GRID, LISTP)                                                    \
    hipLaunchKernelGGL((step_kernel<St, PARITY, WALLS, NTV, LES, FORCE, false, OUTV, PLANAR, LISTV>), \
                       dim3(GRID), dim3(kStepBS), 0, s, (T*)L.f, L.flags, L.link, L.lidm,          \
                       (float4*)L.force, L.eps, p, (const float*)nullptr, (const uint32_t*)(LISTP))
    // The split pays in the even pass (LES sphere 0.474 -> 0.429 ms FP16,
    // 0.899 -> 0.880 FP32, 0.513 -> 0.460 BF16) and in the FP32 odd pass
    // (1.066 -> 1.021 ms); the 16-bit odd passes measured slower split
    // (0.698 -> 0.732 ms FP16, 0.772 -> 0.800 BF16), so they keep one kernel.
    if constexpr (NT && !PLANAR && (PARITY == 0 || sizeof(T) == 4)) {
        if (L.ntSplit) {
            // the blocks without partial cells take the kernel without their
            // path (fewer registers: more waves per SIMD); for cells with
            // eps = 0 both compute the same bits (the golden path)
            if (L.nPlain) { if (out) STRH_STEP(false, true, true, L.nPlain, L.ntSplit); else STRH_STEP(false, false, true, L.nPlain, L.ntSplit); }
            if (L.nNt) { if (out) STRH_STEP(true, true, true, L.nNt, L.ntSplit + L.nPlain); else STRH_STEP(true, false, true, L.nNt, L.ntSplit + L.nPlain); }
            return hipGetLastError();
        }
    }
    if (out) {
        STRH_STEP(NT, true, false, p.blocks, nullptr);
    } else {
        STRH_STEP(NT, false, false, p.blocks, nullptr);
    }
#undef STRH_STEP
    return hipGetLastError();
}

// The odd pass of a lattice one cell deep is instantiated with PLANAR (its z
// shifts wrap onto the cell's own plane); the even pass reads no neighbour.
template <class Conv, int PARITY, bool WALLS, bool NT, bool LES, bool FORCE>
static hipError_t pick_planar(const StepLaunch& L, hipStream_t s) {
    return (PARITY == 1 && L.planar) ? launch_step_t<Conv, PARITY, WALLS, NT, LES, FORCE, PARITY == 1>(L, s)
                                     : launch_step_t<Conv, PARITY, WALLS, NT, LES, FORCE, false>(L, s);
}
template <class Conv, int PARITY, bool WALLS, bool NT, bool LES>
static hipError_t pick_force(const StepLaunch& L, hipStream_t s) {
    return L.has_force ? pick_planar<Conv, PARITY, WALLS, NT, LES, true>(L, s)
                       : pick_planar<Conv, PARITY, WALLS, NT, LES, false>(L, s);
}
template <class Conv, int PARITY, bool WALLS, bool NT>
static hipError_t pick_les(const StepLaunch& L, hipStream_t s) {
    return L.les ? pick_force<Conv, PARITY, WALLS, NT, true>(L, s)
                 : pick_force<Conv, PARITY, WALLS, NT, false>(L, s);
}
template <class Conv, int PARITY, bool WALLS>
static hipError_t pick_nt(const StepLaunch& L, hipStream_t s) {
    return L.nt ? pick_les<Conv, PARITY, WALLS, true>(L, s) : pick_les<Conv, PARITY, WALLS, false>(L, s);
}
template <class Conv, int PARITY>
static hipError_t pick_walls(const StepLaunch& L, hipStream_t s) {
    return L.walls ? pick_nt<Conv, PARITY, true>(L, s) : pick_nt<Conv, PARITY, false>(L, s);
}
template <class Conv>
static hipError_t pick_parity(const StepLaunch& L, hipStream_t s) {
    return L.p.parity == 0 ? pick_walls<Conv, 0>(L, s) : pick_walls<Conv, 1>(L, s);
}


}  // namespace strh
