// D3Q19 lattice-Boltzmann device code for gfx950 (MI355X): what the kernel
// translation units share (storage formats, index helpers, the collision,
// the one-cell AA pass, the step kernels and the LDS row-staged odd pass).
//
// The hot path is `step_kernel`: one AA-pattern time step, collide and stream
// fused, in place (the reference's `step` + `collide`,
// Sources/StrouhalCore/Kernels.metal:105-355). Design for CDNA4:
//  * Blocked SoA DDF layout (DdfLayout in lbm.h): slot i of a block of W cells
//    is one contiguous run, so every wave's load of one slot is 64
//    consecutive values (256 B fp32 / 128 B fp16) and the even pass is
//    perfectly coalesced. The odd pass shifts x by +-1 cell: wall-free
//    lattices stage the shift through LDS by whole x-rows, wall lattices read
//    the shifted runs directly.
//  * One thread per cell, 256-thread workgroups in plain block order, so all
//    XCDs stream neighbouring parts of the lattice at any moment.
//  * Specialised per (storage, parity, walls, NT, LES, forcing): the all-fluid
//    periodic bulk reads and writes exactly 19 + 19 DDFs per cell and
//    nothing else (152 B fp32 / 76 B fp16 per cell).
//  * FP32 arithmetic in the reference's operation order, no contraction
//    (built with -ffp-contract=off); the only fused multiply-add is the one
//    the reference writes (Kernels.metal:181). Terms the reference multiplies
//    by a zero lattice-velocity component are dropped at compile time; the
//    partial sums they would be added to can never be -0 (each starts as
//    +0 + x), so dropping them leaves every result bit unchanged.
//  * No atomics anywhere; reductions run in a fixed order.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"

namespace strh {

// ---- storage -------------------------------------------------------------
// Conversion between the stored DDF and the FP32 arithmetic value
// ((float)f[...] and (FPXX)fh[i] in Kernels.metal:254-257).
struct ConvF32 {
    using T = float;
    __device__ static float ld(T v) { return v; }
    __device__ static T st(float v) { return v; }
};
struct ConvF16 {
    using T = uint16_t;   // binary16 bits
    __device__ static float ld(T v) { return (float)__builtin_bit_cast(_Float16, v); }  // exact
    // IEEE RNE of the binary32 value. The empty asm pins v as a rounded FP32
    // register: without it the compiler fuses the producing fmaf and this
    // conversion into one v_fma_mixlo_f16 (a single rounding to f16), which
    // differs from the reference's (half)(float) double rounding.
    __device__ static T st(float v) {
        asm("" : "+v"(v));
        return __builtin_bit_cast(uint16_t, (_Float16)v);
    }
};
struct ConvBF16 {
    using T = uint16_t;   // bfloat16 bits
    __device__ static float ld(T v) { return __uint_as_float((uint32_t)v << 16); }
    __device__ static T st(float v) {                          // IEEE RNE
        uint32_t u = __float_as_uint(v);
        if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x0040u);
        u += 0x7fffu + ((u >> 16) & 1u);
        return (uint16_t)(u >> 16);
    }
};

// DDF memory access at a 32-bit byte offset from a slot base (a uniform
// pointer), so loads and stores use the scalar-base + 32-bit vector-offset
// addressing form. POL bit 0: nontemporal loads, bit 1: nontemporal stores
// (streaming data touched once per pass; keeps it from displacing lines
// that still have a reader in the same pass).
// The DDFs are always in global memory: the accesses go through global
// address-space pointers, so kernels that read their base pointers from a
// descriptor in memory (the batch) get global instead of FLAT instructions.
#define GLOBAL __attribute__((address_space(1)))
template <class Conv, int POL> struct Store : Conv {
    using T = typename Conv::T;
    __device__ static float rd(const T* base, uint32_t off) {
        const GLOBAL T* p = (const GLOBAL T*)((const char*)base + off);
        return Conv::ld((POL & 1) ? __builtin_nontemporal_load(p) : *p);
    }
    __device__ static void wr(T* base, uint32_t off, float v) { wr_raw(base, off, Conv::st(v)); }
    __device__ static T rd_raw(const T* base, uint32_t off) {
        const GLOBAL T* p = (const GLOBAL T*)((const char*)base + off);
        return (POL & 1) ? __builtin_nontemporal_load(p) : *p;
    }
    __device__ static void wr_raw(T* base, uint32_t off, T t) {
        GLOBAL T* p = (GLOBAL T*)((char*)base + off);
        if (POL & 2) __builtin_nontemporal_store(t, p); else *p = t;
    }
};

// Access policies per pass (measured on MI355X at 256^3): the even pass and
// the row-staged odd pass touch whole aligned lines and run fastest fully
// nontemporal; the direct odd pass writes x-shifted partial lines, which must
// merge in L2, so it keeps the default policy.
constexpr int kPolAligned = 3;
constexpr int kPolShifted = 0;
using StF32 = Store<ConvF32, kPolAligned>;
using StF16 = Store<ConvF16, kPolAligned>;
using StBF16 = Store<ConvBF16, kPolAligned>;

// ---- index helpers ---------------------------------------------------------
__device__ __forceinline__ uint32_t fdiv(uint32_t n, FastDiv d) {
    return (__umulhi(n, d.m) + n) >> d.s;
}

// Element offset of slot 0 of cell n (DdfLayout in lbm.h); slot i adds i * W.
// (n >> lw) * 19 W + (n & mask) = n + 18 (n & ~mask): one AND and two
// shift-adds instead of a quarter-rate 32-bit multiply (the odd pass computes
// 36 of these per cell). A single-block layout has ~mask = 0.
__device__ __forceinline__ uint32_t cell_elem(const DdfLayout& l, uint32_t n) {
    const uint32_t u = n & ~l.mask;
    return n + (u << 4) + (u << 1);
}

// Side arrays (adjoint state, recordings) at a 32-bit byte offset from a
// uniform base, the scalar-base + vector-offset form: one VGPR per access
// instead of a 64-bit address. Offsets stay below 4 GiB for every lattice of
// at most 200M cells (two slots of N floats, or N float4).
template <class T> __device__ __forceinline__ T ld_off(const T* base, uint32_t off) {
    return *(const T*)((const char*)base + off);
}
template <class T> __device__ __forceinline__ void st_off(T* base, uint32_t off, T v) {
    *(T*)((char*)base + off) = v;
}

template <int C> __device__ __forceinline__ int pick(int m, int p) {
    return C < 0 ? m : (C > 0 ? p : 0);
}

// s + c*v for c in {-1, 0, 1}, as the reference's `s += (float)C * v` when
// s can never be -0 (dropping +-0 terms and writing (-1)*v as a subtraction
// are then exact).
template <int C> __device__ __forceinline__ float acc(float s, float v) {
    return C > 0 ? s + v : (C < 0 ? s - v : s);
}

// (cx*ux + cy*uy) + cz*uz with the zero terms dropped (c.u of direction i).
template <int I> __device__ __forceinline__ float cdot(float ux, float uy, float uz) {
    constexpr int a = kCx[I], b = kCy[I], c = kCz[I];
    float s;
    if (a != 0) {
        s = a > 0 ? ux : -ux;
        s = acc<b>(s, uy);
        s = acc<c>(s, uz);
    } else if (b != 0) {
        s = b > 0 ? uy : -uy;
        s = acc<c>(s, uz);
    } else {
        s = c > 0 ? uz : -uz;
    }
    return s;
}

template <int I> __device__ __forceinline__ constexpr float wgt() {
    return I == 0 ? (1.0f / 3.0f) : (I <= 6 ? (1.0f / 18.0f) : (1.0f / 36.0f));
}

// Moving-wall velocity, Kernels.metal:81-87.
__device__ __forceinline__ void wall_vel(uint8_t wf, int wy, const StepParams& p,
                                         float& ux, float& uy, float& uz) {
    if (wf == kLid) { ux = p.ulidX; uy = p.ulidY; uz = p.ulidZ; return; }
    uy = 0.0f; uz = 0.0f;
    if (p.inflowUniform != 0) { ux = p.uin; return; }
    const float H = (float)p.ny - 2.0f;
    const float yd = (float)wy - 0.5f;
    const float v = 4.0f * p.uin * yd * (H - yd) / (H * H);
    ux = v > 0.0f ? v : 0.0f;
}

// ---- collide: Kernels.metal:105-228 -------------------------------------
// Non-equilibrium stress Pi_ab = sum_i c_ia c_ib (f_i - feq_i), Kernels.metal:140-158.
__device__ __forceinline__ void neq_stress(const float (&fh)[19], float rhom1, float rho, float ux,
                                           float uy, float uz, float u2, float& pxx, float& pyy,
                                           float& pzz, float& pxy, float& pxz, float& pyz) {
    pxx = 0.0f; pyy = 0.0f; pzz = 0.0f; pxy = 0.0f; pxz = 0.0f; pyz = 0.0f;
#define STRH_LES_TERM(I)                                                              \
    {                                                                                 \
        const float cu = 3.0f * cdot<I>(ux, uy, uz);                                  \
        const float feq = wgt<I>() * (rhom1 + rho * (cu + 0.5f * cu * cu - 1.5f * u2)); \
        const float d = fh[I] - feq;                                                  \
        pxx = acc<kCx[I] * kCx[I]>(pxx, d);                                           \
        pyy = acc<kCy[I] * kCy[I]>(pyy, d);                                           \
        pzz = acc<kCz[I] * kCz[I]>(pzz, d);                                           \
        pxy = acc<kCx[I] * kCy[I]>(pxy, d);                                           \
        pxz = acc<kCx[I] * kCz[I]>(pxz, d);                                           \
        pyz = acc<kCy[I] * kCz[I]>(pyz, d);                                           \
    }
    STRH_LES_TERM(1) STRH_LES_TERM(2) STRH_LES_TERM(3) STRH_LES_TERM(4) STRH_LES_TERM(5)
    STRH_LES_TERM(6) STRH_LES_TERM(7) STRH_LES_TERM(8) STRH_LES_TERM(9) STRH_LES_TERM(10)
    STRH_LES_TERM(11) STRH_LES_TERM(12) STRH_LES_TERM(13) STRH_LES_TERM(14) STRH_LES_TERM(15)
    STRH_LES_TERM(16) STRH_LES_TERM(17) STRH_LES_TERM(18)
#undef STRH_LES_TERM
}

template <bool LES> __device__ __forceinline__ void les_rates(
    const float (&fh)[19], const StepParams& p, float rhom1, float rho, float inv,
    float ux, float uy, float uz, float u2, float& wp, float& wm) {
    float pxx, pyy, pzz, pxy, pxz, pyz;
    neq_stress(fh, rhom1, rho, ux, uy, uz, u2, pxx, pyy, pzz, pxy, pxz, pyz);
    const float Q = sqrtf(pxx * pxx + pyy * pyy + pzz * pzz
                          + 2.0f * (pxy * pxy + pxz * pxz + pyz * pyz));
    const float tau0 = 1.0f / wp;
    const float tauEff = 0.5f * (tau0 + sqrtf(tau0 * tau0 + 25.455844f * p.cSmago * Q * inv));
    if (p.omegaMinus == p.omega) {
        wp = 1.0f / tauEff;
        wm = wp;
    } else {
        const float lambda = (1.0f / wp - 0.5f) * (1.0f / wm - 0.5f);
        wp = 1.0f / tauEff;
        wm = 1.0f / (0.5f + lambda / (tauEff - 0.5f));
    }
}

// One symmetric/antisymmetric pair (I, I+1) of the eps == 0 path
// (Kernels.metal:183-197).
template <int I, bool FORCE> __device__ __forceinline__ void pair_golden(
    float (&fh)[19], const float* cFv, const float* threeCFv, float rhom1, float rho, float ux,
    float uy, float uz, float u2, float wp, float wm, float ap, float am, float uF) {
    constexpr float W = wgt<I>();
    const float cu = 3.0f * cdot<I>(ux, uy, uz);
    const float feqp = W * (rhom1 + rho * (0.5f * cu * cu - 1.5f * u2));
    const float feqm = W * rho * cu;
    const float fp = 0.5f * (fh[I] + fh[I + 1]);
    const float fm = 0.5f * (fh[I] - fh[I + 1]);
    const float dp = wp * (feqp - fp);
    const float dm = wm * (feqm - fm);
    if (FORCE) {
        const float sm = W * am * threeCFv[I >> 1];
        const float sp = W * ap * (3.0f * cu * cFv[I >> 1] - 3.0f * uF);
        fh[I] += dp + dm + sp + sm;
        fh[I + 1] += dp - dm + sp - sm;
    } else {
        fh[I] += dp + dm;
        fh[I + 1] += dp - dm;
    }
}

// One pair of the Noble-Torczynski path (Kernels.metal:210-227).
template <int I, bool FORCE> __device__ __forceinline__ void pair_nt(
    float (&fh)[19], const float* cFv, const float* threeCFv, float rhom1, float rho, float ux,
    float uy, float uz, float u2, float wp, float wm, float ap, float am, float uF, float B,
    float oneMinusB, float& ntx, float& nty, float& ntz) {
    constexpr float W = wgt<I>();
    const float cu = 3.0f * cdot<I>(ux, uy, uz);
    const float feqp = W * (rhom1 + rho * (0.5f * cu * cu - 1.5f * u2));
    const float feqm = W * rho * cu;
    const float fp = 0.5f * (fh[I] + fh[I + 1]);
    const float fm = 0.5f * (fh[I] - fh[I + 1]);
    const float dp = wp * (feqp - fp);
    const float dm = wm * (feqm - fm);
    const float bounce = B * (fh[I + 1] - fh[I]);
    if (FORCE) {
        const float sm = W * am * threeCFv[I >> 1];
        const float sp = W * ap * (3.0f * cu * cFv[I >> 1] - 3.0f * uF);
        fh[I] += oneMinusB * (dp + dm + sp + sm) + bounce;
        fh[I + 1] += oneMinusB * (dp - dm + sp - sm) - bounce;
    } else {
        fh[I] += oneMinusB * (dp + dm) + bounce;
        fh[I + 1] += oneMinusB * (dp - dm) - bounce;
    }
    // ntF -= 2 c_i bounce; (2*(-1))*b == -(2*b) exactly.
    const float tb = 2.0f * bounce;
    ntx = acc<-kCx[I]>(ntx, tb);
    nty = acc<-kCy[I]>(nty, tb);
    ntz = acc<-kCz[I]>(ntz, tb);
}

// BRINK: Brinkman porosity alpha (no reference counterpart): drag force
// -alpha rho u with u taken implicitly, u = (p + Fb/2) / (rho (1 + alpha/2)); the Guo terms then see the per-cell
// force F = Fb - alpha rho u. With alpha = 0 every value is the same bits as
// the plain forcing path.
template <bool LES, bool FORCE, bool NT, bool BRINK = false>
__device__ __forceinline__ void collide(float (&fh)[19], const StepParams& p, int x,
                                        float eps, float& ntx, float& nty, float& ntz,
                                        float alpha = 0.0f) {
    float wp = p.omega, wm = p.omegaMinus;
    if ((float)x > p.spongeX0) {                                   // 112-123
        const float frac = fminf(((float)x - p.spongeX0) * p.spongeInvW, 1.0f);
        const float tau0 = 1.0f / wp;
        const float tauS = tau0 + frac * (p.spongeTau - tau0);
        if (p.omegaMinus == p.omega) {
            wp = 1.0f / tauS;
            wm = wp;
        } else {
            const float lambda = (1.0f / wp - 0.5f) * (1.0f / wm - 0.5f);
            wp = 1.0f / tauS;
            wm = 1.0f / (0.5f + lambda / (tauS - 0.5f));
        }
    }
    float rhom1 = 0.0f;                                            // 124-128
#pragma unroll
    for (int i = 0; i < 19; i++) rhom1 += fh[i];
    const float rho = 1.0f + rhom1;
    const float inv = 1.0f / rho;
    float px = 0.0f, py = 0.0f, pz = 0.0f;                         // 129-134
#define STRH_MOM(I)                        \
    px = acc<kCx[I]>(px, fh[I]);           \
    py = acc<kCy[I]>(py, fh[I]);           \
    pz = acc<kCz[I]>(pz, fh[I]);
    STRH_MOM(1) STRH_MOM(2) STRH_MOM(3) STRH_MOM(4) STRH_MOM(5) STRH_MOM(6)
    STRH_MOM(7) STRH_MOM(8) STRH_MOM(9) STRH_MOM(10) STRH_MOM(11) STRH_MOM(12)
    STRH_MOM(13) STRH_MOM(14) STRH_MOM(15) STRH_MOM(16) STRH_MOM(17) STRH_MOM(18)
#undef STRH_MOM
    float ux, uy, uz;                                              // 135-138
    if (FORCE) {
        ux = (px + 0.5f * p.fx) * inv;
        uy = (py + 0.5f * p.fy) * inv;
        uz = (pz + 0.5f * p.fz) * inv;
    } else {  // px + 0.5f*0 == px: px is never -0
        ux = px * inv;
        uy = py * inv;
        uz = pz * inv;
    }
    float Fx = p.fx, Fy = p.fy, Fz = p.fz;
    float cFl[9], tCFl[9];
    const float* cFv = p.cF;
    const float* tCFv = p.threeCF;
    if (BRINK) {
        const float g = 1.0f + 0.5f * alpha;
        ux = ux / g;
        uy = uy / g;
        uz = uz / g;
        const float ar = alpha * rho;
        Fx = p.fx - ar * ux;
        Fy = p.fy - ar * uy;
        Fz = p.fz - ar * uz;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const int i = 2 * k + 1;
            cFl[k] = (float)kCx[i] * Fx + (float)kCy[i] * Fy + (float)kCz[i] * Fz;
            tCFl[k] = 3.0f * cFl[k];
        }
        cFv = cFl;
        tCFv = tCFl;
    }
    const float u2 = ux * ux + uy * uy + uz * uz;
    if (LES) {                                                     // 140-167
        if (p.cSmago > 0.0f) les_rates<LES>(fh, p, rhom1, rho, inv, ux, uy, uz, u2, wp, wm);
    }
    float uF = 0.0f, ap = 0.0f, am = 0.0f;                         // 169-171
    if (FORCE) {
        uF = ux * Fx + uy * Fy + uz * Fz;
        ap = 1.0f - 0.5f * wp;
        am = 1.0f - 0.5f * wm;
    }
    const float feq0 = (1.0f / 3.0f) * (rhom1 - 1.5f * rho * u2);
    if (!NT || eps == 0.0f) {                                      // 177-199
        if (FORCE) {
            const float s0 = (1.0f / 3.0f) * ap * (-3.0f * uF);
            fh[0] = fmaf(wp, feq0 - fh[0], fh[0]) + s0;
        } else {
            fh[0] = fmaf(wp, feq0 - fh[0], fh[0]);
        }
#define STRH_PAIR(I) pair_golden<I, FORCE>(fh, cFv, tCFv, rhom1, rho, ux, uy, uz, u2, wp, wm, ap, am, uF);
        STRH_PAIR(1) STRH_PAIR(3) STRH_PAIR(5) STRH_PAIR(7) STRH_PAIR(9)
        STRH_PAIR(11) STRH_PAIR(13) STRH_PAIR(15) STRH_PAIR(17)
#undef STRH_PAIR
        return;
    }
    if (NT) {                                                      // 201-227
        const float tw = 1.0f / wp - 0.5f;
        const float B = eps * tw / ((1.0f - eps) + tw);
        const float oneMinusB = 1.0f - B;
        if (FORCE) {
            const float s0 = (1.0f / 3.0f) * ap * (-3.0f * uF);
            fh[0] += oneMinusB * (wp * (feq0 - fh[0]) + s0);
        } else {
            fh[0] += oneMinusB * (wp * (feq0 - fh[0]));
        }
#define STRH_PAIR(I) pair_nt<I, FORCE>(fh, cFv, tCFv, rhom1, rho, ux, uy, uz, u2, wp, wm, ap, am, uF, B, oneMinusB, ntx, nty, ntz);
        STRH_PAIR(1) STRH_PAIR(3) STRH_PAIR(5) STRH_PAIR(7) STRH_PAIR(9)
        STRH_PAIR(11) STRH_PAIR(13) STRH_PAIR(15) STRH_PAIR(17)
#undef STRH_PAIR
    }
}

// ---- the fused step: Kernels.metal:230-355 ------------------------------
struct Offsets {   // cell-index offsets to the wrapped neighbour in each axis direction
    int xm, xp, ym, yp, zm, zp;
    template <int CX, int CY, int CZ> __device__ __forceinline__ int at() const {
        return pick<CX>(xm, xp) + pick<CY>(ym, yp) + pick<CZ>(zm, zp);
    }
};


// Adjoint hooks of the forward pass (AdjHooks in kernels.h). The collision's
// Jacobian depends on the state only through rho and p, so the reverse step
// needs just these four moments of the pre-collision populations (summed in
// FP64, stored as FP32: rho - 1 keeps the density's small part exact to the
// float's 24 bits). The hook also accumulates the cell's objective term,
//   cp(x) rho + wd alpha J_x,   J_x = (p_x + F_x / 2) / (1 + alpha / 2),
// with cp = +cin on the inlet column and -cout on the outlet column.
// The rate w+ a cell collides with (sponge ramp, then Smagorinsky), for the
// hooks' partial-cell drag; the same float arithmetic as collide.
template <bool LES, bool FORCE>
__device__ __forceinline__ float hook_rate(const float (&fh)[19], const StepParams& p, int x) {
    float wp = p.omega, wm = p.omegaMinus;
    if ((float)x > p.spongeX0) {
        const float frac = fminf(((float)x - p.spongeX0) * p.spongeInvW, 1.0f);
        const float tau0 = 1.0f / wp;
        const float tauS = tau0 + frac * (p.spongeTau - tau0);
        if (p.omegaMinus == p.omega) {
            wp = 1.0f / tauS;
            wm = wp;
        } else {
            const float lambda = (1.0f / wp - 0.5f) * (1.0f / wm - 0.5f);
            wp = 1.0f / tauS;
            wm = 1.0f / (0.5f + lambda / (tauS - 0.5f));
        }
    }
    if (LES && p.cSmago > 0.0f) {
        float rhom1 = 0.0f, px = 0.0f, py = 0.0f, pz = 0.0f;
#pragma unroll
        for (int i = 0; i < 19; i++) {
            rhom1 += fh[i];
            px += (float)kCx[i] * fh[i];
            py += (float)kCy[i] * fh[i];
            pz += (float)kCz[i] * fh[i];
        }
        const float rho = 1.0f + rhom1, inv = 1.0f / rho;
        // the velocity exactly as collide forms it for this kernel
        const float ux = (FORCE ? px + 0.5f * p.fx : px) * inv;
        const float uy = (FORCE ? py + 0.5f * p.fy : py) * inv;
        const float uz = (FORCE ? pz + 0.5f * p.fz : pz) * inv;
        les_rates<true>(fh, p, rhom1, rho, inv, ux, uy, uz, ux * ux + uy * uy + uz * uz, wp, wm);
    }
    return wp;
}

template <bool LES, bool FORCE>
__device__ __forceinline__ void run_hooks(const AdjHooks& hk, const float (&fh)[19], uint32_t n,
                                          int x, float alpha, float eps, const StepParams& p) {
    const float fx = p.fx;
    if (hk.recNt && eps > 0.0f) {
        const int32_t k = hk.ntIdx[n];
#pragma unroll
        for (int i = 0; i < 19; i++) hk.recNt[(size_t)i * hk.M + k] = fh[i];
    }
    if (hk.jacc && hk.wf != 0.0 && eps > 0.0f) {   // partial-cell drag ntF_x = -2 sum c_x B (g_opp - g)
        const double tw = 1.0 / (double)hook_rate<LES, FORCE>(fh, p, x) - 0.5, e = (double)eps;
        const double B = e * tw / ((1.0 - e) + tw);
        double nt = 0.0;
#pragma unroll
        for (int i = 1; i < 19; i += 2) nt -= 2.0 * (double)kCx[i] * B * ((double)fh[i + 1] - (double)fh[i]);
        hk.jacc[n] += hk.wf * nt;
    }
    // the FP64 moments serve the recording and the density / porous-drag terms
    const double cp = (x == hk.inletX ? hk.cin : 0.0) - (x == hk.outletX ? hk.cout : 0.0);
    const bool obj = hk.jacc && (cp != 0.0 || (hk.wd != 0.0 && alpha != 0.0f));
    if (!hk.rec && !obj) return;
    double rm1 = 0.0, px = 0.0, py = 0.0, pz = 0.0;
#pragma unroll
    for (int i = 0; i < 19; i++) {
        const double v = (double)fh[i];
        rm1 += v;
        if (kCx[i]) px += (double)kCx[i] * v;
        if (kCy[i]) py += (double)kCy[i] * v;
        if (kCz[i]) pz += (double)kCz[i] * v;
    }
    if (hk.rec) {
        typedef float f4v __attribute__((ext_vector_type(4)));
        const f4v m = {(float)rm1, (float)px, (float)py, (float)pz};
        __builtin_nontemporal_store(m, reinterpret_cast<f4v*>(hk.rec) + n);
    }
    if (obj) {
        const double a = (double)alpha;
        const double Jx = (px + 0.5 * (double)fx) / (1.0 + 0.5 * a);
        hk.jacc[n] += cp * (1.0 + rm1) + hk.wd * a * Jx;
    }
}

// LES recording: the 19 pre-collision populations of cell n in the storage
// format (exact: each is a stored value before the odd pass's moving-wall
// corrections, which the reverse step re-applies).
template <class St>
__device__ __forceinline__ void record_populations(const AdjHooks& hk, const float (&fh)[19], uint32_t N,
                                                   uint32_t n) {
    using T = typename St::T;
    T* r = (T*)hk.rec19;
#pragma unroll
    for (int i = 0; i < 19; i++)
        __builtin_nontemporal_store(St::st(fh[i]), (T*)((char*)(r + (size_t)i * N) + n * (uint32_t)sizeof(T)));
}

// One cell of one AA pass (the body shared by the single-lattice kernel and
// the multi-lattice batch kernel).
template <class St, int PARITY, bool WALLS, bool NT, bool LES, bool FORCE, bool BRINK = false,
          bool HOOK = false, bool OUT = true, bool PLANAR = false>
__device__ __forceinline__ void step_cell(typename St::T* __restrict__ f,
                                          const uint8_t* __restrict__ flags_,
                                          const uint32_t* __restrict__ link_,
                                          const uint32_t* __restrict__ lidm_,
                                          float4* __restrict__ force,
                                          const float* __restrict__ epsBuf_,
                                          const StepParams& p, uint32_t n,
                                          const float* __restrict__ alphaBuf_ = nullptr,
                                          const AdjHooks* hk = nullptr) {
    // side arrays through global pointers (see Store)
    const GLOBAL uint8_t* __restrict__ flags = (const GLOBAL uint8_t*)flags_;
    const GLOBAL uint32_t* __restrict__ link = (const GLOBAL uint32_t*)link_;
    const GLOBAL uint32_t* __restrict__ lidm = (const GLOBAL uint32_t*)lidm_;
    const GLOBAL float* __restrict__ epsBuf = (const GLOBAL float*)epsBuf_;
    const GLOBAL float* __restrict__ alphaBuf = (const GLOBAL float*)alphaBuf_;
    using T = typename St::T;
    // OUT = false compiles the per-cell force output out (steps that write no
    // force): its six accumulators then do not stay live across the
    // collision, which lowers the pass's VGPRs and raises its occupancy.
    const bool writes = OUT && p.writeForce != 0;
    T* fs[19];
#pragma unroll
    for (int i = 0; i < 19; i++) fs[i] = f + (size_t)i * p.lay.W;
    constexpr int TB = (int)sizeof(T);
    const uint32_t nb = cell_elem(p.lay, n) * (uint32_t)TB;   // byte offset of cell n from the slot bases

    float eps = 0.0f;
    float fh[19];
    float ntx = 0.0f, nty = 0.0f, ntz = 0.0f;

    if (PARITY == 0) {                                             // 252-260
        if (WALLS && flags[n] != kFluid) return;
        if (NT && p.useEps) eps = epsBuf[n];
        const uint32_t t = fdiv(n, p.divx);
        const int x = (int)(n - t * (uint32_t)p.nx);
        // A wave's 64 cells lie in one block of the layout: a uniform block
        // base and a 32-bit offset inside the block reach any image size.
        {
            const uint32_t bk = __builtin_amdgcn_readfirstlane(n >> p.lay.lw);
            T* fb = f + (size_t)bk * p.lay.bstride;
#pragma unroll
            for (int i = 0; i < 19; i++) fs[i] = fb + (size_t)i * p.lay.W;
        }
        const uint32_t nbi = (n & p.lay.mask) * (uint32_t)TB;
#pragma unroll
        for (int i = 0; i < 19; i++) fh[i] = St::rd(fs[i], nbi);
        const float alpha = BRINK ? alphaBuf[n] : 0.0f;
        if (HOOK && hk->rec19) record_populations<St>(*hk, fh, p.N, n);
        if (HOOK) run_hooks<LES, FORCE>(*hk, fh, n, x, alpha, eps, p);
        collide<LES, FORCE, NT, BRINK>(fh, p, x, eps, ntx, nty, ntz, alpha);
        // A separate copy of the offset keeps the stores on the scalar-base
        // form too (otherwise the shared base+offset sum is kept as 19
        // 64-bit VGPR addresses across the collision).
        uint32_t nbs = nbi;
        asm("" : "+v"(nbs));
        St::wr(fs[0], nbs, fh[0]);
#pragma unroll
        for (int i = 1; i < 19; i++) St::wr(fs[kOpp[i]], nbs, fh[i]);
        if (NT && writes && eps > 0.0f) force[n] = make_float4(ntx, nty, ntz, 0.0f);
        return;
    }

    // -------- odd: streaming pass (neighbours + walls), 261-354 --------
    uint32_t sM = 0, lM = 0;
    if (WALLS) {
        const uint32_t lw = link[n];
        if (lw & kNotFluidBit) return;
        sM = lw & 0x3ffffu;
        if (sM != 0u) lM = lidm[n];
    }
    if (NT && p.useEps) eps = epsBuf[n];
    // PLANAR: the lattice is one cell deep (nz = 1), so the z shift of every
    // direction wraps onto the cell's own plane: the 18 neighbours are 8
    // distinct cells (and the cell itself), whose offsets are computed once.
    const uint32_t t = fdiv(n, p.divx);
    const int x = (int)(n - t * (uint32_t)p.nx);
    const uint32_t z = PLANAR ? 0u : fdiv(t, p.divy);
    const int y = (int)(t - z * (uint32_t)p.ny);
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int plane = nx * ny;
    Offsets o;
    o.xm = (x == 0) ? nx - 1 : -1;
    o.xp = (x == nx - 1) ? 1 - nx : 1;
    o.ym = (y == 0) ? (ny - 1) * nx : -nx;
    o.yp = (y == ny - 1) ? -(ny - 1) * nx : nx;
    o.zm = PLANAR ? 0 : ((int)z == 0) ? (nz - 1) * plane : -plane;
    o.zp = PLANAR ? 0 : ((int)z == nz - 1) ? -(nz - 1) * plane : plane;
    const int ni = (int)n;
    // byte offset of the neighbour n + d from the slot bases
    auto nbo = [&](int d) -> uint32_t { return cell_elem(p.lay, (uint32_t)(ni + d)) * (uint32_t)TB; };

    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    uint32_t deferred = 0u;
    // Slots I and opp(I) are adjacent (kOpp pairs 2k-1, 2k), so both reads of
    // direction I are addressed from the lower slot of the pair: the wall read
    // (own cell, slot I) and the pull (neighbour, slot opp(I)) differ only in
    // a per-lane 32-bit offset. The load is then unconditional and the 18
    // pulls issue back to back instead of each waiting inside its own branch.
    // Offsets stay below 4 GiB: a blocked image is under 4 GiB, a single-block
    // one spans at most two slots of N values here.
    const uint32_t WB = p.lay.W * (uint32_t)TB;
    fh[0] = St::rd(fs[0], nb);
#define STRH_PULL(I)                                                                  \
    {                                                                                 \
        constexpr int srcBit = kOpp[I] - 1;                                           \
        constexpr int lo = I < kOpp[I] ? I : kOpp[I];                                 \
        const uint32_t pull = nbo(o.at<-kCx[I], -kCy[I], -kCz[I]>());                 \
        uint32_t off = pull + (kOpp[I] == lo ? 0u : WB);                              \
        if (WALLS && ((sM >> srcBit) & 1u)) off = nb + (I == lo ? 0u : WB);          \
        fh[I] = St::rd(fs[lo], off);                                                  \
    }
    STRH_PULL(1) STRH_PULL(2) STRH_PULL(3) STRH_PULL(4) STRH_PULL(5) STRH_PULL(6)
    STRH_PULL(7) STRH_PULL(8) STRH_PULL(9) STRH_PULL(10) STRH_PULL(11) STRH_PULL(12)
    STRH_PULL(13) STRH_PULL(14) STRH_PULL(15) STRH_PULL(16) STRH_PULL(17) STRH_PULL(18)
#undef STRH_PULL
    if (HOOK && hk->rec19) record_populations<St>(*hk, fh, p.N, n);
    if (WALLS && sM != 0u) {   // wall links: moving-wall corrections to apply, force terms
#define STRH_WALLIN(I)                                                                \
        {                                                                             \
            constexpr int srcBit = kOpp[I] - 1;                                       \
            if ((sM >> srcBit) & 1u) {                                                \
                if ((lM >> srcBit) & 1u) deferred |= 1u << I;                         \
                if (writes) {                                                         \
                    ax = acc<-kCx[I]>(ax, fh[I]);                                     \
                    ay = acc<-kCy[I]>(ay, fh[I]);                                     \
                    az = acc<-kCz[I]>(az, fh[I]);                                     \
                }                                                                     \
            }                                                                         \
        }
        STRH_WALLIN(1) STRH_WALLIN(2) STRH_WALLIN(3) STRH_WALLIN(4) STRH_WALLIN(5) STRH_WALLIN(6)
        STRH_WALLIN(7) STRH_WALLIN(8) STRH_WALLIN(9) STRH_WALLIN(10) STRH_WALLIN(11) STRH_WALLIN(12)
        STRH_WALLIN(13) STRH_WALLIN(14) STRH_WALLIN(15) STRH_WALLIN(16) STRH_WALLIN(17) STRH_WALLIN(18)
#undef STRH_WALLIN
    }

    if (WALLS && deferred != 0u) {                                 // 292-314
#define STRH_DEFER(I)                                                                 \
        if ((deferred >> I) & 1u) {                                                   \
            const int sy = y - kCy[I] < 0 ? ny - 1 : (y - kCy[I] >= ny ? 0 : y - kCy[I]); \
            const uint32_t s = (uint32_t)(ni + o.at<-kCx[I], -kCy[I], -kCz[I]>());      \
            const uint8_t wf = flags[s];                                              \
            if (wf == kOutflow) {                                                     \
                fh[I] = -fh[I];                                                       \
            } else {                                                                  \
                float uwx, uwy, uwz;                                                  \
                wall_vel(wf, sy, p, uwx, uwy, uwz);                                   \
                const float cu = (float)kCx[I] * uwx + (float)kCy[I] * uwy + (float)kCz[I] * uwz; \
                const float corr = 6.0f * wgt<I>() * cu;                              \
                fh[I] += corr;                                                        \
                if (writes) {                                                         \
                    ax = acc<-kCx[I]>(ax, corr);                                      \
                    ay = acc<-kCy[I]>(ay, corr);                                      \
                    az = acc<-kCz[I]>(az, corr);                                      \
                }                                                                     \
            }                                                                         \
        }
        STRH_DEFER(1) STRH_DEFER(2) STRH_DEFER(3) STRH_DEFER(4) STRH_DEFER(5) STRH_DEFER(6)
        STRH_DEFER(7) STRH_DEFER(8) STRH_DEFER(9) STRH_DEFER(10) STRH_DEFER(11) STRH_DEFER(12)
        STRH_DEFER(13) STRH_DEFER(14) STRH_DEFER(15) STRH_DEFER(16) STRH_DEFER(17) STRH_DEFER(18)
#undef STRH_DEFER
    }

    const float alpha = BRINK ? alphaBuf[n] : 0.0f;
    if (HOOK) run_hooks<LES, FORCE>(*hk, fh, n, x, alpha, eps, p);
    collide<LES, FORCE, NT, BRINK>(fh, p, x, eps, ntx, nty, ntz, alpha);  // 315-316
    if (NT) {
        ax += ntx;
        ay += nty;
        az += ntz;
    }
    uint32_t nbs = nb;                                             // 317-350
    asm("" : "+v"(nbs));
    // The push recomputes the neighbour offsets from a pinned copy of n
    // rather than keeping the pull's 18 offsets live across the collision.
    int nip = ni;
    asm("" : "+v"(nip));
    auto nbp = [&](int d) -> uint32_t { return cell_elem(p.lay, (uint32_t)(nip + d)) * (uint32_t)TB; };
    // (the coordinates and the six axis offsets are recomputed from it too)
    const uint32_t tp = fdiv((uint32_t)nip, p.divx);
    const int xp = nip - (int)(tp * (uint32_t)nx);
    const uint32_t zp = PLANAR ? 0u : fdiv(tp, p.divy);
    const int yp = (int)(tp - zp * (uint32_t)ny);
    Offsets op;
    op.xm = (xp == 0) ? nx - 1 : -1;
    op.xp = (xp == nx - 1) ? 1 - nx : 1;
    op.ym = (yp == 0) ? (ny - 1) * nx : -nx;
    op.yp = (yp == ny - 1) ? -(ny - 1) * nx : nx;
    op.zm = PLANAR ? 0 : ((int)zp == 0) ? (nz - 1) * plane : -plane;
    op.zp = PLANAR ? 0 : ((int)zp == nz - 1) ? -(nz - 1) * plane : plane;
    // ... and the slot bases from a pinned copy of f, so the pull's 19 base
    // pointers (38 SGPRs) are not held across the collision either
    T* fpin = f;
    asm("" : "+s"(fpin));
#pragma unroll
    for (int i = 0; i < 19; i++) fs[i] = fpin + (size_t)i * p.lay.W;
    St::wr(fs[0], nbs, fh[0]);
    // Wall links store sign * f - corr into the cell's own opposite slot (the
    // halfway bounce-back, with the moving-wall term or the outflow's sign
    // flip); with sign 1 and corr 0 that is f itself, bit for bit. The
    // corrections are computed in a block only wall cells enter; the store
    // is addressed like the pull, from the lower slot of the pair.
    if (WALLS && sM != 0u) {
#define STRH_WALLOUT(I)                                                               \
        {                                                                             \
            constexpr int dstBit = I - 1;                                             \
            if ((sM >> dstBit) & 1u) {                                                \
                float corr = 0.0f;                                                    \
                float sign = 1.0f;                                                    \
                if ((lM >> dstBit) & 1u) {                                            \
                    const int dy = yp + kCy[I] < 0 ? ny - 1 : (yp + kCy[I] >= ny ? 0 : yp + kCy[I]); \
                    const uint32_t d = (uint32_t)(nip + op.at<kCx[I], kCy[I], kCz[I]>()); \
                    const uint8_t wf = flags[d];                                      \
                    if (wf == kOutflow) {                                             \
                        sign = -1.0f;                                                 \
                    } else {                                                          \
                        float uwx, uwy, uwz;                                          \
                        wall_vel(wf, dy, p, uwx, uwy, uwz);                           \
                        corr = 6.0f * wgt<I>()                                        \
                             * ((float)kCx[I] * uwx + (float)kCy[I] * uwy + (float)kCz[I] * uwz); \
                    }                                                                 \
                }                                                                     \
                if (writes) {                                                         \
                    const float v = sign * fh[I] + 2.0f * wgt<I>();                   \
                    ax = acc<kCx[I]>(ax, v);                                          \
                    ay = acc<kCy[I]>(ay, v);                                          \
                    az = acc<kCz[I]>(az, v);                                          \
                }                                                                     \
                fh[I] = sign * fh[I] - corr;                                          \
            }                                                                         \
        }
        STRH_WALLOUT(1) STRH_WALLOUT(2) STRH_WALLOUT(3) STRH_WALLOUT(4) STRH_WALLOUT(5) STRH_WALLOUT(6)
        STRH_WALLOUT(7) STRH_WALLOUT(8) STRH_WALLOUT(9) STRH_WALLOUT(10) STRH_WALLOUT(11) STRH_WALLOUT(12)
        STRH_WALLOUT(13) STRH_WALLOUT(14) STRH_WALLOUT(15) STRH_WALLOUT(16) STRH_WALLOUT(17) STRH_WALLOUT(18)
#undef STRH_WALLOUT
    }
#define STRH_PUSH(I)                                                                  \
    {                                                                                 \
        constexpr int dstBit = I - 1;                                                 \
        constexpr int lo = I < kOpp[I] ? I : kOpp[I];                                 \
        uint32_t off = nbp(op.at<kCx[I], kCy[I], kCz[I]>()) + (I == lo ? 0u : WB);     \
        if (WALLS && ((sM >> dstBit) & 1u)) off = nbs + (kOpp[I] == lo ? 0u : WB);    \
        St::wr(fs[lo], off, fh[I]);                                                   \
    }
    STRH_PUSH(1) STRH_PUSH(2) STRH_PUSH(3) STRH_PUSH(4) STRH_PUSH(5) STRH_PUSH(6)
    STRH_PUSH(7) STRH_PUSH(8) STRH_PUSH(9) STRH_PUSH(10) STRH_PUSH(11) STRH_PUSH(12)
    STRH_PUSH(13) STRH_PUSH(14) STRH_PUSH(15) STRH_PUSH(16) STRH_PUSH(17) STRH_PUSH(18)
#undef STRH_PUSH
    if (writes && (sM != 0u || eps > 0.0f))                        // 351-353
        force[n] = make_float4(ax, ay, az, 0.0f);
}

// LIST: workgroup i takes block blockList[i] of kStepBS cells (a lattice's
// blocks split by whether they hold partial cells, each part with its own
// instantiation).
template <class St, int PARITY, bool WALLS, bool NT, bool LES, bool FORCE, bool BRINK = false,
          bool OUT = true, bool PLANAR = false, bool LIST = false>
__global__ void __launch_bounds__(kStepBS)
step_kernel(typename St::T* __restrict__ f, const uint8_t* __restrict__ flags,
            const uint32_t* __restrict__ link, const uint32_t* __restrict__ lidm,
            float4* __restrict__ force, const float* __restrict__ epsBuf, StepParams p,
            const float* __restrict__ alphaBuf, const uint32_t* __restrict__ blockList) {
    const uint32_t n = (LIST ? blockList[blockIdx.x] : blockIdx.x) * kStepBS + threadIdx.x;
    if (n >= p.N) return;
    step_cell<St, PARITY, WALLS, NT, LES, FORCE, BRINK, false, OUT, PLANAR>(f, flags, link, lidm, force,
                                                                            epsBuf, p, n, alphaBuf);
}

// Forward step with the adjoint hooks (Guo forcing with Brinkman porosity, or
// partial cells and / or LES). The arithmetic is that of step_kernel: with
// alpha = 0 and no body force the DDFs are the plain update's bits.
template <class St, int PARITY, bool WALLS, bool NT, bool LES, bool BRINK, bool FORCE, bool PLANAR = false>
__global__ void __launch_bounds__(kStepBS)
step_kernel_hooked(typename St::T* __restrict__ f, const uint8_t* __restrict__ flags,
                   const uint32_t* __restrict__ link, const uint32_t* __restrict__ lidm,
                   const float* __restrict__ epsBuf, StepParams p, const float* __restrict__ alphaBuf,
                   AdjHooks hk) {
    const uint32_t n = blockIdx.x * kStepBS + threadIdx.x;
    if (n >= p.N) return;
    // (writes no force: the force pointer is null and writeForce 0; porosity
    // acts through the forcing path, so BRINK comes with FORCE)
    static_assert(FORCE || !BRINK, "porosity needs the forcing path");
    step_cell<St, PARITY, WALLS, NT, LES, FORCE, BRINK, true, false, PLANAR>(f, flags, link, lidm, nullptr,
                                                                             epsBuf, p, n, alphaBuf, &hk);
}

// ---- odd pass with LDS row staging (wall-free lattices) --------------------
// A workgroup of BS threads covers BS/nx whole x-rows (nx a power of two
// dividing BS). The 10 x-shifted directions (c_x != 0) are loaded and stored
// as aligned row segments and shifted by one cell through LDS, with the
// periodic x-wrap applied inside the row; the 8 others are aligned already.
// Every global access is then a whole, aligned run, so the pass streams like
// the even pass. The slot set a workgroup touches belongs to its own cells
// (the AA property), so no other workgroup reads or writes it.
template <int I> __device__ __forceinline__ constexpr int xslot() { return I <= 2 ? I - 1 : I - 5; }

template <class St, int BS, bool NT, bool LES, bool FORCE, bool LARGE>
__global__ void __launch_bounds__(BS)
step_odd_rows(typename St::T* __restrict__ f, float4* __restrict__ force,
              const float* __restrict__ epsBuf, StepParams p) {
    using T = typename St::T;
    __shared__ T row[10][BS];
    const uint32_t t = threadIdx.x;
    const uint32_t n = blockIdx.x * BS + t;   // N % BS == 0
    constexpr int TB = (int)sizeof(T);
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const uint32_t xmask = (uint32_t)nx - 1u;
    const uint32_t x = n & xmask;
    const uint32_t rowbase = t - x;
    // LARGE (images of 4 GiB and more): a wave lies inside one x-row (nx >= 64),
    // so the row and its neighbour rows' first elements are uniform and every
    // access is a 64-bit row base plus the lane's 32-bit x offset. Otherwise
    // each access is a slot base plus the cell's 32-bit offset, which holds
    // fewer registers.
    const uint32_t r = LARGE ? __builtin_amdgcn_readfirstlane(fdiv(n, p.divx)) : fdiv(n, p.divx);
    const uint32_t z = LARGE ? __builtin_amdgcn_readfirstlane(fdiv(r, p.divy)) : fdiv(r, p.divy);
    const int y = (int)(r - z * (uint32_t)ny);
    const int plane = nx * ny;
    Offsets o;
    o.xm = 0; o.xp = 0;
    o.ym = (y == 0) ? (ny - 1) * nx : -nx;
    o.yp = (y == ny - 1) ? -(ny - 1) * nx : nx;
    o.zm = ((int)z == 0) ? (nz - 1) * plane : -plane;
    o.zp = ((int)z == nz - 1) ? -(nz - 1) * plane : plane;
    const uint32_t n0 = r * (uint32_t)nx;
    // slot i of the row n0 + d (rows never straddle a block of the layout)
    auto rowp = [&](int i, int d) -> T* {
        return f + (size_t)i * p.lay.W + cell_elem(p.lay, (uint32_t)((int)n0 + d));
    };
    const uint32_t xo = x * (uint32_t)TB;
    // the same access for images below 4 GiB; the store copy is pinned so the
    // push recomputes its offsets rather than keeping the pull's live
    auto rd = [&](int i, int d) -> float {
        if constexpr (LARGE) return St::rd(rowp(i, d), xo);
        else return St::rd(f + (size_t)i * p.lay.W, cell_elem(p.lay, (uint32_t)((int)n + d)) * (uint32_t)TB);
    };
    auto rd_raw = [&](int i, int d) -> T {
        if constexpr (LARGE) return St::rd_raw(rowp(i, d), xo);
        else return St::rd_raw(f + (size_t)i * p.lay.W, cell_elem(p.lay, (uint32_t)((int)n + d)) * (uint32_t)TB);
    };

    float fh[19];
    fh[0] = rd(0, 0);
#define STRH_RPULL(I)                                                                        \
    if (kCx[I] == 0) fh[I] = rd(kOpp[I], o.at<0, -kCy[I], -kCz[I]>());                       \
    else row[xslot<I>()][t] = rd_raw(kOpp[I], o.at<0, -kCy[I], -kCz[I]>());
    STRH_RPULL(1) STRH_RPULL(2) STRH_RPULL(3) STRH_RPULL(4) STRH_RPULL(5) STRH_RPULL(6)
    STRH_RPULL(7) STRH_RPULL(8) STRH_RPULL(9) STRH_RPULL(10) STRH_RPULL(11) STRH_RPULL(12)
    STRH_RPULL(13) STRH_RPULL(14) STRH_RPULL(15) STRH_RPULL(16) STRH_RPULL(17) STRH_RPULL(18)
#undef STRH_RPULL
    float eps = 0.0f;
    if (NT && p.useEps) eps = epsBuf[n];
    __syncthreads();
#define STRH_RSHIFT(I) \
    if (kCx[I] != 0) fh[I] = St::ld(row[xslot<I>()][rowbase + ((x - (uint32_t)kCx[I]) & xmask)]);
    STRH_RSHIFT(1) STRH_RSHIFT(2) STRH_RSHIFT(7) STRH_RSHIFT(8) STRH_RSHIFT(9) STRH_RSHIFT(10)
    STRH_RSHIFT(11) STRH_RSHIFT(12) STRH_RSHIFT(13) STRH_RSHIFT(14)
#undef STRH_RSHIFT

    float ntx = 0.0f, nty = 0.0f, ntz = 0.0f;
    collide<LES, FORCE, NT>(fh, p, (int)x, eps, ntx, nty, ntz);

    __syncthreads();   // every shifted read of `row` is done
    int nip = (int)n;   // pinned: the push recomputes the row offsets
    asm("" : "+v"(nip));
    auto wr = [&](int i, int d, float v) {
        if constexpr (LARGE) St::wr(rowp(i, d), xo, v);
        else St::wr(f + (size_t)i * p.lay.W, cell_elem(p.lay, (uint32_t)(nip + d)) * (uint32_t)TB, v);
    };
    auto wr_raw = [&](int i, int d, T v) {
        if constexpr (LARGE) St::wr_raw(rowp(i, d), xo, v);
        else St::wr_raw(f + (size_t)i * p.lay.W, cell_elem(p.lay, (uint32_t)(nip + d)) * (uint32_t)TB, v);
    };
    wr(0, 0, fh[0]);
#define STRH_RPUSH(I)                                                                        \
    if (kCx[I] == 0) wr(I, o.at<0, kCy[I], kCz[I]>(), fh[I]);                                \
    else row[xslot<I>()][rowbase + ((x + (uint32_t)kCx[I]) & xmask)] = St::st(fh[I]);
    STRH_RPUSH(1) STRH_RPUSH(2) STRH_RPUSH(3) STRH_RPUSH(4) STRH_RPUSH(5) STRH_RPUSH(6)
    STRH_RPUSH(7) STRH_RPUSH(8) STRH_RPUSH(9) STRH_RPUSH(10) STRH_RPUSH(11) STRH_RPUSH(12)
    STRH_RPUSH(13) STRH_RPUSH(14) STRH_RPUSH(15) STRH_RPUSH(16) STRH_RPUSH(17) STRH_RPUSH(18)
#undef STRH_RPUSH
    __syncthreads();
#define STRH_RSTORE(I) \
    if (kCx[I] != 0) wr_raw(I, o.at<0, kCy[I], kCz[I]>(), row[xslot<I>()][t]);
    STRH_RSTORE(1) STRH_RSTORE(2) STRH_RSTORE(7) STRH_RSTORE(8) STRH_RSTORE(9) STRH_RSTORE(10)
    STRH_RSTORE(11) STRH_RSTORE(12) STRH_RSTORE(13) STRH_RSTORE(14)
#undef STRH_RSTORE
    if (NT && p.writeForce != 0 && eps > 0.0f)                     // 351-353 (no walls)
        force[n] = make_float4(ntx, nty, ntz, 0.0f);
}

// ---- batch probe schedule ----------------------------------------------------
__device__ __forceinline__ bool is_force_step(const BatchLat& B, int64_t t) {
    if (B.fCount <= 0 || t < B.fT0) return false;
    const int64_t d = t - B.fT0;
    return d % B.fPeriod == 0 && d / B.fPeriod < B.fCount;
}

// The force buffer sample j = (t - fT0) / fPeriod is written to: buffer
// j % fRing of the ring, so the samples of up to kProbeSpan steps are reduced
// by one probe launch.
__device__ __forceinline__ float4* force_slot(const BatchLat& B, int64_t t) {
    if (B.fCount <= 0 || t < B.fT0) return B.force;
    return B.force + (size_t)(((t - B.fT0) / B.fPeriod) % B.fRing) * B.p.N;
}

// ---- fixed-order FP64 reductions ------------------------------------------
// The summed range is split into `segs` segments of M elements (the 19 DDF
// slots for the mass sum, one segment otherwise). Stage 1: block (b, g) sums
// elements [b*chunk, min((b+1)*chunk, M)) of segment g; thread t takes every
// 256th element of the chunk in increasing order, then the block folds its
// 256 partials with a fixed binary tree through LDS. Stage 2 folds the
// segment-major partials the same way. The order depends only on the lattice
// size and the fixed chunking — never on timing — so every result is bitwise
// reproducible run to run.
//
// Sum over the 64 lanes of a wave by a fixed butterfly of shuffles: at each
// step a lane adds the value of the lane whose index differs in one bit.
// IEEE addition is commutative, so both lanes of a pair form the same sum,
// and after six steps every lane holds the same bits.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block sum of a 256-thread workgroup: the four wave sums through LDS, added
// in wave order, (w0 + w1) + (w2 + w3). Every thread returns the same bits;
// lds (at least 4 doubles) may be reused as soon as the call returns.
__device__ __forceinline__ double block_tree_sum(double v, double* lds) {
    v = wave_sum(v);
    if ((threadIdx.x & 63u) == 0u) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    const double s = (lds[0] + lds[1]) + (lds[2] + lds[3]);
    __syncthreads();
    return s;
}

// A thread's share of sum e(seg, k), k in [lo, hi), in a 256-thread block:
// the elements lo + t + 256 j, accumulated into four chains by j mod 4 (four
// loads in flight per thread instead of one) and combined as (a0 + a1) +
// (a2 + a3). The order depends only on lo, hi and the thread index.
template <class Elem>
__device__ __forceinline__ double chunk_sum(const Elem& e, uint32_t seg, uint32_t lo, uint32_t hi) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    uint32_t k = lo + threadIdx.x;
    for (; k + 768u < hi; k += 1024u) {
        a0 += e(seg, k);
        a1 += e(seg, k + 256u);
        a2 += e(seg, k + 512u);
        a3 += e(seg, k + 768u);
    }
    if (k < hi) a0 += e(seg, k);
    if (k + 256u < hi) a1 += e(seg, k + 256u);
    if (k + 512u < hi) a2 += e(seg, k + 512u);
    return (a0 + a1) + (a2 + a3);
}

struct BoxForceElem {
    const float4* force;
    uint32_t nx, ny, x0, y0;
    FastDiv bw, bh;
    int comp;
    __device__ double operator()(uint32_t, uint32_t k) const {
        const uint32_t r = fdiv(k, bw);
        const uint32_t xx = k - r * bw.d;
        const uint32_t z = fdiv(r, bh);
        const uint32_t yy = r - z * bh.d;
        const float4 v = force[((size_t)z * ny + (y0 + yy)) * nx + (x0 + xx)];
        return (double)(comp == 0 ? v.x : (comp == 1 ? v.y : v.z));
    }
};

}  // namespace strh
