// D3Q19 lattice-Boltzmann device kernels for gfx950 (MI355X).
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

// BRINK: Brinkman porosity alpha (new in this build, no reference
// counterpart): drag force -alpha rho u with u taken implicitly,
// u = (p + Fb/2) / (rho (1 + alpha/2)); the Guo terms then see the per-cell
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

// Workgroup size of the one-cell-per-thread kernel (64..512 measured within
// 2% of each other at 256^3; 256 is the best or tied).
constexpr uint32_t kStepBS = 256;

// Adjoint hooks of the forward pass (AdjHooks in kernels.h). The collision's
// Jacobian depends on the state only through rho and p, so the reverse step
// needs just these four moments of the pre-collision populations (summed in
// FP64, stored as FP32: rho - 1 keeps the density's small part exact to the
// float's 24 bits). The hook also accumulates the cell's objective term,
//   cp(x) rho + wd alpha J_x,   J_x = (p_x + F_x / 2) / (1 + alpha / 2),
// with cp = +cin on the inlet column and -cout on the outlet column.
// The rate w+ a cell collides with (sponge ramp, then Smagorinsky), for the
// hooks' partial-cell drag; the same float arithmetic as collide.
template <bool LES>
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
        const float ux = (px + 0.5f * p.fx) * inv, uy = (py + 0.5f * p.fy) * inv, uz = (pz + 0.5f * p.fz) * inv;
        les_rates<true>(fh, p, rhom1, rho, inv, ux, uy, uz, ux * ux + uy * uy + uz * uz, wp, wm);
    }
    return wp;
}

template <bool LES>
__device__ __forceinline__ void run_hooks(const AdjHooks& hk, const float (&fh)[19], uint32_t n,
                                          int x, float alpha, float eps, const StepParams& p) {
    const float fx = p.fx;
    if (hk.recNt && eps > 0.0f) {
        const int32_t k = hk.ntIdx[n];
#pragma unroll
        for (int i = 0; i < 19; i++) hk.recNt[(size_t)i * hk.M + k] = fh[i];
    }
    if (hk.jacc && hk.wf != 0.0 && eps > 0.0f) {   // partial-cell drag ntF_x = -2 sum c_x B (g_opp - g)
        const double tw = 1.0 / (double)hook_rate<LES>(fh, p, x) - 0.5, e = (double)eps;
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
          bool HOOK = false>
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
        if (HOOK) run_hooks<LES>(*hk, fh, n, x, alpha, eps, p);
        collide<LES, FORCE, NT, BRINK>(fh, p, x, eps, ntx, nty, ntz, alpha);
        // A separate copy of the offset keeps the stores on the scalar-base
        // form too (otherwise the shared base+offset sum is kept as 19
        // 64-bit VGPR addresses across the collision).
        uint32_t nbs = nbi;
        asm("" : "+v"(nbs));
        St::wr(fs[0], nbs, fh[0]);
#pragma unroll
        for (int i = 1; i < 19; i++) St::wr(fs[kOpp[i]], nbs, fh[i]);
        if (NT && p.writeForce != 0 && eps > 0.0f) force[n] = make_float4(ntx, nty, ntz, 0.0f);
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
    const uint32_t t = fdiv(n, p.divx);
    const int x = (int)(n - t * (uint32_t)p.nx);
    const uint32_t z = fdiv(t, p.divy);
    const int y = (int)(t - z * (uint32_t)p.ny);
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int plane = nx * ny;
    Offsets o;
    o.xm = (x == 0) ? nx - 1 : -1;
    o.xp = (x == nx - 1) ? 1 - nx : 1;
    o.ym = (y == 0) ? (ny - 1) * nx : -nx;
    o.yp = (y == ny - 1) ? -(ny - 1) * nx : nx;
    o.zm = ((int)z == 0) ? (nz - 1) * plane : -plane;
    o.zp = ((int)z == nz - 1) ? -(nz - 1) * plane : plane;
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
                if (p.writeForce != 0) {                                              \
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
                if (p.writeForce != 0) {                                              \
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
    if (HOOK) run_hooks<LES>(*hk, fh, n, x, alpha, eps, p);
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
    const uint32_t zp = fdiv(tp, p.divy);
    const int yp = (int)(tp - zp * (uint32_t)ny);
    Offsets op;
    op.xm = (xp == 0) ? nx - 1 : -1;
    op.xp = (xp == nx - 1) ? 1 - nx : 1;
    op.ym = (yp == 0) ? (ny - 1) * nx : -nx;
    op.yp = (yp == ny - 1) ? -(ny - 1) * nx : nx;
    op.zm = ((int)zp == 0) ? (nz - 1) * plane : -plane;
    op.zp = ((int)zp == nz - 1) ? -(nz - 1) * plane : plane;
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
                if (p.writeForce != 0) {                                              \
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
    if (p.writeForce != 0 && (sM != 0u || eps > 0.0f))               // 351-353
        force[n] = make_float4(ax, ay, az, 0.0f);
}

template <class St, int PARITY, bool WALLS, bool NT, bool LES, bool FORCE, bool BRINK = false>
__global__ void __launch_bounds__(kStepBS)
step_kernel(typename St::T* __restrict__ f, const uint8_t* __restrict__ flags,
            const uint32_t* __restrict__ link, const uint32_t* __restrict__ lidm,
            float4* __restrict__ force, const float* __restrict__ epsBuf, StepParams p,
            const float* __restrict__ alphaBuf) {
    const uint32_t n = blockIdx.x * kStepBS + threadIdx.x;
    if (n >= p.N) return;
    step_cell<St, PARITY, WALLS, NT, LES, FORCE, BRINK>(f, flags, link, lidm, force, epsBuf, p, n,
                                                        alphaBuf);
}

// Forward step with the adjoint hooks (Guo forcing with Brinkman porosity, or
// partial cells and / or LES). The arithmetic is that of step_kernel: with
// alpha = 0 and no body force the DDFs are the plain update's bits.
template <class St, int PARITY, bool WALLS, bool NT, bool LES, bool BRINK>
__global__ void __launch_bounds__(kStepBS)
step_kernel_hooked(typename St::T* __restrict__ f, const uint8_t* __restrict__ flags,
                   const uint32_t* __restrict__ link, const uint32_t* __restrict__ lidm,
                   const float* __restrict__ epsBuf, StepParams p, const float* __restrict__ alphaBuf,
                   AdjHooks hk) {
    const uint32_t n = blockIdx.x * kStepBS + threadIdx.x;
    if (n >= p.N) return;
    step_cell<St, PARITY, WALLS, NT, LES, true, BRINK, true>(f, flags, link, lidm, nullptr, epsBuf, p, n,
                                                             alphaBuf, &hk);
}

// ---- discrete adjoint: one reverse step ------------------------------------
// The forward step is g^{t+1} = P(C(g^t)) in natural slot order: C the
// collision of a fluid cell, P the streaming with halfway bounce-back
//   g^{t+1}_i(x + c_i)  = C_i(x)                    (fluid neighbour)
//   g^{t+1}_opp(i)(x)   = C_i(x) - 6 w_i c_i.u_w    (wall at x + c_i; OUTFLOW: -C_i)
// so mu = P^T lambda^{t+1} gathers lambda from the neighbours (or from the
// opposite slot of the cell itself across a wall), and
// lambda^t = J_C(g^t)^T mu + d(objective_t)/dg^t. With M0 = sum w mu,
// Mc = sum w mu c, Mcc = sum w mu c c^T, J = (p + Fb/2)/(1 + alpha/2),
// F = Fb - alpha J and the collision written through (rho, J, F):
//   E_i = w+ feq+_i + w- feq-_i + S_i,
//   feq+ = w(rho - 1 + 4.5 (c.J)^2/rho - 1.5 J^2/rho), feq- = 3 w c.J,
//   S = w [a+ (9 (c.J)(c.F) - 3 J.F)/rho + 3 a- c.F],  a+- = 1 - w+-/2,
// the partial derivatives of mu.E are
//   dRho = w+ (M0 - 4.5 J.Mcc.J/rho^2 + 1.5 J^2 M0/rho^2) - a+ (9 J.Mcc.F - 3 J.F M0)/rho^2
//   dJ   = w+ (9 Mcc.J - 3 J M0)/rho + 3 w- Mc + a+ (9 Mcc.F - 3 F M0)/rho
//   dF   = a+ (9 Mcc.J - 3 J M0)/rho + 3 a- Mc
// and through J(p, alpha), F(J, alpha): dP = (dJ - alpha dF)/(1 + alpha/2),
//   dJ/dalpha = -J/(2 gamma), dF/dalpha = -J (1 - alpha/(2 gamma)).
// The adjoint state and the collision's vector-Jacobian product are FP32 (the
// precision of the recorded state); the gradients accumulate in FP64.
template <int I> __device__ __forceinline__ float wgt_f() {
    return I == 0 ? 1.0f / 3.0f : (I <= 6 ? 1.0f / 18.0f : 1.0f / 36.0f);
}

// mu = P^T lambda^{t+1} for cell n: the pull from each neighbour, or at a
// wall link the cell's own opposite slot (negated for the anti-bounce-back
// outlet). Returns the cell's share of dJ/d(uin target) from inflow links.
template <bool WALLS>
__device__ __forceinline__ double adj_gather(const AdjStep& q, uint32_t n, int x, int y, int z,
                                             const float* __restrict__ lamN,
                                             const uint8_t* __restrict__ flags, uint32_t sM,
                                             float (&mu)[19]) {
    const uint32_t S = q.S;
    const int nx = (int)q.nx, ny = (int)q.ny, nz = (int)q.nz;
    const int plane = nx * ny;
    Offsets o;   // element offsets to the wrapped neighbours
    o.xm = (x == 0) ? nx - 1 : -1;
    o.xp = (x == nx - 1) ? 1 - nx : 1;
    o.ym = (y == 0) ? (ny - 1) * nx : -nx;
    o.yp = (y == ny - 1) ? -(ny - 1) * nx : nx;
    o.zm = (z == 0) ? (nz - 1) * plane : -plane;
    o.zp = (z == nz - 1) ? -(nz - 1) * plane : plane;
    double gu = 0.0;
    mu[0] = ld_off(lamN, n * 4u);
    // unconditional loads (a wall link reads the cell's own opposite slot),
    // then the rare wall terms, so the 18 loads issue back to back
#define STRH_GATHER(I)                                                                \
    {                                                                                 \
        const uint32_t d = (uint32_t)((int)n + o.at<kCx[I], kCy[I], kCz[I]>());       \
        constexpr int lo = I < kOpp[I] ? I : kOpp[I];                                 \
        uint32_t off = (d + (I == lo ? 0u : S)) * 4u;                                 \
        if (WALLS && ((sM >> (I - 1)) & 1u)) off = (n + (kOpp[I] == lo ? 0u : S)) * 4u; \
        mu[I] = ld_off(lamN + (size_t)lo * S, off);                                   \
    }
    STRH_GATHER(1) STRH_GATHER(2) STRH_GATHER(3) STRH_GATHER(4) STRH_GATHER(5) STRH_GATHER(6)
    STRH_GATHER(7) STRH_GATHER(8) STRH_GATHER(9) STRH_GATHER(10) STRH_GATHER(11) STRH_GATHER(12)
    STRH_GATHER(13) STRH_GATHER(14) STRH_GATHER(15) STRH_GATHER(16) STRH_GATHER(17) STRH_GATHER(18)
#undef STRH_GATHER
    if (WALLS && sM != 0u) {
#define STRH_GWALL(I)                                                                 \
        if ((sM >> (I - 1)) & 1u) {                                                   \
            const uint32_t d = (uint32_t)((int)n + o.at<kCx[I], kCy[I], kCz[I]>());   \
            const uint8_t wf = flags[d];                                              \
            const float ln = mu[I];                                                   \
            if (wf == kOutflow) mu[I] = -ln;                                          \
            if (kCx[I] != 0 && wf == kInflow) {                                       \
                const int dy = y + kCy[I] < 0 ? ny - 1 : (y + kCy[I] >= ny ? 0 : y + kCy[I]); \
                double prof = 1.0;                                                    \
                if (!q.inflowUniform) {                                               \
                    const double H = (double)ny - 2.0, yd = (double)dy - 0.5;         \
                    prof = 4.0 * yd * (H - yd) / (H * H);                             \
                    if (!(q.uinTarget * q.rOdd * prof > 0.0)) prof = 0.0;             \
                }                                                                     \
                gu += (double)ln * (-6.0 * (double)wgt_f<I>() * (double)kCx[I] * q.rOdd * prof); \
            }                                                                         \
        }
        STRH_GWALL(1) STRH_GWALL(2) STRH_GWALL(3) STRH_GWALL(4) STRH_GWALL(5) STRH_GWALL(6)
        STRH_GWALL(7) STRH_GWALL(8) STRH_GWALL(9) STRH_GWALL(10) STRH_GWALL(11) STRH_GWALL(12)
        STRH_GWALL(13) STRH_GWALL(14) STRH_GWALL(15) STRH_GWALL(16) STRH_GWALL(17) STRH_GWALL(18)
#undef STRH_GWALL
    }
    return gu;
}

// Relaxation rates of column x: the outlet sponge ramp (Kernels.metal:112-123).
__device__ __forceinline__ void adj_sponge(const AdjStep& q, int x, float& wp, float& wm) {
    wp = q.wp;
    wm = q.wm;
    if ((float)x > q.spongeX0) {
        const float frac = fminf(((float)x - q.spongeX0) * q.spongeInvW, 1.0f);
        const float tau0 = 1.0f / wp;
        const float tauS = tau0 + frac * (q.spongeTau - tau0);
        if (q.srt) {
            wp = 1.0f / tauS;
            wm = wp;
        } else {
            const float lambda = (1.0f / wp - 0.5f) * (1.0f / wm - 0.5f);
            wp = 1.0f / tauS;
            wm = 1.0f / (0.5f + lambda / (tauS - 0.5f));
        }
    }
}

// Weighted moments of mu: M0 = sum w mu, M_a = sum w c_a mu, M_ab = sum w c_a c_b mu.
__device__ __forceinline__ void mu_moments(const float (&mu)[19], float& M0, float& Mx, float& My,
                                           float& Mz, float& Mxx, float& Myy, float& Mzz, float& Mxy,
                                           float& Mxz, float& Myz) {
    M0 = 0.0f; Mx = 0.0f; My = 0.0f; Mz = 0.0f;
    Mxx = 0.0f; Myy = 0.0f; Mzz = 0.0f; Mxy = 0.0f; Mxz = 0.0f; Myz = 0.0f;
#define STRH_MSUM(I)                                                                  \
    {                                                                                 \
        const float v = wgt_f<I>() * mu[I];                                           \
        M0 += v;                                                                      \
        if (kCx[I]) { Mx += (float)kCx[I] * v; Mxx += v; }                            \
        if (kCy[I]) { My += (float)kCy[I] * v; Myy += v; }                            \
        if (kCz[I]) { Mz += (float)kCz[I] * v; Mzz += v; }                            \
        if (kCx[I] * kCy[I]) Mxy += (float)(kCx[I] * kCy[I]) * v;                     \
        if (kCx[I] * kCz[I]) Mxz += (float)(kCx[I] * kCz[I]) * v;                     \
        if (kCy[I] * kCz[I]) Myz += (float)(kCy[I] * kCz[I]) * v;                     \
    }
    STRH_MSUM(0) STRH_MSUM(1) STRH_MSUM(2) STRH_MSUM(3) STRH_MSUM(4) STRH_MSUM(5) STRH_MSUM(6)
    STRH_MSUM(7) STRH_MSUM(8) STRH_MSUM(9) STRH_MSUM(10) STRH_MSUM(11) STRH_MSUM(12)
    STRH_MSUM(13) STRH_MSUM(14) STRH_MSUM(15) STRH_MSUM(16) STRH_MSUM(17) STRH_MSUM(18)
#undef STRH_MSUM
}

template <bool WALLS, bool NT>
__global__ void __launch_bounds__(256)
adjoint_step_kernel(const float4* __restrict__ rec, const float* __restrict__ lamN,
                    float* __restrict__ lam, const uint8_t* __restrict__ flags,
                    const uint32_t* __restrict__ link, const float* __restrict__ alphaBuf,
                    double* __restrict__ gAlpha, double* __restrict__ gUin,
                    const float* __restrict__ epsBuf, const int32_t* __restrict__ ntIdx,
                    const float* __restrict__ recNt, double* __restrict__ gEps, AdjStep q) {
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= q.N) return;
    uint32_t sM = 0u;   // lattices without walls hold fluid cells only
    if (WALLS) {
        const uint32_t lw = link[n];
        if (lw & kNotFluidBit) return;
        sM = lw & 0x3ffffu;
    }
    const uint32_t S = q.S;
    const uint32_t n4 = n * 4u;
    const uint32_t t = fdiv(n, q.divx);
    const int x = (int)(n - t * q.nx);
    const uint32_t z = fdiv(t, q.divy);
    const int y = (int)(t - z * q.ny);
    float mu[19];
    const double gu = adj_gather<WALLS>(q, n, x, y, (int)z, lamN, flags, sM, mu);

    const float4 m = ld_off(rec, n * 16u);
    const float a = alphaBuf ? alphaBuf[n] : 0.0f;
    const float rho = 1.0f + m.x, ir = 1.0f / rho, ir2 = ir * ir;
    const float gam = 1.0f + 0.5f * a, ig = 1.0f / gam;
    const float fx = q.fx, fy = q.fy, fz = q.fz;
    const float Jx = (m.y + 0.5f * fx) * ig, Jy = (m.z + 0.5f * fy) * ig, Jz = (m.w + 0.5f * fz) * ig;
    const float Fx = fx - a * Jx, Fy = fy - a * Jy, Fz = fz - a * Jz;
    float M0, Mx, My, Mz, Mxx, Myy, Mzz, Mxy, Mxz, Myz;
    mu_moments(mu, M0, Mx, My, Mz, Mxx, Myy, Mzz, Mxy, Mxz, Myz);
    const float MJx = Mxx * Jx + Mxy * Jy + Mxz * Jz;
    const float MJy = Mxy * Jx + Myy * Jy + Myz * Jz;
    const float MJz = Mxz * Jx + Myz * Jy + Mzz * Jz;
    const float MFx = Mxx * Fx + Mxy * Fy + Mxz * Fz;
    const float MFy = Mxy * Fx + Myy * Fy + Myz * Fz;
    const float MFz = Mxz * Fx + Myz * Fy + Mzz * Fz;
    const float JMJ = Jx * MJx + Jy * MJy + Jz * MJz;
    const float JMF = Jx * MFx + Jy * MFy + Jz * MFz;
    const float J2 = Jx * Jx + Jy * Jy + Jz * Jz;
    const float JF = Jx * Fx + Jy * Fy + Jz * Fz;
    float wp, wm;
    adj_sponge(q, x, wp, wm);
    const float ap = 1.0f - 0.5f * wp, am = 1.0f - 0.5f * wm;
    const float gRho = wp * (M0 - 4.5f * JMJ * ir2 + 1.5f * J2 * M0 * ir2)
                     - ap * (9.0f * JMF - 3.0f * JF * M0) * ir2;
    const float gJx = wp * (9.0f * MJx - 3.0f * Jx * M0) * ir + 3.0f * wm * Mx + ap * (9.0f * MFx - 3.0f * Fx * M0) * ir;
    const float gJy = wp * (9.0f * MJy - 3.0f * Jy * M0) * ir + 3.0f * wm * My + ap * (9.0f * MFy - 3.0f * Fy * M0) * ir;
    const float gJz = wp * (9.0f * MJz - 3.0f * Jz * M0) * ir + 3.0f * wm * Mz + ap * (9.0f * MFz - 3.0f * Fz * M0) * ir;
    const float gFx = ap * (9.0f * MJx - 3.0f * Jx * M0) * ir + 3.0f * am * Mx;
    const float gFy = ap * (9.0f * MJy - 3.0f * Jy * M0) * ir + 3.0f * am * My;
    const float gFz = ap * (9.0f * MJz - 3.0f * Jz * M0) * ir + 3.0f * am * Mz;
    const float gPx = (gJx - a * gFx) * ig, gPy = (gJy - a * gFy) * ig, gPz = (gJz - a * gFz) * ig;
    // objective of step t: cp rho + wd alpha J_x
    const float cp = (x == q.inletX ? q.cin : 0.0f) - (x == q.outletX ? q.cout : 0.0f);
    const float wd = q.wd;
    const float dPx = gPx + wd * a * ig;
    const float base = gRho + cp;
    if (NT) {
        const int32_t k = ntIdx[n];
        if (k >= 0) {
            // Noble-Torczynski cell (Kernels.metal:201-227):
            //   f*_j = f_j + (1 - B) Omega_j(f) + B (f_opp(j) - f_j),  f*_0 = f_0 + (1 - B) Omega_0,
            // so lambda_j = mu_j + (1 - B) [Omega^T mu]_j + B (mu_opp(j) - mu_j) + objective,
            // dJ/dB = -mu.Omega + sum_i mu_i (g_opp(i) - g_i) + w_f d(ntF_x)/dB.
            const uint32_t M = q.M;
            float g[19];
#pragma unroll
            for (int i = 0; i < 19; i++) g[i] = recNt[(size_t)i * M + k];
            const float e = epsBuf[n];
            const float tw = 1.0f / wp - 0.5f;
            const float den = (1.0f - e) + tw;
            const float B = e * tw / den, omB = 1.0f - B;
            const float dBde = tw * (1.0f + tw) / (den * den);
            const float ux = Jx * ir, uy = Jy * ir, uz = Jz * ir;
            const float u2 = ux * ux + uy * uy + uz * uz;
            const float uF = ux * fx + uy * fy + uz * fz;
            const float rm1 = m.x;
            float Om[19];
            Om[0] = wp * ((1.0f / 3.0f) * (rm1 - 1.5f * rho * u2) - g[0]) + (1.0f / 3.0f) * ap * (-3.0f * uF);
#define STRH_OMEGA(I)                                                                         \
            {                                                                                 \
                const float W = wgt_f<I>();                                                   \
                const float cu = 3.0f * ((float)kCx[I] * ux + (float)kCy[I] * uy + (float)kCz[I] * uz); \
                const float cF = (float)kCx[I] * fx + (float)kCy[I] * fy + (float)kCz[I] * fz; \
                const float dp = wp * (W * (rm1 + rho * (0.5f * cu * cu - 1.5f * u2)) - 0.5f * (g[I] + g[I + 1])); \
                const float dm = wm * (W * rho * cu - 0.5f * (g[I] - g[I + 1]));              \
                const float sm = W * am * (3.0f * cF);                                        \
                const float sp = W * ap * (3.0f * cu * cF - 3.0f * uF);                       \
                Om[I] = dp + dm + sp + sm;                                                    \
                Om[I + 1] = dp - dm + sp - sm;                                                \
            }
            STRH_OMEGA(1) STRH_OMEGA(3) STRH_OMEGA(5) STRH_OMEGA(7) STRH_OMEGA(9)
            STRH_OMEGA(11) STRH_OMEGA(13) STRH_OMEGA(15) STRH_OMEGA(17)
#undef STRH_OMEGA
            float gB = 0.0f, dNt = 0.0f;
#pragma unroll
            for (int i = 0; i < 19; i++) gB -= mu[i] * Om[i];
#pragma unroll
            for (int i = 1; i < 19; i++) gB += mu[i] * (g[kOpp[i]] - g[i]);
#pragma unroll
            for (int i = 1; i < 19; i += 2) if (kCx[i]) dNt -= 2.0f * (float)kCx[i] * (g[i + 1] - g[i]);
            const float dObjX = q.wf * 2.0f * B;          // d ntF_x / d g_j = 2 B c_jx
            st_off(lam, n4, mu[0] + omB * (-wp * mu[0] + gRho) + cp);
#define STRH_NTLAM(I)                                                                         \
            {                                                                                 \
                const float mp = 0.5f * (mu[I] + mu[kOpp[I]]);                                \
                const float mm = 0.5f * (mu[I] - mu[kOpp[I]]);                                \
                float om = -wp * mp - wm * mm + gRho;                                         \
                if (kCx[I]) om += (float)kCx[I] * gPx;                                        \
                if (kCy[I]) om += (float)kCy[I] * gPy;                                        \
                if (kCz[I]) om += (float)kCz[I] * gPz;                                        \
                float v = mu[I] + omB * om + B * (mu[kOpp[I]] - mu[I]) + cp;                  \
                if (kCx[I]) v += (float)kCx[I] * dObjX;                                       \
                st_off(lam + (size_t)I * S, n4, v);                                            \
            }
            STRH_NTLAM(1) STRH_NTLAM(2) STRH_NTLAM(3) STRH_NTLAM(4) STRH_NTLAM(5) STRH_NTLAM(6)
            STRH_NTLAM(7) STRH_NTLAM(8) STRH_NTLAM(9) STRH_NTLAM(10) STRH_NTLAM(11) STRH_NTLAM(12)
            STRH_NTLAM(13) STRH_NTLAM(14) STRH_NTLAM(15) STRH_NTLAM(16) STRH_NTLAM(17) STRH_NTLAM(18)
#undef STRH_NTLAM
            gEps[k] += (double)((gB + q.wf * dNt) * dBde);
            if (gu != 0.0) gUin[n] += gu;
            return;
        }
    }
    st_off(lam, n4, mu[0] - wp * mu[0] + base);
#define STRH_LAM(I)                                                                   \
    {                                                                                 \
        const float mp = 0.5f * (mu[I] + mu[kOpp[I]]);                                \
        const float mm = 0.5f * (mu[I] - mu[kOpp[I]]);                                \
        float v = mu[I] - wp * mp - wm * mm + base;                                   \
        if (kCx[I]) v += (float)kCx[I] * dPx;                                         \
        if (kCy[I]) v += (float)kCy[I] * gPy;                                         \
        if (kCz[I]) v += (float)kCz[I] * gPz;                                         \
        st_off(lam + (size_t)I * S, n4, v);                                            \
    }
    STRH_LAM(1) STRH_LAM(2) STRH_LAM(3) STRH_LAM(4) STRH_LAM(5) STRH_LAM(6)
    STRH_LAM(7) STRH_LAM(8) STRH_LAM(9) STRH_LAM(10) STRH_LAM(11) STRH_LAM(12)
    STRH_LAM(13) STRH_LAM(14) STRH_LAM(15) STRH_LAM(16) STRH_LAM(17) STRH_LAM(18)
#undef STRH_LAM
    const float gJdotJ = gJx * Jx + gJy * Jy + gJz * Jz;
    const float gFdotJ = gFx * Jx + gFy * Jy + gFz * Jz;
    const float k = 1.0f - 0.5f * a * ig;
    st_off(gAlpha, n * 8u, ld_off(gAlpha, n * 8u) + (double)(-0.5f * gJdotJ * ig - gFdotJ * k + wd * Jx * k));
    if (gu != 0.0) gUin[n] += gu;
}

// Reverse step of a Smagorinsky lattice (Kernels.metal:140-167), optionally
// with partial cells. The rates w+ and w- of a cell depend on its populations
// through Q = |Pi_neq| and rho:
//   tau = (tau0 + s)/2, s = sqrt(tau0^2 + K Q / rho), K = 18 sqrt(2) Cs^2,
//   w+ = 1/tau, w- = w+ (SRT) or 1/(1/2 + Lambda/(tau - 1/2)) (TRT),
// with tau0 = 1/w+ after the sponge ramp. The forward pass records all 19
// pre-collision populations of every cell (rec19, [19][N]); the reverse step
// adds to the fixed-rate product the rate chain
//   lambda_j += dJ/dQ dQ/dg_j + dJ/drho,
//   dJ/dtau = -w+^2 dJ/dw+ + dw-/dtau dJ/dw-,
//   dQ/dg_j = (c.Pi.c - tr(Pi)/3 - 2 c.Pi.J/rho + J.Pi.J/rho^2) / Q,
// where dJ/dw+- = sum_j mu_j dOmega_j/dw+- (scaled by 1 - B on partial cells,
// which add the B(tw(w+)) term). Porosity and LES exclude each other, so F is
// the body force and J = p + F/2.
// The recording holds the populations in the storage format as loaded, before
// the odd pass's moving-wall corrections; the kernel re-applies them from the
// step's parameters sp with the forward's arithmetic.
template <class Conv, bool WALLS, bool NT>
__global__ void __launch_bounds__(256)
adjoint_les_kernel(const typename Conv::T* __restrict__ rec19, const float* __restrict__ lamN,
                   float* __restrict__ lam, const uint8_t* __restrict__ flags,
                   const uint32_t* __restrict__ link, const uint32_t* __restrict__ lidm,
                   double* __restrict__ gUin, const float* __restrict__ epsBuf,
                   const int32_t* __restrict__ ntIdx, double* __restrict__ gEps, AdjStep q,
                   StepParams sp) {
    using T = typename Conv::T;
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= q.N) return;
    uint32_t sM = 0u;   // lattices without walls hold fluid cells only
    if (WALLS) {
        const uint32_t lw = link[n];
        if (lw & kNotFluidBit) return;
        sM = lw & 0x3ffffu;
    }
    const uint32_t S = q.S;
    const uint32_t n4 = n * 4u;
    const uint32_t t = fdiv(n, q.divx);
    const int x = (int)(n - t * q.nx);
    const uint32_t z = fdiv(t, q.divy);
    const int y = (int)(t - z * q.ny);
    float mu[19];
    const double gu = adj_gather<WALLS>(q, n, x, y, (int)z, lamN, flags, sM, mu);

    float g[19];
#pragma unroll
    for (int i = 0; i < 19; i++)
        g[i] = Conv::ld(ld_off(rec19 + (size_t)i * q.N, n * (uint32_t)sizeof(T)));
    if (WALLS && sp.parity == 1 && sM != 0u) {
        const uint32_t dm = sM & lidm[n];   // moving-wall / outflow links (Kernels.metal:292-314)
        if (dm != 0u) {
            const int nx = (int)q.nx, ny = (int)q.ny, nz = (int)q.nz, plane = nx * ny;
            Offsets o;
            o.xm = (x == 0) ? nx - 1 : -1;
            o.xp = (x == nx - 1) ? 1 - nx : 1;
            o.ym = (y == 0) ? (ny - 1) * nx : -nx;
            o.yp = (y == ny - 1) ? -(ny - 1) * nx : nx;
            o.zm = ((int)z == 0) ? (nz - 1) * plane : -plane;
            o.zp = ((int)z == nz - 1) ? -(nz - 1) * plane : plane;
#define STRH_REDEFER(I)                                                                       \
            if ((dm >> (kOpp[I] - 1)) & 1u) {                                                 \
                const int sy = y - kCy[I] < 0 ? ny - 1 : (y - kCy[I] >= ny ? 0 : y - kCy[I]); \
                const uint32_t sc = (uint32_t)((int)n + o.at<-kCx[I], -kCy[I], -kCz[I]>());   \
                const uint8_t wf = flags[sc];                                                 \
                if (wf == kOutflow) {                                                         \
                    g[I] = -g[I];                                                             \
                } else {                                                                      \
                    float uwx, uwy, uwz;                                                      \
                    wall_vel(wf, sy, sp, uwx, uwy, uwz);                                      \
                    const float cu = (float)kCx[I] * uwx + (float)kCy[I] * uwy + (float)kCz[I] * uwz; \
                    g[I] += 6.0f * wgt<I>() * cu;                                             \
                }                                                                             \
            }
            STRH_REDEFER(1) STRH_REDEFER(2) STRH_REDEFER(3) STRH_REDEFER(4) STRH_REDEFER(5)
            STRH_REDEFER(6) STRH_REDEFER(7) STRH_REDEFER(8) STRH_REDEFER(9) STRH_REDEFER(10)
            STRH_REDEFER(11) STRH_REDEFER(12) STRH_REDEFER(13) STRH_REDEFER(14) STRH_REDEFER(15)
            STRH_REDEFER(16) STRH_REDEFER(17) STRH_REDEFER(18)
#undef STRH_REDEFER
        }
    }
    // moments of g: rho - 1, p and the second moments sum c_a c_b g
    float rm1 = 0.0f, px = 0.0f, py = 0.0f, pz = 0.0f;
    float sxx = 0.0f, syy = 0.0f, szz = 0.0f, sxy = 0.0f, sxz = 0.0f, syz = 0.0f;
#define STRH_GMOM(I)                                                                  \
    {                                                                                 \
        const float v = g[I];                                                         \
        rm1 += v;                                                                     \
        if (kCx[I]) { px += (float)kCx[I] * v; sxx += v; }                            \
        if (kCy[I]) { py += (float)kCy[I] * v; syy += v; }                            \
        if (kCz[I]) { pz += (float)kCz[I] * v; szz += v; }                            \
        if (kCx[I] * kCy[I]) sxy += (float)(kCx[I] * kCy[I]) * v;                     \
        if (kCx[I] * kCz[I]) sxz += (float)(kCx[I] * kCz[I]) * v;                     \
        if (kCy[I] * kCz[I]) syz += (float)(kCy[I] * kCz[I]) * v;                     \
    }
    STRH_GMOM(0) STRH_GMOM(1) STRH_GMOM(2) STRH_GMOM(3) STRH_GMOM(4) STRH_GMOM(5) STRH_GMOM(6)
    STRH_GMOM(7) STRH_GMOM(8) STRH_GMOM(9) STRH_GMOM(10) STRH_GMOM(11) STRH_GMOM(12)
    STRH_GMOM(13) STRH_GMOM(14) STRH_GMOM(15) STRH_GMOM(16) STRH_GMOM(17) STRH_GMOM(18)
#undef STRH_GMOM
    const float rho = 1.0f + rm1, ir = 1.0f / rho, ir2 = ir * ir;
    const float fx = q.fx, fy = q.fy, fz = q.fz;
    const float Jx = px + 0.5f * fx, Jy = py + 0.5f * fy, Jz = pz + 0.5f * fz;

    // local rates: sponge, then Smagorinsky with Pi_neq in closed form,
    // Pi_ab = sum c_a c_b g - delta_ab (rho - 1)/3 - J_a J_b / rho
    float wp, wm;
    adj_sponge(q, x, wp, wm);
    const float r3 = rm1 * (1.0f / 3.0f);
    const float pxx = sxx - r3 - Jx * Jx * ir, pyy = syy - r3 - Jy * Jy * ir, pzz = szz - r3 - Jz * Jz * ir;
    const float pxy = sxy - Jx * Jy * ir, pxz = sxz - Jx * Jz * ir, pyz = syz - Jy * Jz * ir;
    const float Q = sqrtf(pxx * pxx + pyy * pyy + pzz * pzz + 2.0f * (pxy * pxy + pxz * pxz + pyz * pyz));
    const float tau0 = 1.0f / wp;
    const float Kc = 25.455844f * q.cSmago;
    const float sq = sqrtf(tau0 * tau0 + Kc * Q * ir);
    const float tau = 0.5f * (tau0 + sq);
    float lambda = 0.0f;
    if (q.srt) {
        wp = 1.0f / tau;
        wm = wp;
    } else {
        lambda = (1.0f / wp - 0.5f) * (1.0f / wm - 0.5f);
        wp = 1.0f / tau;
        wm = 1.0f / (0.5f + lambda / (tau - 0.5f));
    }
    const float ap = 1.0f - 0.5f * wp, am = 1.0f - 0.5f * wm;

    // vector-Jacobian product of E(rho, J) at fixed rates
    float M0, Mx, My, Mz, Mxx, Myy, Mzz, Mxy, Mxz, Myz;
    mu_moments(mu, M0, Mx, My, Mz, Mxx, Myy, Mzz, Mxy, Mxz, Myz);
    const float MJx = Mxx * Jx + Mxy * Jy + Mxz * Jz;
    const float MJy = Mxy * Jx + Myy * Jy + Myz * Jz;
    const float MJz = Mxz * Jx + Myz * Jy + Mzz * Jz;
    const float MFx = Mxx * fx + Mxy * fy + Mxz * fz;
    const float MFy = Mxy * fx + Myy * fy + Myz * fz;
    const float MFz = Mxz * fx + Myz * fy + Mzz * fz;
    const float JMJ = Jx * MJx + Jy * MJy + Jz * MJz;
    const float JMF = Jx * MFx + Jy * MFy + Jz * MFz;
    const float J2 = Jx * Jx + Jy * Jy + Jz * Jz;
    const float JF = Jx * fx + Jy * fy + Jz * fz;
    const float gRho = wp * (M0 - 4.5f * JMJ * ir2 + 1.5f * J2 * M0 * ir2)
                     - ap * (9.0f * JMF - 3.0f * JF * M0) * ir2;
    float gPx = wp * (9.0f * MJx - 3.0f * Jx * M0) * ir + 3.0f * wm * Mx + ap * (9.0f * MFx - 3.0f * fx * M0) * ir;
    float gPy = wp * (9.0f * MJy - 3.0f * Jy * M0) * ir + 3.0f * wm * My + ap * (9.0f * MFy - 3.0f * fy * M0) * ir;
    float gPz = wp * (9.0f * MJz - 3.0f * Jz * M0) * ir + 3.0f * wm * Mz + ap * (9.0f * MFz - 3.0f * fz * M0) * ir;

    // Omega_j = w+ ep_j + w- em_j + W_j (a+ xp_j + a- xm_j) with
    //   ep = feq+ - f+, em = feq- - f-, xp = 9 (c.u)(c.F) - 3 u.F, xm = 3 c.F;
    // their mu-weighted sums from the moments of mu:
    //   sum mu ep = M0 (rho - 1 - 1.5 J^2/rho) + 4.5 J.Mcc.J/rho - sum mu f+,
    //   sum mu em = 3 Mc.J - sum mu f-,
    //   sum mu W xp = 9 J.Mcc.F/rho - 3 (J.F/rho) M0,  sum mu W xm = 3 Mc.F.
    float Sfp = 0.0f, Sfm = 0.0f;   // 2 sum over pairs of mu+ f+ and mu- f- terms
#define STRH_FPM(I)                                                                   \
    Sfp += (mu[I] + mu[I + 1]) * (g[I] + g[I + 1]);                                   \
    Sfm += (mu[I] - mu[I + 1]) * (g[I] - g[I + 1]);
    STRH_FPM(1) STRH_FPM(3) STRH_FPM(5) STRH_FPM(7) STRH_FPM(9)
    STRH_FPM(11) STRH_FPM(13) STRH_FPM(15) STRH_FPM(17)
#undef STRH_FPM
    Sfp = mu[0] * g[0] + 0.5f * Sfp;
    Sfm *= 0.5f;
    const float Sep = M0 * (rm1 - 1.5f * J2 * ir) + 4.5f * JMJ * ir - Sfp;
    const float Sem = 3.0f * (Mx * Jx + My * Jy + Mz * Jz) - Sfm;
    const float Sxp = 9.0f * JMF * ir - 3.0f * JF * ir * M0;
    const float Sxm = 3.0f * (Mx * fx + My * fy + Mz * fz);
    float gwp = Sep - 0.5f * Sxp;
    float gwm = Sem - 0.5f * Sxm;

    const float cp = (x == q.inletX ? q.cin : 0.0f) - (x == q.outletX ? q.cout : 0.0f);
    float B = 0.0f, omB = 1.0f;
    bool nt = false;
    if (NT) {
        const int32_t k = ntIdx[n];
        if (k >= 0) {
            // Noble-Torczynski cell (Kernels.metal:201-227), B = eps tw/((1 - eps) + tw)
            // with tw = 1/w+ - 1/2 at the local rate
            nt = true;
            const float e = epsBuf[n];
            const float tw = 1.0f / wp - 0.5f;
            const float den = (1.0f - e) + tw;
            B = e * tw / den;
            omB = 1.0f - B;
            const float muOm = wp * Sep + wm * Sem + ap * Sxp + am * Sxm;
            float gB = -muOm, dNt = 0.0f;
#pragma unroll
            for (int i = 1; i < 19; i++) gB += mu[i] * (g[kOpp[i]] - g[i]);
#pragma unroll
            for (int i = 1; i < 19; i += 2) if (kCx[i]) dNt -= 2.0f * (float)kCx[i] * (g[i + 1] - g[i]);
            const float gBtot = gB + q.wf * dNt;
            gEps[k] += (double)(gBtot * (tw * (1.0f + tw) / (den * den)));
            gwp *= omB;
            gwm *= omB;
            gwp += gBtot * (e * (1.0f - e) / (den * den)) * (-1.0f / (wp * wp));
        }
    }

    // rate chain through tau(Q, rho):
    //   lambda_j += A (c.Pi.c - 2 c.Pi.J/rho) + A (J.Pi.J/rho^2 - tr(Pi)/3) + dJ/drho,
    // A = dJ/dQ / Q; the c.Pi.J term folds into the momentum part below
    const float dwm = q.srt ? -wp * wp : wm * wm * lambda / ((tau - 0.5f) * (tau - 0.5f));
    const float gTau = gwp * (-wp * wp) + gwm * dwm;
    float A = 0.0f, lesC = 0.0f;
    float Lx = 0.0f, Ly = 0.0f, Lz = 0.0f;
    if (Q > 0.0f) {
        const float gQ = gTau * Kc / (4.0f * rho * sq);
        const float gR = -gTau * Kc * Q * ir / (4.0f * rho * sq);
        A = gQ / Q;
        const float PJx = pxx * Jx + pxy * Jy + pxz * Jz;
        const float PJy = pxy * Jx + pyy * Jy + pyz * Jz;
        const float PJz = pxz * Jx + pyz * Jy + pzz * Jz;
        const float JPJ = Jx * PJx + Jy * PJy + Jz * PJz;
        lesC = A * (JPJ * ir2 - (pxx + pyy + pzz) * (1.0f / 3.0f)) + gR;
        const float A2 = 2.0f * A * ir;
        Lx = -A2 * PJx; Ly = -A2 * PJy; Lz = -A2 * PJz;
    }
    // lambda_j = mu_j + (1 - B)(-w+ mu+_j - w- mu-_j + dRho + c_j.dP) + B (mu_opp - mu_j)
    //          + cp + w_f 2 B c_jx + LES terms
    const float wpB = omB * wp, wmB = omB * wm;
    const float base = omB * gRho + cp + lesC;
    const float dObjX = nt ? q.wf * 2.0f * B : 0.0f;      // d ntF_x / d g_j = 2 B c_jx
    gPx = omB * gPx + Lx + dObjX;
    gPy = omB * gPy + Ly;
    gPz = omB * gPz + Lz;
    const float Dxx = A * pxx, Dyy = A * pyy, Dzz = A * pzz;
    const float Dxy = 2.0f * A * pxy, Dxz = 2.0f * A * pxz, Dyz = 2.0f * A * pyz;
    st_off(lam, n4, mu[0] - wpB * mu[0] + base);
#define STRH_LESPAIR(I)                                                                      \
    {                                                                                        \
        constexpr int cx = kCx[I], cy = kCy[I], cz = kCz[I];                                 \
        const float mp = 0.5f * (mu[I] + mu[I + 1]);                                         \
        const float mm = 0.5f * (mu[I] - mu[I + 1]);                                         \
        float cPc = 0.0f, cgp = 0.0f;                                                        \
        if (cx) { cPc += Dxx; cgp += (float)cx * gPx; }                                      \
        if (cy) { cPc += Dyy; cgp += (float)cy * gPy; }                                      \
        if (cz) { cPc += Dzz; cgp += (float)cz * gPz; }                                      \
        if (cx * cy) cPc += (float)(cx * cy) * Dxy;                                          \
        if (cx * cz) cPc += (float)(cx * cz) * Dxz;                                          \
        if (cy * cz) cPc += (float)(cy * cz) * Dyz;                                          \
        const float common = base - wpB * mp + cPc;                                          \
        const float odd = cgp - wmB * mm;                                                    \
        float v0 = mu[I] + common + odd, v1 = mu[I + 1] + common - odd;                      \
        if (NT) {                                                                            \
            const float bb = B * (mu[I + 1] - mu[I]);                                        \
            v0 += bb;                                                                        \
            v1 -= bb;                                                                        \
        }                                                                                    \
        st_off(lam + (size_t)I * S, n4, v0);                                                  \
        st_off(lam + (size_t)(I + 1) * S, n4, v1);                                     \
    }
    STRH_LESPAIR(1) STRH_LESPAIR(3) STRH_LESPAIR(5) STRH_LESPAIR(7) STRH_LESPAIR(9)
    STRH_LESPAIR(11) STRH_LESPAIR(13) STRH_LESPAIR(15) STRH_LESPAIR(17)
#undef STRH_LESPAIR
    if (gu != 0.0) gUin[n] += gu;
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

// ---- moments probe: Kernels.metal:359-380 (literal loop order) ----------
template <class St>
__global__ void __launch_bounds__(256)
moments_kernel(const typename St::T* __restrict__ f, const uint8_t* __restrict__ flags,
               float4* __restrict__ out, uint32_t N, DdfLayout lay, float fx, float fy, float fz) {
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= N) return;
    if (flags[n] != kFluid) { out[n] = make_float4(0.0f, 0.0f, 0.0f, 0.0f); return; }
    const size_t e = cell_elem(lay, n);
    float rhom1 = 0.0f, px = 0.0f, py = 0.0f, pz = 0.0f;
#pragma unroll
    for (int i = 0; i < 19; i++) {
        const float v = St::ld(f[e + (size_t)i * lay.W]);   // plain load: probe path
        rhom1 += v;
        px += (float)kCx[i] * v;
        py += (float)kCy[i] * v;
        pz += (float)kCz[i] * v;
    }
    const float rho = 1.0f + rhom1;
    out[n] = make_float4((px + 0.5f * fx) / rho, (py + 0.5f * fy) / rho,
                         (pz + 0.5f * fz) / rho, rho);
}

// ---- field initialiser: Kernels.metal:386-453 ----------------------------
template <class St>
__global__ void __launch_bounds__(256)
init_kernel(typename St::T* __restrict__ f, uint32_t nx, uint32_t ny, uint32_t nz,
            uint32_t N, DdfLayout lay, uint32_t mode, float amplitude, float tau) {
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= N) return;
    const size_t e = cell_elem(lay, n);
    const float PI = 3.14159265358979323846f;
    float ux = 0.0f, uy = 0.0f, uz = 0.0f;
    float sxx = 0.0f, rhom1 = 0.0f, sxz = 0.0f, syz = 0.0f;
    if (mode == 1u) {
        const uint32_t x = n % nx, t = n / nx, y = t % ny;
        const float kx = 2.0f * PI / (float)nx;
        const float ky = 2.0f * PI / (float)ny;
        const float xa = (float)x + 0.5f, ya = (float)y + 0.5f;
        ux = amplitude * sinf(kx * xa) * cosf(ky * ya);
        uy = -amplitude * cosf(kx * xa) * sinf(ky * ya);
        sxx = amplitude * kx * cosf(kx * xa) * cosf(ky * ya);
        rhom1 = 0.75f * amplitude * amplitude * (cosf(2.0f * kx * xa) + cosf(2.0f * ky * ya));
    }
    if (mode == 2u) {
        const uint32_t x = n % nx, t = n / nx, y = t % ny, z = t / ny;
        const float kx = 2.0f * PI / (float)nx;
        const float ky = 2.0f * PI / (float)ny;
        const float kz = 2.0f * PI / (float)nz;
        const float xa = kx * ((float)x + 0.5f);
        const float ya = ky * ((float)y + 0.5f);
        const float za = kz * ((float)z + 0.5f);
        ux = amplitude * sinf(xa) * cosf(ya) * cosf(za);
        uy = -amplitude * cosf(xa) * sinf(ya) * cosf(za);
        sxx = amplitude * kx * cosf(xa) * cosf(ya) * cosf(za);
        sxz = -0.5f * amplitude * kz * sinf(xa) * cosf(ya) * sinf(za);
        syz = 0.5f * amplitude * kz * cosf(xa) * sinf(ya) * sinf(za);
        rhom1 = 0.1875f * amplitude * amplitude
              * (cosf(2.0f * xa) + cosf(2.0f * ya)) * (cosf(2.0f * za) + 2.0f);
    }
    const float rho = 1.0f + rhom1;
    const float u2 = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int i = 0; i < 19; i++) {
        const float W = i == 0 ? (1.0f / 3.0f) : (i <= 6 ? (1.0f / 18.0f) : (1.0f / 36.0f));
        const float cu = 3.0f * ((float)kCx[i] * ux + (float)kCy[i] * uy + (float)kCz[i] * uz);
        const float feq = W * (rhom1 + rho * (cu + 0.5f * cu * cu - 1.5f * u2));
        const float cc = (float)(kCx[i] * kCx[i] - kCy[i] * kCy[i]);
        const float scc = sxx * cc + 2.0f * sxz * (float)(kCx[i] * kCz[i])
                        + 2.0f * syz * (float)(kCy[i] * kCz[i]);
        const float fneq = -3.0f * W * tau * scc;
        f[e + (size_t)i * lay.W] = St::st(feq + fneq);
    }
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
__device__ __forceinline__ double block_tree_sum(double v, double* lds) {
    lds[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) lds[threadIdx.x] = lds[threadIdx.x] + lds[threadIdx.x + s];
        __syncthreads();
    }
    return lds[0];
}

template <class Elem>
__global__ void __launch_bounds__(256)
reduce_stage1(Elem e, uint32_t M, uint32_t chunk, double* __restrict__ partials) {
    __shared__ double lds[256];
    const uint32_t lo = blockIdx.x * chunk;
    const uint32_t hi = (M - lo) > chunk ? lo + chunk : M;
    double acc = 0.0;
    for (uint32_t k = lo + threadIdx.x; k < hi; k += 256u) acc += e(blockIdx.y, k);
    const double s = block_tree_sum(acc, lds);
    if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = s;
}

__global__ void __launch_bounds__(256)
reduce_stage2(const double* __restrict__ partials, uint32_t P, double* __restrict__ out) {
    __shared__ double lds[256];
    double acc = 0.0;
    for (uint32_t k = threadIdx.x; k < P; k += 256u) acc += partials[k];
    const double s = block_tree_sum(acc, lds);
    if (threadIdx.x == 0) out[0] = s;
}

template <class St> struct MassElem {
    const typename St::T* f;
    DdfLayout lay;
    __device__ double operator()(uint32_t slot, uint32_t n) const {
        return (double)St::ld(f[cell_elem(lay, n) + (size_t)slot * lay.W]);
    }
};

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

struct KineticElem {
    const float4* mom;
    __device__ double operator()(uint32_t, uint32_t k) const {
        const float4 v = mom[k];
        return (double)(v.x * v.x + v.y * v.y + v.z * v.z);
    }
};

// Viscous dissipation of one cell, 2 nu S:S, with the strain rate from the
// non-equilibrium momentum flux of the natural-order state (Chapman-Enskog:
// Pi_neq = -2 rho c_s^2 tau+ S). With shifted DDFs sum c_a c_b f~ = Pi - delta/3,
// so Pi_neq = sum c_a c_b f~ - rho u_a u_b - (rho - 1)/3 delta_ab. FP64
// throughout; non-fluid cells give 0. No body-force correction (the
// dissipation observable is for force-free flows).
template <class St> struct DissElem {
    const typename St::T* f;
    const uint8_t* flags;     // null: all fluid
    DdfLayout lay;
    double tauP, twoNu;
    __device__ double operator()(uint32_t, uint32_t k) const {
        if (flags && flags[k] != kFluid) return 0.0;
        const size_t e = cell_elem(lay, k);
        double rm1 = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
        double xx = 0.0, yy = 0.0, zz = 0.0, xy = 0.0, xz = 0.0, yz = 0.0;
#pragma unroll
        for (int i = 0; i < 19; i++) {
            const double v = (double)St::ld(f[e + (size_t)i * lay.W]);
            rm1 += v;
            if (kCx[i]) { jx += (double)kCx[i] * v; xx += v; }
            if (kCy[i]) { jy += (double)kCy[i] * v; yy += v; }
            if (kCz[i]) { jz += (double)kCz[i] * v; zz += v; }
            if (kCx[i] * kCy[i]) xy += (double)(kCx[i] * kCy[i]) * v;
            if (kCx[i] * kCz[i]) xz += (double)(kCx[i] * kCz[i]) * v;
            if (kCy[i] * kCz[i]) yz += (double)(kCy[i] * kCz[i]) * v;
        }
        const double rho = 1.0 + rm1, ir = 1.0 / rho;
        const double third = rm1 / 3.0;
        const double pxx = xx - jx * jx * ir - third, pyy = yy - jy * jy * ir - third;
        const double pzz = zz - jz * jz * ir - third;
        const double pxy = xy - jx * jy * ir, pxz = xz - jx * jz * ir, pyz = yz - jy * jz * ir;
        const double c = -1.5 / (rho * tauP);          // S = c Pi_neq
        const double ss = pxx * pxx + pyy * pyy + pzz * pzz + 2.0 * (pxy * pxy + pxz * pxz + pyz * pyz);
        return twoNu * c * c * ss;
    }
};

__global__ void __launch_bounds__(256) fill_zero_kernel(uint32_t* __restrict__ p, uint64_t words) {
    const uint64_t k = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (k < words) p[k] = 0u;
}

// ---- multi-lattice batch ---------------------------------------------------
// One launch advances every lattice of a batch by one step. Workgroup b
// belongs to the lattice whose block range holds it; each lattice keeps its
// own step index (and so its own AA parity), ramp and probe schedule. The
// per-cell arithmetic is the single-lattice kernel's (walls, NT and LES
// enabled are exact no-ops on lattices without them), so a batched lattice
// evolves bit for bit as it would alone.
__device__ __forceinline__ int batch_find(const BatchLat* lats, int L, uint32_t b) {
    int l = 0;
    while (l + 1 < L && b >= lats[l + 1].block0) l++;
    return l;
}

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

// PAR = 0 / 1: the host guarantees that every active lattice is at that
// parity in this launch (equal start parities, even graph lengths), so only
// that pass is compiled in; PAR = -1 branches per lattice. LES = false when no
// lattice of the batch has a Smagorinsky constant, FORCED = false when none has
// a body force (otherwise each lattice takes its own forcing path).
template <class Conv, int BS, int PAR, bool LES, bool FORCED>
__global__ void __launch_bounds__(BS)
step_batch(const BatchLat* __restrict__ lats, int L, const int64_t* __restrict__ base, int k) {
    const uint32_t b = blockIdx.x;
    const int l = batch_find(lats, L, b);
    const BatchLat& B = lats[l];
    const int64_t t = B.step0 + *base + k;
    if (t >= B.tEnd) return;
    const uint32_t n = (b - B.block0) * BS + threadIdx.x;
    if (n >= B.p.N) return;
    StepParams p = B.p;
    p.parity = PAR >= 0 ? PAR : (int)(t & 1);
    const float r = t < B.rampSteps ? B.ramp[t] : 1.0f;      // Sim.swift:258-266
    p.uin = B.uinTarget * r;
    p.ulidX = B.lid[0] * r;
    p.ulidY = B.lid[1] * r;
    p.ulidZ = B.lid[2] * r;
    p.writeForce = is_force_step(B, t) ? 1 : 0;
    float4* fo = p.writeForce ? force_slot(B, t) : B.force;
    using T = typename Conv::T;
    T* f = (T*)B.f;
    if (PAR != 1 && p.parity == 0) {
        // Batched lattices are small enough to stay in the last-level cache
        // between passes: the default policy (measured against nontemporal
        // loads / stores on the street ladder).
        using St = Store<Conv, kPolShifted>;
        if (FORCED && B.hasForce) step_cell<St, 0, true, true, LES, true>(f, B.flags, B.link, B.lidm, fo, B.eps, p, n);
        else step_cell<St, 0, true, true, LES, false>(f, B.flags, B.link, B.lidm, fo, B.eps, p, n);
    } else if (PAR != 0) {
        using St = Store<Conv, kPolShifted>;
        if (FORCED && B.hasForce) step_cell<St, 1, true, true, LES, true>(f, B.flags, B.link, B.lidm, fo, B.eps, p, n);
        else step_cell<St, 1, true, true, LES, false>(f, B.flags, B.link, B.lidm, fo, B.eps, p, n);
    }
}

// After batch steps k .. k + span - 1 (span <= kProbeSpan): for every lattice
// and every force step t among them, sum the per-cell momentum exchange of
// that step's ring buffer over the probe box into the time series. Block
// (l, c) sums component c of lattice l with exactly the chunking of the
// single-lattice probe (reduce_stage1 over Pb chunks, then reduce_stage2), so
// the FP64 values are the same bits.
__global__ void __launch_bounds__(256)
probe_batch(BatchLat* __restrict__ lats, int L, const int64_t* __restrict__ base, int k, int span) {
    __shared__ double lds[256];
    __shared__ double part[kBatchMaxChunks];
    const int l = blockIdx.x / 3, c = blockIdx.x % 3;
    if (l >= L) return;
    const BatchLat& B = lats[l];
    if (B.fCount <= 0) return;
    const int64_t t0 = B.step0 + *base + k;
    int64_t tLast = t0 + span;
    const int64_t tCap = B.fT0 + B.fCount * B.fPeriod;        // one past the last sample's step
    tLast = tLast < B.tEnd ? tLast : B.tEnd;
    tLast = tLast < tCap ? tLast : tCap;
    // first sample step at or after t0
    int64_t t = t0 <= B.fT0 ? B.fT0 : B.fT0 + ((t0 - B.fT0 + B.fPeriod - 1) / B.fPeriod) * B.fPeriod;
    for (; t < tLast; t += B.fPeriod) {
        const BoxForceElem e{force_slot(B, t), (uint32_t)B.p.nx, (uint32_t)B.p.ny, B.bx0, B.by0, B.fbw, B.fbh, c};
        for (uint32_t pb = 0; pb < B.Pb; pb++) {
            const uint32_t lo = pb * B.chunk;
            const uint32_t hi = (B.M - lo) > B.chunk ? lo + B.chunk : B.M;
            double acc = 0.0;
            for (uint32_t q = lo + threadIdx.x; q < hi; q += 256u) acc += e(0, q);
            const double sum = block_tree_sum(acc, lds);
            if (threadIdx.x == 0) part[pb] = sum;
            __syncthreads();
        }
        double acc = 0.0;
        for (uint32_t q = threadIdx.x; q < B.Pb; q += 256u) acc += part[q];
        const double total = block_tree_sum(acc, lds);
        if (threadIdx.x == 0) B.series[3 * ((t - B.fT0) / B.fPeriod) + c] = total;
        __syncthreads();
    }
}

// ---- temporally blocked batch step for lattices one cell deep ------------
// A lattice with nz = 1 advances K steps per launch. A workgroup owns an
// IW x IH patch of cells (IW = TW - 2K, IH = TH - 2K). It loads the patch with
// a K-cell halo, a TW x TH tile wrapped periodically like the step's
// neighbours, into LDS in the storage format, runs the K AA steps there and
// writes the patch back. Step j updates the tile cells at distance >= j from
// the tile edge: an even (in-place) step needs nothing around a cell, an odd
// step pulls from and pushes to the cell's 8 in-plane neighbours, so after
// step j the cells at distance >= j + 1 hold exactly the values the
// lattice-wide pass gives them, and after K steps the patch does. Each cell's
// arithmetic is the one-step batch kernel's (walls, moving walls, partial
// cells, sponge, momentum exchange), and every value passes through the
// storage format between steps as it does in HBM, so the result is the same
// bits. Halo cells of one tile are patch cells of another, so a launch reads
// one image and writes the other (`swap` picks the direction); the host pairs
// the launches so the state is back in `f` after each group.
__device__ __forceinline__ int batch_find_tile(const BatchLat* lats, int L, uint32_t b) {
    int l = 0;
    while (l + 1 < L && b >= lats[l + 1].tile0) l++;
    return l;
}

// Step t of lattice B with the ramped boundary speeds and the probe flag.
__device__ __forceinline__ void batch_params(StepParams& p, const BatchLat& B, int64_t t) {
    const float r = t < B.rampSteps ? B.ramp[t] : 1.0f;      // Sim.swift:258-266
    p.uin = B.uinTarget * r;
    p.ulidX = B.lid[0] * r;
    p.ulidY = B.lid[1] * r;
    p.ulidZ = B.lid[2] * r;
    p.writeForce = is_force_step(B, t) ? 1 : 0;
}

// The per-cell side arrays of a tiled lattice, as global-memory pointers.
struct TileSide {
    const GLOBAL uint8_t* flags;
    const GLOBAL uint32_t* link;
    const GLOBAL uint32_t* lidm;
    const GLOBAL float* eps;
};

// The even pass of one tile cell (step_cell, PARITY 0): own slots in, swapped
// slots out.
template <class Conv, int C>
__device__ __forceinline__ void tile_even(typename Conv::T* __restrict__ lds, const TileSide& sd,
                                          float4* fo, const StepParams& p, uint32_t n, int cl, int gx,
                                          bool own) {
    if (sd.link[n] & kNotFluidBit) return;      // bit 31: the cell is not fluid
    const float eps = p.useEps ? sd.eps[n] : 0.0f;
    float fh[19];
#pragma unroll
    for (int i = 0; i < 19; i++) fh[i] = Conv::ld(lds[i * C + cl]);
    float ntx = 0.0f, nty = 0.0f, ntz = 0.0f;
    collide<false, false, true>(fh, p, gx, eps, ntx, nty, ntz);
    lds[cl] = Conv::st(fh[0]);
#pragma unroll
    for (int i = 1; i < 19; i++) lds[kOpp[i] * C + cl] = Conv::st(fh[i]);
    if (own && p.writeForce != 0 && eps > 0.0f) fo[n] = make_float4(ntx, nty, ntz, 0.0f);
}

// The odd pass of one tile cell (step_cell, PARITY 1): pull from the in-plane
// neighbours (own slot I across a wall link), the moving-wall corrections,
// collide, push (own opposite slot across a wall link), momentum exchange.
// Every access stays inside the tile for cells at distance >= 1 from its edge.
template <class Conv, int TW, int C>
__device__ __forceinline__ void tile_odd(typename Conv::T* __restrict__ lds, const TileSide& sd,
                                         float4* fo, const StepParams& p, uint32_t n, int cl, int gx, int gy,
                                         bool own) {
    const uint32_t lw = sd.link[n];
    if (lw & kNotFluidBit) return;
    const uint32_t sM = lw & 0x3ffffu;
    const uint32_t lM = sM != 0u ? sd.lidm[n] : 0u;
    const float eps = p.useEps ? sd.eps[n] : 0.0f;
    const int nx = p.nx, ny = p.ny;
    float fh[19];
    fh[0] = Conv::ld(lds[cl]);
#define STRH_TPULL(I)                                                                 \
    {                                                                                 \
        constexpr int srcBit = kOpp[I] - 1;                                           \
        const int a = ((sM >> srcBit) & 1u) ? I * C + cl                              \
                                             : kOpp[I] * C + cl - kCx[I] - kCy[I] * TW; \
        fh[I] = Conv::ld(lds[a]);                                                     \
    }
    STRH_TPULL(1) STRH_TPULL(2) STRH_TPULL(3) STRH_TPULL(4) STRH_TPULL(5) STRH_TPULL(6)
    STRH_TPULL(7) STRH_TPULL(8) STRH_TPULL(9) STRH_TPULL(10) STRH_TPULL(11) STRH_TPULL(12)
    STRH_TPULL(13) STRH_TPULL(14) STRH_TPULL(15) STRH_TPULL(16) STRH_TPULL(17) STRH_TPULL(18)
#undef STRH_TPULL
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    if (sM != 0u) {
        uint32_t deferred = 0u;
#define STRH_TWALLIN(I)                                                               \
        {                                                                             \
            constexpr int srcBit = kOpp[I] - 1;                                       \
            if ((sM >> srcBit) & 1u) {                                                \
                if ((lM >> srcBit) & 1u) deferred |= 1u << I;                         \
                if (p.writeForce != 0) {                                              \
                    ax = acc<-kCx[I]>(ax, fh[I]);                                     \
                    ay = acc<-kCy[I]>(ay, fh[I]);                                     \
                    az = acc<-kCz[I]>(az, fh[I]);                                     \
                }                                                                     \
            }                                                                         \
        }
        STRH_TWALLIN(1) STRH_TWALLIN(2) STRH_TWALLIN(3) STRH_TWALLIN(4) STRH_TWALLIN(5) STRH_TWALLIN(6)
        STRH_TWALLIN(7) STRH_TWALLIN(8) STRH_TWALLIN(9) STRH_TWALLIN(10) STRH_TWALLIN(11) STRH_TWALLIN(12)
        STRH_TWALLIN(13) STRH_TWALLIN(14) STRH_TWALLIN(15) STRH_TWALLIN(16) STRH_TWALLIN(17) STRH_TWALLIN(18)
#undef STRH_TWALLIN
        if (deferred != 0u) {
            Offsets o;   // nz = 1: the z neighbour is the cell itself
            o.xm = (gx == 0) ? nx - 1 : -1;
            o.xp = (gx == nx - 1) ? 1 - nx : 1;
            o.ym = (gy == 0) ? (ny - 1) * nx : -nx;
            o.yp = (gy == ny - 1) ? -(ny - 1) * nx : nx;
            o.zm = 0;
            o.zp = 0;
            const int ni = (int)n;
#define STRH_TDEFER(I)                                                                \
            if ((deferred >> I) & 1u) {                                               \
                const int sy = gy - kCy[I] < 0 ? ny - 1 : (gy - kCy[I] >= ny ? 0 : gy - kCy[I]); \
                const uint8_t wf = sd.flags[(uint32_t)(ni + o.at<-kCx[I], -kCy[I], -kCz[I]>())]; \
                if (wf == kOutflow) {                                                 \
                    fh[I] = -fh[I];                                                   \
                } else {                                                              \
                    float uwx, uwy, uwz;                                              \
                    wall_vel(wf, sy, p, uwx, uwy, uwz);                               \
                    const float cu = (float)kCx[I] * uwx + (float)kCy[I] * uwy + (float)kCz[I] * uwz; \
                    const float corr = 6.0f * wgt<I>() * cu;                          \
                    fh[I] += corr;                                                    \
                    if (p.writeForce != 0) {                                          \
                        ax = acc<-kCx[I]>(ax, corr);                                  \
                        ay = acc<-kCy[I]>(ay, corr);                                  \
                        az = acc<-kCz[I]>(az, corr);                                  \
                    }                                                                 \
                }                                                                     \
            }
            STRH_TDEFER(1) STRH_TDEFER(2) STRH_TDEFER(3) STRH_TDEFER(4) STRH_TDEFER(5) STRH_TDEFER(6)
            STRH_TDEFER(7) STRH_TDEFER(8) STRH_TDEFER(9) STRH_TDEFER(10) STRH_TDEFER(11) STRH_TDEFER(12)
            STRH_TDEFER(13) STRH_TDEFER(14) STRH_TDEFER(15) STRH_TDEFER(16) STRH_TDEFER(17) STRH_TDEFER(18)
#undef STRH_TDEFER
        }
    }
    float ntx = 0.0f, nty = 0.0f, ntz = 0.0f;
    collide<false, false, true>(fh, p, gx, eps, ntx, nty, ntz);
    ax += ntx;
    ay += nty;
    az += ntz;
    if (sM != 0u) {
        Offsets o;
        o.xm = (gx == 0) ? nx - 1 : -1;
        o.xp = (gx == nx - 1) ? 1 - nx : 1;
        o.ym = (gy == 0) ? (ny - 1) * nx : -nx;
        o.yp = (gy == ny - 1) ? -(ny - 1) * nx : nx;
        o.zm = 0;
        o.zp = 0;
        const int ni = (int)n;
#define STRH_TWALLOUT(I)                                                              \
        {                                                                             \
            constexpr int dstBit = I - 1;                                             \
            if ((sM >> dstBit) & 1u) {                                                \
                float corr = 0.0f;                                                    \
                float sign = 1.0f;                                                    \
                if ((lM >> dstBit) & 1u) {                                            \
                    const int dy = gy + kCy[I] < 0 ? ny - 1 : (gy + kCy[I] >= ny ? 0 : gy + kCy[I]); \
                    const uint8_t wf = sd.flags[(uint32_t)(ni + o.at<kCx[I], kCy[I], kCz[I]>())]; \
                    if (wf == kOutflow) {                                             \
                        sign = -1.0f;                                                 \
                    } else {                                                          \
                        float uwx, uwy, uwz;                                          \
                        wall_vel(wf, dy, p, uwx, uwy, uwz);                           \
                        corr = 6.0f * wgt<I>()                                        \
                             * ((float)kCx[I] * uwx + (float)kCy[I] * uwy + (float)kCz[I] * uwz); \
                    }                                                                 \
                }                                                                     \
                if (p.writeForce != 0) {                                              \
                    const float v = sign * fh[I] + 2.0f * wgt<I>();                   \
                    ax = acc<kCx[I]>(ax, v);                                          \
                    ay = acc<kCy[I]>(ay, v);                                          \
                    az = acc<kCz[I]>(az, v);                                          \
                }                                                                     \
                fh[I] = sign * fh[I] - corr;                                          \
            }                                                                         \
        }
        STRH_TWALLOUT(1) STRH_TWALLOUT(2) STRH_TWALLOUT(3) STRH_TWALLOUT(4) STRH_TWALLOUT(5) STRH_TWALLOUT(6)
        STRH_TWALLOUT(7) STRH_TWALLOUT(8) STRH_TWALLOUT(9) STRH_TWALLOUT(10) STRH_TWALLOUT(11) STRH_TWALLOUT(12)
        STRH_TWALLOUT(13) STRH_TWALLOUT(14) STRH_TWALLOUT(15) STRH_TWALLOUT(16) STRH_TWALLOUT(17) STRH_TWALLOUT(18)
#undef STRH_TWALLOUT
    }
    // The push addresses are the pull addresses of the opposite directions;
    // a pinned copy of the cell index makes them recomputed here instead of
    // kept across the collision.
    int cp = cl;
    asm("" : "+v"(cp));
    lds[cp] = Conv::st(fh[0]);
#define STRH_TPUSH(I)                                                                 \
    {                                                                                 \
        constexpr int dstBit = I - 1;                                                 \
        const int a = ((sM >> dstBit) & 1u) ? kOpp[I] * C + cp                        \
                                             : I * C + cp + kCx[I] + kCy[I] * TW;     \
        lds[a] = Conv::st(fh[I]);                                                     \
    }
    STRH_TPUSH(1) STRH_TPUSH(2) STRH_TPUSH(3) STRH_TPUSH(4) STRH_TPUSH(5) STRH_TPUSH(6)
    STRH_TPUSH(7) STRH_TPUSH(8) STRH_TPUSH(9) STRH_TPUSH(10) STRH_TPUSH(11) STRH_TPUSH(12)
    STRH_TPUSH(13) STRH_TPUSH(14) STRH_TPUSH(15) STRH_TPUSH(16) STRH_TPUSH(17) STRH_TPUSH(18)
#undef STRH_TPUSH
    if (own && p.writeForce != 0 && (sM != 0u || eps > 0.0f)) fo[n] = make_float4(ax, ay, az, 0.0f);
}

// Step J of a tile launch: the cells at distance >= J from the tile edge,
// row-major over that region, BS at a time.
template <class Conv, int TW, int TH, int K, int BS, int J>
__device__ __forceinline__ void tile_step(typename Conv::T* __restrict__ lds, const BatchLat& B,
                                          const TileSide& sd, int64_t t0, int x0, int y0) {
    constexpr int C = TW * TH;
    constexpr int RW = TW - 2 * J, RH = TH - 2 * J;
    StepParams p = B.p;
    batch_params(p, B, t0 + J);
    float4* fo = p.writeForce ? force_slot(B, t0 + J) : B.force;
    const int nx = p.nx, ny = p.ny;
#pragma unroll 1
    for (int c = threadIdx.x; c < RW * RH; c += BS) {
        const int yl = c / RW + J, xl = c % RW + J;
        const int cl = yl * TW + xl;
        int gx = x0 + xl, gy = y0 + yl;
        // a patch cell (distance >= K) inside the lattice: its force sample is this tile's
        const bool own = xl >= K && xl < TW - K && yl >= K && yl < TH - K && gx < nx && gy < ny;
        gx += gx < 0 ? nx : 0;
        gx -= gx >= nx ? nx : 0;
        gy += gy < 0 ? ny : 0;
        gy -= gy >= ny ? ny : 0;
        const uint32_t n = (uint32_t)(gy * nx + gx);
        if ((J & 1) == 0) tile_even<Conv, C>(lds, sd, fo, p, n, cl, gx, own);
        else tile_odd<Conv, TW, C>(lds, sd, fo, p, n, cl, gx, gy, own);
    }
    __syncthreads();
    if constexpr (J + 1 < K) tile_step<Conv, TW, TH, K, BS, J + 1>(lds, B, sd, t0, x0, y0);
}

// The host launches tiles only for whole K-step blocks inside a batch segment
// (every active lattice at an even step, K even, no LES, no body force,
// nz = 1, nx >= TW, ny >= TH), and reads the state from `f` after an even
// number of launches.
template <class Conv, int TW, int TH, int K, int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4)))
step_tiles(const BatchLat* __restrict__ lats, int L, const int64_t* __restrict__ base, int k, int swap) {
    using T = typename Conv::T;
    constexpr int C = TW * TH;
    constexpr int IW = TW - 2 * K, IH = TH - 2 * K;
    static_assert(K % 2 == 0 && IW > 0 && IH > 0, "tile shape");
    __shared__ T lds[19 * C];
    const uint32_t b = blockIdx.x;
    const BatchLat& B = lats[batch_find_tile(lats, L, b)];
    const int64_t t0 = B.step0 + *base + k;
    if (t0 >= B.tEnd) return;
    const uint32_t tb = b - B.tile0;
    const uint32_t ty = tb / B.tilesX, tx = tb - ty * B.tilesX;
    const int x0 = (int)(tx * IW) - K, y0 = (int)(ty * IH) - K;
    const int nx = B.p.nx, ny = B.p.ny;
    const DdfLayout lay = B.p.lay;
    const TileSide sd{(const GLOBAL uint8_t*)B.flags, (const GLOBAL uint32_t*)B.link,
                      (const GLOBAL uint32_t*)B.lidm, (const GLOBAL float*)B.eps};
    const GLOBAL T* __restrict__ src = (const GLOBAL T*)(swap ? B.f2 : B.f);
    GLOBAL T* __restrict__ dst = (GLOBAL T*)(swap ? B.f : B.f2);
#pragma unroll 1
    for (int c = threadIdx.x; c < C; c += BS) {
        const int yl = c / TW, xl = c % TW;
        int gx = x0 + xl, gy = y0 + yl;
        gx += gx < 0 ? nx : 0;
        gx -= gx >= nx ? nx : 0;
        gy += gy < 0 ? ny : 0;
        gy -= gy >= ny ? ny : 0;
        const uint32_t e = cell_elem(lay, (uint32_t)(gy * nx + gx));
        T v[19];
#pragma unroll
        for (int i = 0; i < 19; i++) v[i] = src[e + (uint32_t)i * lay.W];
#pragma unroll
        for (int i = 0; i < 19; i++) lds[i * C + c] = v[i];
    }
    __syncthreads();
    tile_step<Conv, TW, TH, K, BS, 0>(lds, B, sd, t0, x0, y0);
#pragma unroll 1
    for (int c = threadIdx.x; c < IW * IH; c += BS) {
        const int yi = c / IW, xi = c % IW;
        const int gx = (int)(tx * IW) + xi, gy = (int)(ty * IH) + yi;
        if (gx >= nx || gy >= ny) continue;
        const int cl = (yi + K) * TW + xi + K;
        const uint32_t e = cell_elem(lay, (uint32_t)(gy * nx + gx));
        T v[19];
#pragma unroll
        for (int i = 0; i < 19; i++) v[i] = lds[i * C + cl];
#pragma unroll
        for (int i = 0; i < 19; i++) dst[e + (uint32_t)i * lay.W] = v[i];
    }
}

__global__ void advance_base(int64_t* base, int64_t by) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *base += by;
}

__global__ void set_base(int64_t* base, int64_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *base = v;
}

// ---- launchers -------------------------------------------------------------
template <class Conv, int PARITY, bool WALLS, bool NT, bool LES, bool FORCE>
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
        hipLaunchKernelGGL((step_kernel<St, PARITY, WALLS, false, false, true, true>), dim3(p.blocks),
                           dim3(kStepBS), 0, s, (T*)L.f, L.flags, L.link, L.lidm,
                           (float4*)L.force, L.eps, p, L.alpha);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((step_kernel<St, PARITY, WALLS, NT, LES, FORCE>), dim3(p.blocks),
                       dim3(kStepBS), 0, s, (T*)L.f, L.flags, L.link, L.lidm,
                       (float4*)L.force, L.eps, p, (const float*)nullptr);
    return hipGetLastError();
}

template <class Conv, int PARITY, bool WALLS, bool NT, bool LES>
static hipError_t pick_force(const StepLaunch& L, hipStream_t s) {
    return L.has_force ? launch_step_t<Conv, PARITY, WALLS, NT, LES, true>(L, s)
                       : launch_step_t<Conv, PARITY, WALLS, NT, LES, false>(L, s);
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

hipError_t launch_step(const StepLaunch& L, hipStream_t s) {
    switch (L.precision) {
        case kFP32: return pick_parity<ConvF32>(L, s);
        case kFP16S: return pick_parity<ConvF16>(L, s);
        case kBF16: return pick_parity<ConvBF16>(L, s);
    }
    return hipErrorInvalidValue;
}

template <class Conv, int PARITY, bool WALLS>
static hipError_t launch_hooked_t(const StepLaunch& L, const AdjHooks& hk, hipStream_t s) {
    using T = typename Conv::T;
    using St = Store<Conv, PARITY == 0 ? kPolAligned : kPolShifted>;
    StepParams p = L.p;
    p.blocks = (p.N + kStepBS - 1) / kStepBS;
    // porosity excludes partial cells and LES
#define STRH_HOOKED(NTV, LESV, BR) hipLaunchKernelGGL((step_kernel_hooked<St, PARITY, WALLS, NTV, LESV, BR>), \
        dim3(p.blocks), dim3(kStepBS), 0, s, (T*)L.f, L.flags, L.link, L.lidm, L.eps, p, L.alpha, hk)
    if (L.alpha) STRH_HOOKED(false, false, true);
    else if (L.nt && L.les) STRH_HOOKED(true, true, false);
    else if (L.nt) STRH_HOOKED(true, false, false);
    else STRH_HOOKED(false, true, false);
#undef STRH_HOOKED
    return hipGetLastError();
}

template <class Conv>
static hipError_t launch_hooked_c(const StepLaunch& L, const AdjHooks& hk, hipStream_t s) {
    if (L.p.parity == 0)
        return L.walls ? launch_hooked_t<Conv, 0, true>(L, hk, s) : launch_hooked_t<Conv, 0, false>(L, hk, s);
    return L.walls ? launch_hooked_t<Conv, 1, true>(L, hk, s) : launch_hooked_t<Conv, 1, false>(L, hk, s);
}

hipError_t launch_step_hooked(const StepLaunch& L, const AdjHooks& hk, hipStream_t s) {
    // porosity alone, or partial cells and/or LES
    if (L.alpha ? (L.nt || L.les) : !(L.nt || L.les)) return hipErrorInvalidValue;
    switch (L.precision) {
        case kFP32: return launch_hooked_c<ConvF32>(L, hk, s);
        case kFP16S: return launch_hooked_c<ConvF16>(L, hk, s);
        case kBF16: return launch_hooked_c<ConvBF16>(L, hk, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_adjoint_step(const AdjStep& q0, bool walls, const float4* rec, const float* lamNext,
                               float* lam, const uint8_t* flags, const uint32_t* link,
                               const float* alpha, double* gAlpha, double* gUin,
                               const float* eps, const int32_t* ntIdx, const float* recNt,
                               double* gEps, const void* rec19, const uint32_t* lidm,
                               const StepParams& sp, hipStream_t s) {
    AdjStep q = q0;
    q.blocks = (q.N + 255u) / 256u;
    const dim3 g(q.blocks), b(256);
    if (rec19) {
#define STRH_ADJL(C, W, NTV) hipLaunchKernelGGL((adjoint_les_kernel<C, W, NTV>), g, b, 0, s,              \
                                                (const typename C::T*)rec19, lamNext, lam, flags, link,   \
                                                lidm, gUin, eps, ntIdx, gEps, q, sp)
#define STRH_ADJC(C)                                                                                  \
        if (ntIdx) { if (walls) STRH_ADJL(C, true, true); else STRH_ADJL(C, false, true); }           \
        else { if (walls) STRH_ADJL(C, true, false); else STRH_ADJL(C, false, false); }
        switch (q.precision) {
            case kFP32: STRH_ADJC(ConvF32); break;
            case kFP16S: STRH_ADJC(ConvF16); break;
            case kBF16: STRH_ADJC(ConvBF16); break;
            default: return hipErrorInvalidValue;
        }
#undef STRH_ADJC
#undef STRH_ADJL
        return hipGetLastError();
    }
#define STRH_ADJ(W, NTV) hipLaunchKernelGGL((adjoint_step_kernel<W, NTV>), g, b, 0, s, rec, lamNext, lam, \
                                            flags, link, alpha, gAlpha, gUin, eps, ntIdx, recNt, gEps, q)
    if (ntIdx) {
        if (walls) STRH_ADJ(true, true); else STRH_ADJ(false, true);
    } else {
        if (walls) STRH_ADJ(true, false); else STRH_ADJ(false, false);
    }
#undef STRH_ADJ
    return hipGetLastError();
}

// Row staging applies when nx is a power of two that tiles a 256- or
// 512-thread workgroup and the lattice is a whole number of workgroups.
uint32_t rows_block_size(uint32_t nx, uint64_t N) {
    if (nx < 64 || (nx & (nx - 1)) != 0) return 0;
    const uint32_t bs = nx <= 256 ? 256u : (nx == 512 ? 512u : 0u);
    if (bs == 0 || N % bs != 0) return 0;
    return bs;
}

static inline uint32_t grid_for(uint64_t n) { return (uint32_t)((n + 255u) / 256u); }

hipError_t launch_moments(int precision, const void* f, const uint8_t* flags, float* out,
                          uint32_t N, DdfLayout lay, float fx, float fy, float fz, hipStream_t s) {
    const dim3 g(grid_for(N)), b(256);
    switch (precision) {
        case kFP32: hipLaunchKernelGGL(moments_kernel<StF32>, g, b, 0, s, (const float*)f, flags, (float4*)out, N, lay, fx, fy, fz); break;
        case kFP16S: hipLaunchKernelGGL(moments_kernel<StF16>, g, b, 0, s, (const uint16_t*)f, flags, (float4*)out, N, lay, fx, fy, fz); break;
        case kBF16: hipLaunchKernelGGL(moments_kernel<StBF16>, g, b, 0, s, (const uint16_t*)f, flags, (float4*)out, N, lay, fx, fy, fz); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_init(int precision, void* f, uint32_t nx, uint32_t ny, uint32_t nz,
                       DdfLayout lay, uint32_t mode, float amplitude, float tau, hipStream_t s) {
    const uint32_t N = nx * ny * nz;
    const dim3 g(grid_for(N)), b(256);
    switch (precision) {
        case kFP32: hipLaunchKernelGGL(init_kernel<StF32>, g, b, 0, s, (float*)f, nx, ny, nz, N, lay, mode, amplitude, tau); break;
        case kFP16S: hipLaunchKernelGGL(init_kernel<StF16>, g, b, 0, s, (uint16_t*)f, nx, ny, nz, N, lay, mode, amplitude, tau); break;
        case kBF16: hipLaunchKernelGGL(init_kernel<StBF16>, g, b, 0, s, (uint16_t*)f, nx, ny, nz, N, lay, mode, amplitude, tau); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

void box_reduce_plan(uint32_t M, uint32_t max_partials, uint32_t* Pb, uint32_t* chunk);

template <class Elem>
static hipError_t reduce(const Elem& e, uint32_t M, uint32_t segs, double* partials,
                         uint32_t max_partials, double* out, hipStream_t s) {
    uint32_t Pb, chunk;                              // >= 16 elements per thread
    box_reduce_plan(M, max_partials / segs, &Pb, &chunk);
    hipLaunchKernelGGL(reduce_stage1<Elem>, dim3(Pb, segs), dim3(256), 0, s, e, M, chunk, partials);
    hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(256), 0, s, partials, Pb * segs, out);
    return hipGetLastError();
}

hipError_t launch_mass_sum(int precision, const void* f, uint32_t N, DdfLayout lay,
                           double* partials, uint32_t max_partials, double* out, hipStream_t s) {
    switch (precision) {
        case kFP32: return reduce(MassElem<StF32>{(const float*)f, lay}, N, 19, partials, max_partials, out, s);
        case kFP16S: return reduce(MassElem<StF16>{(const uint16_t*)f, lay}, N, 19, partials, max_partials, out, s);
        case kBF16: return reduce(MassElem<StBF16>{(const uint16_t*)f, lay}, N, 19, partials, max_partials, out, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_box_force_sum(const float* force, uint32_t nx, uint32_t ny, uint32_t nz,
                                uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1,
                                double* partials, uint32_t max_partials, double* out3,
                                hipStream_t s) {
    const uint32_t bw = x1 - x0 + 1, bh = y1 - y0 + 1;
    const uint32_t M = bw * bh * nz;
    for (int c = 0; c < 3; c++) {
        BoxForceElem e{(const float4*)force, nx, ny, x0, y0, make_fastdiv(bw), make_fastdiv(bh), c};
        hipError_t err = reduce(e, M, 1, partials, max_partials, out3 + c, s);
        if (err != hipSuccess) return err;
    }
    return hipSuccess;
}

struct SumElem {
    const double* v;
    __device__ double operator()(uint32_t, uint32_t k) const { return v[k]; }
};

hipError_t launch_sum_doubles(const double* v, uint32_t M, double* partials, uint32_t max_partials,
                              double* out, hipStream_t s) {
    return reduce(SumElem{v}, M, 1, partials, max_partials, out, s);
}

hipError_t launch_dissipation_sum(int precision, const void* f, const uint8_t* flags, uint32_t N,
                                  DdfLayout lay, double tauP, double* partials, uint32_t max_partials,
                                  double* out, hipStream_t s) {
    const double twoNu = 2.0 * (tauP - 0.5) / 3.0;
    switch (precision) {
        case kFP32: return reduce(DissElem<StF32>{(const float*)f, flags, lay, tauP, twoNu}, N, 1, partials, max_partials, out, s);
        case kFP16S: return reduce(DissElem<StF16>{(const uint16_t*)f, flags, lay, tauP, twoNu}, N, 1, partials, max_partials, out, s);
        case kBF16: return reduce(DissElem<StBF16>{(const uint16_t*)f, flags, lay, tauP, twoNu}, N, 1, partials, max_partials, out, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_kinetic_sum(const float* moments, uint32_t N, double* partials,
                              uint32_t max_partials, double* out, hipStream_t s) {
    return reduce(KineticElem{(const float4*)moments}, N, 1, partials, max_partials, out, s);
}

hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t s) {
    const uint64_t words = bytes / 4;
    if (words == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((uint32_t)((words + 255) / 256)), dim3(256), 0, s,
                       (uint32_t*)p, words);
    return hipGetLastError();
}

template <class Conv, int BS, int PAR>
static void launch_batch_par(bool les, bool forced, const BatchLat* lats, int L, uint32_t totalBlocks,
                             const int64_t* base, int k, hipStream_t s) {
#define STRH_BATCH(LESV, FV) hipLaunchKernelGGL((step_batch<Conv, BS, PAR, LESV, FV>), dim3(totalBlocks), \
                                                dim3(BS), 0, s, lats, L, base, k)
    if (les) { if (forced) STRH_BATCH(true, true); else STRH_BATCH(true, false); }
    else { if (forced) STRH_BATCH(false, true); else STRH_BATCH(false, false); }
#undef STRH_BATCH
}
template <class Conv, int BS>
static void launch_batch_conv(int parity, bool les, bool forced, const BatchLat* lats, int L,
                              uint32_t totalBlocks, const int64_t* base, int k, hipStream_t s) {
    if (parity == 0) launch_batch_par<Conv, BS, 0>(les, forced, lats, L, totalBlocks, base, k, s);
    else if (parity == 1) launch_batch_par<Conv, BS, 1>(les, forced, lats, L, totalBlocks, base, k, s);
    else hipLaunchKernelGGL((step_batch<Conv, BS, -1, true, true>), dim3(totalBlocks), dim3(BS), 0, s,
                            lats, L, base, k);
}
template <int BS>
static hipError_t launch_batch_bs(int precision, int parity, bool les, bool forced, const BatchLat* lats,
                                  int L, uint32_t totalBlocks, const int64_t* base, int k, hipStream_t s) {
    switch (precision) {
        case kFP32: launch_batch_conv<ConvF32, BS>(parity, les, forced, lats, L, totalBlocks, base, k, s); break;
        case kFP16S: launch_batch_conv<ConvF16, BS>(parity, les, forced, lats, L, totalBlocks, base, k, s); break;
        case kBF16: launch_batch_conv<ConvBF16, BS>(parity, les, forced, lats, L, totalBlocks, base, k, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_batch_step(int precision, int bs, int parity, bool les, bool forced, const BatchLat* lats,
                             int L, uint32_t totalBlocks, const int64_t* base, int k, hipStream_t s) {
    switch (bs) {
        case 128: return launch_batch_bs<128>(precision, parity, les, forced, lats, L, totalBlocks, base, k, s);
        case 256: return launch_batch_bs<256>(precision, parity, les, forced, lats, L, totalBlocks, base, k, s);
        case 512: return launch_batch_bs<512>(precision, parity, les, forced, lats, L, totalBlocks, base, k, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_batch_probe(BatchLat* lats, int L, const int64_t* base, int k, int span, hipStream_t s) {
    hipLaunchKernelGGL(probe_batch, dim3(3 * L), dim3(256), 0, s, lats, L, base, k, span);
    return hipGetLastError();
}

// Tiles of 76 KiB of LDS (two workgroups per CU): 32 x 32 cells in FP32,
// 64 x 32 in the 16-bit formats.
constexpr int kTileBS = 512;

void batch_tile_shape(int precision, int mode, int* tw, int* th, int* K) {
    *tw = precision == kFP32 ? 32 : 64;
    *th = 32;
    *K = mode == 2 ? 2 : 4;
}

template <class Conv, int TW>
static void launch_tiles_t(int mode, const BatchLat* lats, int L, uint32_t tileBlocks, const int64_t* base,
                           int k, int swap, hipStream_t s) {
    if (mode == 2)
        hipLaunchKernelGGL((step_tiles<Conv, TW, 32, 2, kTileBS>), dim3(tileBlocks), dim3(kTileBS), 0, s,
                           lats, L, base, k, swap);
    else
        hipLaunchKernelGGL((step_tiles<Conv, TW, 32, 4, kTileBS>), dim3(tileBlocks), dim3(kTileBS), 0, s,
                           lats, L, base, k, swap);
}

hipError_t launch_batch_tiles(int precision, int mode, const BatchLat* lats, int L, uint32_t tileBlocks,
                              const int64_t* base, int k, int swap, hipStream_t s) {
    if (mode != 1 && mode != 2) return hipErrorInvalidValue;
    switch (precision) {
        case kFP32: launch_tiles_t<ConvF32, 32>(mode, lats, L, tileBlocks, base, k, swap, s); break;
        case kFP16S: launch_tiles_t<ConvF16, 64>(mode, lats, L, tileBlocks, base, k, swap, s); break;
        case kBF16: launch_tiles_t<ConvBF16, 64>(mode, lats, L, tileBlocks, base, k, swap, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_set_base(int64_t* base, int64_t v, hipStream_t s) {
    hipLaunchKernelGGL(set_base, dim3(1), dim3(64), 0, s, base, v);
    return hipGetLastError();
}

hipError_t launch_advance_base(int64_t* base, int64_t by, hipStream_t s) {
    hipLaunchKernelGGL(advance_base, dim3(1), dim3(64), 0, s, base, by);
    return hipGetLastError();
}

void box_reduce_plan(uint32_t M, uint32_t max_partials, uint32_t* Pb, uint32_t* chunk) {
    uint32_t p = (M + 4095u) / 4096u;
    if (p > max_partials) p = max_partials;
    if (p == 0) p = 1;
    *Pb = p;
    *chunk = (uint32_t)(((uint64_t)M + p - 1) / p);
}

}  // namespace strh
