// Recording forward step and reverse (adjoint) collide-stream kernels.
#include "device.h"

namespace strh {

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
// PLANAR: the lattice is one cell deep, so every z shift is the cell's own
// plane and the 18 neighbours are 8 cells.
template <bool WALLS, bool PLANAR>
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
    o.zm = PLANAR ? 0 : (z == 0) ? (nz - 1) * plane : -plane;
    o.zp = PLANAR ? 0 : (z == nz - 1) ? -(nz - 1) * plane : plane;
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

template <bool WALLS, bool NT, bool PLANAR>
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
    const uint32_t z = PLANAR ? 0u : fdiv(t, q.divy);
    const int y = (int)(t - z * q.ny);
    float mu[19];
    const double gu = adj_gather<WALLS, PLANAR>(q, n, x, y, (int)z, lamN, flags, sM, mu);

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
// FORCE = false: the lattice has no body force, and the terms it would carry
// (J = p + F/2, the Guo source's derivatives) are exact zeros left out.
template <class Conv, bool WALLS, bool NT, bool FORCE>
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
    const double gu = adj_gather<WALLS, false>(q, n, x, y, (int)z, lamN, flags, sM, mu);

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
    const float Jx = FORCE ? px + 0.5f * fx : px;
    const float Jy = FORCE ? py + 0.5f * fy : py;
    const float Jz = FORCE ? pz + 0.5f * fz : pz;

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
    const float JMJ = Jx * MJx + Jy * MJy + Jz * MJz;
    const float J2 = Jx * Jx + Jy * Jy + Jz * Jz;
    float gRho = wp * (M0 - 4.5f * JMJ * ir2 + 1.5f * J2 * M0 * ir2);
    float gPx = wp * (9.0f * MJx - 3.0f * Jx * M0) * ir + 3.0f * wm * Mx;
    float gPy = wp * (9.0f * MJy - 3.0f * Jy * M0) * ir + 3.0f * wm * My;
    float gPz = wp * (9.0f * MJz - 3.0f * Jz * M0) * ir + 3.0f * wm * Mz;
    float JMF = 0.0f, JF = 0.0f;
    if (FORCE) {
        const float MFx = Mxx * fx + Mxy * fy + Mxz * fz;
        const float MFy = Mxy * fx + Myy * fy + Myz * fz;
        const float MFz = Mxz * fx + Myz * fy + Mzz * fz;
        JMF = Jx * MFx + Jy * MFy + Jz * MFz;
        JF = Jx * fx + Jy * fy + Jz * fz;
        gRho = gRho - ap * (9.0f * JMF - 3.0f * JF * M0) * ir2;
        gPx = gPx + ap * (9.0f * MFx - 3.0f * fx * M0) * ir;
        gPy = gPy + ap * (9.0f * MFy - 3.0f * fy * M0) * ir;
        gPz = gPz + ap * (9.0f * MFz - 3.0f * fz * M0) * ir;
    }

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
    // the partial-cell sums over the recorded populations, formed here so that
    // the populations are not held through the rate chain below:
    // sum_i mu_i (g_opp(i) - g_i) and the x momentum exchange's -2 c_x (g_i+1 - g_i)
    float sumB = 0.0f, dNt = 0.0f;
    if (NT) {
#pragma unroll
        for (int i = 1; i < 19; i++) sumB += mu[i] * (g[kOpp[i]] - g[i]);
#pragma unroll
        for (int i = 1; i < 19; i += 2) if (kCx[i]) dNt -= 2.0f * (float)kCx[i] * (g[i + 1] - g[i]);
        asm("" : "+v"(sumB), "+v"(dNt));   // formed here, not sunk to their use
    }
    const float Sep = M0 * (rm1 - 1.5f * J2 * ir) + 4.5f * JMJ * ir - Sfp;
    const float Sem = 3.0f * (Mx * Jx + My * Jy + Mz * Jz) - Sfm;
    const float Sxp = FORCE ? 9.0f * JMF * ir - 3.0f * JF * ir * M0 : 0.0f;
    const float Sxm = FORCE ? 3.0f * (Mx * fx + My * fy + Mz * fz) : 0.0f;
    float gwp = FORCE ? Sep - 0.5f * Sxp : Sep;
    float gwm = FORCE ? Sem - 0.5f * Sxm : Sem;

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
            // two divisions (1/w+, 1/den) feed every quotient of the block
            const float rwp = 1.0f / wp;
            const float tw = rwp - 0.5f;
            const float den = (1.0f - e) + tw;
            const float rden = 1.0f / den;
            B = e * tw * rden;
            omB = 1.0f - B;
            const float muOm = FORCE ? wp * Sep + wm * Sem + ap * Sxp + am * Sxm : wp * Sep + wm * Sem;
            const float gB = sumB - muOm;
            const float gBtot = gB + q.wf * dNt;
            gEps[k] += (double)(gBtot * (tw * (1.0f + tw) * (rden * rden)));
            gwp *= omB;
            gwm *= omB;
            gwp -= gBtot * (e * (1.0f - e) * (rden * rden)) * (rwp * rwp);
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


template <class Conv, int PARITY, bool WALLS, bool PLANAR>
static hipError_t launch_hooked_t(const StepLaunch& L, const AdjHooks& hk, hipStream_t s) {
    using T = typename Conv::T;
    using St = Store<Conv, PARITY == 0 ? kPolAligned : kPolShifted>;
    StepParams p = L.p;
    p.blocks = (p.N + kStepBS - 1) / kStepBS;
    // porosity excludes partial cells and LES
#define STRH_HOOKED(NTV, LESV, BR, FO) hipLaunchKernelGGL((step_kernel_hooked<St, PARITY, WALLS, NTV, LESV, BR, FO, PLANAR>), \
        dim3(p.blocks), dim3(kStepBS), 0, s, (T*)L.f, L.flags, L.link, L.lidm, L.eps, p, L.alpha, hk)
    // the forcing path only with a body force, as the plain step's kernel
    // choice (porosity always uses it)
    if (L.alpha) STRH_HOOKED(false, false, true, true);
    else if (L.nt && L.les) { if (L.has_force) STRH_HOOKED(true, true, false, true); else STRH_HOOKED(true, true, false, false); }
    else if (L.nt) { if (L.has_force) STRH_HOOKED(true, false, false, true); else STRH_HOOKED(true, false, false, false); }
    else { if (L.has_force) STRH_HOOKED(false, true, false, true); else STRH_HOOKED(false, true, false, false); }
#undef STRH_HOOKED
    return hipGetLastError();
}

template <class Conv>
static hipError_t launch_hooked_c(const StepLaunch& L, const AdjHooks& hk, hipStream_t s) {
    if (L.p.parity == 0)
        return L.walls ? launch_hooked_t<Conv, 0, true, false>(L, hk, s) : launch_hooked_t<Conv, 0, false, false>(L, hk, s);
    // the odd pass of a lattice one cell deep: 8 neighbour cells
    if (L.planar)
        return L.walls ? launch_hooked_t<Conv, 1, true, true>(L, hk, s) : launch_hooked_t<Conv, 1, false, true>(L, hk, s);
    return L.walls ? launch_hooked_t<Conv, 1, true, false>(L, hk, s) : launch_hooked_t<Conv, 1, false, false>(L, hk, s);
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
#define STRH_ADJL(C, W, NTV, FO) hipLaunchKernelGGL((adjoint_les_kernel<C, W, NTV, FO>), g, b, 0, s,      \
                                                    (const typename C::T*)rec19, lamNext, lam, flags, link, \
                                                    lidm, gUin, eps, ntIdx, gEps, q, sp)
#define STRH_ADJF(C, W, NTV) if (forced) STRH_ADJL(C, W, NTV, true); else STRH_ADJL(C, W, NTV, false);
#define STRH_ADJC(C)                                                                                  \
        if (ntIdx) { if (walls) STRH_ADJF(C, true, true) else STRH_ADJF(C, false, true) }             \
        else { if (walls) STRH_ADJF(C, true, false) else STRH_ADJF(C, false, false) }
        const bool forced = !(q.fx == 0.0f && q.fy == 0.0f && q.fz == 0.0f);
        switch (q.precision) {
            case kFP32: STRH_ADJC(ConvF32); break;
            case kFP16S: STRH_ADJC(ConvF16); break;
            case kBF16: STRH_ADJC(ConvBF16); break;
            default: return hipErrorInvalidValue;
        }
#undef STRH_ADJC
#undef STRH_ADJF
#undef STRH_ADJL
        return hipGetLastError();
    }
#define STRH_ADJ(W, NTV, PL) hipLaunchKernelGGL((adjoint_step_kernel<W, NTV, PL>), g, b, 0, s, rec, lamNext, lam, \
                                                flags, link, alpha, gAlpha, gUin, eps, ntIdx, recNt, gEps, q)
    const bool planar = q.nz == 1;
    if (ntIdx) {
        if (walls) { if (planar) STRH_ADJ(true, true, true); else STRH_ADJ(true, true, false); }
        else { if (planar) STRH_ADJ(false, true, true); else STRH_ADJ(false, true, false); }
    } else {
        if (walls) { if (planar) STRH_ADJ(true, false, true); else STRH_ADJ(true, false, false); }
        else { if (planar) STRH_ADJ(false, false, true); else STRH_ADJ(false, false, false); }
    }
#undef STRH_ADJ
    return hipGetLastError();
}

}  // namespace strh
