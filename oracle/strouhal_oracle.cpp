// Strouhal CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// A scalar restatement of the reference's fused D3Q19 collide-and-stream
// update, used only by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg as the checker. Nothing in the product links it.
//
// It follows, line by line:
//   Sources/StrouhalCore/Kernels.metal:39-47   direction tables
//   Sources/StrouhalCore/Kernels.metal:71-87   fidx / wrap / wallVel
//   Sources/StrouhalCore/Kernels.metal:105-228 collide
//   Sources/StrouhalCore/Kernels.metal:230-355 step (AA even / odd pass)
//   Sources/StrouhalCore/Kernels.metal:359-380 momentsEven
//   Sources/StrouhalCore/Kernels.metal:393-453 initField
//   Sources/StrouhalCore/Sim.swift:188-193     trtOmegas (in Python host code)
//   Sources/StrouhalCore/Sim.swift:227-255     neighbour-mask builder
//   Sources/StrouhalCore/Sim.swift:258-281     ramp / params
//   Sources/StrouhalCore/Sim.swift:288-331     run (step order, force steps)
//   Sources/StrouhalCore/Sim.swift:359-372     probeForce (fixed-order FP64 sum)
//   Sources/StrouhalCore/Sim.swift:447-458     massSum
//
// Two arithmetic modes: FP64 (storage f64) is "the FP64 oracle"; FP32 with
// f32 / f16 / bf16 storage is "the FP32 oracle". Built with
// -ffp-contract=off so the only fused multiply-add is the explicit one the
// reference writes (Kernels.metal:181).
//
// Parity pins (each function has at least one; see tests/test_oracle_*.py):
//   streaming / AA slots: lone-DDF propagation, bit-exact (Cases.swift:85-125)
//   rest fixed points: bit-exact (Cases.swift:58-83)
//   collide SRT/TRT/Guo: mass/momentum invariants, Poiseuille closed form
//     (Cases.swift:286-317), Taylor-Green decay order (Cases.swift:329-370)
//   walls / moving walls / inflow: Couette closed form, Poiseuille wall
//     force balance, channel flux (Cases.swift:406-437)
//   NT partial cells: eps=0 identity, eps=1 full bounce, force balance
//   LES: Cs=0 identity, closed-form tau_eff on a constructed cell
//   sponge: uniform-sponge Taylor-Green decay rate
//   initField: moments equal the analytic fields, Pi_neq = -(2/3) tau S
//   f16 / bf16 rounding: numpy float16 / torch bfloat16 (IEEE RNE)
//   massSum / moments / force sum: closed forms on constructed states

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <memory>
#include <algorithm>
#include <type_traits>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// Kernels.metal:38-47 — pair order: for i >= 1, opp(i) = i odd ? i+1 : i-1.
static const int CX[19] = {0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 1,-1, 1,-1, 0, 0, 0, 0};
static const int CY[19] = {0, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 0, 0, 0, 0, 1,-1, 1,-1};
static const int CZ[19] = {0, 0, 0, 0, 0, 1,-1, 0, 0, 0, 0, 1,-1,-1, 1, 1,-1,-1, 1};
static const int OPP[19] = {0, 2,1, 4,3, 6,5, 8,7, 10,9, 12,11, 14,13, 16,15, 18,17};

// 18*sqrt(2) as the literal the reference writes (Kernels.metal:156-157).
template <class R> static R smagoConst();
template <> float smagoConst<float>() { return 25.455844f; }
template <> double smagoConst<double>() { return 25.455844; }

template <class R> static R weight(int i) {
    if (i == 0) return R(1) / R(3);
    if (i <= 6) return R(1) / R(18);
    return R(1) / R(36);
}

// Cell flags, Kernels.metal:65-69.
enum : uint8_t { FLUID = 0, SOLID = 1, LID = 2, INFLOW = 3, OUTFLOW = 4 };

// ---- IEEE round-to-nearest-even storage conversions -----------------------
static inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// binary32 -> binary16, round to nearest even. Normal range rounds by adding
// 0xfff plus the kept-mantissa LSB before truncating 13 bits; results that
// are binary16 subnormals are produced by an FPU addition against 2^-1 scaled
// so the hardware rounding (RNE) does the work.
static uint16_t f32_to_f16(float v) {
    uint32_t u = fbits(v);
    const uint32_t sign = (u >> 16) & 0x8000u;
    u &= 0x7fffffffu;
    if (u >= 0x47800000u) {                       // |v| >= 65536 or inf/nan
        if (u > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);
        return (uint16_t)(sign | 0x7c00u);
    }
    if (u < 0x38800000u) {                        // below 2^-14: subnormal half
        // 2^-14 * 2^-10 is the half ulp there; adding 0.5 aligns that ulp to
        // the float LSB of 0.5's mantissa, so the FPU rounds RNE for us.
        const float magic = bitsf(0x3f000000u);   // 0.5
        const float s = bitsf(u) + magic;
        return (uint16_t)(sign | (fbits(s) - 0x3f000000u));
    }
    const uint32_t lsb = (u >> 13) & 1u;
    u += 0xc8000fffu + lsb;                       // rebias exponent (-112<<23) + round
    return (uint16_t)(sign | (u >> 13));
}

static float f16_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu;
    const uint32_t m = h & 0x3ffu;
    if (e == 0) {
        // subnormal or zero: m * 2^-24, exact in binary32
        const float r = (float)m * bitsf(0x33800000u);
        return bitsf(fbits(r) | sign);
    }
    if (e == 31) return bitsf(sign | 0x7f800000u | (m << 13));
    return bitsf(sign | ((e + 112u) << 23) | (m << 13));
}

static uint16_t f32_to_bf16(float v) {
    uint32_t u = fbits(v);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x0040u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static float bf16_to_f32(uint16_t b) { return bitsf((uint32_t)b << 16); }

// Storage traits: how a DDF is held and how it widens to / narrows from the
// arithmetic type ((float)f[...] and (FPXX)fh[i] in Kernels.metal:254-257).
struct StF64 { using T = double;   using R = double; static R ld(T v) { return v; } static T st(R v) { return v; } };
struct StF32 { using T = float;    using R = float;  static R ld(T v) { return v; } static T st(R v) { return v; } };
struct StF16 { using T = uint16_t; using R = float;  static R ld(T v) { return f16_to_f32(v); } static T st(R v) { return f32_to_f16(v); } };
struct StB16 { using T = uint16_t; using R = float;  static R ld(T v) { return bf16_to_f32(v); } static T st(R v) { return f32_to_bf16(v); } };

// Host-side solver state (Sim.swift:164-172), held as double and narrowed to
// the arithmetic type exactly as Swift narrows Double to Float.
struct HostParams {
    double omega, omega_minus;
    double lid[3];
    double uin_target;
    int64_t ramp_steps;
    double force[3];
    double c_smago;
    int32_t inflow_uniform;
    int32_t sponge_on;
    double sponge_x0, sponge_width, sponge_tau;
};

// Time-averaged objective of the differentiable channel (new; no reference
// counterpart): J = (1/T) sum over the T steps after `spinup` of
//   w_p * (mean rho(inlet column) - mean rho(outlet column)) / 3
// + w_d * sum_x alpha(x) J_x(x)       (the Brinkman drag, J = rho u)
// with rho, J the pre-collision moments of each step.
//   + w_f * sum_x ntF_x(x)                (the partial-cell momentum exchange,
//                                          the drag on a Noble-Torczynski body)
struct Objective {
    int32_t inlet_x, outlet_x;
    double w_p, w_d;
    int64_t spinup, avg;
    double w_f;
};

// Per-dispatch block, Kernels.metal:49-63.
template <class R> struct Params {
    int nx, ny, nz, parity;
    R omega, omegaMinus, uin, cSmago, fx, fy, fz;
    int writeForce;
    R ulidX, ulidY, ulidZ;
    int inflowUniform;
    R spongeX0, spongeInvW, spongeTau;
    int useEps;
};

// Kernels.metal:73-77
static inline int wrap(int v, int n) {
    v += (v < 0) ? n : 0;
    v -= (v >= n) ? n : 0;
    return v;
}

struct Base {
    virtual ~Base() {}
    int nx = 0, ny = 0, nz = 0;
    size_t N = 0;
    int64_t stepsDone = 0;
    HostParams hp{};
    std::vector<uint8_t> flags;
    std::vector<uint32_t> smask, lmask;
    bool wantsForces = false;
    int threads = 1;
    virtual void setEps(const float* eps) = 0;
    virtual void setAlpha(const float* alpha) = 0;
    virtual void initField(int mode, double amplitude) = 0;
    virtual void run(int64_t steps, int64_t forceStep) = 0;
    virtual int moments(double* out) = 0;
    virtual void forceField(double* out) = 0;
    virtual void zeroForce() = 0;
    virtual double massSum() = 0;
    virtual size_t ddfBytes() = 0;
    virtual void download(void* out) = 0;
    virtual void upload(const void* in) = 0;
    virtual void collideCell(const double* fin, double eps, int x, double* fout,
                             double* rho, double* ntf) = 0;
    virtual double vjp(const Objective& o, double* gAlpha, double* gEps, double* gUin,
                       int64_t checkpointEvery) = 0;
    virtual double objective(const Objective& o) = 0;
};

template <class S> struct Lattice final : Base {
    using R = typename S::R;
    using T = typename S::T;
    std::vector<T> f;
    std::vector<R> force;   // (x, y, z, 0) per cell, float4 in the reference
    std::vector<R> eps;     // NT solid fraction (Sim.swift:413-421)
    bool useEps = false;
    std::vector<R> alpha;   // Brinkman porosity (new)
    bool useAlpha = false;

    Lattice(int nx_, int ny_, int nz_, const uint8_t* fl, bool wants) {
        nx = nx_; ny = ny_; nz = nz_;
        N = (size_t)nx * ny * nz;
        f.assign(19 * N, S::st(R(0)));      // shifted rest state is exactly 0 (Sim.swift:225)
        wantsForces = wants;
        force.assign(wants ? 4 * N : 4, R(0));
        flags.assign(fl, fl + N);
        buildMasks();
    }

    // Sim.swift:235-255 — bit (i-1): neighbour n + c_i (periodic) is not fluid;
    // lidMask marks the moving-wall family (LID / INFLOW / OUTFLOW).
    void buildMasks() {
        smask.assign(N, 0u);
        lmask.assign(N, 0u);
        for (int z = 0; z < nz; z++) for (int y = 0; y < ny; y++) for (int x = 0; x < nx; x++) {
            const size_t idx = ((size_t)z * ny + y) * nx + x;
            uint32_t s = 0, l = 0;
            if (flags[idx] == FLUID) {
                for (int i = 1; i < 19; i++) {
                    const int xn = (x + CX[i] + nx) % nx;
                    const int yn = (y + CY[i] + ny) % ny;
                    const int zn = (z + CZ[i] + nz) % nz;
                    const uint8_t c = flags[((size_t)zn * ny + yn) * nx + xn];
                    if (c == SOLID) s |= 1u << (i - 1);
                    else if (c == LID || c == INFLOW || c == OUTFLOW) {
                        s |= 1u << (i - 1); l |= 1u << (i - 1);
                    }
                }
            }
            smask[idx] = s; lmask[idx] = l;
        }
    }

    void setAlpha(const float* a) override {
        alpha.resize(N);
        for (size_t n = 0; n < N; n++) alpha[n] = (R)a[n];
        useAlpha = true;
    }

    void setEps(const float* e) override {
        eps.resize(N);
        for (size_t n = 0; n < N; n++) eps[n] = (R)e[n];
        useEps = true;
    }

    // Sim.swift:258-266 — C1 cosine ramp, evaluated in double then narrowed.
    R ramp(int64_t t) const {
        if (hp.ramp_steps <= 0) return R(1);
        const double x = std::fmin(1.0, (double)t / (double)hp.ramp_steps);
        return (R)(0.5 * (1.0 - std::cos(M_PI * x)));
    }

    // Sim.swift:268-281
    Params<R> params(int64_t t, bool writeForce) const {
        Params<R> p;
        const R r = ramp(t);
        p.nx = nx; p.ny = ny; p.nz = nz; p.parity = (int)(t & 1);
        p.omega = (R)hp.omega; p.omegaMinus = (R)hp.omega_minus;
        p.uin = (R)hp.uin_target * r;
        p.cSmago = (R)hp.c_smago;
        p.fx = (R)hp.force[0]; p.fy = (R)hp.force[1]; p.fz = (R)hp.force[2];
        p.writeForce = writeForce ? 1 : 0;
        p.ulidX = (R)hp.lid[0] * r; p.ulidY = (R)hp.lid[1] * r; p.ulidZ = (R)hp.lid[2] * r;
        p.inflowUniform = hp.inflow_uniform ? 1 : 0;
        if (hp.sponge_on) {
            p.spongeX0 = (R)hp.sponge_x0;
            p.spongeInvW = R(1) / (R)hp.sponge_width;
            p.spongeTau = (R)hp.sponge_tau;
        } else {
            p.spongeX0 = (R)1e9; p.spongeInvW = R(0); p.spongeTau = R(1);
        }
        p.useEps = useEps ? 1 : 0;
        return p;
    }

    // Kernels.metal:81-87
    static void wallVel(uint8_t wf, int wy, const Params<R>& p, R& ux, R& uy, R& uz) {
        if (wf == LID) { ux = p.ulidX; uy = p.ulidY; uz = p.ulidZ; return; }
        if (p.inflowUniform != 0) { ux = p.uin; uy = R(0); uz = R(0); return; }
        const R H = (R)p.ny - R(2);
        const R yd = (R)wy - R(0.5);
        const R v = R(4) * p.uin * yd * (H - yd) / (H * H);
        ux = v > R(0) ? v : R(0); uy = R(0); uz = R(0);
    }

    // Kernels.metal:105-228
    // alphav / brink: Brinkman porosity (no reference counterpart): the cell
    // feels the drag force -alpha rho u, with u taken
    // implicitly, u = (p + Fb/2) / (rho (1 + alpha/2)), so the force enters
    // the Guo terms as F = Fb - alpha rho u. With alpha = 0 every value is
    // the same bits as without it.
    static void collide(R* fh, const Params<R>& p, int x, R epsv, R alphav, bool brink,
                        R& rhoOut, R& ntx, R& nty, R& ntz) {
        R wp = p.omega, wm = p.omegaMinus;
        if ((R)x > p.spongeX0) {                                   // 112-123
            const R frac = std::fmin(((R)x - p.spongeX0) * p.spongeInvW, R(1));
            const R tau0 = R(1) / wp;
            const R tauS = tau0 + frac * (p.spongeTau - tau0);
            if (p.omegaMinus == p.omega) {
                wp = R(1) / tauS; wm = wp;
            } else {
                const R lambda = (R(1) / wp - R(0.5)) * (R(1) / wm - R(0.5));
                wp = R(1) / tauS;
                wm = R(1) / (R(0.5) + lambda / (tauS - R(0.5)));
            }
        }
        R rhom1 = R(0);                                            // 124-126
        for (int i = 0; i < 19; i++) rhom1 += fh[i];
        const R rho = R(1) + rhom1;
        rhoOut = rho;
        const R inv = R(1) / rho;
        R px = R(0), py = R(0), pz = R(0);                         // 129-134
        for (int i = 1; i < 19; i++) {
            px += (R)CX[i] * fh[i];
            py += (R)CY[i] * fh[i];
            pz += (R)CZ[i] * fh[i];
        }
        R ux = (px + R(0.5) * p.fx) * inv;                         // 135-138
        R uy = (py + R(0.5) * p.fy) * inv;
        R uz = (pz + R(0.5) * p.fz) * inv;
        R Fx = p.fx, Fy = p.fy, Fz = p.fz;
        if (brink) {
            const R g = R(1) + R(0.5) * alphav;
            ux = ux / g; uy = uy / g; uz = uz / g;
            const R ar = alphav * rho;
            Fx = p.fx - ar * ux; Fy = p.fy - ar * uy; Fz = p.fz - ar * uz;
        }
        const R u2 = ux * ux + uy * uy + uz * uz;

        if (p.cSmago > R(0)) {                                     // 140-167
            R pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
            for (int i = 1; i < 19; i++) {
                const R cu = R(3) * ((R)CX[i] * ux + (R)CY[i] * uy + (R)CZ[i] * uz);
                const R feq = weight<R>(i) * (rhom1 + rho * (cu + R(0.5) * cu * cu - R(1.5) * u2));
                const R d = fh[i] - feq;
                pxx += (R)(CX[i] * CX[i]) * d;
                pyy += (R)(CY[i] * CY[i]) * d;
                pzz += (R)(CZ[i] * CZ[i]) * d;
                pxy += (R)(CX[i] * CY[i]) * d;
                pxz += (R)(CX[i] * CZ[i]) * d;
                pyz += (R)(CY[i] * CZ[i]) * d;
            }
            const R Q = std::sqrt(pxx * pxx + pyy * pyy + pzz * pzz
                                  + R(2) * (pxy * pxy + pxz * pxz + pyz * pyz));
            const R tau0 = R(1) / wp;
            const R tauEff = R(0.5) * (tau0 + std::sqrt(tau0 * tau0 + smagoConst<R>() * p.cSmago * Q * inv));
            if (p.omegaMinus == p.omega) {
                wp = R(1) / tauEff; wm = wp;
            } else {
                const R lambda = (R(1) / wp - R(0.5)) * (R(1) / wm - R(0.5));
                wp = R(1) / tauEff;
                wm = R(1) / (R(0.5) + lambda / (tauEff - R(0.5)));
            }
        }

        const R uF = ux * Fx + uy * Fy + uz * Fz;                   // 169-171
        const R ap = R(1) - R(0.5) * wp;
        const R am = R(1) - R(0.5) * wm;

        if (epsv == R(0)) {                                        // 177-199
            {
                const R feq0 = weight<R>(0) * (rhom1 - R(1.5) * rho * u2);
                const R s0 = weight<R>(0) * ap * (R(-3) * uF);
                fh[0] = std::fma(wp, feq0 - fh[0], fh[0]) + s0;
            }
            for (int i = 1; i < 19; i += 2) {
                const R W = weight<R>(i);
                const R cu = R(3) * ((R)CX[i] * ux + (R)CY[i] * uy + (R)CZ[i] * uz);
                const R cF = (R)CX[i] * Fx + (R)CY[i] * Fy + (R)CZ[i] * Fz;
                const R feqp = W * (rhom1 + rho * (R(0.5) * cu * cu - R(1.5) * u2));
                const R feqm = W * rho * cu;
                const R fp = R(0.5) * (fh[i] + fh[i + 1]);
                const R fm = R(0.5) * (fh[i] - fh[i + 1]);
                const R dp = wp * (feqp - fp);
                const R dm = wm * (feqm - fm);
                const R sm = W * am * (R(3) * cF);
                const R sp = W * ap * (R(3) * cu * cF - R(3) * uF);
                fh[i] += dp + dm + sp + sm;
                fh[i + 1] += dp - dm + sp - sm;
            }
            return;
        }

        const R tw = R(1) / wp - R(0.5);                           // 201-227
        const R B = epsv * tw / ((R(1) - epsv) + tw);
        const R oneMinusB = R(1) - B;
        {
            const R feq0 = weight<R>(0) * (rhom1 - R(1.5) * rho * u2);
            const R s0 = weight<R>(0) * ap * (R(-3) * uF);
            fh[0] += oneMinusB * (wp * (feq0 - fh[0]) + s0);
        }
        for (int i = 1; i < 19; i += 2) {
            const R W = weight<R>(i);
            const R cu = R(3) * ((R)CX[i] * ux + (R)CY[i] * uy + (R)CZ[i] * uz);
            const R cF = (R)CX[i] * Fx + (R)CY[i] * Fy + (R)CZ[i] * Fz;
            const R feqp = W * (rhom1 + rho * (R(0.5) * cu * cu - R(1.5) * u2));
            const R feqm = W * rho * cu;
            const R fp = R(0.5) * (fh[i] + fh[i + 1]);
            const R fm = R(0.5) * (fh[i] - fh[i + 1]);
            const R dp = wp * (feqp - fp);
            const R dm = wm * (feqm - fm);
            const R sm = W * am * (R(3) * cF);
            const R sp = W * ap * (R(3) * cu * cF - R(3) * uF);
            const R bounce = B * (fh[i + 1] - fh[i]);
            fh[i] += oneMinusB * (dp + dm + sp + sm) + bounce;
            fh[i + 1] += oneMinusB * (dp - dm + sp - sm) - bounce;
            ntx -= R(2) * (R)CX[i] * bounce;
            nty -= R(2) * (R)CY[i] * bounce;
            ntz -= R(2) * (R)CZ[i] * bounce;
        }
    }

    inline size_t fidx(int i, size_t n) const { return (size_t)i * N + n; }

    // Kernels.metal:230-355, one cell.
    void stepCell(size_t n, const Params<R>& p) {
        if (flags[n] != FLUID) return;
        const R epsv = p.useEps ? eps[n] : R(0);
        const R alv = useAlpha ? alpha[n] : R(0);
        const int x = (int)(n % (size_t)nx);
        const size_t t = n / (size_t)nx;
        const int y = (int)(t % (size_t)ny);
        const int z = (int)(t / (size_t)ny);
        R fh[19];
        R rho = R(1);
        R ntx = 0, nty = 0, ntz = 0;
        if (p.parity == 0) {                                       // 252-260
            for (int i = 0; i < 19; i++) fh[i] = S::ld(f[fidx(i, n)]);
            collide(fh, p, x, epsv, alv, useAlpha, rho, ntx, nty, ntz);
            f[fidx(0, n)] = S::st(fh[0]);
            for (int i = 1; i < 19; i++) f[fidx(OPP[i], n)] = S::st(fh[i]);
            if (p.writeForce != 0 && epsv > R(0)) {
                force[4 * n + 0] = ntx; force[4 * n + 1] = nty;
                force[4 * n + 2] = ntz; force[4 * n + 3] = R(0);
            }
            return;
        }
        const uint32_t sM = smask[n], lM = lmask[n];               // 268-291
        R ax = 0, ay = 0, az = 0;
        uint32_t deferred = 0;
        fh[0] = S::ld(f[fidx(0, n)]);
        for (int i = 1; i < 19; i++) {
            const int srcBit = OPP[i] - 1;
            if ((sM >> srcBit) & 1u) {
                fh[i] = S::ld(f[fidx(i, n)]);
                if ((lM >> srcBit) & 1u) deferred |= 1u << i;
                if (p.writeForce != 0) {
                    ax -= (R)CX[i] * fh[i];
                    ay -= (R)CY[i] * fh[i];
                    az -= (R)CZ[i] * fh[i];
                }
            } else {
                const int sx = wrap(x - CX[i], nx);
                const int sy = wrap(y - CY[i], ny);
                const int sz = wrap(z - CZ[i], nz);
                const size_t s = ((size_t)sz * ny + sy) * nx + sx;
                fh[i] = S::ld(f[fidx(OPP[i], s)]);
            }
        }
        if (deferred != 0) {                                       // 292-314
            for (int i = 1; i < 19; i++) {
                if ((deferred >> i) & 1u) {
                    const int sy = wrap(y - CY[i], ny);
                    const size_t s = ((size_t)wrap(z - CZ[i], nz) * ny + sy) * nx + wrap(x - CX[i], nx);
                    const uint8_t wf = flags[s];
                    if (wf == OUTFLOW) {
                        fh[i] = -fh[i];
                    } else {
                        R uwx, uwy, uwz;
                        wallVel(wf, sy, p, uwx, uwy, uwz);
                        const R cu = (R)CX[i] * uwx + (R)CY[i] * uwy + (R)CZ[i] * uwz;
                        const R corr = R(6) * weight<R>(i) * cu;
                        fh[i] += corr;
                        if (p.writeForce != 0) {
                            ax -= (R)CX[i] * corr;
                            ay -= (R)CY[i] * corr;
                            az -= (R)CZ[i] * corr;
                        }
                    }
                }
            }
        }
        collide(fh, p, x, epsv, alv, useAlpha, rho, ntx, nty, ntz);  // 315-316
        ax += ntx; ay += nty; az += ntz;
        f[fidx(0, n)] = S::st(fh[0]);                              // 317-350
        for (int i = 1; i < 19; i++) {
            const int dstBit = i - 1;
            if ((sM >> dstBit) & 1u) {
                R corr = R(0);
                R sign = R(1);
                if ((lM >> dstBit) & 1u) {
                    const int dy = wrap(y + CY[i], ny);
                    const size_t d = ((size_t)wrap(z + CZ[i], nz) * ny + dy) * nx + wrap(x + CX[i], nx);
                    const uint8_t wf = flags[d];
                    if (wf == OUTFLOW) {
                        sign = R(-1);
                    } else {
                        R uwx, uwy, uwz;
                        wallVel(wf, dy, p, uwx, uwy, uwz);
                        corr = R(6) * weight<R>(i)
                             * ((R)CX[i] * uwx + (R)CY[i] * uwy + (R)CZ[i] * uwz);
                    }
                }
                f[fidx(OPP[i], n)] = S::st(sign * fh[i] - corr);
                if (p.writeForce != 0) {
                    const R v = sign * fh[i] + R(2) * weight<R>(i);
                    ax += (R)CX[i] * v;
                    ay += (R)CY[i] * v;
                    az += (R)CZ[i] * v;
                }
            } else {
                const int dx = wrap(x + CX[i], nx);
                const int dy = wrap(y + CY[i], ny);
                const int dz = wrap(z + CZ[i], nz);
                const size_t d = ((size_t)dz * ny + dy) * nx + dx;
                f[fidx(i, d)] = S::st(fh[i]);
            }
        }
        if (p.writeForce != 0 && (sM != 0u || epsv > R(0))) {      // 351-353
            force[4 * n + 0] = ax; force[4 * n + 1] = ay;
            force[4 * n + 2] = az; force[4 * n + 3] = R(0);
        }
    }

    // Sim.swift:288-331 (chunking and in-flight limits do not change results).
    void run(int64_t steps, int64_t forceStep) override {
        for (int64_t s = 0; s < steps; s++) {
            const int64_t t = stepsDone + s;
            const Params<R> p = params(t, t == forceStep);
#ifdef _OPENMP
            #pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
#endif
            for (long long n = 0; n < (long long)N; n++) stepCell((size_t)n, p);
        }
        stepsDone += steps;
    }

    // Kernels.metal:359-380 — (ux, uy, uz, rho); valid after an even count.
    int moments(double* out) override {
        if (stepsDone % 2 != 0) return -1;
        const Params<R> p = params(stepsDone, false);
        for (size_t n = 0; n < N; n++) {
            if (flags[n] != FLUID) { out[4 * n] = out[4 * n + 1] = out[4 * n + 2] = out[4 * n + 3] = 0.0; continue; }
            R rhom1 = 0, px = 0, py = 0, pz = 0;
            for (int i = 0; i < 19; i++) {
                const R v = S::ld(f[fidx(i, n)]);
                rhom1 += v;
                px += (R)CX[i] * v;
                py += (R)CY[i] * v;
                pz += (R)CZ[i] * v;
            }
            const R rho = R(1) + rhom1;
            out[4 * n + 0] = (double)((px + R(0.5) * p.fx) / rho);
            out[4 * n + 1] = (double)((py + R(0.5) * p.fy) / rho);
            out[4 * n + 2] = (double)((pz + R(0.5) * p.fz) / rho);
            out[4 * n + 3] = (double)rho;
        }
        return 0;
    }

    // Kernels.metal:393-453 (tau = 1/omega, Sim.swift:432).
    void initField(int mode, double amplitude) override {
        const R A = (R)amplitude;
        const R tau = R(1) / (R)hp.omega;
        const R PI = (R)M_PI;
#ifdef _OPENMP
        #pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
#endif
        for (long long nn = 0; nn < (long long)N; nn++) {
            const size_t n = (size_t)nn;
            R ux = 0, uy = 0, uz = 0, sxx = 0, rhom1 = 0, sxz = 0, syz = 0;
            if (mode == 1) {
                const size_t x = n % (size_t)nx, t = n / (size_t)nx, y = t % (size_t)ny;
                const R kx = R(2) * PI / (R)nx;
                const R ky = R(2) * PI / (R)ny;
                const R xa = (R)x + R(0.5), ya = (R)y + R(0.5);
                ux = A * std::sin(kx * xa) * std::cos(ky * ya);
                uy = -A * std::cos(kx * xa) * std::sin(ky * ya);
                sxx = A * kx * std::cos(kx * xa) * std::cos(ky * ya);
                rhom1 = R(0.75) * A * A * (std::cos(R(2) * kx * xa) + std::cos(R(2) * ky * ya));
            }
            if (mode == 2) {
                const size_t x = n % (size_t)nx, t = n / (size_t)nx, y = t % (size_t)ny, z = t / (size_t)ny;
                const R kx = R(2) * PI / (R)nx;
                const R ky = R(2) * PI / (R)ny;
                const R kz = R(2) * PI / (R)nz;
                const R xa = kx * ((R)x + R(0.5));
                const R ya = ky * ((R)y + R(0.5));
                const R za = kz * ((R)z + R(0.5));
                ux = A * std::sin(xa) * std::cos(ya) * std::cos(za);
                uy = -A * std::cos(xa) * std::sin(ya) * std::cos(za);
                sxx = A * kx * std::cos(xa) * std::cos(ya) * std::cos(za);
                sxz = R(-0.5) * A * kz * std::sin(xa) * std::cos(ya) * std::sin(za);
                syz = R(0.5) * A * kz * std::cos(xa) * std::sin(ya) * std::sin(za);
                rhom1 = R(0.1875) * A * A * (std::cos(R(2) * xa) + std::cos(R(2) * ya)) * (std::cos(R(2) * za) + R(2));
            }
            const R rho = R(1) + rhom1;
            const R u2 = ux * ux + uy * uy + uz * uz;
            for (int i = 0; i < 19; i++) {
                const R cu = R(3) * ((R)CX[i] * ux + (R)CY[i] * uy + (R)CZ[i] * uz);
                const R feq = weight<R>(i) * (rhom1 + rho * (cu + R(0.5) * cu * cu - R(1.5) * u2));
                const R cc = (R)(CX[i] * CX[i] - CY[i] * CY[i]);
                const R scc = sxx * cc + R(2) * sxz * (R)(CX[i] * CZ[i]) + R(2) * syz * (R)(CY[i] * CZ[i]);
                const R fneq = R(-3) * weight<R>(i) * tau * scc;
                f[fidx(i, n)] = S::st(feq + fneq);
            }
        }
    }

    void forceField(double* out) override {
        if (!wantsForces) return;
        for (size_t k = 0; k < 4 * N; k++) out[k] = (double)force[k];
    }
    void zeroForce() override { std::fill(force.begin(), force.end(), R(0)); }

    // Sim.swift:447-458 — double sum over all 19N shifted DDFs in index order.
    double massSum() override {
        double total = 0.0;
        for (size_t k = 0; k < 19 * N; k++) total += (double)S::ld(f[k]);
        return total;
    }
    size_t ddfBytes() override { return 19 * N * sizeof(T); }
    void download(void* out) override { std::memcpy(out, f.data(), ddfBytes()); }
    void upload(const void* in) override { std::memcpy(f.data(), in, ddfBytes()); }

    // ---- differentiable objective and its discrete adjoint (FP64 only) -----
    // Pre-collision populations of every fluid cell for the step about to run
    // (absolute step t): the in-cell slots after an even count, the odd-pass
    // gather (neighbour pull, bounce-back, moving-wall corrections) otherwise.
    void gatherState(const Params<R>& p, std::vector<double>& g) const {
        g.assign(19 * N, 0.0);
#ifdef _OPENMP
        #pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
#endif
        for (long long nn = 0; nn < (long long)N; nn++) {
            const size_t n = (size_t)nn;
            if (flags[n] != FLUID) continue;
            const int x = (int)(n % (size_t)nx);
            const size_t t = n / (size_t)nx;
            const int y = (int)(t % (size_t)ny), z = (int)(t / (size_t)ny);
            if (p.parity == 0) {
                for (int i = 0; i < 19; i++) g[i * N + n] = (double)S::ld(f[fidx(i, n)]);
                continue;
            }
            g[n] = (double)S::ld(f[fidx(0, n)]);
            for (int i = 1; i < 19; i++) {
                const int srcBit = OPP[i] - 1;
                R v;
                if ((smask[n] >> srcBit) & 1u) {
                    v = S::ld(f[fidx(i, n)]);
                    if ((lmask[n] >> srcBit) & 1u) {
                        const int sy = wrap(y - CY[i], ny);
                        const size_t s = ((size_t)wrap(z - CZ[i], nz) * ny + sy) * nx + wrap(x - CX[i], nx);
                        if (flags[s] == OUTFLOW) {
                            v = -v;
                        } else {
                            R uwx, uwy, uwz;
                            wallVel(flags[s], sy, p, uwx, uwy, uwz);
                            v += R(6) * weight<R>(i) * ((R)CX[i] * uwx + (R)CY[i] * uwy + (R)CZ[i] * uwz);
                        }
                    }
                } else {
                    const size_t s = ((size_t)wrap(z - CZ[i], nz) * ny + wrap(y - CY[i], ny)) * nx + wrap(x - CX[i], nx);
                    v = S::ld(f[fidx(OPP[i], s)]);
                }
                g[i * N + n] = (double)v;
            }
        }
    }

    // Column weights of the pressure-drop term: +1/(3 n_in), -1/(3 n_out).
    void pressureWeights(const Objective& o, std::vector<double>& cp) const {
        cp.assign(N, 0.0);
        size_t nin = 0, nout = 0;
        for (size_t n = 0; n < N; n++) {
            if (flags[n] != FLUID) continue;
            const int x = (int)(n % (size_t)nx);
            if (x == o.inlet_x) nin++;
            if (x == o.outlet_x) nout++;
        }
        for (size_t n = 0; n < N; n++) {
            if (flags[n] != FLUID) continue;
            const int x = (int)(n % (size_t)nx);
            if (x == o.inlet_x && nin) cp[n] += 1.0 / (3.0 * (double)nin);
            if (x == o.outlet_x && nout) cp[n] -= 1.0 / (3.0 * (double)nout);
        }
    }

    // Partial-cell bounce weight B = eps tw / ((1 - eps) + tw), tw = 1/w+ - 1/2
    // (Kernels.metal:201-203), and dB/deps.
    static double ntB(double e, double wp) {
        const double tw = 1.0 / wp - 0.5;
        return e * tw / ((1.0 - e) + tw);
    }
    static double ntdB(double e, double wp) {
        const double tw = 1.0 / wp - 0.5;
        const double d = (1.0 - e) + tw;
        return tw * (1.0 + tw) / (d * d);
    }

    // The relaxation rates a cell collides with (Kernels.metal:112-167): the
    // sponge ramp by x, then the Smagorinsky rate from Q = |Pi_neq|, with the
    // pieces the adjoint differentiates. g: the 19 populations, stride `st`.
    struct LocalRates {
        double wp, wm;
        bool les = false, srt = true;
        double tau = 0.0, s = 0.0, K = 0.0, Q = 0.0, lambda = 0.0;
        double Pi[6] = {0, 0, 0, 0, 0, 0};   // xx yy zz xy xz yz
    };
    LocalRates localRates(const double* g, size_t st, int x, const Params<R>& p) const {
        LocalRates r;
        double wp = (double)p.omega, wm = (double)p.omegaMinus;
        r.srt = p.omegaMinus == p.omega;
        if ((double)x > (double)p.spongeX0) {
            const double frac = std::fmin(((double)x - (double)p.spongeX0) * (double)p.spongeInvW, 1.0);
            const double tau0 = 1.0 / wp;
            const double tauS = tau0 + frac * ((double)p.spongeTau - tau0);
            if (r.srt) { wp = 1.0 / tauS; wm = wp; }
            else {
                const double lam = (1.0 / wp - 0.5) * (1.0 / wm - 0.5);
                wp = 1.0 / tauS;
                wm = 1.0 / (0.5 + lam / (tauS - 0.5));
            }
        }
        if ((double)p.cSmago > 0.0) {
            double rho = 1.0, pv[3] = {0, 0, 0};
            for (int i = 0; i < 19; i++) {
                const double v = g[(size_t)i * st];
                rho += v; pv[0] += CX[i] * v; pv[1] += CY[i] * v; pv[2] += CZ[i] * v;
            }
            const double rm1 = rho - 1.0;
            const double ux = (pv[0] + 0.5 * (double)p.fx) / rho, uy = (pv[1] + 0.5 * (double)p.fy) / rho,
                         uz = (pv[2] + 0.5 * (double)p.fz) / rho;
            const double u2 = ux * ux + uy * uy + uz * uz;
            double* P = r.Pi;
            for (int i = 1; i < 19; i++) {
                const double cu = 3.0 * (CX[i] * ux + CY[i] * uy + CZ[i] * uz);
                const double d = g[(size_t)i * st] - weight<double>(i) * (rm1 + rho * (cu + 0.5 * cu * cu - 1.5 * u2));
                P[0] += CX[i] * CX[i] * d; P[1] += CY[i] * CY[i] * d; P[2] += CZ[i] * CZ[i] * d;
                P[3] += CX[i] * CY[i] * d; P[4] += CX[i] * CZ[i] * d; P[5] += CY[i] * CZ[i] * d;
            }
            r.Q = std::sqrt(P[0] * P[0] + P[1] * P[1] + P[2] * P[2] + 2.0 * (P[3] * P[3] + P[4] * P[4] + P[5] * P[5]));
            r.K = smagoConst<double>() * (double)p.cSmago;
            const double tau0 = 1.0 / wp;
            r.s = std::sqrt(tau0 * tau0 + r.K * r.Q / rho);
            r.tau = 0.5 * (tau0 + r.s);
            r.les = true;
            if (r.srt) { wp = 1.0 / r.tau; wm = wp; }
            else {
                r.lambda = (1.0 / wp - 0.5) * (1.0 / wm - 0.5);
                wp = 1.0 / r.tau;
                wm = 1.0 / (0.5 + r.lambda / (r.tau - 0.5));
            }
        }
        r.wp = wp; r.wm = wm;
        return r;
    }

    double cellObjective(const double* g, double a, double e, const Params<R>& p, double cpn,
                         const Objective& o, int x) const {
        double rho = 1.0, px = 0.0;
        for (int i = 0; i < 19; i++) { rho += g[i]; px += CX[i] * g[i]; }
        const double gam = 1.0 + 0.5 * a;
        const double Jx = (px + 0.5 * (double)p.fx) / gam;
        double v = o.w_p * cpn * rho + o.w_d * a * Jx;
        if (e != 0.0 && o.w_f != 0.0) {                 // ntF_x = -2 sum c_x B (g_opp - g)
            const double B = ntB(e, localRates(g, 1, x, p).wp);
            double nt = 0.0;
            for (int i = 1; i < 19; i += 2) nt -= 2.0 * CX[i] * B * (g[i + 1] - g[i]);
            v += o.w_f * nt;
        }
        return v;
    }

    bool adjointSupported() const {
        return !(useAlpha && (useEps || hp.c_smago > 0.0));
    }

    // Objective terms of step k's pre-collision state g, summed over the
    // fluid cells in index order.
    double windowSum(const std::vector<double>& g, const Params<R>& p, const std::vector<double>& cp,
                     const Objective& o) const {
        double sum = 0.0;
        double gc[19];
        for (size_t n = 0; n < N; n++) {
            if (flags[n] != FLUID) continue;
            for (int i = 0; i < 19; i++) gc[i] = g[i * N + n];
            sum += cellObjective(gc, useAlpha ? (double)alpha[n] : 0.0, useEps ? (double)eps[n] : 0.0,
                                 p, cp[n], o, (int)(n % (size_t)nx));
        }
        return sum;
    }

    double objective(const Objective& o) override {
        if (!std::is_same<R, double>::value || !adjointSupported()) return std::nan("");
        std::vector<double> g, cp;
        pressureWeights(o, cp);
        double J = 0.0;
        const int64_t total = o.spinup + o.avg;
        for (int64_t k = 0; k < total; k++) {
            const Params<R> p = params(stepsDone, false);
            if (k >= o.spinup) {
                gatherState(p, g);
                J += windowSum(g, p, cp, o) / (double)o.avg;
            }
            run(1, -1);
        }
        return J;
    }

    // Discrete adjoint of `total = spinup + avg` steps from the current state.
    // The forward pass keeps every pre-collision state; the backward pass runs
    //   mu   = P^T lambda^{t+1}            (reverse streaming over walls)
    //   lambda^t = J_C(g^t)^T mu + dj_t/dg (collision vector-Jacobian product)
    // and accumulates dJ/dalpha per cell and dJ/d(uin target) from the inflow
    // links. Collision map: f* = f - w+ f^+ - w- f^- + E(rho, J) with
    //   E_i = w+ feq+_i + w- feq-_i + S_i(Guo),  J = (p + Fb/2)/(1 + alpha/2),
    //   F = Fb - alpha J, feq+ = w(rho - 1 + 4.5 (c.J)^2/rho - 1.5 J^2/rho),
    //   feq- = 3 w c.J, S = w[a+ (9 (c.J)(c.F) - 3 J.F)/rho + 3 a- c.F].
    double vjp(const Objective& o, double* gAlpha, double* gEps, double* gUin, int64_t K) override {
        if (!std::is_same<R, double>::value || !adjointSupported()) return std::nan("");
        const int64_t total = o.spinup + o.avg;
        if (K <= 0) K = (int64_t)std::ceil(std::sqrt((double)total));
        if (K > total) K = total;
        const int64_t C = (total + K - 1) / K;
        const int64_t t0 = stepsDone;
        std::vector<double> g, cp;
        pressureWeights(o, cp);
        // forward pass: objective and a DDF checkpoint every K steps
        std::vector<std::vector<T>> ck((size_t)C);
        double J = 0.0;
        for (int64_t k = 0; k < total; k++) {
            if (k % K == 0) ck[(size_t)(k / K)] = f;
            const Params<R> p = params(stepsDone, false);
            if (k >= o.spinup) {
                gatherState(p, g);
                J += windowSum(g, p, cp, o) / (double)o.avg;
            }
            run(1, -1);
        }
        std::vector<T> fEnd = f;
        const int64_t tEnd = stepsDone;
        std::vector<double> G((size_t)K * 19 * N);
        std::vector<double> lamNext(19 * N, 0.0), lam(19 * N, 0.0);
        std::vector<double> ga(N, 0.0), ge(N, 0.0), guc(N, 0.0);
        for (int64_t seg = C - 1; seg >= 0; seg--) {
          // replay the segment from its checkpoint, keeping every state
          const int64_t k0 = seg * K, len = std::min(K, total - k0);
          f = ck[(size_t)seg];
          stepsDone = t0 + k0;
          for (int64_t s2 = 0; s2 < len; s2++) {
              gatherState(params(stepsDone, false), g);
              std::copy(g.begin(), g.end(), G.begin() + (size_t)s2 * 19 * N);
              run(1, -1);
          }
          for (int64_t k = k0 + len - 1; k >= k0; k--) {
            const int64_t t = t0 + k;
            const Params<R> p = params(t, false);
            const int64_t todd = (t & 1) ? t : t + 1;
            const double rOdd = (double)ramp(todd);
            const double wobj = k >= o.spinup ? 1.0 / (double)o.avg : 0.0;
            const double Fb[3] = {(double)p.fx, (double)p.fy, (double)p.fz};
            std::fill(lam.begin(), lam.end(), 0.0);
            const double* gk = &G[(size_t)(k - k0) * 19 * N];
#ifdef _OPENMP
            #pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
#endif
            for (long long nn = 0; nn < (long long)N; nn++) {
                const size_t n = (size_t)nn;
                if (flags[n] != FLUID) continue;
                const int x = (int)(n % (size_t)nx);
                const size_t tt = n / (size_t)nx;
                const int y = (int)(tt % (size_t)ny), z = (int)(tt / (size_t)ny);
                double mu[19];
                mu[0] = lamNext[n];
                for (int i = 1; i < 19; i++) {
                    const int dx = wrap(x + CX[i], nx), dy = wrap(y + CY[i], ny), dz = wrap(z + CZ[i], nz);
                    const size_t d = ((size_t)dz * ny + dy) * nx + dx;
                    if ((smask[n] >> (i - 1)) & 1u) {
                        const double ln = lamNext[OPP[i] * N + n];
                        mu[i] = flags[d] == OUTFLOW ? -ln : ln;
                        if (flags[d] == INFLOW) {
                            // f^{t+1}_{opp i}(x) = f*_i(x) - 6 w_i c_i.u_w, u_w,x = uin r prof(y_d)
                            // wallVel clamps the parabolic profile at 0
                            double prof = 1.0;
                            if (!hp.inflow_uniform) {
                                const double H = (double)ny - 2.0, yd = (double)dy - 0.5;
                                prof = 4.0 * yd * (H - yd) / (H * H);
                                if (!(hp.uin_target * rOdd * prof > 0.0)) prof = 0.0;
                            }
                            guc[n] += ln * (-6.0 * weight<double>(i) * CX[i] * rOdd * prof);
                        }
                    } else {
                        mu[i] = lamNext[i * N + d];
                    }
                }
                const double a = useAlpha ? (double)alpha[n] : 0.0;
                const double* gc = gk + n;
                auto G = [&](int i) { return gc[(size_t)i * N]; };
                double rho = 1.0, pv[3] = {0, 0, 0};
                for (int i = 0; i < 19; i++) {
                    const double v = G(i);
                    rho += v; pv[0] += CX[i] * v; pv[1] += CY[i] * v; pv[2] += CZ[i] * v;
                }
                const double gam = 1.0 + 0.5 * a;
                double Jv[3], Fv[3];
                for (int c = 0; c < 3; c++) { Jv[c] = (pv[c] + 0.5 * Fb[c]) / gam; Fv[c] = Fb[c] - a * Jv[c]; }
                const double rm1 = rho - 1.0;
                // local relaxation rates: sponge ramp, then Smagorinsky (Kernels.metal:112-167)
                const LocalRates lr = localRates(gc, N, x, p);
                const double wp = lr.wp, wm = lr.wm;
                const double ap = 1.0 - 0.5 * wp, am = 1.0 - 0.5 * wm;
                double M0 = 0, Mc[3] = {0, 0, 0}, Mcc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                for (int i = 0; i < 19; i++) {
                    const double wm_ = mu[i] * weight<double>(i);
                    const int cc[3] = {CX[i], CY[i], CZ[i]};
                    M0 += wm_;
                    for (int a1 = 0; a1 < 3; a1++) {
                        Mc[a1] += wm_ * cc[a1];
                        for (int b1 = 0; b1 < 3; b1++) Mcc[a1][b1] += wm_ * cc[a1] * cc[b1];
                    }
                }
                double MJ[3], MF[3], JMJ = 0, JMF = 0, J2 = 0, JF = 0;
                for (int a1 = 0; a1 < 3; a1++) {
                    MJ[a1] = Mcc[a1][0] * Jv[0] + Mcc[a1][1] * Jv[1] + Mcc[a1][2] * Jv[2];
                    MF[a1] = Mcc[a1][0] * Fv[0] + Mcc[a1][1] * Fv[1] + Mcc[a1][2] * Fv[2];
                }
                for (int a1 = 0; a1 < 3; a1++) {
                    JMJ += Jv[a1] * MJ[a1]; JMF += Jv[a1] * MF[a1];
                    J2 += Jv[a1] * Jv[a1]; JF += Jv[a1] * Fv[a1];
                }
                const double r2 = rho * rho;
                const double gRho = wp * (M0 - 4.5 * JMJ / r2 + 1.5 * J2 * M0 / r2)
                                  - ap * (9.0 * JMF - 3.0 * JF * M0) / r2;
                double gJ[3], gF[3], gP[3];
                for (int a1 = 0; a1 < 3; a1++) {
                    gJ[a1] = wp * (9.0 * MJ[a1] - 3.0 * Jv[a1] * M0) / rho + 3.0 * wm * Mc[a1]
                           + ap * (9.0 * MF[a1] - 3.0 * Fv[a1] * M0) / rho;
                    gF[a1] = ap * (9.0 * MJ[a1] - 3.0 * Jv[a1] * M0) / rho + 3.0 * am * Mc[a1];
                    gP[a1] = (gJ[a1] - a * gF[a1]) / gam;
                }
                // the increment Omega of the golden path and its rate derivatives
                const double ux = Jv[0] / rho, uy = Jv[1] / rho, uz = Jv[2] / rho;
                const double u2 = ux * ux + uy * uy + uz * uz;
                const double uF = ux * Fv[0] + uy * Fv[1] + uz * Fv[2];
                double Om[19], dOwp[19], dOwm[19];
                {
                    const double feq0 = weight<double>(0) * (rm1 - 1.5 * rho * u2);
                    const double x0 = -3.0 * uF;
                    Om[0] = wp * (feq0 - G(0)) + weight<double>(0) * ap * x0;
                    dOwp[0] = (feq0 - G(0)) - 0.5 * weight<double>(0) * x0;
                    dOwm[0] = 0.0;
                }
                for (int i = 1; i < 19; i += 2) {
                    const double W = weight<double>(i);
                    const double cu = 3.0 * (CX[i] * ux + CY[i] * uy + CZ[i] * uz);
                    const double cF = CX[i] * Fv[0] + CY[i] * Fv[1] + CZ[i] * Fv[2];
                    const double ep = W * (rm1 + rho * (0.5 * cu * cu - 1.5 * u2)) - 0.5 * (G(i) + G(i + 1));
                    const double em = W * rho * cu - 0.5 * (G(i) - G(i + 1));
                    const double xp = 3.0 * cu * cF - 3.0 * uF, xm = 3.0 * cF;
                    Om[i] = wp * ep + wm * em + W * ap * xp + W * am * xm;
                    Om[i + 1] = wp * ep - wm * em + W * ap * xp - W * am * xm;
                    dOwp[i] = dOwp[i + 1] = ep - 0.5 * W * xp;
                    dOwm[i] = em - 0.5 * W * xm;
                    dOwm[i + 1] = -dOwm[i];
                }
                // objective: w_p cp rho + w_d alpha J_x (+ w_f ntF_x on partial cells)
                const double dObjRho = wobj * o.w_p * cp[n];
                const double dObjPx = wobj * o.w_d * a / gam;
                const double e = useEps ? (double)eps[n] : 0.0;
                double gwp = 0.0, gwm = 0.0;
                if (e == 0.0) {
                    for (int j = 0; j < 19; j++) {
                        const double mp = j == 0 ? mu[0] : 0.5 * (mu[j] + mu[OPP[j]]);
                        const double mm = j == 0 ? 0.0 : 0.5 * (mu[j] - mu[OPP[j]]);
                        lam[j * N + n] = mu[j] - wp * mp - wm * mm + gRho
                                       + CX[j] * gP[0] + CY[j] * gP[1] + CZ[j] * gP[2]
                                       + dObjRho + dObjPx * CX[j];
                        gwp += mu[j] * dOwp[j];
                        gwm += mu[j] * dOwm[j];
                    }
                } else {
                    // Noble-Torczynski cell (Kernels.metal:201-227):
                    //   f*_j = f_j + (1 - B) Omega_j(f) + B (f_opp(j) - f_j),  j >= 1,
                    //   f*_0 = f_0 + (1 - B) Omega_0(f),
                    // B = eps tw / ((1 - eps) + tw), tw = 1/w+ - 1/2 at the local rate.
                    const double B = ntB(e, wp), omB = 1.0 - B;
                    double gB = 0.0, dNt = 0.0;
                    for (int i = 0; i < 19; i++) gB -= mu[i] * Om[i];
                    for (int i = 1; i < 19; i++) gB += mu[i] * (G(OPP[i]) - G(i));
                    for (int i = 1; i < 19; i += 2) dNt -= 2.0 * CX[i] * (G(i + 1) - G(i));
                    const double dObjNt = wobj * o.w_f * 2.0 * B;     // d ntF_x / d g_j = 2 B c_jx
                    for (int j = 0; j < 19; j++) {
                        const double mp = j == 0 ? mu[0] : 0.5 * (mu[j] + mu[OPP[j]]);
                        const double mm = j == 0 ? 0.0 : 0.5 * (mu[j] - mu[OPP[j]]);
                        const double om = -wp * mp - wm * mm + gRho + CX[j] * gP[0] + CY[j] * gP[1] + CZ[j] * gP[2];
                        const double bounce = j == 0 ? 0.0 : B * (mu[OPP[j]] - mu[j]);
                        lam[j * N + n] = mu[j] + omB * om + bounce + dObjRho + (dObjPx + dObjNt) * CX[j];
                        gwp += omB * mu[j] * dOwp[j];
                        gwm += omB * mu[j] * dOwm[j];
                    }
                    const double gBtot = gB + wobj * o.w_f * dNt;
                    ge[n] += gBtot * ntdB(e, wp);
                    // B depends on the rate through tw = 1/w+ - 1/2
                    const double tw = 1.0 / wp - 0.5, den = (1.0 - e) + tw;
                    gwp += gBtot * (e * (1.0 - e) / (den * den)) * (-1.0 / (wp * wp));
                }
                if (lr.les) {
                    // w+ = 1/tau, tau = (tau0 + s)/2, s = sqrt(tau0^2 + K Q / rho); w- = w+ (SRT)
                    // or 1/(1/2 + Lambda/(tau - 1/2)) (TRT); Q = |Pi_neq| with
                    // Pi_ab = sum_i c_ia c_ib f_i - delta_ab (rho-1)/3 - J_a J_b / rho.
                    const double dwm = lr.srt ? -wp * wp : wm * wm * lr.lambda / ((lr.tau - 0.5) * (lr.tau - 0.5));
                    const double gTau = gwp * (-wp * wp) + gwm * dwm;
                    if (lr.Q > 0.0) {
                        const double gQ = gTau * lr.K / (4.0 * rho * lr.s);
                        const double gR = -gTau * lr.K * lr.Q / (4.0 * rho * rho * lr.s);
                        const double* P3 = lr.Pi;   // xx yy zz xy xz yz
                        const double Pm[3][3] = {{P3[0], P3[3], P3[4]}, {P3[3], P3[1], P3[5]}, {P3[4], P3[5], P3[2]}};
                        const double tr = P3[0] + P3[1] + P3[2];
                        double PJ[3], JPJ = 0.0;
                        for (int a1 = 0; a1 < 3; a1++) {
                            PJ[a1] = Pm[a1][0] * Jv[0] + Pm[a1][1] * Jv[1] + Pm[a1][2] * Jv[2];
                            JPJ += Jv[a1] * PJ[a1];
                        }
                        for (int j = 0; j < 19; j++) {
                            const int cc[3] = {CX[j], CY[j], CZ[j]};
                            double cPc = 0.0, cPJ = 0.0;
                            for (int a1 = 0; a1 < 3; a1++) {
                                cPJ += cc[a1] * PJ[a1];
                                for (int b1 = 0; b1 < 3; b1++) cPc += cc[a1] * Pm[a1][b1] * cc[b1];
                            }
                            const double dQ = (cPc - tr / 3.0 - 2.0 * cPJ / rho + JPJ / (rho * rho)) / lr.Q;
                            lam[j * N + n] += gQ * dQ + gR;
                        }
                    }
                }
                const double gJdotJ = gJ[0] * Jv[0] + gJ[1] * Jv[1] + gJ[2] * Jv[2];
                const double gFdotJ = gF[0] * Jv[0] + gF[1] * Jv[1] + gF[2] * Jv[2];
                ga[n] += -gJdotJ / (2.0 * gam) - gFdotJ * (1.0 - a / (2.0 * gam))
                       + wobj * o.w_d * Jv[0] * (1.0 - a / (2.0 * gam));
            }
            lamNext.swap(lam);
          }
        }
        f.swap(fEnd);
        stepsDone = tEnd;
        double gu = 0.0;
        for (size_t n = 0; n < N; n++) gu += guc[n];
        if (gAlpha) for (size_t n = 0; n < N; n++) gAlpha[n] = ga[n];
        if (gEps) for (size_t n = 0; n < N; n++) gEps[n] = ge[n];
        if (gUin) *gUin = gu;
        return J;
    }

    // Single-cell collide for unit tests: fin are the 19 (widened) DDFs; the
    // step-0 params of this lattice are used with parity ignored.
    void collideCell(const double* fin, double e, int x, double* fout, double* rho, double* ntf) override {
        const Params<R> p = params(stepsDone, false);
        R fh[19];
        for (int i = 0; i < 19; i++) fh[i] = (R)fin[i];
        R r = 1, a = 0, b = 0, c = 0;
        collide(fh, p, x, (R)e, R(0), false, r, a, b, c);
        for (int i = 0; i < 19; i++) fout[i] = (double)fh[i];
        *rho = (double)r; ntf[0] = (double)a; ntf[1] = (double)b; ntf[2] = (double)c;
    }
};

}  // namespace orc

using namespace orc;

#pragma GCC visibility push(default)
extern "C" {

// arith: 0 = FP64 (storage must be 0 = f64); 1 = FP32 with storage
// 1 = f32, 2 = f16 (shifted, RNE), 3 = bf16 (RNE).
void* orc_create(int arith, int storage, int nx, int ny, int nz,
                 const uint8_t* flags, int wants_forces) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return nullptr;
    if (arith == 0 && storage == 0) return new Lattice<StF64>(nx, ny, nz, flags, wants_forces != 0);
    if (arith == 1 && storage == 1) return new Lattice<StF32>(nx, ny, nz, flags, wants_forces != 0);
    if (arith == 1 && storage == 2) return new Lattice<StF16>(nx, ny, nz, flags, wants_forces != 0);
    if (arith == 1 && storage == 3) return new Lattice<StB16>(nx, ny, nz, flags, wants_forces != 0);
    return nullptr;
}
void orc_destroy(void* h) { delete (Base*)h; }
void orc_set_host_params(void* h, const HostParams* p) { ((Base*)h)->hp = *p; }
void orc_set_threads(void* h, int t) { ((Base*)h)->threads = t > 0 ? t : 1; }
void orc_set_solid_fractions(void* h, const float* eps) { ((Base*)h)->setEps(eps); }
void orc_set_porosity(void* h, const float* alpha) { ((Base*)h)->setAlpha(alpha); }
// FP64 lattices only (NaN otherwise, or with LES / sponge / NT cells).
double orc_objective(void* h, const Objective* o) { return ((Base*)h)->objective(*o); }
double orc_vjp(void* h, const Objective* o, double* grad_alpha, double* grad_eps, double* grad_uin,
               int64_t checkpoint_every) {
    return ((Base*)h)->vjp(*o, grad_alpha, grad_eps, grad_uin, checkpoint_every);
}
void orc_init_field(void* h, int mode, double amplitude) { ((Base*)h)->initField(mode, amplitude); }
void orc_run(void* h, int64_t steps) { ((Base*)h)->run(steps, -1); }
int64_t orc_steps_done(void* h) { return ((Base*)h)->stepsDone; }
int orc_probe_moments(void* h, double* out) { return ((Base*)h)->moments(out); }

// Sim.swift:359-372: zero, one even+odd pair with the odd pass writing
// per-cell momentum exchange, then a double sum in z, y, x nested order.
int orc_probe_force(void* h, int x0, int x1, int y0, int y1, double* out3) {
    Base* b = (Base*)h;
    if (b->stepsDone % 2 != 0 || !b->wantsForces) return -1;
    b->zeroForce();
    b->run(2, b->stepsDone + 1);
    std::vector<double> ff(4 * b->N);
    b->forceField(ff.data());
    double tx = 0, ty = 0, tz = 0;
    for (int z = 0; z < b->nz; z++) for (int y = y0; y <= y1; y++) for (int x = x0; x <= x1; x++) {
        const size_t n = ((size_t)z * b->ny + y) * b->nx + x;
        tx += ff[4 * n]; ty += ff[4 * n + 1]; tz += ff[4 * n + 2];
    }
    out3[0] = tx; out3[1] = ty; out3[2] = tz;
    return 0;
}
void orc_force_field(void* h, double* out) { ((Base*)h)->forceField(out); }
double orc_mass_sum(void* h) { return ((Base*)h)->massSum(); }
size_t orc_ddf_bytes(void* h) { return ((Base*)h)->ddfBytes(); }
void orc_download_ddf(void* h, void* out) { ((Base*)h)->download(out); }
void orc_upload_ddf(void* h, const void* in) { ((Base*)h)->upload(in); }
void orc_masks(void* h, uint32_t* smask, uint32_t* lmask) {
    Base* b = (Base*)h;
    std::memcpy(smask, b->smask.data(), 4 * b->N);
    std::memcpy(lmask, b->lmask.data(), 4 * b->N);
}
void orc_collide_cell(void* h, const double* fin, double eps, int x,
                      double* fout, double* rho, double* ntf) {
    ((Base*)h)->collideCell(fin, eps, x, fout, rho, ntf);
}
// Storage conversions exposed for their own pins.
uint16_t orc_f32_to_f16(float v) { return f32_to_f16(v); }
float orc_f16_to_f32(uint16_t v) { return f16_to_f32(v); }
uint16_t orc_f32_to_bf16(float v) { return f32_to_bf16(v); }
float orc_bf16_to_f32(uint16_t v) { return bf16_to_f32(v); }
void orc_f32_to_f16_array(const float* in, uint16_t* out, size_t n) { for (size_t k = 0; k < n; k++) out[k] = f32_to_f16(in[k]); }
void orc_f32_to_bf16_array(const float* in, uint16_t* out, size_t n) { for (size_t k = 0; k < n; k++) out[k] = f32_to_bf16(in[k]); }
void orc_f16_to_f32_array(const uint16_t* in, float* out, size_t n) { for (size_t k = 0; k < n; k++) out[k] = f16_to_f32(in[k]); }

}  // extern "C"
#pragma GCC visibility pop
