// Observables and utilities: moments probe, field initialiser, fixed-order
// FP64 reductions, fills; and the storage dispatch of the step launch.
#include "device.h"

namespace strh {

// ---- moments probe: Kernels.metal:359-380 (literal loop order) ----------
// (ux, uy, uz, rho) of fluid cell n; shared by the probe and the fused
// kinetic-energy reduction so both give the same bits.
template <class St>
__device__ __forceinline__ float4 cell_moments(const typename St::T* __restrict__ f, const DdfLayout& lay,
                                               uint32_t n, float fx, float fy, float fz) {
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
    return make_float4((px + 0.5f * fx) / rho, (py + 0.5f * fy) / rho, (pz + 0.5f * fz) / rho, rho);
}

template <class St>
__global__ void __launch_bounds__(256)
moments_kernel(const typename St::T* __restrict__ f, const uint8_t* __restrict__ flags,
               float4* __restrict__ out, uint32_t N, DdfLayout lay, float fx, float fy, float fz) {
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= N) return;
    out[n] = flags[n] != kFluid ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : cell_moments<St>(f, lay, n, fx, fy, fz);
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


template <class Elem>
__global__ void __launch_bounds__(256)
reduce_stage1(Elem e, uint32_t M, uint32_t chunk, double* __restrict__ partials) {
    __shared__ double lds[4];
    const uint32_t lo = blockIdx.x * chunk;
    const uint32_t hi = (M - lo) > chunk ? lo + chunk : M;
    const double s = block_tree_sum(chunk_sum(e, blockIdx.y, lo, hi), lds);
    if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = s;
}

struct PartialElem {
    const double* p;
    __device__ double operator()(uint32_t, uint32_t k) const { return p[k]; }
};

__global__ void __launch_bounds__(256)
reduce_stage2(const double* __restrict__ partials, uint32_t P, double* __restrict__ out) {
    __shared__ double lds[4];
    const double s = block_tree_sum(chunk_sum(PartialElem{partials}, 0u, 0u, P), lds);
    if (threadIdx.x == 0) out[0] = s;
}

template <class St> struct MassElem {
    const typename St::T* f;
    DdfLayout lay;
    __device__ double operator()(uint32_t slot, uint32_t n) const {
        return (double)St::ld(f[cell_elem(lay, n) + (size_t)slot * lay.W]);
    }
};

struct KineticElem {
    const float4* mom;
    __device__ double operator()(uint32_t, uint32_t k) const {
        const float4 v = mom[k];
        return (double)(v.x * v.x + v.y * v.y + v.z * v.z);
    }
};

// The same per-cell term straight from the DDFs, without the moments field:
// cell_moments then KineticElem's arithmetic, so the sum is the same bits.
template <class St> struct KineticCellElem {
    const typename St::T* f;
    const uint8_t* flags;
    DdfLayout lay;
    float fx, fy, fz;
    __device__ double operator()(uint32_t, uint32_t k) const {
        if (flags[k] != kFluid) return 0.0;    // the probe's zero moments
        const float4 v = cell_moments<St>(f, lay, k, fx, fy, fz);
        return (double)(v.x * v.x + v.y * v.y + v.z * v.z);
    }
};

// Reduction scratch of the allocation-free kinetic energy
// (strh_kinetic_energy_noalloc): kReducePartials block partials, then the
// result. One copy per GPU, held until the process ends; the host reaches it
// only through the address scratch_address() returns on that device. Every
// partial a reduction reads is written by the same reduction's first stage,
// so nothing has to start at zero; calls that overlap on one GPU race.
static __device__ __align__(256) double scratch[kScratchDoubles];

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

hipError_t launch_step(const StepLaunch& L, hipStream_t s) {
    switch (L.precision) {
        case kFP32: return launch_step_fp32(L, s);
        case kFP16S: return launch_step_fp16(L, s);
        case kBF16: return launch_step_bf16(L, s);
    }
    return hipErrorInvalidValue;
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

hipError_t scratch_address(double** out) {
    return hipGetSymbolAddress((void**)out, HIP_SYMBOL(scratch));
}

hipError_t launch_kinetic_sum_cells(int precision, const void* f, const uint8_t* flags, uint32_t N,
                                    DdfLayout lay, float fx, float fy, float fz, double* partials,
                                    uint32_t max_partials, double* out, hipStream_t s) {
    switch (precision) {
        case kFP32: return reduce(KineticCellElem<StF32>{(const float*)f, flags, lay, fx, fy, fz}, N, 1, partials, max_partials, out, s);
        case kFP16S: return reduce(KineticCellElem<StF16>{(const uint16_t*)f, flags, lay, fx, fy, fz}, N, 1, partials, max_partials, out, s);
        case kBF16: return reduce(KineticCellElem<StBF16>{(const uint16_t*)f, flags, lay, fx, fy, fz}, N, 1, partials, max_partials, out, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t s) {
    const uint64_t words = bytes / 4;
    if (words == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((uint32_t)((words + 255) / 256)), dim3(256), 0, s,
                       (uint32_t*)p, words);
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
