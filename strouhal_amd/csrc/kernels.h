// Host-side launch interface of the D3Q19 kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "lbm.h"

namespace strh {

// Everything one step launch needs. The boolean features select the kernel
// specialisation; the host sets them from the lattice's contents so the
// all-fluid periodic bulk runs the mask-free, eps-free, force-free kernel.
struct StepLaunch {
    int precision;
    void* f;
    const uint8_t* flags;
    const uint32_t* link;
    const uint32_t* lidm;
    float* force;
    const float* eps;
    bool walls;      // lattice has non-fluid cells
    bool nt;         // Noble-Torczynski field installed
    bool les;        // cSmago > 0
    bool has_force;  // body force non-zero
    uint32_t rows_bs;  // 256 / 512: odd pass uses LDS row staging; 0: direct
    StepParams p;
};

hipError_t launch_step(const StepLaunch& L, hipStream_t s);
// Workgroup size of the LDS row-staged odd pass for this lattice, or 0.
uint32_t rows_block_size(uint32_t nx, uint64_t N);
hipError_t launch_moments(int precision, const void* f, const uint8_t* flags, float* out,
                          uint32_t N, uint32_t S, float fx, float fy, float fz, hipStream_t s);
hipError_t launch_init(int precision, void* f, uint32_t nx, uint32_t ny, uint32_t nz,
                       uint32_t S, uint32_t mode, float amplitude, float tau, hipStream_t s);
hipError_t launch_mass_sum(int precision, const void* f, uint32_t N, uint32_t S,
                           double* partials, uint32_t max_partials, double* out, hipStream_t s);
hipError_t launch_box_force_sum(const float* force, uint32_t nx, uint32_t ny, uint32_t nz,
                                uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1,
                                double* partials, uint32_t max_partials, double* out3,
                                hipStream_t s);
hipError_t launch_kinetic_sum(const float* moments, uint32_t N, double* partials,
                              uint32_t max_partials, double* out, hipStream_t s);
hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t s);

}  // namespace strh
