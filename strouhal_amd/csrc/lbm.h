// Shared host/device definitions of the D3Q19 solver.
#pragma once

#include <stdint.h>

namespace strh {

// D3Q19 in the reference's pair order (Sources/StrouhalCore/Kernels.metal:39-47):
// i = 0 rest; pairs (2k-1, 2k) are opposites. The order is load-bearing: the
// AA slot swap, the neighbour-mask bits and the digest byte order follow it.
constexpr int kCx[19] = {0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 1,-1, 1,-1, 0, 0, 0, 0};
constexpr int kCy[19] = {0, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 0, 0, 0, 0, 1,-1, 1,-1};
constexpr int kCz[19] = {0, 0, 0, 0, 0, 1,-1, 0, 0, 0, 0, 1,-1,-1, 1, 1,-1,-1, 1};
constexpr int kOpp[19] = {0, 2,1, 4,3, 6,5, 8,7, 10,9, 12,11, 14,13, 16,15, 18,17};

enum : uint8_t { kFluid = 0, kSolid = 1, kLid = 2, kInflow = 3, kOutflow = 4 };
enum : int { kFP32 = 0, kFP16S = 1, kBF16 = 2 };

// Bit 31 of a cell's link word: the cell is not fluid (the kernel skips it).
// Bits 0..17: neighbour n + c_i is not fluid (bit i-1), Sim.swift:235-255.
constexpr uint32_t kNotFluidBit = 0x80000000u;

// Division by an invariant 32-bit divisor d >= 1 for dividends n < 2^31
// (Granlund & Montgomery 1994, "Division by invariant integers using
// multiplication", round-up variant): q = (umulhi(n, m) + n) >> s with
// s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1. umulhi(n, m) <= n, so
// the 32-bit sum cannot overflow while n < 2^31.
struct FastDiv {
    uint32_t d, m, s;
};

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f;
    uint32_t s = 0;
    while ((1ull << s) < d) s++;
    f.d = d;
    f.s = s;
    f.m = (uint32_t)(((1ull << 32) * ((1ull << s) - d)) / d + 1ull);
    return f;
}

// Storage of the DDFs. Cells are grouped in blocks of W = 2^lw cells, and a
// block keeps the 19 slots of its cells one after another: slot i of cell n
// lives at element (n >> lw) * 19 W + i * W + (n & (W - 1)). A wavefront's 19
// loads and 19 stores then fall in one window of 19 W elements instead of 19
// arrays a lattice apart; with W * B = 32 KiB this lifts the even pass at
// 256^3 FP32 from 69% to 78% of the HBM peak on MI355X. The kernels address
// the DDFs as a 64-bit slot base plus a 32-bit byte offset, so a lattice whose
// blocked image would exceed 4 GiB, or that fits in one block, keeps a single
// block (lw = 31, mask = ~0: slot i of cell n at i * W + n, W = N rounded up
// to 64), the plain slot-major layout.
struct DdfLayout {
    uint32_t W;        // slot stride inside a block, elements
    uint32_t lw;       // log2 of the block size; 31 for a single block
    uint32_t mask;     // W - 1, or all ones for a single block
    uint32_t bstride;  // 19 W: elements from one block to the next
    uint32_t blocks;   // number of blocks
};

// large: the image may exceed 4 GiB. Only kernels that address the DDFs from
// per-workgroup block or per-row bases (the even pass and the row-staged odd
// pass of wall-free lattices) reach such an image; the direct odd pass's
// 32-bit offsets from the slot bases do not.
inline DdfLayout make_ddf_layout(uint64_t N, uint32_t bytes, bool large = false) {
    DdfLayout l;
    const uint32_t lw = bytes == 4 ? 13u : 14u;
    const uint64_t W = 1ull << lw;
    const uint64_t nb = (N + W - 1) / W;
    if (N > W && (large || nb * 19 * W * bytes < (1ull << 32))) {
        l.W = (uint32_t)W;
        l.lw = lw;
        l.mask = (uint32_t)W - 1u;
        l.bstride = 19u * (uint32_t)W;
        l.blocks = (uint32_t)nb;
    } else {
        l.W = (uint32_t)((N + 63u) & ~63ull);
        l.lw = 31u;
        l.mask = 0xffffffffu;
        l.bstride = 0u;
        l.blocks = 1u;
    }
    return l;
}

inline uint64_t ddf_elements(const DdfLayout& l) { return (uint64_t)l.blocks * 19u * l.W; }

// Per-launch block of the step kernel: the reference's Params
// (Kernels.metal:49-63) plus the layout facts the kernel needs.
struct StepParams {
    int32_t nx, ny, nz, parity;
    uint32_t N;          // cells
    uint32_t S;          // N rounded up to 64 (pitch of per-cell side arrays)
    DdfLayout lay;       // where the DDFs of a cell live
    float omega, omegaMinus, uin, cSmago, fx, fy, fz;
    int32_t writeForce;
    float ulidX, ulidY, ulidZ;
    int32_t inflowUniform;
    float spongeX0, spongeInvW, spongeTau;
    int32_t useEps;
    // c_i . F and 3 (c_i . F) for the pair leaders i = 1, 3, ..., 17, rounded
    // as the reference rounds cF = (cx*fx + cy*fy) + cz*fz and 3.0f * cF
    // (Kernels.metal:185, 193). Uniform over the lattice, so the host
    // evaluates them once in the same IEEE binary32 arithmetic.
    float cF[9];
    float threeCF[9];
    FastDiv divx, divy;  // n -> (x, y, z)
    uint32_t blocks;     // grid size of the launch
};

}  // namespace strh
