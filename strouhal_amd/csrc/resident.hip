// Resident batch launch: lattices one cell deep (nz = 1) held in LDS across
// many steps.
//
// The streaming batch kernel (batch.hip) reads and writes every DDF of every
// lattice from memory at each step. Each lattice of the resolution ladder is
// small enough to live on chip instead: its cells are cut into rectangular
// tiles, one workgroup per CU holds a tile (plus a one-cell halo ring) in LDS,
// and the workgroups of a lattice advance it together for up to kProbeSpan
// steps per launch. Between steps a tile trades only the populations that
// cross its edges with the tiles around it:
//   * after an even pass, each tile publishes the slots of its edge cells
//     that a neighbour's odd pass will pull (slot opp(i) of cell b when
//     b + c_i lies outside the tile) and loads the same slots of its halo
//     cells;
//   * after an odd pass, it publishes what its edge cells pushed into halo
//     cells (slot i of halo cell h when h - c_i lies inside the tile) and
//     loads what the neighbours pushed into its own edge cells.
// Only links between two fluid cells carry a word: along a wall link a cell
// bounces its populations back into its own slots, which no neighbour may
// overwrite (and no neighbour reads a non-fluid cell's slots).
// The exchange goes through a per-lattice array X[19][N] of 64-bit words,
// (tag << 32) | stored bits, written and read with agent-scope relaxed
// atomics: the tag names the step whose result the word holds, so a reader
// spins until its tag appears and needs no fence or flag. A writer can be at
// most one step ahead of any reader (it needs the reader's previous export
// first), so a word is never overwritten before it was read. The array is
// zeroed at the start of every batch run and tags count from 1 within the
// run.
//
// Each cell's update is the single-lattice kernel's (step_cell in device.h),
// with the DDF accesses redirected to LDS: the same collide, the same
// wall-link reads and bounce-back stores, the same moving-wall and outflow
// corrections and the same momentum-exchange sums in the same order, so every
// DDF and every force sample is the same bits as the streaming path. The AA
// pattern keeps the update race-free in place: in either pass a cell reads
// and writes one set of locations that no other cell touches.
//
// Each step computes the tile's edge ring first, publishes it, computes the
// interior while the exchange is in flight, then loads its halo. Spins are
// bounded: after a second without progress a wave raises the batch's abort
// flag, every workgroup that waits sees it and returns, and the host reports
// the failure. The host launches at most as many workgroups as the device
// holds at once, so every tile's neighbours run concurrently.
#include <stdint.h>

#include <type_traits>

#include "device.h"

namespace strh {

// Local LDS index of tile cell (lx, ly), lx in [-1, w], ly in [-1, h].
__device__ __forceinline__ int rl_index(int lx, int ly, int pitch) { return (ly + 1) * pitch + (lx + 1); }

__device__ __forceinline__ int wrap_coord(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Edge ring cell r of a w x h tile (w, h >= 3): top row, bottom row, then the
// left and right columns without their corners.
__device__ __forceinline__ void ring_cell(int r, int w, int h, int& lx, int& ly) {
    if (r < w) { lx = r; ly = 0; return; }
    if (r < 2 * w) { lx = r - w; ly = h - 1; return; }
    const int q = r - 2 * w;
    if (q < h - 2) { lx = 0; ly = 1 + q; return; }
    lx = w - 1; ly = 1 + (q - (h - 2));
}

// Halo cell r of a w x h tile: the row above, the row below (both w + 2
// long), then the left and right columns.
__device__ __forceinline__ void halo_cell(int r, int w, int h, int& lx, int& ly) {
    if (r < w + 2) { lx = r - 1; ly = -1; return; }
    if (r < 2 * (w + 2)) { lx = r - (w + 2) - 1; ly = h; return; }
    const int q = r - 2 * (w + 2);
    if (q < h) { lx = -1; ly = q; return; }
    lx = w; ly = q - h;
}

__device__ __forceinline__ bool inside(int lx, int ly, int w, int h) {
    return lx >= 0 && lx < w && ly >= 0 && ly < h;
}

// One cell of one AA pass with the DDFs in LDS: step_cell's arithmetic and
// boundary handling. The tile is cell-major in LDS: slot i of local cell li
// is element 19 li + i, so every access is one register base plus an
// immediate offset. The odd pass reads and writes all 18 neighbour slots
// unconditionally and then, in cells with wall links only, takes the wall
// reads from the cell's own slots and stores the bounce-back into them. The
// neighbour slots touched along a wall link belong to non-fluid cells, which
// no fluid cell reads, the exchange skips and the write-back leaves alone.
// Wall velocity (Kernels.metal:81-87) of a wall cell dy rows from the cell
// whose row's parabolic inflow value is prow[0]: the same arithmetic as
// wall_vel, with the division done once per tile row and step.
__device__ __forceinline__ void wall_vel_row(uint8_t wf, const float* prow, int dy, const StepParams& p,
                                             float& ux, float& uy, float& uz) {
    if (wf == kLid) { ux = p.ulidX; uy = p.ulidY; uz = p.ulidZ; return; }
    uy = 0.0f; uz = 0.0f;
    ux = p.inflowUniform != 0 ? p.uin : prow[dy];
}

// The parabolic inflow value max(4 uin yd (H - yd) / H^2, 0) of every row of a
// tile and its halo (wall_vel's expression, bit for bit) for one step.
__device__ __forceinline__ void inflow_rows(float* prof, int rows, int y0, int ny, float uin, int tid, int bsz) {
    for (int r = tid; r < rows; r += bsz) {
        const int wy = wrap_coord(y0 + r - 1, ny);
        const float H = (float)ny - 2.0f;
        const float yd = (float)wy - 0.5f;
        const float v = 4.0f * uin * yd * (H - yd) / (H * H);
        prof[r] = v > 0.0f ? v : 0.0f;
    }
}

// lw and eps are the cell's link word and solid fraction (0 without one), lM
// its moving-wall mask (0 without wall links); flL holds the flags of the tile
// and its halo in LDS (index li + dx + dy pitch for the neighbour (dx, dy)),
// and prow the parabolic inflow value of the cell's row in LDS (the rows
// above and below at prow[1], prow[-1]), so a moving-wall correction waits on
// no global load and divides nothing.
template <class Conv, int PARITY, bool LES, bool FORCE>
__device__ __forceinline__ void resident_cell(typename Conv::T* __restrict__ s, int li, int pitch,
                                              const uint8_t* __restrict__ flL, const float* __restrict__ prow,
                                              uint32_t lM, const StepParams& p, uint32_t n, int x,
                                              uint32_t lw, float eps, float4* __restrict__ force) {
    using T = typename Conv::T;
    constexpr uint32_t TB = (uint32_t)sizeof(T), CB = 19u * TB;
    char* sb = (char*)s;
    auto ldb = [&](uint32_t off) -> float { return Conv::ld(*(const T*)(sb + off)); };
    auto stb = [&](uint32_t off, float v) { *(T*)(sb + off) = Conv::st(v); };
    const uint32_t own = (uint32_t)li * CB;
    float fh[19];
    float ntx = 0.0f, nty = 0.0f, ntz = 0.0f;

    if (PARITY == 0) {                                             // Kernels.metal:252-260
        if (lw & kNotFluidBit) return;
#pragma unroll
        for (int i = 0; i < 19; i++) fh[i] = ldb(own + (uint32_t)i * TB);
        collide<LES, FORCE, true>(fh, p, x, eps, ntx, nty, ntz);
        uint32_t ows = own;
        asm("" : "+v"(ows));   // one base for the stores too (not 19 addresses kept live)
        stb(ows, fh[0]);
#pragma unroll
        for (int i = 1; i < 19; i++) stb(ows + (uint32_t)kOpp[i] * TB, fh[i]);
        return;   // probe steps are odd: the even pass writes no force
    }

    // -------- odd: streaming pass (neighbours + walls), 261-354 --------
    if (lw & kNotFluidBit) return;
    const uint32_t sM = lw & 0x3ffffu;
    const bool writes = p.writeForce != 0;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    // byte offset of the row below / of / above the cell, one cell to the left
    const uint32_t rb0 = own - CB, rbm = rb0 - (uint32_t)pitch * CB, rbp = rb0 + (uint32_t)pitch * CB;
    // neighbour (dx, dy), slot i: a row base plus a non-negative immediate
#define STRH_NB(DX, DY, SLOT) \
    ((DY) < 0 ? rbm : ((DY) > 0 ? rbp : rb0)) + (uint32_t)(((DX) + 1) * 19 + (SLOT)) * TB
    fh[0] = ldb(own);
#define STRH_RPULL(I) fh[I] = ldb(STRH_NB(-kCx[I], -kCy[I], kOpp[I]));
    STRH_RPULL(1) STRH_RPULL(2) STRH_RPULL(3) STRH_RPULL(4) STRH_RPULL(5) STRH_RPULL(6)
    STRH_RPULL(7) STRH_RPULL(8) STRH_RPULL(9) STRH_RPULL(10) STRH_RPULL(11) STRH_RPULL(12)
    STRH_RPULL(13) STRH_RPULL(14) STRH_RPULL(15) STRH_RPULL(16) STRH_RPULL(17) STRH_RPULL(18)
#undef STRH_RPULL
    uint32_t deferred = 0u;
    if (sM != 0u) {   // wall links: reads from the own slots, moving-wall corrections, force terms
#define STRH_RWALLIN(I)                                                               \
        {                                                                             \
            constexpr int srcBit = kOpp[I] - 1;                                       \
            if ((sM >> srcBit) & 1u) {                                                \
                fh[I] = ldb(own + (uint32_t)(I) * TB);                                \
                if ((lM >> srcBit) & 1u) deferred |= 1u << I;                         \
                if (writes) {                                                         \
                    ax = acc<-kCx[I]>(ax, fh[I]);                                     \
                    ay = acc<-kCy[I]>(ay, fh[I]);                                     \
                    az = acc<-kCz[I]>(az, fh[I]);                                     \
                }                                                                     \
            }                                                                         \
        }
        STRH_RWALLIN(1) STRH_RWALLIN(2) STRH_RWALLIN(3) STRH_RWALLIN(4) STRH_RWALLIN(5) STRH_RWALLIN(6)
        STRH_RWALLIN(7) STRH_RWALLIN(8) STRH_RWALLIN(9) STRH_RWALLIN(10) STRH_RWALLIN(11) STRH_RWALLIN(12)
        STRH_RWALLIN(13) STRH_RWALLIN(14) STRH_RWALLIN(15) STRH_RWALLIN(16) STRH_RWALLIN(17) STRH_RWALLIN(18)
#undef STRH_RWALLIN
    }
    if (deferred != 0u) {                                          // 292-314
#define STRH_RDEFER(I)                                                                \
        if ((deferred >> I) & 1u) {                                                   \
            const uint8_t wf = flL[li - kCx[I] - kCy[I] * pitch];                     \
            if (wf == kOutflow) {                                                     \
                fh[I] = -fh[I];                                                       \
            } else {                                                                  \
                float uwx, uwy, uwz;                                                  \
                wall_vel_row(wf, prow, -kCy[I], p, uwx, uwy, uwz);                    \
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
        STRH_RDEFER(1) STRH_RDEFER(2) STRH_RDEFER(3) STRH_RDEFER(4) STRH_RDEFER(5) STRH_RDEFER(6)
        STRH_RDEFER(7) STRH_RDEFER(8) STRH_RDEFER(9) STRH_RDEFER(10) STRH_RDEFER(11) STRH_RDEFER(12)
        STRH_RDEFER(13) STRH_RDEFER(14) STRH_RDEFER(15) STRH_RDEFER(16) STRH_RDEFER(17) STRH_RDEFER(18)
#undef STRH_RDEFER
    }
    collide<LES, FORCE, true>(fh, p, x, eps, ntx, nty, ntz);      // 315-316
    ax += ntx;
    ay += nty;
    az += ntz;
    // the push from a pinned copy of the cell's offset (not 19 live addresses)
    uint32_t op = own;
    asm("" : "+v"(op));
    const uint32_t pb0 = op - CB, pbm = pb0 - (uint32_t)pitch * CB, pbp = pb0 + (uint32_t)pitch * CB;
    stb(op, fh[0]);                                                // 317-350
#define STRH_PB(DX, DY, SLOT) \
    ((DY) < 0 ? pbm : ((DY) > 0 ? pbp : pb0)) + (uint32_t)(((DX) + 1) * 19 + (SLOT)) * TB
#define STRH_RPUSH(I) stb(STRH_PB(kCx[I], kCy[I], I), fh[I]);
    STRH_RPUSH(1) STRH_RPUSH(2) STRH_RPUSH(3) STRH_RPUSH(4) STRH_RPUSH(5) STRH_RPUSH(6)
    STRH_RPUSH(7) STRH_RPUSH(8) STRH_RPUSH(9) STRH_RPUSH(10) STRH_RPUSH(11) STRH_RPUSH(12)
    STRH_RPUSH(13) STRH_RPUSH(14) STRH_RPUSH(15) STRH_RPUSH(16) STRH_RPUSH(17) STRH_RPUSH(18)
#undef STRH_RPUSH
#undef STRH_PB
#undef STRH_NB
    // Wall links store sign * f - corr into the cell's own opposite slot (the
    // halfway bounce-back, with the moving-wall term or the outflow's sign flip).
    if (sM != 0u) {
#define STRH_RWALLOUT(I)                                                              \
        {                                                                             \
            constexpr int dstBit = I - 1;                                             \
            if ((sM >> dstBit) & 1u) {                                                \
                float corr = 0.0f;                                                    \
                float sign = 1.0f;                                                    \
                if ((lM >> dstBit) & 1u) {                                            \
                    const uint8_t wf = flL[li + kCx[I] + kCy[I] * pitch];             \
                    if (wf == kOutflow) {                                             \
                        sign = -1.0f;                                                 \
                    } else {                                                          \
                        float uwx, uwy, uwz;                                          \
                        wall_vel_row(wf, prow, kCy[I], p, uwx, uwy, uwz);             \
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
                stb(op + (uint32_t)kOpp[I] * TB, sign * fh[I] - corr);               \
            }                                                                         \
        }
        STRH_RWALLOUT(1) STRH_RWALLOUT(2) STRH_RWALLOUT(3) STRH_RWALLOUT(4) STRH_RWALLOUT(5) STRH_RWALLOUT(6)
        STRH_RWALLOUT(7) STRH_RWALLOUT(8) STRH_RWALLOUT(9) STRH_RWALLOUT(10) STRH_RWALLOUT(11) STRH_RWALLOUT(12)
        STRH_RWALLOUT(13) STRH_RWALLOUT(14) STRH_RWALLOUT(15) STRH_RWALLOUT(16) STRH_RWALLOUT(17) STRH_RWALLOUT(18)
#undef STRH_RWALLOUT
    }
    if (writes && (sM != 0u || eps > 0.0f))                        // 351-353
        force[n] = make_float4(ax, ay, az, 0.0f);
}

// Direction sets as bit masks over i = 0..18 (bit i): the moving directions
// with a component in the xy plane, and those with c_x = -1 / +1, c_y = -1 / +1.
constexpr uint32_t dir_mask(int ax, int v) {
    uint32_t m = 0;
    for (int i = 1; i < 19; i++) {
        const int c = ax == 0 ? kCx[i] : kCy[i];
        if (c == v && (kCx[i] != 0 || kCy[i] != 0)) m |= 1u << i;
    }
    return m;
}
constexpr uint32_t kPlanar = dir_mask(0, -1) | dir_mask(0, 1) | dir_mask(1, -1) | dir_mask(1, 1);
constexpr uint32_t kXm = dir_mask(0, -1), kXp = dir_mask(0, 1), kYm = dir_mask(1, -1), kYp = dir_mask(1, 1);

// ---- the edge exchange -----------------------------------------------------
// Word of (slot, cell (x, y)): slots whose direction moves along x are kept
// x-major (index x ny + y), the others y-major (y nx + x), so the words a
// tile trades across a vertical edge, and the straight ones across a
// horizontal edge, lie next to each other in memory.
__device__ __forceinline__ size_t x_word(uint32_t slot, uint32_t N, int x, int y, int nx, int ny) {
    const bool alongX = ((kXm | kXp) >> slot) & 1u;
    return (size_t)slot * N + (alongX ? (uint32_t)x * (uint32_t)ny + (uint32_t)y : (uint32_t)y * (uint32_t)nx + (uint32_t)x);
}

__device__ __forceinline__ void x_put(uint64_t* X, size_t word, uint32_t tag, uint32_t bits) {
    __hip_atomic_store(X + word, ((uint64_t)tag << 32) | bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T> __device__ __forceinline__ uint32_t raw_bits(T v) {
    if constexpr (sizeof(T) == 4) return __builtin_bit_cast(uint32_t, v);
    else return (uint32_t)v;
}
template <class T> __device__ __forceinline__ T from_bits(uint32_t b) {
    if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, b);
    else return (T)b;
}

// The planar directions i with (lx + c_ix, ly + c_iy) inside the w x h tile.
__device__ __forceinline__ uint32_t stays_in(int lx, int ly, int w, int h) {
    uint32_t m = kPlanar;
    if (lx == -1) m &= kXp;
    else if (lx == w) m &= kXm;
    else { if (lx == 0) m &= ~kXm; if (lx == w - 1) m &= ~kXp; }
    if (ly == -1) m &= kYp;
    else if (ly == h) m &= kYm;
    else { if (ly == 0) m &= ~kYm; if (ly == h - 1) m &= ~kYp; }
    return m;
}
// The set of the opposite directions (pairs 2k - 1, 2k swap).
__device__ __forceinline__ uint32_t opposite_set(uint32_t m) {
    constexpr uint32_t odd = 0x2aaaau, even = 0x55554u;   // bits 1, 3, .., 17 and 2, 4, .., 18
    return ((m & odd) << 1) | ((m & even) >> 1);
}
__device__ __forceinline__ int opp_dir(int i) { return ((i - 1) ^ 1) + 1; }

// After a pass of parity `par`: publish this tile's share of the exchange.
// lwRing / lwHalo: the link words of the ring and halo cells in their
// ring_cell / halo_cell order, held in LDS for the launch.
template <class T, int BS>
__device__ __forceinline__ void x_publish(const T* __restrict__ s, const ResTile& tl,
                                          const BatchLat& B, uint64_t* X, int par, uint32_t tag,
                                          const uint32_t* __restrict__ lwRing,
                                          const uint32_t* __restrict__ lwHalo) {
    const int w = tl.w, h = tl.h, pitch = w + 2, nx = B.p.nx, ny = B.p.ny;
    const uint32_t N = B.p.N;
    // even: edge cell b publishes slot opp(i) for each i that leaves the tile;
    // odd: halo cell h publishes slot i for each i pushed into it from the tile
    const int cnt = par == 0 ? 2 * w + 2 * (h - 2) : 2 * (w + 2) + 2 * h;
    for (int r = threadIdx.x; r < cnt; r += (int)blockDim.x) {
        int lx, ly;
        if (par == 0) ring_cell(r, w, h, lx, ly); else halo_cell(r, w, h, lx, ly);
        const uint32_t lw = par == 0 ? lwRing[r] : lwHalo[r];
        if (lw & kNotFluidBit) continue;
        const int gx = wrap_coord(tl.x0 + lx, nx), gy = wrap_coord(tl.y0 + ly, ny);
        const uint32_t li = (uint32_t)rl_index(lx, ly, pitch);
        // only links between fluid cells carry a population across: along a
        // wall link the cell bounces its own populations back instead
        const uint32_t open = ~((lw & 0x3ffffu) << 1);
        const uint32_t in = stays_in(lx, ly, w, h) & open;
        // slots: even opp(i) for each i leaving the tile to a fluid cell;
        // odd: i for each i pushed in from a fluid cell of the tile
        uint32_t m = par == 0 ? opposite_set(kPlanar & ~stays_in(lx, ly, w, h) & open) : opposite_set(in);
#pragma unroll 1
        while (m != 0u) {
            const int slot = __builtin_ctz(m);
            m &= m - 1u;
            x_put(X, x_word((uint32_t)slot, N, gx, gy, nx, ny), tag, raw_bits<T>(s[li * 19u + (uint32_t)slot]));
        }
    }
}

// Before the next pass: load the neighbours' share (the same words, chosen
// from the same cell's link word). False if the batch aborted.
template <class T, int BS>
__device__ __forceinline__ bool x_collect(T* __restrict__ s, const ResTile& tl, const BatchLat& B,
                                          const uint64_t* X, int par, uint32_t tag, int* abortFlag,
                                          uint64_t spinTicks, const uint32_t* __restrict__ lwRing,
                                          const uint32_t* __restrict__ lwHalo) {
    const int w = tl.w, h = tl.h, pitch = w + 2, nx = B.p.nx, ny = B.p.ny;
    const uint32_t N = B.p.N;
    // even: halo cell h loads slot opp(i) for each i with h + c_i in the tile;
    // odd: edge cell b loads slot i for each i with b - c_i outside the tile
    const int cnt = par == 0 ? 2 * (w + 2) + 2 * h : 2 * w + 2 * (h - 2);
    for (int r = threadIdx.x; r < cnt; r += (int)blockDim.x) {
        int lx, ly;
        if (par == 0) halo_cell(r, w, h, lx, ly); else ring_cell(r, w, h, lx, ly);
        const uint32_t lw = par == 0 ? lwHalo[r] : lwRing[r];
        if (lw & kNotFluidBit) continue;
        const int gx = wrap_coord(tl.x0 + lx, nx), gy = wrap_coord(tl.y0 + ly, ny);
        const uint32_t li = (uint32_t)rl_index(lx, ly, pitch);
        const uint32_t open = ~((lw & 0x3ffffu) << 1);
        const uint32_t in = stays_in(lx, ly, w, h) & open;
        uint32_t m = par == 0 ? opposite_set(in) : opposite_set(kPlanar & ~stays_in(lx, ly, w, h) & open);
        // every word of the cell in flight at once (a cell has at most 9);
        // while some are not there yet, all of those are polled together
        uint64_t v[9];
        int slot[9];
        size_t word[9];
        uint32_t pending = 0u;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            slot[k] = m != 0u ? __builtin_ctz(m) : -1;
            m &= m - 1u;
            v[k] = 0u;
            word[k] = slot[k] >= 0 ? x_word((uint32_t)slot[k], N, gx, gy, nx, ny) : 0u;
            if (slot[k] >= 0) {
                v[k] = __hip_atomic_load(X + word[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                pending |= 1u << k;
            }
        }
        const uint64_t t0 = pending != 0u ? wall_clock64() : 0u;
        for (;;) {
#pragma unroll
            for (int k = 0; k < 9; k++)
                if (((pending >> k) & 1u) && (uint32_t)(v[k] >> 32) == tag) pending &= ~(1u << k);
            if (pending == 0u) break;
            if (__hip_atomic_load(abortFlag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
            if (wall_clock64() - t0 > spinTicks) {
                __hip_atomic_store(abortFlag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
            __builtin_amdgcn_s_sleep(1);
#pragma unroll
            for (int k = 0; k < 9; k++)
                if ((pending >> k) & 1u)
                    v[k] = __hip_atomic_load(X + word[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
#pragma unroll
        for (int k = 0; k < 9; k++)
            if (slot[k] >= 0) s[li * 19u + (uint32_t)slot[k]] = from_bits<T>((uint32_t)v[k]);
    }
    return true;
}

// One of four per-thread values by a runtime index below 4 (selects, so the
// values stay in registers).
template <class V> __device__ __forceinline__ V pick4(const V (&a)[4], int j) {
    return j == 0 ? a[0] : (j == 1 ? a[1] : (j == 2 ? a[2] : a[3]));
}

// Steps k0 .. k0 + K - 1 of the tiles of one group of lattices (K <= kProbeSpan).
// A workgroup of blockDim.x <= BS threads; thread j takes the tile's cells
// j, j + blockDim.x, ... (at most kResidentMaxRounds of them) in the order
// edge ring, interior, and keeps their link words and solid fractions in
// registers for the whole launch.
template <class Conv, int BS, bool LES, bool FORCE>
__global__ void __launch_bounds__(BS)
step_resident(const BatchLat* __restrict__ lats, const ResTile* __restrict__ tiles,
              const int64_t* __restrict__ base, int k0, int K, int* __restrict__ abortFlag,
              uint64_t spinTicks) {
    static_assert(kResidentMaxRounds == 4, "pick4 holds four rounds");
    using T = typename Conv::T;
    extern __shared__ __align__(16) unsigned char smem[];
    T* s = (T*)smem;
    __shared__ int stop;
    __shared__ float rampL[kProbeSpan];   // the launch's ramp factors, read once
    const ResTile tl = tiles[blockIdx.x];
    const BatchLat& B = lats[tl.lat];
    const int w = tl.w, h = tl.h, pitch = w + 2, nx = B.p.nx, ny = B.p.ny;
    const int bsz = (int)blockDim.x, tid = (int)threadIdx.x;
    const uint32_t cs = (uint32_t)((w + 2) * (h + 2));
    GLOBAL T* f = (GLOBAL T*)B.f;
    const GLOBAL uint8_t* flg = (const GLOBAL uint8_t*)B.flags;
    const GLOBAL uint32_t* lidm = (const GLOBAL uint32_t*)B.lidm;
    const DdfLayout lay = B.p.lay;
    if (tid == 0) stop = 0;
    const int R = 2 * w + 2 * (h - 2), C = w * h;
    const int ringRounds = (R + bsz - 1) / bsz, nRounds = (C + bsz - 1) / bsz;
    const GLOBAL uint32_t* link = (const GLOBAL uint32_t*)B.link;
    // Work order: the ring in the first rounds, then the interior; inside each
    // part plain fluid cells first, then partial cells, cells with wall
    // links, cells with moving-wall links (the corrections) and non-fluid
    // cells, so that cells taking the same branches share waves. (A cell's update does
    // not depend on the order.) The order is built in the LDS area the tile
    // occupies next.
    __shared__ int bucket[20];
    uint16_t* order = (uint16_t*)smem;
    const bool useEps = B.p.useEps != 0;
    const GLOBAL float* epsG = (const GLOBAL float*)B.eps;
    auto cell_of = [&](int e, int& lx, int& ly) {
        if (e < R) {
            ring_cell(e, w, h, lx, ly);
        } else {
            const uint32_t q = fdiv((uint32_t)(e - R), tl.divi);
            lx = 1 + (int)((uint32_t)(e - R) - q * tl.divi.d);
            ly = 1 + (int)q;
        }
    };
    auto kind_of = [&](int e) -> int {
        int lx, ly;
        cell_of(e, lx, ly);
        const uint32_t n = (uint32_t)(tl.y0 + ly) * (uint32_t)nx + (uint32_t)(tl.x0 + lx);
        const uint32_t lw = link[n];
        int k = 0;
        if (lw & kNotFluidBit) k = 4;
        else if ((lw & 0x3ffffu) != 0u) k = lidm[n] != 0u ? 3 : 2;
        else if (useEps && epsG[n] > 0.0f) k = 1;
        return (e < R ? 0 : 5) + k;
    };
    if (tid < 20) bucket[tid] = 0;
    __syncthreads();
    for (int e = tid; e < C; e += bsz) atomicAdd(&bucket[kind_of(e)], 1);
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int b = 0; b < 10; b++) { const int c = bucket[b]; bucket[b] = acc; bucket[10 + b] = 0; acc += c; }
    }
    __syncthreads();
    for (int e = tid; e < C; e += bsz) {
        const int b = kind_of(e);
        order[bucket[b] + atomicAdd(&bucket[10 + b], 1)] = (uint16_t)e;
    }
    __syncthreads();
    // this thread's cells: tile coordinates, link word, solid fraction
    int cpos[4];
    uint32_t clw[4], clm[4];
    float ceps[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int q = j * bsz + tid;
        cpos[j] = 0;
        clw[j] = kNotFluidBit;   // no cell: skipped like a non-fluid one
        clm[j] = 0u;
        ceps[j] = 0.0f;
        if (j < nRounds && q < C) {
            int lx, ly;
            cell_of((int)order[q], lx, ly);
            const uint32_t n = (uint32_t)(tl.y0 + ly) * (uint32_t)nx + (uint32_t)(tl.x0 + lx);
            cpos[j] = lx | (ly << 16);
            clw[j] = link[n];
            if (!(clw[j] & kNotFluidBit) && (clw[j] & 0x3ffffu) != 0u) clm[j] = lidm[n];
            if (useEps) ceps[j] = epsG[n];
        }
    }
    __syncthreads();   // the order is read: the tile may overwrite it
    // the tile and its halo ring from the image, their flags after them, then
    // two buffers of per-row inflow values (steps of even / odd k)
    uint8_t* flL = smem + (size_t)cs * 19u * sizeof(T);
    float* profL = (float*)(flL + ((cs + 3u) & ~3u));
    const int rows = h + 2;
    const bool parabolic = B.inflowRows != 0;
    // the link words of the ring and halo cells, which the exchange reads every step
    uint32_t* lwRing = (uint32_t*)(profL + 2 * rows);
    uint32_t* lwHalo = lwRing + R;
    const int Hc = 2 * (w + 2) + 2 * h;
    for (int r = tid; r < R + Hc; r += bsz) {
        int lx, ly;
        if (r < R) ring_cell(r, w, h, lx, ly); else halo_cell(r - R, w, h, lx, ly);
        lwRing[r] = link[(uint32_t)wrap_coord(tl.y0 + ly, ny) * (uint32_t)nx + (uint32_t)wrap_coord(tl.x0 + lx, nx)];
    }
    for (uint32_t c = (uint32_t)tid; c < cs; c += (uint32_t)bsz) {
        const int ly = (int)(c / (uint32_t)pitch) - 1, lx = (int)(c % (uint32_t)pitch) - 1;
        const uint32_t n = (uint32_t)wrap_coord(tl.y0 + ly, ny) * (uint32_t)nx + (uint32_t)wrap_coord(tl.x0 + lx, nx);
        const uint32_t e = cell_elem(lay, n);
        flL[c] = flg[n];
        T v[19];   // all 19 loads in flight, then the LDS writes
#pragma unroll
        for (int i = 0; i < 19; i++) v[i] = f[(size_t)e + (size_t)i * lay.W];
#pragma unroll
        for (int i = 0; i < 19; i++) s[c * 19u + (uint32_t)i] = v[i];
    }
    const int64_t tBase = B.step0 + *base + k0;
    for (int q = tid; q < K; q += bsz) {
        const int64_t t = tBase + q;                                   // Sim.swift:258-266
        rampL[q] = t < B.rampSteps ? ((const GLOBAL float*)B.ramp)[t] : 1.0f;
    }
    if (parabolic) {   // the first step's rows (its ramp factor read as rampL[0] is)
        const float r0 = tBase < B.rampSteps ? ((const GLOBAL float*)B.ramp)[tBase] : 1.0f;
        inflow_rows(profL, rows, tl.y0, ny, B.uinTarget * r0, tid, bsz);
    }
    __syncthreads();
    // the probe schedule followed step by step (one division per launch)
    int64_t tf = INT64_MAX, jf = 0;   // next force step and its sample index
    if (B.fCount > 0) {
        jf = tBase <= B.fT0 ? 0 : (tBase - B.fT0 + B.fPeriod - 1) / B.fPeriod;
        tf = B.fT0 + jf * B.fPeriod;
    }
    int64_t ringSlot = B.fCount > 0 ? jf % B.fRing : 0;
    for (int k = 0; k < K; k++) {
        const int64_t t = tBase + k;
        if (t >= B.tEnd) break;   // the lattice is done: every tile of it stops here
        StepParams p = B.p;
        const int par = (int)(t & 1);
        p.parity = par;
        const float r = rampL[k];
        p.uin = B.uinTarget * r;
        p.ulidX = B.lid[0] * r;
        p.ulidY = B.lid[1] * r;
        p.ulidZ = B.lid[2] * r;
        const float* prof = profL + (k & 1) * rows;   // this step's inflow rows
        // the batch kernel's is_force_step / force_slot, incrementally
        float4* fo = B.force;
        p.writeForce = 0;
        if (t == tf) {
            if (jf < B.fCount) {
                p.writeForce = 1;
                fo = B.force + (size_t)ringSlot * B.p.N;
            }
            tf += B.fPeriod;
            jf++;
            ringSlot = ringSlot + 1 == B.fRing ? 0 : ringSlot + 1;
        }
        const uint32_t tag = (uint32_t)(t - B.tagBase + 1);
        // The rounds that hold the ring come first, then the barrier and the
        // ring's exchange words, then the remaining rounds while the words
        // travel.
        auto pass = [&](auto parc) {
            constexpr int PAR = decltype(parc)::value;
#pragma unroll 1
            for (int j = 0; j < nRounds; j++) {
                if (j == ringRounds) {
                    __syncthreads();
                    x_publish<T, BS>(s, tl, B, B.xchg, PAR, tag, lwRing, lwHalo);
                }
                const uint32_t lw = pick4(clw, j);
                if (!(lw & kNotFluidBit)) {
                    const int pos = pick4(cpos, j);
                    const int lx = pos & 0xffff, ly = pos >> 16;
                    const int x = tl.x0 + lx, y = tl.y0 + ly;
                    resident_cell<Conv, PAR, LES, FORCE>(s, rl_index(lx, ly, pitch), pitch, flL,
                                                         prof + ly + 1, pick4(clm, j), p,
                                                         (uint32_t)y * (uint32_t)nx + (uint32_t)x, x, lw,
                                                         pick4(ceps, j), fo);
                }
            }
            if (ringRounds >= nRounds) {
                __syncthreads();
                x_publish<T, BS>(s, tl, B, B.xchg, PAR, tag, lwRing, lwHalo);
            }
        };
        if (par == 0) pass(std::integral_constant<int, 0>{});
        else pass(std::integral_constant<int, 1>{});
        // the next step's rows into the other buffer: the last reads of it
        // were in step k - 1, before the barrier that ended that step
        if (parabolic && k + 1 < K)
            inflow_rows(profL + ((k + 1) & 1) * rows, rows, tl.y0, ny, B.uinTarget * rampL[k + 1], tid, bsz);
        if (!x_collect<T, BS>(s, tl, B, B.xchg, par, tag, abortFlag, spinTicks, lwRing, lwHalo)) stop = 1;
        __syncthreads();
        if (stop) return;
    }
    // the tile's own cells back to the image (non-fluid cells are never
    // updated, and their LDS slots may hold discarded pushes)
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (clw[j] & kNotFluidBit) continue;
        const int lx = cpos[j] & 0xffff, ly = cpos[j] >> 16;
        const uint32_t n = (uint32_t)(tl.y0 + ly) * (uint32_t)nx + (uint32_t)(tl.x0 + lx);
        const uint32_t e = cell_elem(lay, n);
        const uint32_t li = (uint32_t)rl_index(lx, ly, pitch);
        T v[19];
#pragma unroll
        for (int i = 0; i < 19; i++) v[i] = s[li * 19u + (uint32_t)i];
#pragma unroll
        for (int i = 0; i < 19; i++) f[(size_t)e + (size_t)i * lay.W] = v[i];
    }
}

template <class Conv, int BS>
static void launch_resident_conv(bool les, bool forced, const BatchLat* lats, const ResTile* tiles, int nTiles,
                                 int bs, size_t lds, const int64_t* base, int k0, int K, int* abortFlag,
                                 uint64_t spinTicks, hipStream_t s) {
#define STRH_RES(LESV, FV) hipLaunchKernelGGL((step_resident<Conv, BS, LESV, FV>), dim3(nTiles), dim3(bs), lds, s, \
                                              lats, tiles, base, k0, K, abortFlag, spinTicks)
    if (les) { if (forced) STRH_RES(true, true); else STRH_RES(true, false); }
    else { if (forced) STRH_RES(false, true); else STRH_RES(false, false); }
#undef STRH_RES
}

hipError_t launch_resident(int precision, bool les, bool forced, const BatchLat* lats, const ResTile* tiles,
                           int nTiles, int bs, size_t lds, const int64_t* base, int k0, int K, int* abortFlag,
                           uint64_t spinTicks, hipStream_t s) {
    if (bs < 64 || bs > kResidentBS || bs % 64 != 0) return hipErrorInvalidValue;
    switch (precision) {
        case kFP32: launch_resident_conv<ConvF32, kResidentBS>(les, forced, lats, tiles, nTiles, bs, lds, base, k0, K, abortFlag, spinTicks, s); break;
        case kFP16S: launch_resident_conv<ConvF16, kResidentBS>(les, forced, lats, tiles, nTiles, bs, lds, base, k0, K, abortFlag, spinTicks, s); break;
        case kBF16: launch_resident_conv<ConvBF16, kResidentBS>(les, forced, lats, tiles, nTiles, bs, lds, base, k0, K, abortFlag, spinTicks, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class Conv>
static hipError_t prepare_conv(size_t lds, int* perCU) {
    hipError_t e = hipSuccess;
#define STRH_PREP(LESV, FV)                                                                                  \
    if (e == hipSuccess) e = hipFuncSetAttribute((const void*)step_resident<Conv, kResidentBS, LESV, FV>,      \
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);      \
    if (e == hipSuccess) {                                                                                   \
        int o = 0;                                                                                           \
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, step_resident<Conv, kResidentBS, LESV, FV>,     \
                                                         kResidentBS, lds);                                  \
        if (e == hipSuccess && o < *perCU) *perCU = o;                                                       \
    }
    STRH_PREP(false, false) STRH_PREP(false, true) STRH_PREP(true, false) STRH_PREP(true, true)
#undef STRH_PREP
    return e;
}

hipError_t resident_prepare(int precision, size_t lds, int* perCU) {
    *perCU = 1 << 20;
    switch (precision) {
        case kFP32: return prepare_conv<ConvF32>(lds, perCU);
        case kFP16S: return prepare_conv<ConvF16>(lds, perCU);
        case kBF16: return prepare_conv<ConvBF16>(lds, perCU);
    }
    return hipErrorInvalidValue;
}

}  // namespace strh
