// Host-side launch interface of the D3Q19 kernels (the .hip translation units
// beside this header; device code they share is in device.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "lbm.h"

namespace strh {

// Everything one step launch needs. The boolean features select the kernel
// specialisation; the host sets them from the lattice's contents so the
// all-fluid periodic bulk runs the mask-free, eps-free, force-free kernel.
// Workgroup size of the one-cell-per-thread kernel (64..512 measured within
// 2% of each other at 256^3; 256 is the best or tied).
constexpr uint32_t kStepBS = 256;

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
    bool large;        // blocked image of 4 GiB or more (row-staged pass with row bases)
    bool planar;       // one cell deep (nz = 1): the direct odd pass's 18 neighbours are 8 cells
    // partial-cell lattices of many blocks whose partial cells sit in few of
    // them: the kStepBS-cell blocks without one (nPlain, then the nNt with
    // one) in ntSplit, run by the kernel without the partial-cell path and
    // the kernel with it; null: one kernel for every block
    const uint32_t* ntSplit;
    uint32_t nPlain, nNt;
    const float* alpha;  // Brinkman porosity per cell, or null
    StepParams p;
};

// One lattice of a multi-lattice batch (device-resident descriptor).
constexpr uint32_t kBatchMaxChunks = 1024;
struct BatchLat {
    StepParams p;            // per-lattice constants; per-step fields set in-kernel
    void* f;
    const uint8_t* flags;
    const uint32_t* link;
    const uint32_t* lidm;
    float4* force;           // force buffers: a ring of fRing buffers of N cells when probed
    const float* eps;
    int32_t fRing;           // force buffers in the ring (sample j writes buffer j % fRing)
    uint32_t block0;         // first workgroup of this lattice in the batch grid
    uint32_t nblocks;
    int64_t step0;           // step count at the start of the batch run
    int64_t tEnd;            // absolute step at which this lattice stops
    float uinTarget;
    float lid[3];
    const float* ramp;       // ramp factor for t < rampSteps (host-evaluated)
    int64_t rampSteps;
    int32_t hasForce;
    int32_t inflowRows;      // INFLOW cells with the parabolic profile: resident tiles tabulate it per row
    // momentum-exchange probe schedule: odd steps fT0 + j*fPeriod, j < fCount
    int64_t fT0, fPeriod, fCount;
    uint32_t bx0, by0;
    FastDiv fbw, fbh;        // probe box width / height
    uint32_t M, Pb, chunk;   // box cells, reduction chunks (as the single probe)
    double* series;          // 3 doubles per sample
    // resident launches (resident.hip): the edge-exchange words [19][N] and
    // the step whose exchange tag is 1 (the step the batch run started at)
    uint64_t* xchg;
    int64_t tagBase;
};

// One tile of a resident launch: cells [x0, x0 + w) x [y0, y0 + h) of lattice
// `lat` (an index into the launch's BatchLat array; nz = 1, w, h >= 3), held
// in LDS with a one-cell halo ring; divi divides by w - 2 (interior rows).
struct ResTile {
    int32_t lat, x0, y0, w, h;
    FastDiv divi;
};
// Largest workgroup of the resident launch (one workgroup per CU); a launch
// uses a multiple of 64 up to it, and each thread takes at most
// kResidentMaxRounds cells of its tile.
constexpr int kResidentBS = 768;
constexpr int kResidentMaxRounds = 4;
// A wave that waits this long for a neighbour's exchange word aborts the
// batch by default (ticks of the 100 MHz wall clock: one second).
constexpr uint64_t kResidentSpinTicks = 100000000ull;
// LDS bytes a resident tile may use (the workgroup's static LDS, the ramp
// factors of kProbeSpan steps among it, stays below 1,536 bytes).
constexpr size_t kResidentMaxLds = 163840 - 1536;

// Hooks of the recording forward step of the adjoint (walls, Guo forcing,
// Brinkman porosity, partial cells, LES, sponge).
struct AdjHooks {
    float4* rec;             // [N] pre-collision moments (rho - 1, p_x, p_y, p_z), or null
    double* jacc;            // per-cell objective accumulator, or null
    int32_t inletX, outletX; // columns of the pressure-drop term
    double cin, cout;        // w_p / (3 n_in T), w_p / (3 n_out T)
    double wd;               // w_d / T
    // partial cells (lattices with solid fractions)
    const int32_t* ntIdx;    // [N] index of the cell among the M cells with eps > 0, or -1
    float* recNt;            // [19][M] pre-collision populations of those cells, or null
    uint32_t M;
    double wf;               // w_f / T (partial-cell drag)
    void* rec19;             // [19][N] pre-collision populations of every cell (LES), in the
                             // storage format, before the moving-wall corrections; or null
};

// One reverse step: lambda^t from lambda^{t+1} and the recorded moments of g^t.
struct AdjStep {
    uint32_t nx, ny, nz, N, S, blocks;
    FastDiv divx, divy;
    float wp, wm, fx, fy, fz;    // collision rates and body force of step t
    double rOdd;                 // ramp of the odd step that streams after step t
    double uinTarget;
    int32_t inflowUniform;
    int32_t inletX, outletX;     // objective of step t (weights 0 outside the window)
    float cin, cout, wd;
    float wf;                    // partial-cell drag weight of step t
    uint32_t M;                  // cells with a solid fraction
    float spongeX0, spongeInvW, spongeTau;   // Params' sponge (x0 = 1e9: none)
    float cSmago;                // Smagorinsky Cs^2 (LES lattices)
    int32_t srt;                 // omega- == omega+ (the reference's SRT test)
    int32_t precision;           // storage of the recorded populations (LES lattices)
};

hipError_t launch_step(const StepLaunch& L, hipStream_t s);
// launch_step per storage format (one translation unit each)
hipError_t launch_step_fp32(const StepLaunch& L, hipStream_t s);
hipError_t launch_step_fp16(const StepLaunch& L, hipStream_t s);
hipError_t launch_step_bf16(const StepLaunch& L, hipStream_t s);
// Forward step with the adjoint hooks; L.alpha must be set (zeros for a
// lattice without porosity) unless the lattice has solid fractions (L.nt) or
// Smagorinsky LES (L.les), which exclude porosity.
hipError_t launch_step_hooked(const StepLaunch& L, const AdjHooks& hk, hipStream_t s);
// Partial-cell lattices pass eps, ntIdx, recNt (this step's [19][M]) and gEps ([M]).
// LES lattices pass rec19 (this step's [19][N] populations of every cell, storage
// format, before the moving-wall corrections), lidm and the step's parameters sp
// instead of rec / recNt.
hipError_t launch_adjoint_step(const AdjStep& q, bool walls, const float4* rec, const float* lamNext,
                               float* lam, const uint8_t* flags, const uint32_t* link,
                               const float* alpha, double* gAlpha, double* gUin,
                               const float* eps, const int32_t* ntIdx, const float* recNt,
                               double* gEps, const void* rec19, const uint32_t* lidm,
                               const StepParams& sp, hipStream_t s);
hipError_t launch_sum_doubles(const double* v, uint32_t M, double* partials, uint32_t max_partials,
                              double* out, hipStream_t s);
// parity: 0 / 1 when every active lattice is at that parity in this launch,
// -1 to branch per lattice (then les and forced are taken as true); les: some
// lattice has a Smagorinsky constant; forced: some lattice has a body force.
hipError_t launch_batch_step(int precision, int bs, int parity, bool les, bool forced, bool planar,
                             const BatchLat* lats,
                             int L, uint32_t totalBlocks, const int64_t* base, int k, hipStream_t s);
// Probes every force step among steps k .. k + span - 1 (span <= kProbeSpan).
// A batch reduces its force samples, and advances a resident group per
// launch, over spans of kProbeSpan steps, or kProbeSpanSmall where a
// lattice's ring of force buffers for the longer span would be too large.
constexpr int kProbeSpan = 256;
constexpr int kProbeSpanSmall = 64;
// parts: 3 L maxSamples maxPb doubles (maxPb: the largest Pb of the lattices,
// maxSamples: the most force steps any of them has in one span).
hipError_t launch_batch_probe(BatchLat* lats, int L, const int64_t* base, int k, int span, double* parts,
                              int maxPb, int maxSamples, hipStream_t s);
hipError_t launch_advance_base(int64_t* base, int64_t by, hipStream_t s);
// Steps k0 .. k0 + K - 1 (K <= kProbeSpan) of every tile of one resident group;
// lds is the largest tile's (w + 2)(h + 2) 19 sizeof(T) bytes.
hipError_t launch_resident(int precision, bool les, bool forced, const BatchLat* lats, const ResTile* tiles,
                           int nTiles, int bs, size_t lds, const int64_t* base, int k0, int K, int* abortFlag,
                           uint64_t spinTicks, hipStream_t s);
// Allows kResidentMaxLds of dynamic LDS for every resident kernel of this
// precision and returns how many of its workgroups one CU holds at that size.
hipError_t resident_prepare(int precision, size_t lds, int* perCU);
hipError_t launch_set_base(int64_t* base, int64_t v, hipStream_t s);
void box_reduce_plan(uint32_t M, uint32_t max_partials, uint32_t* Pb, uint32_t* chunk);
// Workgroup size of the LDS row-staged odd pass for this lattice, or 0.
uint32_t rows_block_size(uint32_t nx, uint64_t N);
hipError_t launch_moments(int precision, const void* f, const uint8_t* flags, float* out,
                          uint32_t N, DdfLayout lay, float fx, float fy, float fz, hipStream_t s);
hipError_t launch_init(int precision, void* f, uint32_t nx, uint32_t ny, uint32_t nz,
                       DdfLayout lay, uint32_t mode, float amplitude, float tau, hipStream_t s);
hipError_t launch_mass_sum(int precision, const void* f, uint32_t N, DdfLayout lay,
                           double* partials, uint32_t max_partials, double* out, hipStream_t s);
hipError_t launch_box_force_sum(const float* force, uint32_t nx, uint32_t ny, uint32_t nz,
                                uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1,
                                double* partials, uint32_t max_partials, double* out3,
                                hipStream_t s);
hipError_t launch_dissipation_sum(int precision, const void* f, const uint8_t* flags, uint32_t N,
                                  DdfLayout lay, double tauP, double* partials, uint32_t max_partials,
                                  double* out, hipStream_t s);
// Block partials of the fixed-order FP64 reductions (stage 1 writes at most
// this many; box_reduce_plan caps the plan at it).
constexpr uint32_t kReducePartials = 19 * 512;
// The allocation-free kinetic energy's static device scratch: the partials and
// the result (77,856 bytes).
constexpr uint32_t kScratchDoubles = kReducePartials + 4;
// Device address of that scratch on the current device.
hipError_t scratch_address(double** out);
// sum over cells of |u|^2 straight from the DDFs (no moments field); the same
// bits as launch_moments followed by launch_kinetic_sum.
hipError_t launch_kinetic_sum_cells(int precision, const void* f, const uint8_t* flags, uint32_t N,
                                    DdfLayout lay, float fx, float fy, float fz, double* partials,
                                    uint32_t max_partials, double* out, hipStream_t s);
hipError_t launch_kinetic_sum(const float* moments, uint32_t N, double* partials,
                              uint32_t max_partials, double* out, hipStream_t s);
hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t s);

}  // namespace strh
