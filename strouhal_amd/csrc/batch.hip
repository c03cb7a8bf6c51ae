// Multi-lattice batch kernels: one launch per step of every lattice, the
// device-side force probes, the step base.
#include "device.h"

namespace strh {

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

// PAR = 0 / 1: the host guarantees that every active lattice is at that
// parity in this launch (equal start parities, even graph lengths), so only
// that pass is compiled in; PAR = -1 branches per lattice. LES = false when no
// lattice of the batch has a Smagorinsky constant, FORCED = false when none has
// a body force (otherwise each lattice takes its own forcing path). PLANAR: every
// lattice is one cell deep, so the odd pass's 18 neighbours are 8 cells.
template <class Conv, int BS, int PAR, bool LES, bool FORCED, bool PLANAR>
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
        // Probe steps are odd (strh_batch_set_probe refuses others), so the
        // even pass never writes a force: its output is compiled out, which
        // frees its accumulators' registers.
        using St = Store<Conv, kPolShifted>;
        if (FORCED && B.hasForce)
            step_cell<St, 0, true, true, LES, true, false, false, false>(f, B.flags, B.link, B.lidm, fo, B.eps, p, n);
        else
            step_cell<St, 0, true, true, LES, false, false, false, false>(f, B.flags, B.link, B.lidm, fo, B.eps, p, n);
    } else if (PAR != 0) {
        using St = Store<Conv, kPolShifted>;
        if (FORCED && B.hasForce)
            step_cell<St, 1, true, true, LES, true, false, false, true, PLANAR>(f, B.flags, B.link, B.lidm, fo, B.eps, p, n);
        else
            step_cell<St, 1, true, true, LES, false, false, false, true, PLANAR>(f, B.flags, B.link, B.lidm, fo, B.eps, p, n);
    }
}

// After batch steps k .. k + span - 1 (span <= kProbeSpan): for every lattice
// and every force step t among them, sum the per-cell momentum exchange of
// that step's ring buffer over the probe box into the time series, with
// exactly the chunking of the single-lattice probe (reduce_stage1 over Pb
// chunks, then reduce_stage2), so the FP64 values are the same bits. Two
// launches: probe_chunks sums chunk pb of component c of lattice l for the
// span's j-th force step into parts[((l 3 + c) maxSamples + j) maxPb + pb];
// probe_totals adds each sample's Pb chunk sums in stage 2's order.
__device__ __forceinline__ int64_t first_sample(const BatchLat& B, int64_t t0) {
    return t0 <= B.fT0 ? B.fT0 : B.fT0 + ((t0 - B.fT0 + B.fPeriod - 1) / B.fPeriod) * B.fPeriod;
}
__device__ __forceinline__ int64_t span_end(const BatchLat& B, int64_t t0, int span) {
    int64_t tLast = t0 + span;
    const int64_t tCap = B.fT0 + B.fCount * B.fPeriod;        // one past the last sample's step
    tLast = tLast < B.tEnd ? tLast : B.tEnd;
    return tLast < tCap ? tLast : tCap;
}

// Block (l, c, j0, pb) of 3 L jb maxPb takes samples j = j0, j0 + jb, ...
__global__ void __launch_bounds__(256)
probe_chunks(const BatchLat* __restrict__ lats, int L, const int64_t* __restrict__ base, int k, int span,
             double* __restrict__ parts, int maxPb, int maxSamples, int jb) {
    __shared__ double lds[4];
    const int pb = (int)(blockIdx.x % (uint32_t)maxPb);
    const int j0 = (int)((blockIdx.x / (uint32_t)maxPb) % (uint32_t)jb);
    const int lc = (int)(blockIdx.x / ((uint32_t)maxPb * (uint32_t)jb));
    const int l = lc / 3, c = lc % 3;
    if (l >= L) return;
    const BatchLat& B = lats[l];
    if (B.fCount <= 0 || (uint32_t)pb >= B.Pb) return;
    const int64_t t0 = B.step0 + *base + k, tLast = span_end(B, t0, span), tFirst = first_sample(B, t0);
    const uint32_t lo = (uint32_t)pb * B.chunk;
    const uint32_t hi = (B.M - lo) > B.chunk ? lo + B.chunk : B.M;
    for (int j = j0; j < maxSamples; j += jb) {
        const int64_t t = tFirst + (int64_t)j * B.fPeriod;
        if (t >= tLast) return;
        const BoxForceElem e{force_slot(B, t), (uint32_t)B.p.nx, (uint32_t)B.p.ny, B.bx0, B.by0, B.fbw, B.fbh, c};
        const double sum = block_tree_sum(chunk_sum(e, 0u, lo, hi), lds);
        if (threadIdx.x == 0) parts[((size_t)lc * maxSamples + j) * maxPb + pb] = sum;
    }
}

// The chunk sums of one sample, read in stage 2's order.
struct GlobalPartials {
    const double* p;
    __device__ double operator()(uint32_t, uint32_t k) const { return p[k]; }
};

// Block (l, c, j) of 3 L maxSamples.
__global__ void __launch_bounds__(256)
probe_totals(BatchLat* __restrict__ lats, int L, const int64_t* __restrict__ base, int k, int span,
             const double* __restrict__ parts, int maxPb, int maxSamples) {
    __shared__ double lds[4];
    const int j = (int)(blockIdx.x % (uint32_t)maxSamples);
    const int lc = (int)(blockIdx.x / (uint32_t)maxSamples);
    const int l = lc / 3, c = lc % 3;
    if (l >= L) return;
    const BatchLat& B = lats[l];
    if (B.fCount <= 0) return;
    const int64_t t0 = B.step0 + *base + k, tLast = span_end(B, t0, span);
    const int64_t t = first_sample(B, t0) + (int64_t)j * B.fPeriod;
    if (t >= tLast) return;
    const GlobalPartials g{parts + ((size_t)lc * maxSamples + j) * maxPb};
    const double total = block_tree_sum(chunk_sum(g, 0u, 0u, B.Pb), lds);
    if (threadIdx.x == 0) B.series[3 * ((t - B.fT0) / B.fPeriod) + c] = total;
}

__global__ void advance_base(int64_t* base, int64_t by) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *base += by;
}

__global__ void set_base(int64_t* base, int64_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *base = v;
}

template <class Conv, int BS, int PAR, bool PLANAR>
static void launch_batch_par(bool les, bool forced, const BatchLat* lats, int L, uint32_t totalBlocks,
                             const int64_t* base, int k, hipStream_t s) {
#define STRH_BATCH(LESV, FV) hipLaunchKernelGGL((step_batch<Conv, BS, PAR, LESV, FV, PLANAR>), dim3(totalBlocks), \
                                                dim3(BS), 0, s, lats, L, base, k)
    if (les) { if (forced) STRH_BATCH(true, true); else STRH_BATCH(true, false); }
    else { if (forced) STRH_BATCH(false, true); else STRH_BATCH(false, false); }
#undef STRH_BATCH
}
template <class Conv, int BS>
static void launch_batch_conv(int parity, bool les, bool forced, bool planar, const BatchLat* lats, int L,
                              uint32_t totalBlocks, const int64_t* base, int k, hipStream_t s) {
    if (parity == 0) launch_batch_par<Conv, BS, 0, false>(les, forced, lats, L, totalBlocks, base, k, s);
    else if (parity == 1 && planar) launch_batch_par<Conv, BS, 1, true>(les, forced, lats, L, totalBlocks, base, k, s);
    else if (parity == 1) launch_batch_par<Conv, BS, 1, false>(les, forced, lats, L, totalBlocks, base, k, s);
    else hipLaunchKernelGGL((step_batch<Conv, BS, -1, true, true, false>), dim3(totalBlocks), dim3(BS), 0, s,
                            lats, L, base, k);
}
template <int BS>
static hipError_t launch_batch_bs(int precision, int parity, bool les, bool forced, bool planar,
                                  const BatchLat* lats, int L, uint32_t totalBlocks, const int64_t* base, int k,
                                  hipStream_t s) {
    switch (precision) {
        case kFP32: launch_batch_conv<ConvF32, BS>(parity, les, forced, planar, lats, L, totalBlocks, base, k, s); break;
        case kFP16S: launch_batch_conv<ConvF16, BS>(parity, les, forced, planar, lats, L, totalBlocks, base, k, s); break;
        case kBF16: launch_batch_conv<ConvBF16, BS>(parity, les, forced, planar, lats, L, totalBlocks, base, k, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_batch_step(int precision, int bs, int parity, bool les, bool forced, bool planar,
                             const BatchLat* lats, int L, uint32_t totalBlocks, const int64_t* base, int k,
                             hipStream_t s) {
    switch (bs) {
        case 128: return launch_batch_bs<128>(precision, parity, les, forced, planar, lats, L, totalBlocks, base, k, s);
        case 256: return launch_batch_bs<256>(precision, parity, les, forced, planar, lats, L, totalBlocks, base, k, s);
        case 512: return launch_batch_bs<512>(precision, parity, les, forced, planar, lats, L, totalBlocks, base, k, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_batch_probe(BatchLat* lats, int L, const int64_t* base, int k, int span, double* parts,
                              int maxPb, int maxSamples, hipStream_t s) {
    if (L <= 0 || maxPb <= 0 || maxSamples <= 0 || !parts) return hipErrorInvalidValue;
    // at most about 2^20 blocks of chunk sums; each takes every jb-th sample
    int jb = maxSamples;
    while (jb > 1 && (int64_t)3 * L * jb * maxPb > ((int64_t)1 << 20)) jb = (jb + 1) / 2;
    if ((int64_t)3 * L * jb * maxPb > ((int64_t)1 << 30) || (int64_t)3 * L * maxSamples > ((int64_t)1 << 30))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(probe_chunks, dim3((uint32_t)(3 * L * jb * maxPb)), dim3(256), 0, s, lats, L, base, k, span,
                       parts, maxPb, maxSamples, jb);
    hipLaunchKernelGGL(probe_totals, dim3((uint32_t)(3 * L * maxSamples)), dim3(256), 0, s, lats, L, base, k, span,
                       parts, maxPb, maxSamples);
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

}  // namespace strh
