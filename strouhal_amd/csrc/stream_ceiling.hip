// Streaming ceiling of the step kernels' access pattern (a measurement
// library, not part of the solver's C ABI): every cell's 19 slots are read
// and 19 written in the blocked DDF layout (a block of W cells keeps its 19
// slots of W values one after another, W * B = 32 KiB), with nontemporal
// loads and stores and no arithmetic. The slot order of the stores is
// reversed, as the even pass writes slot i's value to slot opp(i) of the same
// cell. The time of one such pass bounds a step that moves the same bytes in
// the same pattern.
#include <hip/hip_runtime.h>

#include <stdint.h>

namespace {

typedef uint32_t v2u __attribute__((ext_vector_type(2)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));

constexpr size_t kBlockSlotBytes = 32768;

// V: the bytes one lane moves per access (4, 8 or 16). nvec = vectors per
// slot over the whole image, Wv = vectors per slot of one block.
template <class V>
__global__ void __launch_bounds__(256) pass19(const V* __restrict__ src, V* __restrict__ dst,
                                              uint32_t nvec, uint32_t Wv) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= nvec) return;
    const uint32_t blk = i / Wv, r = i - blk * Wv;
    const size_t base = (size_t)blk * 19u * Wv + r;
    V v[19];
#pragma unroll
    for (int s = 0; s < 19; s++) v[s] = __builtin_nontemporal_load(src + base + (size_t)s * Wv);
#pragma unroll
    for (int s = 0; s < 19; s++) __builtin_nontemporal_store(v[18 - s], dst + base + (size_t)s * Wv);
}

template <class V>
hipError_t launch(const void* src, void* dst, size_t slotBytes) {
    const uint32_t nvec = (uint32_t)(slotBytes / sizeof(V));
    const uint32_t Wv = (uint32_t)(kBlockSlotBytes / sizeof(V));
    hipLaunchKernelGGL(pass19<V>, dim3((nvec + 255u) / 256u), dim3(256), 0, 0, (const V*)src, (V*)dst,
                       nvec, Wv);
    return hipGetLastError();
}

}  // namespace

extern "C" {

// Time `reps` passes over an image of `cells` cells of `storage_bytes` per
// value (4 or 2) moved `lane_bytes` (4, 8, 16) per lane access; *ms_per_pass
// is the HIP-event mean. cells * storage_bytes must be a multiple of 32 KiB
// (whole blocks). Returns 0, or -1 on bad arguments, or the HIP error.
__attribute__((visibility("default")))
int strh_stream_ceiling(uint64_t cells, int storage_bytes, int lane_bytes, int reps, float* ms_per_pass) {
    if (!ms_per_pass || reps <= 0 || (storage_bytes != 4 && storage_bytes != 2) ||
        (lane_bytes != 4 && lane_bytes != 8 && lane_bytes != 16))
        return -1;
    const size_t slotBytes = (size_t)cells * (size_t)storage_bytes;
    if (slotBytes == 0 || slotBytes % kBlockSlotBytes != 0 || slotBytes / 4 > 0xffffffffu) return -1;
    const size_t bytes = 19 * slotBytes;
    void *src = nullptr, *dst = nullptr;
    hipError_t e = hipMalloc(&src, bytes);
    if (e == hipSuccess) e = hipMalloc(&dst, bytes);
    if (e == hipSuccess) e = hipMemset(src, 0, bytes);
    if (e == hipSuccess) e = hipMemset(dst, 0, bytes);
    hipEvent_t a = nullptr, b = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    auto pass = [&]() {
        return lane_bytes == 4 ? launch<uint32_t>(src, dst, slotBytes)
             : lane_bytes == 8 ? launch<v2u>(src, dst, slotBytes) : launch<v4u>(src, dst, slotBytes);
    };
    for (int k = 0; k < 3 && e == hipSuccess; k++) e = pass();                 // warm-up
    if (e == hipSuccess) e = hipEventRecord(a, 0);
    for (int k = 0; k < reps && e == hipSuccess; k++) e = pass();
    if (e == hipSuccess) e = hipEventRecord(b, 0);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (e == hipSuccess) *ms_per_pass = ms / (float)reps;
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    if (src) hipFree(src);
    if (dst) hipFree(dst);
    return (int)e;
}

}  // extern "C"
