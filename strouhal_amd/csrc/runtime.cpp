// Host runtime of the solver and its C ABI (include/strouhal_hip.h).
//
// The MI355X counterpart of the reference's `GPU` + `Simulation`
// (Sources/StrouhalCore/Sim.swift:60-459): lattice construction with the
// neighbour-mask builder, the per-step parameter block and cosine ramp, the
// stepping loop, the probes and the SHA-256 state digest. Every call reports
// failure through a status code and a thread-local message; nothing aborts.
#include <hip/hip_runtime.h>

#include <math.h>

#include <cmath>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <string>
#include <thread>
#include <vector>

#include "../../include/strouhal_hip.h"
#include "kernels.h"
#include "lbm.h"
#include "sha256.h"

using namespace strh;

namespace {

thread_local std::string g_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(STRH_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                        \
    do {                                           \
        hipError_t _e = (expr);                    \
        if (_e != hipSuccess) return hip_fail(_e, what); \
    } while (0)

constexpr uint32_t kMaxPartials = kReducePartials;
// The largest lattice: 19*N must index the mass reduction in 32 bits and the
// per-slot byte offsets of the step kernel must fit 32 bits.
constexpr uint64_t kMaxCells = 200000000ull;

}  // namespace

struct strh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    std::string name;
    // device address of this device's static reduction scratch (observe.hip),
    // or null with the reason in scratchError
    double* scratch = nullptr;
    std::string scratchError;
};

struct strh_lattice {
    strh_ctx* ctx = nullptr;
    int precision = kFP32;
    int nx = 0, ny = 0, nz = 0;
    uint32_t N = 0, S = 0;      // cells; N rounded up to 64
    size_t B = 4;               // bytes per stored DDF
    DdfLayout lay;              // block layout of f (lbm.h)
    void* f = nullptr;          // the DDFs, ddf_elements(lay) elements
    uint8_t* flags = nullptr;   // N raw Cell values
    uint32_t* link = nullptr;   // N link words (walls only)
    uint32_t* lidm = nullptr;   // N moving-wall masks (walls only)
    float* force = nullptr;     // 4N floats (wants_forces)
    float* eps = nullptr;       // N solid fractions (after set_solid_fractions)
    uint32_t* ntSplit = nullptr;   // step blocks without, then with, partial cells (StepLaunch)
    uint32_t nPlainBlocks = 0, nNtBlocks = 0;
    float* alpha = nullptr;     // N Brinkman porosities (after set_porosity)
    float* moments = nullptr;   // 4N floats, allocated on first probe
    double* partials = nullptr; // reduction scratch
    double* result = nullptr;   // 3 doubles
    void* snap = nullptr;       // snapshot of f
    int64_t snapSteps = -1;
    bool walls = false, useEps = false, wantsForces = false;
    bool hasInflow = false;     // some cell is INFLOW
    // Sim.swift:164-172
    float omega = 1.0f, omegaMinus = 1.0f;
    float lid[3] = {0, 0, 0};
    float uinTarget = 0.0f;
    int64_t rampSteps = 0;
    float bodyForce[3] = {0, 0, 0};
    float cSmago = 0.0f;
    bool inflowUniform = false;
    bool spongeOn = false;
    float spongeX0 = 0, spongeWidth = 1, spongeTau = 1;
    int64_t stepsDone = 0;
    double gpuSeconds = 0.0;
};

namespace {

size_t ddf_bytes(const strh_lattice* L) { return (size_t)ddf_elements(L->lay) * L->B; }

// A blocked image of 4 GiB or more: only the even pass and the row-staged odd
// pass reach it (per-workgroup block / per-row bases); porosity, batches and
// the adjoint, which run the direct odd pass, refuse such lattices.
bool large_image(const strh_lattice* L) { return L->lay.blocks > 1 && ddf_bytes(L) >= (1ull << 32); }

// A lattice the resident batch launch can hold: one cell deep, at least 3 x 3.
bool resident_candidate(const strh_lattice* L) { return L->nz == 1 && L->nx >= 3 && L->ny >= 3; }

// Copy cells [c0, c0 + count) of DDF slot `slot` between the device image
// and the contiguous buffer `lin` (count elements), in at most three copies:
// the rest of a partly covered first block, one pitched copy of the whole
// blocks, and the tail.
hipError_t slot_copy(const strh_lattice* L, uint32_t slot, uint64_t c0, uint64_t count, void* lin,
                     bool toDevice, hipStream_t s) {
    const DdfLayout& l = L->lay;
    const uint64_t B = L->B, W = l.W;
    char* h = (char*)lin;
    auto dptr = [&](uint64_t c) {
        return (char*)L->f + ((c >> l.lw) * (uint64_t)l.bstride + slot * W + (c & l.mask)) * B;
    };
    const hipMemcpyKind kind = toDevice ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    auto flat = [&](uint64_t c, uint64_t n) {
        return toDevice ? hipMemcpyAsync(dptr(c), h, n * B, kind, s) : hipMemcpyAsync(h, dptr(c), n * B, kind, s);
    };
    hipError_t e = hipSuccess;
    if (l.lw == 31u) return count ? flat(c0, count) : hipSuccess;   // one block: plain slot-major
    uint64_t c = c0, left = count;
    const uint64_t w = c & (W - 1);
    if (w != 0 && left != 0) {
        const uint64_t take = std::min(left, W - w);
        if ((e = flat(c, take)) != hipSuccess) return e;
        c += take; h += take * B; left -= take;
    }
    const uint64_t full = left / W;
    if (full != 0) {
        const size_t pitch = (size_t)l.bstride * B, row = (size_t)(W * B);
        e = toDevice ? hipMemcpy2DAsync(dptr(c), pitch, h, row, row, full, kind, s)
                     : hipMemcpy2DAsync(h, row, dptr(c), pitch, row, full, kind, s);
        if (e != hipSuccess) return e;
        c += full * W; h += full * W * B; left -= full * W;
    }
    if (left != 0) e = flat(c, left);
    return e;
}

int set_device(strh_ctx* c) {
    HIP_TRY(hipSetDevice(c->device), "hipSetDevice");
    return STRH_OK;
}

// Sim.swift:258-266 — C1 cosine ramp in double, narrowed to float.
float ramp(const strh_lattice* L, int64_t t) {
    if (L->rampSteps <= 0) return 1.0f;
    const double x = std::min(1.0, (double)t / (double)L->rampSteps);
    return (float)(0.5 * (1.0 - cos(M_PI * x)));
}

// Sim.swift:268-281 plus the launch geometry.
StepParams make_params(const strh_lattice* L, int64_t t, bool writeForce) {
    StepParams p;
    memset(&p, 0, sizeof(p));
    const float r = ramp(L, t);
    p.nx = L->nx; p.ny = L->ny; p.nz = L->nz;
    p.parity = (int)(t & 1);
    p.N = L->N; p.S = L->S;
    p.lay = L->lay;
    p.omega = L->omega; p.omegaMinus = L->omegaMinus;
    p.uin = L->uinTarget * r;
    p.cSmago = L->cSmago;
    p.fx = L->bodyForce[0]; p.fy = L->bodyForce[1]; p.fz = L->bodyForce[2];
    p.writeForce = writeForce ? 1 : 0;
    p.ulidX = L->lid[0] * r; p.ulidY = L->lid[1] * r; p.ulidZ = L->lid[2] * r;
    p.inflowUniform = L->inflowUniform ? 1 : 0;
    if (L->spongeOn) {
        p.spongeX0 = L->spongeX0;
        p.spongeInvW = 1.0f / L->spongeWidth;
        p.spongeTau = L->spongeTau;
    } else {
        p.spongeX0 = 1e9f; p.spongeInvW = 0.0f; p.spongeTau = 1.0f;
    }
    p.useEps = L->useEps ? 1 : 0;
    for (int k = 0; k < 9; k++) {
        const int i = 2 * k + 1;
        const float cF = (float)kCx[i] * p.fx + (float)kCy[i] * p.fy + (float)kCz[i] * p.fz;
        p.cF[k] = cF;
        p.threeCF[k] = 3.0f * cF;
    }
    p.divx = make_fastdiv((uint32_t)L->nx);
    p.divy = make_fastdiv((uint32_t)L->ny);
    p.blocks = (L->N + 255u) / 256u;
    return p;
}

StepLaunch make_launch(const strh_lattice* L, int64_t t, bool writeForce) {
    StepLaunch s;
    s.precision = L->precision;
    s.f = L->f;
    s.flags = L->flags;
    s.link = L->link;
    s.lidm = L->lidm;
    s.force = L->force;
    s.eps = L->eps;
    s.walls = L->walls;
    s.nt = L->useEps;
    s.les = L->cSmago > 0.0f;
    s.has_force = !(L->bodyForce[0] == 0.0f && L->bodyForce[1] == 0.0f && L->bodyForce[2] == 0.0f);
    s.rows_bs = L->walls ? 0u : rows_block_size((uint32_t)L->nx, L->N);
    s.large = large_image(L);
    s.planar = L->nz == 1;
    s.ntSplit = L->useEps ? L->ntSplit : nullptr;
    s.nPlain = L->nPlainBlocks;
    s.nNt = L->nNtBlocks;
    s.alpha = L->alpha;
    s.p = make_params(L, t, writeForce);
    return s;
}

// Sim.swift:235-255 over z-slabs on host threads: bit (i-1) of the link word
// is set when neighbour n + c_i (periodic) is not fluid; the moving-wall
// family (LID / INFLOW / OUTFLOW) also sets it in the lid mask.
void build_masks(const uint8_t* fl, int nx, int ny, int nz, std::vector<uint32_t>& link,
                 std::vector<uint32_t>& lidm) {
    const size_t N = (size_t)nx * ny * nz;
    link.assign(N, 0u);
    lidm.assign(N, 0u);
    auto work = [&](int z0, int z1) {
        for (int z = z0; z < z1; z++) for (int y = 0; y < ny; y++) for (int x = 0; x < nx; x++) {
            const size_t idx = ((size_t)z * ny + y) * nx + x;
            if (fl[idx] != kFluid) { link[idx] = kNotFluidBit; continue; }
            uint32_t s = 0, l = 0;
            for (int i = 1; i < 19; i++) {
                const int xn = (x + kCx[i] + nx) % nx;
                const int yn = (y + kCy[i] + ny) % ny;
                const int zn = (z + kCz[i] + nz) % nz;
                const uint8_t c = fl[((size_t)zn * ny + yn) * nx + xn];
                if (c == kSolid) {
                    s |= 1u << (i - 1);
                } else if (c == kLid || c == kInflow || c == kOutflow) {
                    s |= 1u << (i - 1);
                    l |= 1u << (i - 1);
                }
            }
            link[idx] = s;
            lidm[idx] = l;
        }
    };
    const int nt = std::max(1, std::min<int>(nz, (int)std::min(16u, std::thread::hardware_concurrency())));
    if (nt == 1 || N < (1u << 18)) { work(0, nz); return; }
    std::vector<std::thread> ts;
    for (int k = 0; k < nt; k++) ts.emplace_back(work, nz * k / nt, nz * (k + 1) / nt);
    for (auto& t : ts) t.join();
}

void free_lattice(strh_lattice* L) {
    if (!L) return;
    if (L->ctx) hipSetDevice(L->ctx->device);
    void* ptrs[] = {L->f, L->flags, L->link, L->lidm, L->force, L->eps, L->alpha, L->moments,
                    L->partials, L->result, L->snap, L->ntSplit};
    for (void* p : ptrs) if (p) hipFree(p);
    delete L;
}

int sync(strh_lattice* L, const char* what) {
    HIP_TRY(hipStreamSynchronize(L->ctx->stream), what);
    return STRH_OK;
}

// Launch `steps` steps starting at stepsDone; force accumulation on absolute
// step `forceStep` (or none when < 0). Sim.swift:288-331.
int run_steps(strh_lattice* L, int64_t steps, int64_t forceStep) {
    if (steps < 0) return fail(STRH_ERR_INVALID, "negative step count");
    if (steps == 0) return STRH_OK;
    if (L->alpha && (L->useEps || L->cSmago > 0.0f))
        return fail(STRH_ERR_INVALID, "porosity cannot be combined with solid fractions or LES");
    strh_ctx* c = L->ctx;
    HIP_TRY(hipEventRecord(c->ev0, c->stream), "hipEventRecord");
    for (int64_t s = 0; s < steps; s++) {
        const int64_t t = L->stepsDone + s;
        const StepLaunch sl = make_launch(L, t, t == forceStep);
        HIP_TRY(launch_step(sl, c->stream), "step kernel launch");
    }
    HIP_TRY(hipEventRecord(c->ev1, c->stream), "hipEventRecord");
    HIP_TRY(hipEventSynchronize(c->ev1), "step kernels");
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1), "hipEventElapsedTime");
    L->gpuSeconds += (double)ms * 1e-3;
    L->stepsDone += steps;
    return STRH_OK;
}

int ensure_moments(strh_lattice* L) {
    if (L->moments) return STRH_OK;
    HIP_TRY(hipMalloc(&L->moments, (size_t)L->N * 16), "hipMalloc(moments)");
    return STRH_OK;
}

int compute_moments(strh_lattice* L) {
    int rc = ensure_moments(L);
    if (rc) return rc;
    HIP_TRY(launch_moments(L->precision, L->f, L->flags, L->moments, L->N, L->lay,
                           L->bodyForce[0], L->bodyForce[1], L->bodyForce[2], L->ctx->stream),
            "moments kernel launch");
    return STRH_OK;
}

int check_lat(strh_lattice* L) {
    if (!L || !L->ctx) return fail(STRH_ERR_INVALID, "null lattice handle");
    return set_device(L->ctx);
}

}  // namespace

extern "C" {

const char* strh_last_error(void) { return g_error.c_str(); }

int strh_ctx_create(int device_ordinal, strh_ctx** out) {
    if (!out) return fail(STRH_ERR_INVALID, "null output pointer");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(STRH_ERR_NO_DEVICE, "no HIP device");
    if (device_ordinal < 0 || device_ordinal >= count)
        return fail(STRH_ERR_NO_DEVICE, "device ordinal %d out of range (%d devices)", device_ordinal, count);
    HIP_TRY(hipSetDevice(device_ordinal), "hipSetDevice");
    strh_ctx* c = new strh_ctx();
    c->device = device_ordinal;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_ordinal) == hipSuccess) c->name = prop.gcnArchName;
    // Each device holds its own copy of the static scratch at its own address:
    // resolved here, with this device current, and kept with the context.
    const hipError_t se = scratch_address(&c->scratch);
    if (se != hipSuccess) {
        c->scratch = nullptr;
        c->scratchError = hipGetErrorString(se);
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->ev2) != hipSuccess || hipEventCreate(&c->ev3) != hipSuccess) {
        delete c;
        return fail(STRH_ERR_DEVICE, "stream/event creation failed");
    }
    *out = c;
    return STRH_OK;
}

int strh_ctx_destroy(strh_ctx* c) {
    if (!c) return STRH_OK;
    hipSetDevice(c->device);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->ev2) hipEventDestroy(c->ev2);
    if (c->ev3) hipEventDestroy(c->ev3);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return STRH_OK;
}

int strh_device_name(strh_ctx* c, char* out, size_t cap) {
    if (!c || !out || cap == 0) return fail(STRH_ERR_INVALID, "bad arguments");
    snprintf(out, cap, "%s", c->name.c_str());
    return STRH_OK;
}

int strh_lattice_create(strh_ctx* c, const strh_lattice_desc* d, const uint8_t* flags,
                        strh_lattice** out) {
    if (!c || !d || !flags || !out) return fail(STRH_ERR_INVALID, "null argument");
    *out = nullptr;
    if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0) return fail(STRH_ERR_INVALID, "lattice dimensions must be positive");
    const uint64_t N = (uint64_t)d->nx * d->ny * d->nz;
    if (N > kMaxCells) return fail(STRH_ERR_INVALID, "lattice of %llu cells exceeds the %llu-cell limit",
                                   (unsigned long long)N, (unsigned long long)kMaxCells);
    if (d->precision != STRH_FP32 && d->precision != STRH_FP16S && d->precision != STRH_BF16)
        return fail(STRH_ERR_INVALID, "unknown precision %d", d->precision);
    for (uint64_t n = 0; n < N; n++)
        if (flags[n] > STRH_OUTFLOW) return fail(STRH_ERR_INVALID, "invalid cell flag %d at cell %llu", flags[n], (unsigned long long)n);
    int rc = set_device(c);
    if (rc) return rc;

    strh_lattice* L = new strh_lattice();
    L->ctx = c;
    L->precision = d->precision;
    L->nx = d->nx; L->ny = d->ny; L->nz = d->nz;
    L->N = (uint32_t)N;
    L->S = (uint32_t)((N + 63u) & ~63ull);
    L->B = d->precision == STRH_FP32 ? 4 : 2;
    bool anyWall = false;
    for (uint64_t n = 0; n < N && !anyWall; n++) anyWall = flags[n] != STRH_FLUID;
    // blocked beyond 4 GiB only where every pass addresses from block / row bases
    L->lay = make_ddf_layout(N, (uint32_t)L->B, !anyWall && rows_block_size((uint32_t)d->nx, N) != 0);
    L->omega = d->omega;
    L->omegaMinus = d->omega_minus < 0.0f ? d->omega : d->omega_minus;
    for (int k = 0; k < 3; k++) { L->lid[k] = d->lid[k]; L->bodyForce[k] = d->force[k]; }
    L->uinTarget = d->uin;
    L->rampSteps = d->ramp_steps;
    L->cSmago = d->c_smago;
    L->wantsForces = d->wants_forces != 0;

    auto bail = [&](hipError_t e, const char* what) {
        free_lattice(L);
        return e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation
                   ? fail(STRH_ERR_ALLOC, "%s: %s", what, hipGetErrorString(e))
                   : hip_fail(e, what);
    };
    hipError_t e;
    const size_t fbytes = ddf_bytes(L);
    if ((e = hipMalloc(&L->f, fbytes)) != hipSuccess) return bail(e, "hipMalloc(f)");
    // Shifted rest state is exactly 0 in every precision (Sim.swift:225).
    if ((e = hipMemsetAsync(L->f, 0, fbytes, c->stream)) != hipSuccess) return bail(e, "hipMemset(f)");
    if ((e = hipMalloc(&L->flags, N)) != hipSuccess) return bail(e, "hipMalloc(flags)");
    if ((e = hipMemcpy(L->flags, flags, N, hipMemcpyHostToDevice)) != hipSuccess) return bail(e, "copy flags");
    L->walls = anyWall;
    for (uint64_t n = 0; n < N && !L->hasInflow; n++) L->hasInflow = flags[n] == STRH_INFLOW;
    if (anyWall) {
        std::vector<uint32_t> link, lidm;
        build_masks(flags, d->nx, d->ny, d->nz, link, lidm);
        if ((e = hipMalloc(&L->link, N * 4)) != hipSuccess) return bail(e, "hipMalloc(link)");
        if ((e = hipMalloc(&L->lidm, N * 4)) != hipSuccess) return bail(e, "hipMalloc(lidm)");
        if ((e = hipMemcpy(L->link, link.data(), N * 4, hipMemcpyHostToDevice)) != hipSuccess) return bail(e, "copy link");
        if ((e = hipMemcpy(L->lidm, lidm.data(), N * 4, hipMemcpyHostToDevice)) != hipSuccess) return bail(e, "copy lidm");
    }
    if (L->wantsForces) {
        if ((e = hipMalloc(&L->force, N * 16)) != hipSuccess) return bail(e, "hipMalloc(force)");
        if ((e = hipMemsetAsync(L->force, 0, N * 16, c->stream)) != hipSuccess) return bail(e, "hipMemset(force)");
    }
    if ((e = hipMalloc(&L->partials, kMaxPartials * sizeof(double))) != hipSuccess) return bail(e, "hipMalloc(partials)");
    if ((e = hipMalloc(&L->result, 4 * sizeof(double))) != hipSuccess) return bail(e, "hipMalloc(result)");
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail(e, "lattice init");
    *out = L;
    return STRH_OK;
}

int strh_lattice_destroy(strh_lattice* L) {
    free_lattice(L);
    return STRH_OK;
}

int strh_lattice_cells(strh_lattice* L, int64_t* cells) {
    if (!L || !cells) return fail(STRH_ERR_INVALID, "null argument");
    *cells = L->N;
    return STRH_OK;
}

int strh_ddf_bytes(strh_lattice* L, size_t* bytes) {
    if (!L || !bytes) return fail(STRH_ERR_INVALID, "null argument");
    *bytes = 19 * (size_t)L->N * L->B;
    return STRH_OK;
}

int strh_set_omega(strh_lattice* L, float omega, float omega_minus) {
    if (!L) return fail(STRH_ERR_INVALID, "null lattice handle");
    L->omega = omega;
    L->omegaMinus = omega_minus < 0.0f ? omega : omega_minus;
    return STRH_OK;
}

int strh_set_lid_velocity(strh_lattice* L, float ux, float uy, float uz) {
    if (!L) return fail(STRH_ERR_INVALID, "null lattice handle");
    L->lid[0] = ux; L->lid[1] = uy; L->lid[2] = uz;
    return STRH_OK;
}

int strh_set_uin(strh_lattice* L, float uin) {
    if (!L) return fail(STRH_ERR_INVALID, "null lattice handle");
    L->uinTarget = uin;
    return STRH_OK;
}

int strh_set_ramp_steps(strh_lattice* L, int64_t r) {
    if (!L) return fail(STRH_ERR_INVALID, "null lattice handle");
    L->rampSteps = r;
    return STRH_OK;
}

int strh_set_force(strh_lattice* L, float fx, float fy, float fz) {
    if (!L) return fail(STRH_ERR_INVALID, "null lattice handle");
    L->bodyForce[0] = fx; L->bodyForce[1] = fy; L->bodyForce[2] = fz;
    return STRH_OK;
}

int strh_set_c_smago(strh_lattice* L, float c) {
    if (!L) return fail(STRH_ERR_INVALID, "null lattice handle");
    L->cSmago = c;
    return STRH_OK;
}

int strh_set_inflow_uniform(strh_lattice* L, int enabled) {
    if (!L) return fail(STRH_ERR_INVALID, "null lattice handle");
    L->inflowUniform = enabled != 0;
    return STRH_OK;
}

int strh_set_sponge(strh_lattice* L, int enabled, float x0, float width, float tau) {
    if (!L) return fail(STRH_ERR_INVALID, "null lattice handle");
    if (enabled && !(width > 0.0f)) return fail(STRH_ERR_INVALID, "sponge width must be positive");
    L->spongeOn = enabled != 0;
    L->spongeX0 = x0; L->spongeWidth = width; L->spongeTau = tau;
    return STRH_OK;
}

int strh_set_solid_fractions(strh_lattice* L, const float* eps) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!eps) return fail(STRH_ERR_INVALID, "null eps");
    if (!L->eps) HIP_TRY(hipMalloc(&L->eps, (size_t)L->N * 4), "hipMalloc(eps)");
    HIP_TRY(hipMemcpy(L->eps, eps, (size_t)L->N * 4, hipMemcpyHostToDevice), "copy eps");
    L->useEps = true;
    // Lattices of at least 2^22 cells (3-D) whose partial cells lie in at most
    // half of the kStepBS-cell blocks run those blocks with the partial-cell
    // kernel and the rest without it (two launches per pass: on smaller
    // lattices the second launch would cost more than it saves).
    if (L->ntSplit) HIP_TRY(hipFree(L->ntSplit), "hipFree(split)");
    L->ntSplit = nullptr;
    L->nPlainBlocks = L->nNtBlocks = 0;
    if (L->nz > 1 && L->N >= (1u << 22)) {
        const uint32_t nb = (L->N + kStepBS - 1) / kStepBS;
        std::vector<uint32_t> plain, nt;
        for (uint32_t b = 0; b < nb; b++) {
            bool any = false;
            for (uint32_t n = b * kStepBS; n < std::min<uint32_t>(L->N, (b + 1) * kStepBS) && !any; n++)
                any = eps[n] > 0.0f;
            (any ? nt : plain).push_back(b);
        }
        if (nt.size() * 2 <= nb) {
            plain.insert(plain.end(), nt.begin(), nt.end());
            HIP_TRY(hipMalloc(&L->ntSplit, plain.size() * 4), "hipMalloc(split)");
            HIP_TRY(hipMemcpy(L->ntSplit, plain.data(), plain.size() * 4, hipMemcpyHostToDevice), "copy split");
            L->nPlainBlocks = nb - (uint32_t)nt.size();
            L->nNtBlocks = (uint32_t)nt.size();
        }
    }
    return STRH_OK;
}

int strh_set_porosity(strh_lattice* L, const float* alpha) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!alpha) {   // back to a plain lattice
        if (L->alpha) HIP_TRY(hipFree(L->alpha), "hipFree(alpha)");
        L->alpha = nullptr;
        return STRH_OK;
    }
    if (large_image(L))
        return fail(STRH_ERR_UNSUPPORTED, "porosity needs the direct odd pass, which addresses images "
                                          "below 4 GiB; this lattice's image is %zu bytes", ddf_bytes(L));
    for (int64_t n = 0; n < L->N; n++)
        if (!(alpha[n] >= 0.0f) || !std::isfinite(alpha[n]))
            return fail(STRH_ERR_INVALID, "porosity at cell %lld is not a finite value >= 0",
                        (long long)n);
    if (!L->alpha) HIP_TRY(hipMalloc(&L->alpha, (size_t)L->N * 4), "hipMalloc(alpha)");
    HIP_TRY(hipMemcpy(L->alpha, alpha, (size_t)L->N * 4, hipMemcpyHostToDevice), "copy alpha");
    return STRH_OK;
}

int strh_init_field(strh_lattice* L, uint32_t mode, float amplitude) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (mode > 2) return fail(STRH_ERR_INVALID, "unknown init mode %u", mode);
    const float tau = 1.0f / L->omega;
    HIP_TRY(launch_init(L->precision, L->f, (uint32_t)L->nx, (uint32_t)L->ny, (uint32_t)L->nz,
                        L->lay, mode, amplitude, tau, L->ctx->stream), "init kernel launch");
    return sync(L, "init kernel");
}

int strh_run(strh_lattice* L, int64_t steps) {
    int rc = check_lat(L);
    if (rc) return rc;
    return run_steps(L, steps, -1);
}

int strh_run_timed(strh_lattice* L, int64_t steps, float* ms) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!ms || steps < 0 || steps > (1 << 20)) return fail(STRH_ERR_INVALID, "bad arguments (steps in [0, 2^20])");
    if (L->alpha && (L->useEps || L->cSmago > 0.0f))
        return fail(STRH_ERR_INVALID, "porosity cannot be combined with solid fractions or LES");
    if (steps == 0) return STRH_OK;
    strh_ctx* c = L->ctx;
    std::vector<hipEvent_t> ev((size_t)steps + 1, nullptr);
    auto release = [&]() { for (hipEvent_t e : ev) if (e) hipEventDestroy(e); };
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) { e = nullptr; release(); return fail(STRH_ERR_DEVICE, "hipEventCreate"); }
    hipError_t e = hipEventRecord(ev[0], c->stream);
    for (int64_t s = 0; s < steps && e == hipSuccess; s++) {
        const int64_t t = L->stepsDone + s;
        e = launch_step(make_launch(L, t, false), c->stream);
        if (e == hipSuccess) e = hipEventRecord(ev[(size_t)s + 1], c->stream);
    }
    if (e == hipSuccess) e = hipEventSynchronize(ev[(size_t)steps]);
    double total = 0.0;
    for (int64_t s = 0; s < steps && e == hipSuccess; s++) {
        e = hipEventElapsedTime(&ms[s], ev[(size_t)s], ev[(size_t)s + 1]);
        total += ms[s];
    }
    release();
    if (e != hipSuccess) return hip_fail(e, "timed steps");
    L->gpuSeconds += total * 1e-3;
    L->stepsDone += steps;
    return STRH_OK;
}

int strh_steps_done(strh_lattice* L, int64_t* steps) {
    if (!L || !steps) return fail(STRH_ERR_INVALID, "null argument");
    *steps = L->stepsDone;
    return STRH_OK;
}

int strh_gpu_seconds(strh_lattice* L, double* seconds) {
    if (!L || !seconds) return fail(STRH_ERR_INVALID, "null argument");
    *seconds = L->gpuSeconds;
    return STRH_OK;
}

int strh_probe_moments(strh_lattice* L, float* out) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!out) return fail(STRH_ERR_INVALID, "null output");
    if (L->stepsDone % 2 != 0) return fail(STRH_ERR_ODD_STEPS, "moments probe requires even step count");
    if ((rc = compute_moments(L))) return rc;
    HIP_TRY(hipMemcpyAsync(out, L->moments, (size_t)L->N * 16, hipMemcpyDeviceToHost, L->ctx->stream), "copy moments");
    return sync(L, "moments probe");
}

int strh_probe_force(strh_lattice* L, int x0, int x1, int y0, int y1, double out[3]) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!out) return fail(STRH_ERR_INVALID, "null output");
    if (!L->wantsForces) return fail(STRH_ERR_NO_FORCES, "lattice created without wants_forces");
    if (L->stepsDone % 2 != 0) return fail(STRH_ERR_ODD_STEPS, "force probe requires even step count");
    if (x0 < 0 || x1 >= L->nx || x0 > x1 || y0 < 0 || y1 >= L->ny || y0 > y1)
        return fail(STRH_ERR_INVALID, "probe box x[%d,%d] y[%d,%d] outside %dx%d", x0, x1, y0, y1, L->nx, L->ny);
    HIP_TRY(launch_fill_zero(L->force, (size_t)L->N * 16, L->ctx->stream), "zero force");
    if ((rc = run_steps(L, 2, L->stepsDone + 1))) return rc;
    HIP_TRY(launch_box_force_sum(L->force, (uint32_t)L->nx, (uint32_t)L->ny, (uint32_t)L->nz,
                                 (uint32_t)x0, (uint32_t)x1, (uint32_t)y0, (uint32_t)y1,
                                 L->partials, kMaxPartials, L->result, L->ctx->stream),
            "force reduction launch");
    HIP_TRY(hipMemcpyAsync(out, L->result, 3 * sizeof(double), hipMemcpyDeviceToHost, L->ctx->stream), "copy force");
    return sync(L, "force probe");
}

int strh_mass_sum(strh_lattice* L, double* out) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!out) return fail(STRH_ERR_INVALID, "null output");
    HIP_TRY(launch_mass_sum(L->precision, L->f, L->N, L->lay, L->partials, kMaxPartials, L->result,
                            L->ctx->stream), "mass reduction launch");
    HIP_TRY(hipMemcpyAsync(out, L->result, sizeof(double), hipMemcpyDeviceToHost, L->ctx->stream), "copy mass");
    return sync(L, "mass sum");
}

int strh_kinetic_energy(strh_lattice* L, double* out) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!out) return fail(STRH_ERR_INVALID, "null output");
    if (L->stepsDone % 2 != 0) return fail(STRH_ERR_ODD_STEPS, "kinetic energy requires even step count");
    if ((rc = compute_moments(L))) return rc;
    HIP_TRY(launch_kinetic_sum(L->moments, L->N, L->partials, kMaxPartials, L->result, L->ctx->stream),
            "kinetic reduction launch");
    HIP_TRY(hipMemcpyAsync(out, L->result, sizeof(double), hipMemcpyDeviceToHost, L->ctx->stream), "copy energy");
    return sync(L, "kinetic energy");
}

int strh_kinetic_energy_noalloc(strh_lattice* L, double* out) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!out) return fail(STRH_ERR_INVALID, "null output");
    if (L->stepsDone % 2 != 0) return fail(STRH_ERR_ODD_STEPS, "kinetic energy requires even step count");
    double* scr = L->ctx->scratch;
    if (!scr) return fail(STRH_ERR_DEVICE, "static scratch address: %s", L->ctx->scratchError.c_str());
    uint32_t Pb = 0, chunk = 0;
    box_reduce_plan(L->N, kReducePartials, &Pb, &chunk);
    if (Pb > kReducePartials)
        return fail(STRH_ERR_INVALID, "lattice of %u cells needs %u partial sums; the static scratch holds %u",
                    L->N, Pb, kReducePartials);
    HIP_TRY(launch_kinetic_sum_cells(L->precision, L->f, L->flags, L->N, L->lay, L->bodyForce[0],
                                     L->bodyForce[1], L->bodyForce[2], scr, kReducePartials,
                                     scr + kReducePartials, L->ctx->stream),
            "kinetic reduction launch");
    HIP_TRY(hipMemcpyAsync(out, scr + kReducePartials, sizeof(double), hipMemcpyDeviceToHost, L->ctx->stream),
            "copy energy");
    return sync(L, "kinetic energy");
}

int strh_dissipation(strh_lattice* L, double* out) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!out) return fail(STRH_ERR_INVALID, "null output");
    if (L->stepsDone % 2 != 0) return fail(STRH_ERR_ODD_STEPS, "dissipation requires even step count");
    const double tauP = 1.0 / (double)L->omega;
    HIP_TRY(launch_dissipation_sum(L->precision, L->f, L->walls ? L->flags : nullptr, L->N, L->lay, tauP,
                                   L->partials, kMaxPartials, L->result, L->ctx->stream),
            "dissipation reduction launch");
    HIP_TRY(hipMemcpyAsync(out, L->result, sizeof(double), hipMemcpyDeviceToHost, L->ctx->stream), "copy dissipation");
    return sync(L, "dissipation");
}

// Canonical image: 19 slots of N values, slot-major (Sim.swift:441-444 hashes
// exactly these bytes), gathered from / scattered to the block layout.
int strh_download_ddf(strh_lattice* L, void* out, size_t bytes) {
    int rc = check_lat(L);
    if (rc) return rc;
    const size_t row = (size_t)L->N * L->B;
    if (!out || bytes != 19 * row) return fail(STRH_ERR_INVALID, "download needs exactly %zu bytes", 19 * row);
    for (uint32_t i = 0; i < 19; i++)
        HIP_TRY(slot_copy(L, i, 0, L->N, (char*)out + i * row, false, L->ctx->stream), "download ddf");
    return sync(L, "download ddf");
}

int strh_upload_ddf(strh_lattice* L, const void* in, size_t bytes) {
    int rc = check_lat(L);
    if (rc) return rc;
    const size_t row = (size_t)L->N * L->B;
    if (!in || bytes != 19 * row) return fail(STRH_ERR_INVALID, "upload needs exactly %zu bytes", 19 * row);
    for (uint32_t i = 0; i < 19; i++)
        HIP_TRY(slot_copy(L, i, 0, L->N, (char*)in + i * row, true, L->ctx->stream), "upload ddf");
    return sync(L, "upload ddf");
}

int strh_download_force(strh_lattice* L, float* out) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!out) return fail(STRH_ERR_INVALID, "null output");
    if (!L->wantsForces) return fail(STRH_ERR_NO_FORCES, "lattice created without wants_forces");
    HIP_TRY(hipMemcpyAsync(out, L->force, (size_t)L->N * 16, hipMemcpyDeviceToHost, L->ctx->stream), "copy force field");
    return sync(L, "download force");
}

}  // extern "C"

namespace {

// Stream the canonical image (19 slots of N values, slot-major) through two
// pinned staging buffers of up to 32 MiB: the copy of the next chunk overlaps
// `consume(bytes, len)` on the current one. `consume` returns false to stop.
template <class F> int stream_canonical(strh_lattice* L, F&& consume) {
    const size_t row = (size_t)L->N * L->B;
    const size_t total = 19 * row;
    const size_t chunk = std::min<size_t>(total, (size_t)32 << 20);
    void* stage[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    auto cleanup = [&]() {
        for (int k = 0; k < 2; k++) {
            if (stage[k]) hipHostFree(stage[k]);
            if (done[k]) hipEventDestroy(done[k]);
        }
    };
    for (int k = 0; k < 2; k++) {
        if (hipHostMalloc(&stage[k], chunk, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&done[k], hipEventDisableTiming) != hipSuccess) {
            cleanup();
            return fail(STRH_ERR_ALLOC, "pinned staging allocation failed");
        }
    }
    // Canonical byte offset o maps to slot o / row, in-slot offset o % row.
    auto issue = [&](size_t off, size_t len, int k) -> hipError_t {
        size_t done_bytes = 0;
        while (done_bytes < len) {
            const size_t o = off + done_bytes;
            const size_t slot = o / row, in = o % row;
            const size_t take = std::min(len - done_bytes, row - in);
            hipError_t e = slot_copy(L, (uint32_t)slot, in / L->B, take / L->B,
                                     (char*)stage[k] + done_bytes, false, L->ctx->stream);
            if (e != hipSuccess) return e;
            done_bytes += take;
        }
        return hipEventRecord(done[k], L->ctx->stream);
    };
    size_t off = 0;
    int k = 0;
    bool ok = true;
    hipError_t e = issue(0, std::min(chunk, total), 0);
    while (e == hipSuccess && off < total && ok) {
        const size_t len = std::min(chunk, total - off);
        const size_t next = off + len;
        if (next < total) e = issue(next, std::min(chunk, total - next), k ^ 1);
        if (e != hipSuccess) break;
        e = hipEventSynchronize(done[k]);
        if (e != hipSuccess) break;
        ok = consume((const uint8_t*)stage[k], len);
        off = next;
        k ^= 1;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(L->ctx->stream);
    cleanup();
    if (e != hipSuccess) return hip_fail(e, "canonical image copy");
    return ok ? STRH_OK : STRH_ERR_INVALID;
}

// State file: a 128-byte little-endian header, then the canonical image.
struct StateHeader {
    char magic[8];          // "STRHDDF1"
    uint32_t version;       // 1
    uint32_t precision;     // strh_precision
    int32_t nx, ny, nz;
    uint32_t bytes_per_value;
    int64_t steps_done;
    uint64_t image_bytes;   // 19 * N * bytes_per_value
    uint8_t digest[32];     // SHA-256 of the image (the state digest)
    uint8_t reserved[48];
};
static_assert(sizeof(StateHeader) == 128, "state header is 128 bytes");

}  // namespace

extern "C" {

// SHA-256 over the canonical image (Sim.swift:441-444).
int strh_state_digest(strh_lattice* L, uint8_t out[32]) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!out) return fail(STRH_ERR_INVALID, "null output");
    Sha256 sha;
    if ((rc = stream_canonical(L, [&](const uint8_t* b, size_t n) { sha.update(b, n); return true; }))) return rc;
    sha.finish(out);
    return STRH_OK;
}

int strh_save_state(strh_lattice* L, const char* path) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!path) return fail(STRH_ERR_INVALID, "null path");
    FILE* fp = fopen(path, "wb");
    if (!fp) return fail(STRH_ERR_INVALID, "cannot open %s for writing", path);
    StateHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, "STRHDDF1", 8);
    h.version = 1;
    h.precision = (uint32_t)L->precision;
    h.nx = L->nx; h.ny = L->ny; h.nz = L->nz;
    h.bytes_per_value = (uint32_t)L->B;
    h.steps_done = L->stepsDone;
    h.image_bytes = 19ull * L->N * L->B;
    bool ok = fwrite(&h, sizeof(h), 1, fp) == 1;
    Sha256 sha;
    if (ok) rc = stream_canonical(L, [&](const uint8_t* b, size_t n) {
        sha.update(b, n);
        return fwrite(b, 1, n, fp) == n;
    });
    if (ok && rc == STRH_OK) {
        sha.finish(h.digest);
        ok = fseek(fp, 0, SEEK_SET) == 0 && fwrite(&h, sizeof(h), 1, fp) == 1;
    }
    ok = (fclose(fp) == 0) && ok;
    if (rc) return rc == STRH_ERR_INVALID ? fail(STRH_ERR_INVALID, "write to %s failed", path) : rc;
    if (!ok) return fail(STRH_ERR_INVALID, "write to %s failed", path);
    return STRH_OK;
}

int strh_load_state(strh_lattice* L, const char* path) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!path) return fail(STRH_ERR_INVALID, "null path");
    FILE* fp = fopen(path, "rb");
    if (!fp) return fail(STRH_ERR_INVALID, "cannot open %s", path);
    StateHeader h;
    const size_t row = (size_t)L->N * L->B;
    auto bad = [&](const char* why) { fclose(fp); return fail(STRH_ERR_INVALID, "%s: %s", path, why); };
    if (fread(&h, sizeof(h), 1, fp) != 1) return bad("truncated header");
    if (memcmp(h.magic, "STRHDDF1", 8) != 0 || h.version != 1) return bad("not a state file (magic / version)");
    if (h.precision != (uint32_t)L->precision || h.bytes_per_value != L->B)
        return bad("precision differs from the lattice's");
    if (h.nx != L->nx || h.ny != L->ny || h.nz != L->nz) return bad("dimensions differ from the lattice's");
    if (h.image_bytes != 19 * row) return bad("image size does not match");
    // pass 1: the image must hash to the header's digest before anything is uploaded
    std::vector<uint8_t> buf((size_t)std::min<uint64_t>(h.image_bytes, 32ull << 20));
    Sha256 sha;
    for (uint64_t left = h.image_bytes; left;) {
        const size_t n = (size_t)std::min<uint64_t>(left, buf.size());
        if (fread(buf.data(), 1, n, fp) != n) return bad("truncated image");
        sha.update(buf.data(), n);
        left -= n;
    }
    uint8_t dg[32];
    sha.finish(dg);
    if (memcmp(dg, h.digest, 32) != 0) return bad("image does not match its digest");
    // pass 2: upload slot by slot
    if (fseek(fp, (long)sizeof(h), SEEK_SET) != 0) return bad("seek failed");
    for (uint32_t i = 0; i < 19; i++) {
        for (size_t done = 0; done < row;) {
            const size_t n = std::min(row - done, buf.size());
            if (fread(buf.data(), 1, n, fp) != n) return bad("truncated image");
            hipError_t e = slot_copy(L, i, done / L->B, n / L->B, buf.data(), true, L->ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(L->ctx->stream);
            if (e != hipSuccess) { fclose(fp); return hip_fail(e, "upload state"); }
            done += n;
        }
    }
    fclose(fp);
    L->stepsDone = h.steps_done;
    return STRH_OK;
}

int strh_snapshot(strh_lattice* L) {
    int rc = check_lat(L);
    if (rc) return rc;
    const size_t fbytes = ddf_bytes(L);
    if (!L->snap) HIP_TRY(hipMalloc(&L->snap, fbytes), "hipMalloc(snapshot)");
    HIP_TRY(hipMemcpyAsync(L->snap, L->f, fbytes, hipMemcpyDeviceToDevice, L->ctx->stream), "snapshot");
    L->snapSteps = L->stepsDone;
    return sync(L, "snapshot");
}

int strh_restore(strh_lattice* L) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!L->snap || L->snapSteps < 0) return fail(STRH_ERR_INVALID, "no snapshot taken");
    HIP_TRY(hipMemcpyAsync(L->f, L->snap, ddf_bytes(L), hipMemcpyDeviceToDevice, L->ctx->stream), "restore");
    L->stepsDone = L->snapSteps;
    return sync(L, "restore");
}

}  // extern "C"

// ---- multi-lattice batch -----------------------------------------------------
struct strh_batch {
    strh_ctx* ctx = nullptr;
    int precision = 0;
    std::vector<strh_lattice*> lats;
    std::vector<BatchLat> host;          // per-lattice descriptors (host)
    BatchLat* dev = nullptr;             // per-segment compacted copies (device)
    int64_t* base = nullptr;             // device step offset read by the kernels
    std::vector<float*> ramp;            // per lattice, device
    std::vector<int64_t> rampLen;
    std::vector<double*> series;         // per lattice, device
    std::vector<int64_t> taken;          // samples taken so far
    std::vector<int64_t> lastSteps;      // stepsDone of each lattice after its last batch run
    std::vector<void*> fring;            // per probed lattice, ring of force buffers (device)
    std::vector<int> fringLen;           // buffers in that ring
    std::vector<int> fringSpan;          // steps whose samples the ring holds (kProbeSpan or kProbeSpanSmall)
    size_t devCap = 0;                   // descriptors `dev` can hold
    int bs = 256;
    int graphSteps = kProbeSpan;
    double gpuSeconds = 0.0;
    // resident launches (resident.hip)
    int resident = 3;                    // 0 never, 1 one launch per segment, 2 required, 3 measured
    hipEvent_t trialEv[3] = {nullptr, nullptr, nullptr};   // mode 3's timing of the two launch kinds
    int residentSlots = 0;               // workgroups the device holds at once at the largest tile
    std::vector<uint64_t*> xchg;         // per lattice, [19][N] exchange words, or null
    int* abortFlag = nullptr;            // raised by a resident launch whose exchange stalled
    uint64_t residentSpin = kResidentSpinTicks;   // wall-clock ticks a wave waits before aborting
    ResTile* tiles = nullptr;            // every resident group's tiles of one run (device)
    size_t tilesCap = 0;
    // the probe's chunk sums: 3 lattices x probeSamples x probePb doubles (device)
    double* probeParts = nullptr;
    size_t probePartsCap = 0;
    int probePb = 0;                     // most chunks of any probe box
    int probeSamples = 0;                // most force samples of any lattice in one span
};

namespace {

// Fill the per-run fields of every lattice descriptor (host side).
int batch_prepare(strh_batch* B, const int64_t* steps) {
    for (size_t i = 0; i < B->lats.size(); i++) {
        strh_lattice* L = B->lats[i];
        if (L->alpha) return fail(STRH_ERR_INVALID, "lattice %zu has porosity: batches run plain lattices", i);
        BatchLat& d = B->host[i];
        d.p = make_params(L, L->stepsDone, false);
        d.f = L->f; d.flags = L->flags; d.link = L->link; d.lidm = L->lidm;
        d.force = B->fring[i] ? (float4*)B->fring[i] : (float4*)L->force;
        d.fRing = B->fring[i] ? B->fringLen[i] : 1;
        d.eps = L->eps;
        d.step0 = L->stepsDone;
        d.tEnd = L->stepsDone + (steps ? steps[i] : 0);
        d.uinTarget = L->uinTarget;
        for (int k = 0; k < 3; k++) d.lid[k] = L->lid[k];
        d.hasForce = !(L->bodyForce[0] == 0.0f && L->bodyForce[1] == 0.0f && L->bodyForce[2] == 0.0f);
        d.inflowRows = L->hasInflow && !L->inflowUniform ? 1 : 0;
        // ramp factors of Sim.swift:258-266, evaluated by the same host code
        const int64_t rl = L->rampSteps > 0 ? L->rampSteps : 0;
        if (rl != B->rampLen[i]) {
            if (B->ramp[i]) hipFree(B->ramp[i]);
            B->ramp[i] = nullptr;
            if (rl > 0) HIP_TRY(hipMalloc(&B->ramp[i], rl * sizeof(float)), "hipMalloc(ramp)");
            B->rampLen[i] = rl;
        }
        if (rl > 0) {
            std::vector<float> r(rl);
            for (int64_t t = 0; t < rl; t++) r[t] = ramp(L, t);
            HIP_TRY(hipMemcpy(B->ramp[i], r.data(), rl * sizeof(float), hipMemcpyHostToDevice), "copy ramp");
        }
        d.ramp = B->ramp[i];
        d.rampSteps = rl;
        if (B->resident != 0 && resident_candidate(L) && !B->xchg[i])
            HIP_TRY(hipMalloc(&B->xchg[i], (size_t)19 * L->N * sizeof(uint64_t)), "hipMalloc(exchange)");
        d.xchg = B->xchg[i];
        d.tagBase = L->stepsDone;
    }
    return STRH_OK;
}

// One resident launch: the tiles [first, first + count) of the run's tile
// array, covering a group of the segment's lattices that fits in LDS at once.
struct ResGroup {
    size_t first;
    int count;
    int bs;                  // workgroup size: the largest tile in as few full rounds as possible
    size_t lds;              // dynamic LDS of the largest tile
    bool les, forced;        // forced: every lattice of the group has a body force, else none
};

struct Segment {
    int64_t k0, k1;          // lock-step batch steps [k0, k1)
    int L;                   // active lattices
    uint32_t blocks;         // workgroups of the active lattices
    BatchLat* dev;           // their descriptors (device)
    int parity0;             // parity of every active lattice at k0, or -1 if they differ
    bool les;                // some active lattice has a Smagorinsky constant
    bool forced;             // some active lattice has a body force
    bool planar;             // every active lattice is one cell deep
    bool probes;             // some active lattice has force samples
    std::vector<ResGroup> groups;   // resident launches, or empty for the streaming kernel
    bool trial = false;      // time-sliced groups: time both launch kinds, keep the faster (mode 3)
};

// ---- resident tiling ---------------------------------------------------------
// A lattice one cell deep is cut into gx x gy tiles of near-equal size (each
// at least 3 x 3 cells); a group of lattices shares the device's workgroup
// slots in proportion to their cells, and groups are time-sliced: each runs
// its steps in a launch of its own.
struct TilePlan {
    int gx = 0, gy = 0;
    int wmax = 0, hmax = 0;
};

// The tile and its halo ring (19 slots per cell), their flags (one byte per
// cell, rounded up to four), two buffers of one inflow value per row, then the
// link words of the edge ring and the halo ring.
size_t tile_lds(int w, int h, int bytes) {
    const size_t cs = (size_t)(w + 2) * (size_t)(h + 2);
    const size_t rings = (size_t)(2 * w + 2 * (h - 2)) + (size_t)(2 * (w + 2) + 2 * h);
    return cs * 19u * (size_t)bytes + ((cs + 3u) & ~(size_t)3u) + 2u * (size_t)(h + 2) * sizeof(float) +
           rings * sizeof(uint32_t);
}

// The split of nx x ny into at most g tiles with the fewest cells in its
// largest tile counted with its halo ring (then the fewest without), or gx = 0
// when none fits in LDS.
TilePlan plan_lattice(int nx, int ny, int g, int bytes) {
    TilePlan best;
    int64_t bestCells = -1, bestEdge = 0;
    for (int gx = 1; gx <= g && gx <= nx / 3; gx++) {
        const int gy = std::min(g / gx, ny / 3);
        if (gy < 1) continue;
        const int w = (nx + gx - 1) / gx, h = (ny + gy - 1) / gy;
        if (tile_lds(w, h, bytes) > kResidentMaxLds || w * h > kResidentMaxRounds * kResidentBS) continue;
        // tile and halo cells: the cells a step updates plus the edge the
        // exchange trades (long thin tiles trade more words per cell)
        const int64_t cells = (int64_t)(w + 2) * (h + 2), edge = (int64_t)w * h;
        if (bestCells < 0 || cells < bestCells || (cells == bestCells && edge < bestEdge)) {
            best.gx = gx; best.gy = gy; best.wmax = w; best.hmax = h;
            bestCells = cells; bestEdge = edge;
        }
    }
    return best;
}


// Tiles of the lattices `members` (indices into lats) sharing `slots` workgroups;
// false when some lattice gets no split that fits in LDS.
// The fewest tiles that split lattice L with at most r * kResidentBS cells in
// its largest tile (r = 1 .. kResidentMaxRounds rounds of the workgroup),
// index r; 0 where no split into at most `slots` tiles does.
using RoundTiles = std::array<int, kResidentMaxRounds + 1>;
RoundTiles round_tiles(const strh_lattice* L, int slots, int bytes) {
    RoundTiles t{};
    for (int r = 1; r <= kResidentMaxRounds; r++) {
        const int64_t cap = (int64_t)r * kResidentBS;
        for (int g = std::max<int64_t>(1, ((int64_t)L->N + cap - 1) / cap); g <= slots; g++) {
            const TilePlan tp = plan_lattice(L->nx, L->ny, g, bytes);
            if (tp.gx != 0 && (int64_t)tp.wmax * tp.hmax <= cap) { t[r] = g; break; }
        }
    }
    return t;
}

// rt: round_tiles of each lattice (indexed like lats), or null. With it the
// group takes the fewest rounds r whose tile counts fit the slots: each
// lattice starts from its fewest tiles of at most r rounds, and the slots left
// go one by one to the lattice furthest below its share in proportion to its
// cells (a lattice keeps the fewest tiles if the larger count splits it into
// tiles of more than r rounds); without it (and when no r fits) slots are
// shared in proportion alone.
bool plan_group(const std::vector<strh_lattice*>& lats, const std::vector<int>& members, int slots, int bytes,
                std::vector<TilePlan>* plans, const std::vector<RoundTiles>* rt = nullptr) {
    double cells = 0.0;
    for (int i : members) cells += (double)lats[i]->N;
    if (rt) {
        const int r0 = std::max(1, (int)std::ceil(cells / ((double)slots * kResidentBS)));
        for (int r = r0; r <= kResidentMaxRounds; r++) {
            int used = 0;
            bool ok = true;
            for (int i : members) {
                const int g = (*rt)[(size_t)i][(size_t)r];
                if (g == 0) { ok = false; break; }
                used += g;
            }
            if (!ok || used > slots) continue;
            if (plans) {
                std::vector<int> gs;
                for (int i : members) gs.push_back((*rt)[(size_t)i][(size_t)r]);
                for (; used < slots; used++) {
                    int pick = -1;
                    double deficit = 0.0;
                    for (size_t q = 0; q < members.size(); q++) {
                        const double d = (double)slots * (double)lats[members[q]]->N / cells - (double)gs[q];
                        if (d > deficit) { deficit = d; pick = (int)q; }
                    }
                    if (pick < 0) break;
                    gs[(size_t)pick]++;
                }
                plans->clear();
                for (size_t q = 0; q < members.size(); q++) {
                    const strh_lattice* L = lats[members[q]];
                    TilePlan tp = plan_lattice(L->nx, L->ny, gs[q], bytes);
                    if (tp.gx == 0 || (int64_t)tp.wmax * tp.hmax > (int64_t)r * kResidentBS)
                        tp = plan_lattice(L->nx, L->ny, (*rt)[(size_t)members[q]][(size_t)r], bytes);
                    plans->push_back(tp);
                }
            }
            return true;
        }
    }
    int used = 0;
    std::vector<TilePlan> out;
    for (int i : members) {
        const strh_lattice* L = lats[i];
        int g = std::max(1, (int)std::floor((double)slots * (double)L->N / cells));
        // at most one tile per kResidentBS cells: smaller tiles leave threads idle
        g = std::min(g, std::max(1, (int)((L->N + kResidentBS - 1) / kResidentBS)));
        TilePlan tp = plan_lattice(L->nx, L->ny, g, bytes);
        if (tp.gx == 0) return false;
        used += tp.gx * tp.gy;
        out.push_back(tp);
    }
    if (used > slots) return false;
    if (plans) *plans = out;
    return true;
}

// The batch's span: the shortest span of its lattices' force rings.
int batch_span(const strh_batch* B) {
    int span = kProbeSpan;
    for (size_t i = 0; i < B->lats.size(); i++)
        if (B->fring[i]) span = std::min(span, B->fringSpan[i]);
    return span;
}

// Launches steps k = kOff .. kOff+count-1 from base; parity0 is the common
// parity at k = kOff (or -1), so step k runs the pass (parity0 + k - kOff) & 1 only.
// The force samples of every span of steps (and of the last few) are
// reduced by one probe launch; each lattice's ring holds that many samples.
int batch_launch_steps(strh_batch* B, const Segment& sg, int bs, int count, int parity0, int kOff = 0) {
    hipStream_t s = B->ctx->stream;
    const int span = batch_span(B);
    if (!sg.groups.empty()) {
        // resident: each group advances a span of steps per launch, then the probes
        for (int j = 0; j < count; j += span) {
            const int K = std::min(span, count - j);
            for (const ResGroup& g : sg.groups)
                HIP_TRY(launch_resident(B->precision, g.les, g.forced, sg.dev, B->tiles + g.first, g.count, g.bs, g.lds,
                                        B->base, kOff + j, K, B->abortFlag, B->residentSpin, s),
                        "resident launch");
            if (sg.probes)
                HIP_TRY(launch_batch_probe(sg.dev, sg.L, B->base, kOff + j, K, B->probeParts, B->probePb,
                                           B->probeSamples, s),
                        "batch probe launch");
        }
        return STRH_OK;
    }
    int k0 = kOff;
    for (int j = 0; j < count; j++) {
        const int k = kOff + j;
        const int par = parity0 < 0 ? -1 : (parity0 + j) & 1;
        HIP_TRY(launch_batch_step(B->precision, bs, par, sg.les, sg.forced, sg.planar, sg.dev, sg.L, sg.blocks, B->base, k, s),
                "batch step launch");
        if (sg.probes && (k + 1 - k0 == span || j == count - 1)) {
            HIP_TRY(launch_batch_probe(sg.dev, sg.L, B->base, k0, k + 1 - k0, B->probeParts, B->probePb,
                                       B->probeSamples, s),
                    "batch probe launch");
            k0 = k + 1;
        }
    }
    return STRH_OK;
}

int batch_capture(strh_batch* B, const Segment& sg, hipGraphExec_t* out) {
    hipStream_t s = B->ctx->stream;
    hipGraph_t g = nullptr;
    // replays advance base by graphSteps: fixed parities need an even length
    const int par0 = (B->graphSteps % 2 == 0) ? sg.parity0 : -1;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
    int rc = batch_launch_steps(B, sg, B->bs, B->graphSteps, par0);
    hipError_t e2 = launch_advance_base(B->base, B->graphSteps, s);
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) { if (g) hipGraphDestroy(g); return rc; }
    if (e2 != hipSuccess) { if (g) hipGraphDestroy(g); return hip_fail(e2, "capture advance"); }
    if (e != hipSuccess) return hip_fail(e, "end capture");
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) { *out = nullptr; return hip_fail(e, "graph instantiate"); }
    return STRH_OK;
}

// Advance the batch by steps[i] per lattice. The run is split where lattices
// finish: each segment launches only the still-active lattices (compacted
// descriptors and grid), as whole captured graphs of graphSteps steps plus
// single launches for the remainder. Kernels read t = step0 + base + k, base
// being set at each segment start and advanced by every graph replay.
// Group the segment's lattices `members` (indices into B->lats; compact
// descriptor index = position in members) into resident launches and append
// their tiles to `tiles`. Returns false when some lattice does not qualify.
bool plan_segment(strh_batch* B, const std::vector<int>& members, Segment& sg, std::vector<ResTile>& tiles) {
    const int bytes = (int)B->lats[0]->B;
    for (int i : members)
        if (!resident_candidate(B->lats[i]) || !B->xchg[i]) return false;
    std::vector<int> order(members.size());
    for (size_t j = 0; j < order.size(); j++) order[j] = (int)j;
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        return B->lats[members[a]]->N > B->lats[members[b]]->N;
    });
    auto forced_of = [&](int j) { return B->host[members[j]].hasForce != 0; };
    std::vector<RoundTiles> rt(B->lats.size());
    for (int i : members) rt[(size_t)i] = round_tiles(B->lats[i], B->residentSlots, bytes);
    std::vector<std::vector<int>> groups;   // positions in members
    auto lat_ids = [&](const std::vector<int>& g) {
        std::vector<int> v;
        for (int j : g) v.push_back(members[j]);
        return v;
    };
    // First fit in decreasing size: the fewest groups.
    for (int j : order) {
        bool placed = false;
        for (auto& g : groups) {
            if (forced_of(g[0]) != forced_of(j)) continue;
            std::vector<int> cand = g;
            cand.push_back(j);
            if (plan_group(B->lats, lat_ids(cand), B->residentSlots, bytes, nullptr, &rt)) { g = cand; placed = true; break; }
        }
        if (placed) continue;
        if (!plan_group(B->lats, lat_ids({j}), B->residentSlots, bytes, nullptr, &rt)) return false;
        groups.push_back({j});
    }
    // A group's step costs about one round of cells per thread plus one more
    // for the exchange and barriers (measured 2.3-2.7 us per unit on the
    // ladder's rung sets), so first fit, which fills the first groups to
    // three or four rounds and leaves a small last group, is compared with
    // balanced plans: for each group count, the lattices in decreasing size
    // each join the least-filled group (of its forcing kind) that still
    // plans. The cheapest plan is kept (first fit on ties).
    auto cost_of = [&](const std::vector<std::vector<int>>& gs) -> int64_t {
        int64_t c = 0;
        for (const auto& g : gs) {
            std::vector<TilePlan> plans;
            if (g.empty() || !plan_group(B->lats, lat_ids(g), B->residentSlots, bytes, &plans, &rt)) return -1;
            int cmax = 0;
            for (const TilePlan& tp : plans) cmax = std::max(cmax, tp.wmax * tp.hmax);
            c += (cmax + kResidentBS - 1) / kResidentBS + 1;
        }
        return c;
    };
    int64_t best = cost_of(groups);
    for (size_t G = 1; G <= order.size(); G++) {
        std::vector<std::vector<int>> gs(G);
        std::vector<double> fill(G, 0.0);
        bool ok = true;
        for (int j : order) {
            int pick = -1;
            for (size_t q = 0; q < G; q++) {
                if (!gs[q].empty() && forced_of(gs[q][0]) != forced_of(j)) continue;
                if (pick >= 0 && fill[q] >= fill[(size_t)pick]) continue;
                std::vector<int> cand = gs[q];
                cand.push_back(j);
                if (plan_group(B->lats, lat_ids(cand), B->residentSlots, bytes, nullptr, &rt)) pick = (int)q;
            }
            if (pick < 0) { ok = false; break; }
            gs[(size_t)pick].push_back(j);
            fill[(size_t)pick] += (double)B->lats[members[j]]->N;
        }
        if (!ok) continue;
        const int64_t c = cost_of(gs);
        if (c >= 0 && c < best) { best = c; groups = gs; }
    }
    for (const auto& g : groups) {
        std::vector<TilePlan> plans;
        plan_group(B->lats, lat_ids(g), B->residentSlots, bytes, &plans, &rt);
        ResGroup rg;
        rg.first = tiles.size();
        rg.lds = 0;
        rg.les = false;
        rg.forced = forced_of(g[0]);
        for (size_t q = 0; q < g.size(); q++) {
            const strh_lattice* L = B->lats[members[g[q]]];
            rg.les = rg.les || L->cSmago > 0.0f;
            const TilePlan& tp = plans[q];
            for (int b = 0; b < tp.gy; b++) {
                for (int a = 0; a < tp.gx; a++) {
                    ResTile t;
                    t.lat = g[q];
                    t.x0 = (int)((int64_t)a * L->nx / tp.gx);
                    t.w = (int)((int64_t)(a + 1) * L->nx / tp.gx) - t.x0;
                    t.y0 = (int)((int64_t)b * L->ny / tp.gy);
                    t.h = (int)((int64_t)(b + 1) * L->ny / tp.gy) - t.y0;
                    t.divi = make_fastdiv((uint32_t)(t.w - 2));
                    rg.lds = std::max(rg.lds, tile_lds(t.w, t.h, bytes));
                    tiles.push_back(t);
                }
            }
        }
        rg.count = (int)(tiles.size() - rg.first);
        int cmax = 0;
        for (size_t q = rg.first; q < tiles.size(); q++) cmax = std::max(cmax, tiles[q].w * tiles[q].h);
        const int rounds = (cmax + kResidentBS - 1) / kResidentBS;
        rg.bs = ((cmax + rounds - 1) / rounds + 63) / 64 * 64;
        sg.groups.push_back(rg);
    }
    return true;
}

int batch_execute(strh_batch* B, const int64_t* steps, bool useGraph, bool probes, double* seconds) {
    hipStream_t s = B->ctx->stream;
    const int n = (int)B->lats.size();
    std::vector<int64_t> ends;
    for (int i = 0; i < n; i++) if (steps[i] > 0) ends.push_back(steps[i]);
    std::sort(ends.begin(), ends.end());
    ends.erase(std::unique(ends.begin(), ends.end()), ends.end());
    if (ends.empty()) { if (seconds) *seconds = 0.0; return STRH_OK; }
    if (B->resident != 0 && B->residentSlots == 0) {
        int perCU = 0, cus = 0;
        HIP_TRY(resident_prepare(B->precision, kResidentMaxLds, &perCU), "resident kernel attributes");
        HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, B->ctx->device), "CU count");
        B->residentSlots = std::max(0, cus * perCU);
        if (B->residentSlots == 0) return fail(STRH_ERR_UNSUPPORTED, "the resident kernel does not fit on a CU");
    }
    // scratch for the probes' chunk sums, sized for every lattice's box and period
    if (probes) {
        const int span = batch_span(B);
        int pb = 0, ns = 0;
        for (int i = 0; i < n; i++) {
            const BatchLat& d = B->host[(size_t)i];
            if (d.fCount <= 0) continue;
            pb = std::max(pb, (int)d.Pb);
            ns = std::max(ns, (int)((span + d.fPeriod - 1) / d.fPeriod));
        }
        const size_t need = (size_t)3 * (size_t)n * (size_t)std::max(pb, 1) * (size_t)std::max(ns, 1);
        if (need > B->probePartsCap) {
            if (B->probeParts) hipFree(B->probeParts);
            B->probeParts = nullptr;
            B->probePartsCap = 0;
            HIP_TRY(hipMalloc(&B->probeParts, need * sizeof(double)), "hipMalloc(probe chunk sums)");
            B->probePartsCap = need;
        }
        B->probePb = std::max(pb, 1);
        B->probeSamples = std::max(ns, 1);
    }
    // compact descriptor arrays of every segment, uploaded once
    std::vector<Segment> segs;
    std::vector<BatchLat> up;
    std::vector<ResTile> tiles;
    bool anyResident = false;
    int64_t k0 = 0;
    for (int64_t e : ends) {
        Segment sg;
        sg.k0 = k0; sg.k1 = e; sg.L = 0; sg.blocks = 0; sg.dev = nullptr;
        sg.parity0 = -2;
        sg.les = false;
        sg.forced = false;
        sg.planar = true;
        sg.probes = false;
        std::vector<int> members;
        for (int i = 0; i < n; i++) {
            if (steps[i] <= k0) continue;
            BatchLat d = B->host[i];
            const int par = (int)((d.step0 + k0) & 1);
            sg.parity0 = sg.parity0 == -2 ? par : (sg.parity0 == par ? par : -1);
            sg.les = sg.les || B->lats[i]->cSmago > 0.0f;
            sg.forced = sg.forced || d.hasForce != 0;
            sg.planar = sg.planar && B->lats[i]->nz == 1;
            d.block0 = sg.blocks;
            d.nblocks = (B->lats[i]->N + (uint32_t)B->bs - 1) / (uint32_t)B->bs;
            if (!probes) d.fCount = 0;
            sg.probes = sg.probes || d.fCount > 0;
            sg.blocks += d.nblocks;
            up.push_back(d);
            members.push_back(i);
            sg.L++;
        }
        if (B->resident != 0) {
            // modes 1 and 3 take the resident launch where the segment's
            // lattices fit on chip in one launch; where they need time-sliced
            // groups mode 1 streams and mode 3 measures both (the faster
            // depends on the storage and the lattices); mode 2 takes it
            // wherever they qualify
            const size_t mark = tiles.size();
            const bool ok = plan_segment(B, members, sg, tiles);
            sg.trial = ok && B->resident == 3 && sg.groups.size() > 1;
            if (!ok || (B->resident == 1 && sg.groups.size() > 1)) {
                sg.groups.clear();
                tiles.resize(mark);
                if (B->resident == 2)
                    return fail(STRH_ERR_UNSUPPORTED, "resident launches need lattices one cell deep (nz = 1, "
                                "nx, ny >= 3) that fit in LDS");
            }
            anyResident = anyResident || !sg.groups.empty();
        }
        segs.push_back(sg);
        k0 = e;
    }
    if (up.size() > B->devCap) {
        if (B->dev) hipFree(B->dev);
        B->dev = nullptr;
        B->devCap = 0;
        HIP_TRY(hipMalloc(&B->dev, up.size() * sizeof(BatchLat)), "hipMalloc(batch descriptors)");
        B->devCap = up.size();
    }
    if (tiles.size() > B->tilesCap) {
        if (B->tiles) hipFree(B->tiles);
        B->tiles = nullptr;
        B->tilesCap = 0;
        HIP_TRY(hipMalloc(&B->tiles, tiles.size() * sizeof(ResTile)), "hipMalloc(resident tiles)");
        B->tilesCap = tiles.size();
    }
    HIP_TRY(hipStreamSynchronize(s), "batch descriptors");
    HIP_TRY(hipMemcpy(B->dev, up.data(), up.size() * sizeof(BatchLat), hipMemcpyHostToDevice),
            "upload batch descriptors");
    if (!tiles.empty())
        HIP_TRY(hipMemcpy(B->tiles, tiles.data(), tiles.size() * sizeof(ResTile), hipMemcpyHostToDevice),
                "upload resident tiles");
    size_t off = 0;
    for (Segment& sg : segs) { sg.dev = B->dev + off; off += sg.L; }
    if (anyResident) {
        // exchange tags count from 1 in every run: clear the words and the abort flag
        for (int i = 0; i < n; i++)
            if (B->xchg[i] && steps[i] > 0)
                HIP_TRY(launch_fill_zero(B->xchg[i], (size_t)19 * B->lats[i]->N * sizeof(uint64_t), s), "clear exchange");
        HIP_TRY(launch_fill_zero(B->abortFlag, sizeof(int), s), "clear abort flag");
    }

    HIP_TRY(hipEventRecord(B->ctx->ev0, s), "hipEventRecord");
    for (const Segment& seg : segs) {
        HIP_TRY(launch_set_base(B->base, seg.k0, s), "set base");
        const int64_t K = seg.k1 - seg.k0;
        int64_t done = 0;
        Segment sg = seg;
        if (seg.trial) sg.groups.clear();   // streams unless the trial below picks the groups
        // Mode 3 with time-sliced groups: one untimed replay of each kind of
        // graph, two timed replays of each, then the faster for the rest of
        // the segment. Both kinds give the same bits, so the choice changes
        // only the time.
        const bool trial = seg.trial && useGraph && B->graphSteps > 0 && K >= 8 * (int64_t)B->graphSteps;
        hipGraphExec_t kept = nullptr;   // the trial's faster graph, replayed for the rest
        if (trial) {
            for (hipEvent_t& ev : B->trialEv)
                if (!ev) HIP_TRY(hipEventCreate(&ev), "hipEventCreate");
            hipGraphExec_t gk[2] = {nullptr, nullptr};   // streaming, resident
            int rc = batch_capture(B, sg, &gk[0]);
            if (!rc) rc = batch_capture(B, seg, &gk[1]);
            float ms[2] = {0.0f, 0.0f};
            hipError_t e = hipSuccess;
            for (int k = 0; k < 2 && !rc && e == hipSuccess; k++) { e = hipGraphLaunch(gk[k], s); done += B->graphSteps; }
            for (int k = 0; k < 2 && !rc && e == hipSuccess; k++) {
                e = hipEventRecord(B->trialEv[k], s);
                for (int r = 0; r < 2 && e == hipSuccess; r++) { e = hipGraphLaunch(gk[k], s); done += B->graphSteps; }
            }
            if (!rc && e == hipSuccess) e = hipEventRecord(B->trialEv[2], s);
            const hipError_t es = hipStreamSynchronize(s);   // before any exec is destroyed
            if (!rc && e == hipSuccess) e = es;
            if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms[0], B->trialEv[0], B->trialEv[1]);
            if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms[1], B->trialEv[1], B->trialEv[2]);
            const int pick = ms[1] < ms[0] ? 1 : 0;
            for (int k = 0; k < 2; k++)
                if (gk[k] && (rc || e != hipSuccess || k != pick)) hipGraphExecDestroy(gk[k]);
            if (rc) return rc;
            if (e != hipSuccess) return hip_fail(e, "resident trial");
            if (pick == 1) sg.groups = seg.groups;
            kept = gk[pick];
        }
        if (useGraph && B->graphSteps > 0 && K - done >= B->graphSteps) {
            hipGraphExec_t g = kept;
            kept = nullptr;
            if (!g) {
                int rc = batch_capture(B, sg, &g);
                if (rc) return rc;
            }
            while (K - done >= B->graphSteps) {
                hipError_t e = hipGraphLaunch(g, s);
                if (e != hipSuccess) {
                    (void)hipStreamSynchronize(s);   // earlier replays of g may still run
                    hipGraphExecDestroy(g);
                    return hip_fail(e, "graph launch");
                }
                done += B->graphSteps;
            }
            // the exec may be destroyed once its launches are queued
            hipError_t e = hipStreamSynchronize(s);
            hipGraphExecDestroy(g);
            if (e != hipSuccess) return hip_fail(e, "graph replays");
        }
        if (kept) hipGraphExecDestroy(kept);   // (the stream was synchronised in the trial)
        if (done < K) {
            const int par0 = sg.parity0 < 0 ? -1 : (int)((sg.parity0 + done) & 1);
            int rc = batch_launch_steps(B, sg, B->bs, (int)(K - done), par0);
            if (rc) return rc;
        }
    }
    HIP_TRY(hipEventRecord(B->ctx->ev1, s), "hipEventRecord");
    HIP_TRY(hipEventSynchronize(B->ctx->ev1), "batch steps");
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, B->ctx->ev0, B->ctx->ev1), "hipEventElapsedTime");
    if (seconds) *seconds = ms * 1e-3;
    if (anyResident) {
        int aborted = 0;
        HIP_TRY(hipMemcpy(&aborted, B->abortFlag, sizeof(int), hipMemcpyDeviceToHost), "read abort flag");
        if (aborted)
            return fail(STRH_ERR_DEVICE, "a resident launch stalled waiting for a neighbour tile's exchange; "
                        "the batch's lattices are left in an undefined state");
    }
    return STRH_OK;
}

}  // namespace

extern "C" {

int strh_batch_create(strh_ctx* c, strh_lattice* const* lats, int count, strh_batch** out) {
    if (!c || !lats || !out || count <= 0) return fail(STRH_ERR_INVALID, "bad batch arguments");
    *out = nullptr;
    for (int i = 0; i < count; i++) {
        if (!lats[i] || lats[i]->ctx != c) return fail(STRH_ERR_INVALID, "lattice %d is not on this context", i);
        if (lats[i]->precision != lats[0]->precision) return fail(STRH_ERR_INVALID, "lattices must share one precision");
        if (lats[i]->alpha) return fail(STRH_ERR_INVALID, "lattice %d has porosity: batches run plain lattices", i);
        if (large_image(lats[i]))
            return fail(STRH_ERR_UNSUPPORTED, "lattice %d: batches address images below 4 GiB", i);
        for (int j = 0; j < i; j++)
            if (lats[j] == lats[i]) return fail(STRH_ERR_INVALID, "lattice %d appears twice", i);
    }
    int rc = set_device(c);
    if (rc) return rc;
    strh_batch* B = new strh_batch();
    B->ctx = c;
    B->precision = lats[0]->precision;
    B->lats.assign(lats, lats + count);
    B->host.assign(count, BatchLat{});
    B->ramp.assign(count, nullptr);
    B->rampLen.assign(count, 0);
    B->series.assign(count, nullptr);
    B->taken.assign(count, 0);
    B->lastSteps.resize(count);
    for (int i = 0; i < count; i++) B->lastSteps[i] = lats[i]->stepsDone;
    B->fring.assign(count, nullptr);
    B->fringLen.assign(count, 0);
    B->fringSpan.assign(count, kProbeSpan);
    B->xchg.assign(count, nullptr);
    auto bail = [&](hipError_t e, const char* what) {
        strh_batch_destroy(B);
        return hip_fail(e, what);
    };
    hipError_t e;
    if ((e = hipMalloc(&B->base, sizeof(int64_t))) != hipSuccess) return bail(e, "hipMalloc(base)");
    if ((e = hipMalloc(&B->abortFlag, sizeof(int))) != hipSuccess) return bail(e, "hipMalloc(abort flag)");
    // The batch kernel runs the walls path everywhere: wall-free lattices get
    // all-zero link words (identical arithmetic, every neighbour fluid).
    for (int i = 0; i < count; i++) {
        strh_lattice* L = lats[i];
        if (!L->link) {
            if ((e = hipMalloc(&L->link, (size_t)L->N * 4)) != hipSuccess) return bail(e, "hipMalloc(link)");
            if ((e = hipMemset(L->link, 0, (size_t)L->N * 4)) != hipSuccess) return bail(e, "zero link");
        }
    }
    *out = B;
    return STRH_OK;
}

int strh_batch_destroy(strh_batch* B) {
    if (!B) return STRH_OK;
    hipSetDevice(B->ctx->device);
    for (float* r : B->ramp) if (r) hipFree(r);
    for (double* d : B->series) if (d) hipFree(d);
    for (void* p : B->fring) if (p) hipFree(p);
    for (uint64_t* p : B->xchg) if (p) hipFree(p);
    if (B->tiles) hipFree(B->tiles);
    if (B->probeParts) hipFree(B->probeParts);
    for (hipEvent_t ev : B->trialEv) if (ev) hipEventDestroy(ev);
    if (B->abortFlag) hipFree(B->abortFlag);
    if (B->dev) hipFree(B->dev);
    if (B->base) hipFree(B->base);
    delete B;
    return STRH_OK;
}

int strh_batch_set_probe(strh_batch* B, int index, int64_t first, int64_t period, int64_t count,
                         int x0, int x1, int y0, int y1) {
    if (!B || index < 0 || index >= (int)B->lats.size()) return fail(STRH_ERR_INVALID, "bad batch index");
    int rc = set_device(B->ctx);
    if (rc) return rc;
    strh_lattice* L = B->lats[index];
    if (!L->wantsForces) return fail(STRH_ERR_NO_FORCES, "lattice %d created without wants_forces", index);
    if (count < 0 || (count > 0 && (period <= 0 || (period & 1) || !(first & 1))))
        return fail(STRH_ERR_INVALID, "probe steps must be odd with an even period");
    if (x0 < 0 || x1 >= L->nx || x0 > x1 || y0 < 0 || y1 >= L->ny || y0 > y1)
        return fail(STRH_ERR_INVALID, "probe box outside the lattice");
    BatchLat& d = B->host[index];
    const uint32_t bw = (uint32_t)(x1 - x0 + 1), bh = (uint32_t)(y1 - y0 + 1);
    d.M = bw * bh * (uint32_t)L->nz;
    box_reduce_plan(d.M, kMaxPartials, &d.Pb, &d.chunk);
    if (d.Pb > kBatchMaxChunks) return fail(STRH_ERR_INVALID, "probe box too large for a batched probe");
    d.bx0 = (uint32_t)x0; d.by0 = (uint32_t)y0;
    d.fbw = make_fastdiv(bw); d.fbh = make_fastdiv(bh);
    d.fT0 = first; d.fPeriod = period; d.fCount = count;
    if (B->series[index]) hipFree(B->series[index]);
    B->series[index] = nullptr;
    if (B->fring[index]) hipFree(B->fring[index]);
    B->fring[index] = nullptr;
    B->fringLen[index] = 0;
    B->fringSpan[index] = kProbeSpan;
    if (count > 0) HIP_TRY(hipMalloc(&B->series[index], (size_t)count * 3 * sizeof(double)), "hipMalloc(series)");
    d.series = B->series[index];
    B->taken[index] = 0;
    B->lastSteps[index] = L->stepsDone;
    // unwritten samples read back as NaN, never as stale device memory
    if (count > 0)
        HIP_TRY(hipMemsetD32Async((hipDeviceptr_t)B->series[index], 0x7ff80000u, (size_t)count * 6,
                                  B->ctx->stream), "fill series");
    if (count > 0) {
        // One force buffer per sample of a probe launch's span, plus one: the
        // long span where that ring stays within 1 GiB, else the short one.
        int span = kProbeSpan;
        int ring = (int)((span + period - 1) / period) + 1;
        if ((size_t)ring * L->N * 16 > ((size_t)1 << 30)) {
            span = kProbeSpanSmall;
            ring = (int)((span + period - 1) / period) + 1;
        }
        const size_t bytes = (size_t)ring * L->N * 16;
        B->fringSpan[index] = span;
        HIP_TRY(hipMalloc(&B->fring[index], bytes), "hipMalloc(force ring)");
        B->fringLen[index] = ring;
        // Cells outside the momentum-exchange set are never written on force
        // steps; one zero fill here stands for the per-probe clear of Sim.swift:361.
        HIP_TRY(launch_fill_zero(B->fring[index], bytes, B->ctx->stream), "zero force");
        HIP_TRY(hipStreamSynchronize(B->ctx->stream), "zero force");
    }
    return STRH_OK;
}

int strh_batch_run(strh_batch* B, const int64_t* steps) {
    if (!B || !steps) return fail(STRH_ERR_INVALID, "null argument");
    int rc = set_device(B->ctx);
    if (rc) return rc;
    int64_t K = 0;
    for (size_t i = 0; i < B->lats.size(); i++) {
        if (steps[i] < 0) return fail(STRH_ERR_INVALID, "negative step count");
        K = std::max(K, steps[i]);
    }
    if (K == 0) return STRH_OK;
    for (size_t i = 0; i < B->lats.size(); i++)
        if (B->host[i].fCount > 0 && B->lats[i]->stepsDone != B->lastSteps[i])
            return fail(STRH_ERR_INVALID, "lattice %zu was stepped outside the batch since its last batch run "
                        "(%lld steps done, %lld expected): its probe samples would be skipped; set its probe "
                        "schedule again", i, (long long)B->lats[i]->stepsDone, (long long)B->lastSteps[i]);
    if ((rc = batch_prepare(B, steps))) return rc;
    double sec = 0.0;
    if ((rc = batch_execute(B, steps, true, true, &sec))) return rc;
    B->gpuSeconds += sec;
    for (size_t i = 0; i < B->lats.size(); i++) {
        strh_lattice* L = B->lats[i];
        const BatchLat& d = B->host[i];
        if (d.fCount > 0) {
            // samples whose step fell inside [step0, tEnd)
            const int64_t last = d.tEnd - 1;
            int64_t n = last < d.fT0 ? 0 : (last - d.fT0) / d.fPeriod + 1;
            B->taken[i] = std::min<int64_t>(std::max<int64_t>(n, 0), d.fCount);
        }
        L->stepsDone += steps[i];
        L->gpuSeconds += sec;
        B->lastSteps[i] = L->stepsDone;
    }
    return STRH_OK;
}

int strh_batch_probe_series(strh_batch* B, int index, double* out, int64_t capacity, int64_t* taken) {
    if (!B || index < 0 || index >= (int)B->lats.size() || !taken) return fail(STRH_ERR_INVALID, "bad arguments");
    int rc = set_device(B->ctx);
    if (rc) return rc;
    *taken = B->taken[index];
    const int64_t n = std::min(capacity, B->taken[index]);
    if (n > 0 && out)
        HIP_TRY(hipMemcpy(out, B->series[index], (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost), "copy series");
    return STRH_OK;
}

int strh_batch_set_block_size(strh_batch* B, int bs) {
    if (!B || (bs != 128 && bs != 256 && bs != 512)) return fail(STRH_ERR_INVALID, "block size must be 128, 256 or 512");
    B->bs = bs;
    return STRH_OK;
}

int strh_batch_set_graph_steps(strh_batch* B, int steps) {
    if (!B || steps < 0 || steps > 4096) return fail(STRH_ERR_INVALID, "graph steps must be in [0, 4096]");
    B->graphSteps = steps;
    return STRH_OK;
}

int strh_batch_set_resident(strh_batch* B, int mode) {
    if (!B || mode < 0 || mode > 3) return fail(STRH_ERR_INVALID, "resident mode must be 0, 1, 2 or 3");
    B->resident = mode;
    return STRH_OK;
}

int strh_batch_set_resident_timeout(strh_batch* B, double seconds) {
    if (!B || !(seconds > 0.0) || seconds > 60.0) return fail(STRH_ERR_INVALID, "timeout must be in (0, 60] seconds");
    int rc = set_device(B->ctx);
    if (rc) return rc;
    int khz = 0;
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, B->ctx->device), "wall clock rate");
    if (khz <= 0) khz = 100000;
    B->residentSpin = std::max<uint64_t>(1, (uint64_t)(seconds * 1e3 * (double)khz));
    return STRH_OK;
}

int strh_batch_gpu_seconds(strh_batch* B, double* seconds) {
    if (!B || !seconds) return fail(STRH_ERR_INVALID, "null argument");
    *seconds = B->gpuSeconds;
    return STRH_OK;
}

// Times 16 steps of the batch at each workgroup size without probes, and
// restores every lattice's DDFs afterwards (the state is copied aside on the
// device). Keeps the fastest; ties go to the smaller size.
int strh_batch_autotune(strh_batch* B, int* chosen) {
    if (!B) return fail(STRH_ERR_INVALID, "null batch");
    int rc = set_device(B->ctx);
    if (rc) return rc;
    std::vector<void*> saved(B->lats.size(), nullptr);
    auto release = [&]() { for (void* p : saved) if (p) hipFree(p); };
    for (size_t i = 0; i < B->lats.size(); i++) {
        const size_t bytes = ddf_bytes(B->lats[i]);
        if (hipMalloc(&saved[i], bytes) != hipSuccess) { release(); return fail(STRH_ERR_ALLOC, "autotune copy"); }
        if (hipMemcpy(saved[i], B->lats[i]->f, bytes, hipMemcpyDeviceToDevice) != hipSuccess) {
            release(); return fail(STRH_ERR_DEVICE, "autotune copy");
        }
    }
    const int keep = B->bs;
    std::vector<int64_t> steps(B->lats.size(), 16);
    std::vector<int64_t> warm(B->lats.size(), 8);
    double best = 1e30;
    int bestBs = keep;
    auto time_it = [&](double* sec) {
        int r = batch_prepare(B, warm.data());
        if (!r) r = batch_execute(B, warm.data(), false, false, sec);
        if (!r) r = batch_prepare(B, steps.data());
        if (!r) r = batch_execute(B, steps.data(), false, false, sec);
        return r;
    };
    for (int bs : {128, 256, 512}) {
        B->bs = bs;
        double sec = 0.0;
        if ((rc = time_it(&sec))) break;
        if (sec < best * 0.98) { best = sec; bestBs = bs; }
    }
    for (size_t i = 0; i < B->lats.size(); i++)
        hipMemcpy(B->lats[i]->f, saved[i], ddf_bytes(B->lats[i]), hipMemcpyDeviceToDevice);
    release();
    if (rc) { B->bs = keep; return rc; }
    B->bs = bestBs;
    if (chosen) *chosen = bestBs;
    return STRH_OK;
}

}  // extern "C"

// ---- discrete adjoint ----------------------------------------------------
namespace {

struct AdjWork {
    std::vector<void*> bufs;
    uint64_t bytes = 0;
    ~AdjWork() { for (void* p : bufs) hipFree(p); }
    template <class T> int get(T** out, size_t bytes_, const char* what) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes_ ? bytes_ : 1);
        if (e != hipSuccess) return fail(STRH_ERR_ALLOC, "hipMalloc(%s, %zu bytes): %s", what, bytes_,
                                         hipGetErrorString(e));
        bufs.push_back(p);
        bytes += bytes_;
        *out = (T*)p;
        return STRH_OK;
    }
};

struct AdjPlan {
    int64_t T = 0;
    double cin = 0.0, cout = 0.0, wd = 0.0, wf = 0.0;   // weights of one window step
    // partial cells: the fluid cells with eps > 0, numbered 0..M-1
    std::vector<int32_t> ntIdx;
    std::vector<uint32_t> ntCells;
};

int adj_plan(strh_lattice* L, const strh_objective* o, AdjPlan& P) {
    if (!o) return fail(STRH_ERR_INVALID, "null objective");
    if (o->spinup < 0 || o->avg <= 0) return fail(STRH_ERR_INVALID, "objective needs spinup >= 0 and avg > 0");
    if (o->inlet_x < -1 || o->inlet_x >= L->nx || o->outlet_x < -1 || o->outlet_x >= L->nx)
        return fail(STRH_ERR_INVALID, "objective columns must lie in [-1, nx)");
    if (!std::isfinite(o->w_p) || !std::isfinite(o->w_d) || !std::isfinite(o->w_f))
        return fail(STRH_ERR_INVALID, "objective weights must be finite");
    if (L->alpha && (L->useEps || L->cSmago > 0.0f))
        return fail(STRH_ERR_UNSUPPORTED, "the adjoint does not combine Brinkman porosity with partial "
                                          "cells or LES");
    if (large_image(L))
        return fail(STRH_ERR_UNSUPPORTED, "the adjoint's recording step addresses images below 4 GiB");
    P.T = o->spinup + o->avg;
    std::vector<uint8_t> fl(L->N);
    HIP_TRY(hipMemcpy(fl.data(), L->flags, L->N, hipMemcpyDeviceToHost), "copy flags");
    if (L->useEps) {
        std::vector<float> e(L->N);
        HIP_TRY(hipMemcpy(e.data(), L->eps, (size_t)L->N * 4, hipMemcpyDeviceToHost), "copy eps");
        P.ntIdx.assign(L->N, -1);
        for (uint32_t n = 0; n < L->N; n++)
            if (fl[n] == kFluid && e[n] > 0.0f) {
                P.ntIdx[n] = (int32_t)P.ntCells.size();
                P.ntCells.push_back(n);
            }
    }
    int64_t nin = 0, nout = 0;
    for (uint32_t n = 0; n < L->N; n++) {
        if (fl[n] != kFluid) continue;
        const int x = (int)(n % (uint32_t)L->nx);
        if (x == o->inlet_x) nin++;
        if (x == o->outlet_x) nout++;
    }
    const double inv = 1.0 / (double)o->avg;
    P.cin = nin ? o->w_p * inv / (3.0 * (double)nin) : 0.0;
    P.cout = nout ? o->w_p * inv / (3.0 * (double)nout) : 0.0;
    P.wd = o->w_d * inv;
    P.wf = o->w_f * inv;
    return STRH_OK;
}

// Hooked forward steps k0 .. k0+count-1 of the objective window on buffer f
// (absolute steps t0 + k): pre-collision moments recorded into rec (count * N
// float4) when given, objective accumulated into jacc when given.
// Partial-cell lattices also record the 19 populations of their M cells with
// eps > 0 into recNt (count * 19 * M floats) when given.
// LES lattices record all 19 populations of every cell into rec19 instead
// (count * 19 * N values in the storage format).
int adj_forward(strh_lattice* L, void* f, const float* alpha, int64_t t0, int64_t k0, int64_t count,
                const strh_objective* o, const AdjPlan& P, float4* rec, double* jacc,
                const int32_t* ntIdx = nullptr, float* recNt = nullptr, uint8_t* rec19 = nullptr,
                hipStream_t stream = nullptr) {
    if (!stream) stream = L->ctx->stream;
    const uint32_t M = (uint32_t)P.ntCells.size();
    for (int64_t s = 0; s < count; s++) {
        const int64_t k = k0 + s;
        StepLaunch sl = make_launch(L, t0 + k, false);
        sl.f = f;
        sl.alpha = (L->useEps || L->cSmago > 0.0f) ? nullptr : alpha;
        AdjHooks hk;
        hk.rec = rec ? rec + (size_t)s * L->N : nullptr;
        const bool win = k >= o->spinup;
        hk.jacc = (jacc && win) ? jacc : nullptr;
        hk.inletX = o->inlet_x;
        hk.outletX = o->outlet_x;
        hk.cin = win ? P.cin : 0.0;
        hk.cout = win ? P.cout : 0.0;
        hk.wd = win ? P.wd : 0.0;
        hk.ntIdx = ntIdx;
        hk.recNt = (recNt && ntIdx) ? recNt + (size_t)s * 19 * M : nullptr;
        hk.M = M;
        hk.wf = win ? P.wf : 0.0;
        hk.rec19 = rec19 ? rec19 + (size_t)s * 19 * L->N * L->B : nullptr;
        HIP_TRY(launch_step_hooked(sl, hk, stream), "hooked step launch");
    }
    return STRH_OK;
}

int reduce_to_host(strh_lattice* L, const double* v, double* out) {
    HIP_TRY(launch_sum_doubles(v, L->N, L->partials, kMaxPartials, L->result, L->ctx->stream), "sum launch");
    HIP_TRY(hipMemcpyAsync(out, L->result, sizeof(double), hipMemcpyDeviceToHost, L->ctx->stream), "copy sum");
    HIP_TRY(hipStreamSynchronize(L->ctx->stream), "sum");
    return STRH_OK;
}

double elapsed(strh_ctx* c) {
    float ms = 0.0f;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    return (double)ms * 1e-3;
}

}  // namespace

extern "C" {

int strh_objective_eval(strh_lattice* L, const strh_objective* o, double* J) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!J) return fail(STRH_ERR_INVALID, "null J");
    AdjPlan P;
    if ((rc = adj_plan(L, o, P))) return rc;
    AdjWork W;
    double* jacc = nullptr;
    float* zeroAlpha = nullptr;
    if ((rc = W.get(&jacc, (size_t)L->N * 8, "objective accumulator"))) return rc;
    HIP_TRY(launch_fill_zero(jacc, (size_t)L->N * 8, L->ctx->stream), "clear");
    const float* alpha = L->alpha;
    if (!alpha && !L->useEps && !(L->cSmago > 0.0f)) {
        if ((rc = W.get(&zeroAlpha, (size_t)L->N * 4, "zero porosity"))) return rc;
        HIP_TRY(launch_fill_zero(zeroAlpha, (size_t)L->N * 4, L->ctx->stream), "clear");
        alpha = zeroAlpha;
    }
    int32_t* ntIdx = nullptr;
    if (L->useEps) {
        if ((rc = W.get(&ntIdx, (size_t)L->N * 4, "partial-cell index"))) return rc;
        HIP_TRY(hipMemcpy(ntIdx, P.ntIdx.data(), (size_t)L->N * 4, hipMemcpyHostToDevice), "copy index");
    }
    strh_ctx* c = L->ctx;
    HIP_TRY(hipEventRecord(c->ev0, c->stream), "hipEventRecord");
    if ((rc = adj_forward(L, L->f, alpha, L->stepsDone, 0, P.T, o, P, nullptr, jacc, ntIdx, nullptr))) return rc;
    HIP_TRY(hipEventRecord(c->ev1, c->stream), "hipEventRecord");
    HIP_TRY(hipEventSynchronize(c->ev1), "objective steps");
    L->gpuSeconds += elapsed(c);
    L->stepsDone += P.T;
    if ((rc = reduce_to_host(L, jacc, J))) return rc;
    if (!std::isfinite(*J)) return fail(STRH_ERR_NOT_FINITE, "objective is not finite");
    return STRH_OK;
}

}  // extern "C"

namespace {

// Memory of a VJP plan beyond the fixed buffers: K2 recorded steps, the outer
// checkpoints (one DDF image per outer segment but the last) and the inner
// ones (one per inner segment of an outer segment but the last).
struct VjpPlan {
    int64_t K1 = 0, K2 = 0;
};

double vjp_plan_bytes(int64_t T, int64_t K1, int64_t K2, double perStep, double ddf) {
    const int64_t C1 = (T + K1 - 1) / K1, C2 = (K1 + K2 - 1) / K2;
    return (double)K2 * perStep + (double)(C1 - 1 + C2 - 1) * ddf;
}

// The adjoint over T steps with two checkpoint levels. The forward pass keeps
// the DDFs at the start of every outer segment of K1 steps. The reverse pass
// takes the outer segments last to first; each is replayed once, keeping the
// DDFs at the start of every inner segment of K2 steps and recording the last
// inner segment; the others are replayed from their checkpoints, recording,
// last to first, and the reverse steps run over each recording. K2 = K1 is
// the single-level scheme (one replay per step); K2 < K1 replays each step
// twice and needs ceil(T/K1) - 1 + ceil(K1/K2) - 1 images and K2 recorded
// steps instead of ceil(T/K1) - 1 images and K1 recorded steps. The last
// outer segment's inner checkpoints come from the forward pass itself. Every
// replay runs the forward pass's kernel from the forward pass's bits, so the
// recordings, and with them J and the gradients, do not depend on K1 or K2.
int vjp_run(strh_lattice* L, const strh_objective* o, const AdjPlan& P, VjpPlan plan, double* J,
            double* gradAlpha, double* gradEps, double* gradUin, strh_vjp_info* info) {
    int rc = STRH_OK;
    const int64_t T = P.T;
    int64_t K1 = std::min(plan.K1, T);
    int64_t K2 = std::min(plan.K2, K1);
    const size_t N = L->N, S = L->S;
    const size_t ddf = ddf_bytes(L);
    const size_t M = P.ntCells.size();
    const bool les = L->cSmago > 0.0f;
    const int64_t C1 = (T + K1 - 1) / K1;
    const int64_t C2 = (K1 + K2 - 1) / K2;
    strh_ctx* c = L->ctx;
    hipStream_t st = c->stream;

    AdjWork W;
    double *jacc = nullptr, *gA = nullptr, *gU = nullptr;
    float4* rec = nullptr;
    uint8_t* rec19 = nullptr;
    float *lam0 = nullptr, *lam1 = nullptr, *zeroAlpha = nullptr;
    uint8_t *ck = nullptr, *ick = nullptr, *work = nullptr;
    if ((rc = W.get(&jacc, N * 8, "objective accumulator"))) return rc;
    if ((rc = W.get(&gA, N * 8, "porosity gradient"))) return rc;
    if ((rc = W.get(&gU, N * 8, "inlet gradient"))) return rc;
    if ((rc = W.get(&lam0, 19 * S * 4, "adjoint state"))) return rc;
    if ((rc = W.get(&lam1, 19 * S * 4, "adjoint state"))) return rc;
    if (les) {
        if ((rc = W.get(&rec19, (size_t)K2 * 19 * N * L->B, "recorded populations"))) return rc;
    } else {
        if ((rc = W.get(&rec, (size_t)K2 * N * 16, "recorded moments"))) return rc;
    }
    if ((rc = W.get(&ck, (size_t)(C1 > 1 ? C1 - 1 : 1) * ddf, "checkpoints"))) return rc;
    if (C2 > 1 && (rc = W.get(&ick, (size_t)(C2 - 1) * ddf, "inner checkpoints"))) return rc;
    if ((rc = W.get(&work, ddf, "replay DDFs"))) return rc;
    // One level with replays on a lattice whose step kernels leave most of the
    // GPU idle (at most 2^20 cells: latency-bound launches), when a second
    // recording and replay image fit: segment s - 1 is replayed on a second
    // stream while the reverse steps of segment s run (pipelined); otherwise
    // the two alternate. (On bandwidth-bound lattices the overlap measured
    // within +-5% of alternating: both kinds of kernel want the whole HBM.)
    const size_t recBytes = les ? (size_t)K2 * 19 * N * L->B : (size_t)K2 * N * 16;
    const size_t recNtBytes = L->useEps && !les ? (size_t)K2 * 19 * std::max<size_t>(M, 1) * 4 : 0;
    bool pipe = false;
    if (C2 == 1 && C1 > 1 && N <= ((size_t)1 << 20)) {
        size_t freeB = 0, totalB = 0;
        HIP_TRY(hipMemGetInfo(&freeB, &totalB), "hipMemGetInfo");
        pipe = freeB > recBytes + recNtBytes + ddf + ((size_t)1 << 30);
    }
    float4* recX[2] = {nullptr, nullptr};
    uint8_t* rec19X[2] = {nullptr, nullptr};
    float* recNtX[2] = {nullptr, nullptr};
    uint8_t* workX[2] = {work, nullptr};
    if (pipe) {
        if (les) { if ((rc = W.get(&rec19X[1], recBytes, "recorded populations"))) return rc; }
        else if ((rc = W.get(&recX[1], recBytes, "recorded moments"))) return rc;
        if ((rc = W.get(&workX[1], ddf, "replay DDFs"))) return rc;
    }
    const float* alpha = L->alpha;
    if (!alpha && !L->useEps && !les) {
        if ((rc = W.get(&zeroAlpha, N * 4, "zero porosity"))) return rc;
        HIP_TRY(launch_fill_zero(zeroAlpha, N * 4, st), "clear");
        alpha = zeroAlpha;
    }
    int32_t* ntIdx = nullptr;
    float* recNt = nullptr;
    double* gE = nullptr;
    if (L->useEps) {
        const size_t m = std::max<size_t>(M, 1);
        if ((rc = W.get(&ntIdx, N * 4, "partial-cell index"))) return rc;
        if (!les && (rc = W.get(&recNt, (size_t)K2 * 19 * m * 4, "recorded partial cells"))) return rc;
        if (pipe && !les && (rc = W.get(&recNtX[1], (size_t)K2 * 19 * m * 4, "recorded partial cells"))) return rc;
        if ((rc = W.get(&gE, m * 8, "solid-fraction gradient"))) return rc;
        HIP_TRY(hipMemcpy(ntIdx, P.ntIdx.data(), N * 4, hipMemcpyHostToDevice), "copy index");
        HIP_TRY(launch_fill_zero(gE, m * 8, st), "clear");
    }
    for (void* p : {(void*)jacc, (void*)gA, (void*)gU})
        HIP_TRY(launch_fill_zero(p, N * 8, st), "clear");
    HIP_TRY(launch_fill_zero(lam0, 19 * S * 4, st), "clear");
    HIP_TRY(launch_fill_zero(lam1, 19 * S * 4, st), "clear");
    const int64_t t0 = L->stepsDone;

    // Steps [k0, k0 + len) of an outer segment on the DDFs f: the inner
    // checkpoints at the start of every inner segment but the last, which is
    // recorded (jacc: accumulate the objective, first forward pass only).
    auto advance_outer = [&](void* f, int64_t k0, int64_t len, double* acc) -> int {
        const int64_t n2 = (len + K2 - 1) / K2;
        for (int64_t j = 0; j < n2; j++) {
            const int64_t kk = k0 + j * K2, l2 = std::min(K2, len - j * K2);
            const bool lastInner = j == n2 - 1;
            if (!lastInner)
                HIP_TRY(hipMemcpyAsync(ick + (size_t)j * ddf, f, ddf, hipMemcpyDeviceToDevice, st),
                        "inner checkpoint");
            int r = adj_forward(L, f, alpha, t0, kk, l2, o, P, lastInner ? rec : nullptr, acc, ntIdx,
                                lastInner ? recNt : nullptr, lastInner ? rec19 : nullptr);
            if (r) return r;
        }
        return STRH_OK;
    };

    recX[0] = rec; rec19X[0] = rec19; recNtX[0] = recNt;
    if (pipe) {   // the forward pass records the last segment into buffer (C1 - 1) & 1
        rec = recX[(C1 - 1) & 1]; rec19 = rec19X[(C1 - 1) & 1]; recNt = recNtX[(C1 - 1) & 1];
    }

    // forward pass: outer checkpoints at the start of outer segments
    // 0 .. C1-2; the last outer segment lays down its inner checkpoints and
    // records its last inner segment
    HIP_TRY(hipEventRecord(c->ev0, st), "hipEventRecord");
    for (int64_t seg = 0; seg < C1; seg++) {
        const int64_t k0 = seg * K1, len = std::min(K1, T - k0);
        if (seg != C1 - 1) {
            HIP_TRY(hipMemcpyAsync(ck + (size_t)seg * ddf, L->f, ddf, hipMemcpyDeviceToDevice, st), "checkpoint");
            if ((rc = adj_forward(L, L->f, alpha, t0, k0, len, o, P, nullptr, jacc, ntIdx, nullptr, nullptr)))
                return rc;
        } else if ((rc = advance_outer(L->f, k0, len, jacc))) {
            return rc;
        }
    }
    HIP_TRY(hipEventRecord(c->ev1, st), "hipEventRecord");
    HIP_TRY(hipEventSynchronize(c->ev1), "forward pass");
    double tf = elapsed(c), tr = 0.0, tb = 0.0;
    L->gpuSeconds += tf;
    L->stepsDone += T;

    // reverse pass
    float* lamN = lam0;
    float* lamT = lam1;
    // the reverse steps kk + l2 - 1 .. kk over the recording (recB, recNtB,
    // rec19B) of steps kk .. kk + l2 - 1, on stream sr
    auto reverse_steps = [&](int64_t kk, int64_t l2, float4* recB, float* recNtB, uint8_t* rec19B,
                             hipStream_t sr) -> int {
        for (int64_t k = kk + l2 - 1; k >= kk; k--) {
            const int64_t t = t0 + k;
            const StepParams sp = make_params(L, t, false);
            AdjStep q;
            memset(&q, 0, sizeof(q));
            q.nx = (uint32_t)L->nx; q.ny = (uint32_t)L->ny; q.nz = (uint32_t)L->nz;
            q.N = L->N; q.S = L->S;
            q.divx = sp.divx; q.divy = sp.divy;
            q.wp = sp.omega; q.wm = sp.omegaMinus;
            q.fx = sp.fx; q.fy = sp.fy; q.fz = sp.fz;
            q.rOdd = (double)ramp(L, (t & 1) ? t : t + 1);
            q.uinTarget = L->uinTarget;
            q.inflowUniform = L->inflowUniform ? 1 : 0;
            const bool win = k >= o->spinup;
            q.inletX = o->inlet_x; q.outletX = o->outlet_x;
            q.cin = win ? (float)P.cin : 0.0f;
            q.cout = win ? (float)P.cout : 0.0f;
            q.wd = win ? (float)P.wd : 0.0f;
            q.wf = win ? (float)P.wf : 0.0f;
            q.M = (uint32_t)M;
            q.spongeX0 = sp.spongeX0; q.spongeInvW = sp.spongeInvW; q.spongeTau = sp.spongeTau;
            q.cSmago = sp.cSmago;
            q.srt = sp.omegaMinus == sp.omega ? 1 : 0;
            q.precision = L->precision;
            HIP_TRY(launch_adjoint_step(q, L->walls, recB ? recB + (size_t)(k - kk) * N : nullptr, lamN, lamT,
                                        L->flags, L->link, alpha, gA, gU, L->useEps ? L->eps : nullptr,
                                        ntIdx, recNtB ? recNtB + (size_t)(k - kk) * 19 * M : nullptr, gE,
                                        rec19B ? rec19B + (size_t)(k - kk) * 19 * N * L->B : nullptr, L->lidm,
                                        sp, sr),
                    "adjoint step launch");
            std::swap(lamN, lamT);
        }
        return STRH_OK;
    };
    double tp = 0.0;   // wall time of the reverse pass
    if (pipe) {
        // Segment s's reverse steps (stream st, buffer s & 1) wait for its
        // replay; the replay of segment s - 1 (stream sb, buffer (s - 1) & 1)
        // waits for the reverse steps of segment s + 1, the last reader of that
        // buffer. Nothing else is shared: the forward pass has completed.
        hipStream_t sb = nullptr;
        HIP_TRY(hipStreamCreateWithFlags(&sb, hipStreamNonBlocking), "replay stream");
        std::vector<hipEvent_t> ev(4 * (size_t)C1, nullptr);   // replay begin / end, reverse begin / end
        auto cleanup = [&]() {
            (void)hipStreamSynchronize(sb);
            (void)hipStreamSynchronize(st);
            for (hipEvent_t e : ev) if (e) hipEventDestroy(e);
            hipStreamDestroy(sb);
        };
        for (hipEvent_t& e : ev)
            if (hipEventCreate(&e) != hipSuccess) { cleanup(); return fail(STRH_ERR_DEVICE, "hipEventCreate"); }
        auto repB = [&](int64_t s) { return ev[4 * s]; };
        auto repE = [&](int64_t s) { return ev[4 * s + 1]; };
        auto revB = [&](int64_t s) { return ev[4 * s + 2]; };
        auto revE = [&](int64_t s) { return ev[4 * s + 3]; };
        hipError_t e = hipEventRecord(c->ev0, st);
        for (int64_t seg = C1 - 1; seg >= 0 && e == hipSuccess && !rc; seg--) {
            if (seg >= 1) {
                const int64_t sr = seg - 1, k0r = sr * K1, lr = std::min(K1, T - k0r);
                const int b = (int)(sr & 1);
                if (sr + 2 <= C1 - 1) e = hipStreamWaitEvent(sb, revE(sr + 2), 0);
                if (e == hipSuccess) e = hipEventRecord(repB(sr), sb);
                if (e == hipSuccess) e = hipMemcpyAsync(workX[b], ck + (size_t)sr * ddf, ddf, hipMemcpyDeviceToDevice, sb);
                if (e == hipSuccess)
                    rc = adj_forward(L, workX[b], alpha, t0, k0r, lr, o, P, recX[b], nullptr, ntIdx, recNtX[b],
                                     rec19X[b], sb);
                if (e == hipSuccess && !rc) e = hipEventRecord(repE(sr), sb);
            }
            if (e != hipSuccess || rc) break;
            const int a = (int)(seg & 1);
            const int64_t k0s = seg * K1, ls = std::min(K1, T - k0s);
            if (seg != C1 - 1) e = hipStreamWaitEvent(st, repE(seg), 0);
            if (e == hipSuccess) e = hipEventRecord(revB(seg), st);
            if (e == hipSuccess) rc = reverse_steps(k0s, ls, recX[a], recNtX[a], rec19X[a], st);
            if (e == hipSuccess && !rc) e = hipEventRecord(revE(seg), st);
        }
        if (e == hipSuccess && !rc) e = hipEventRecord(c->ev1, st);
        if (e == hipSuccess && !rc) e = hipEventSynchronize(c->ev1);
        if (e == hipSuccess && !rc) {
            tp = elapsed(c);
            float ms = 0.0f;
            for (int64_t sg2 = 0; sg2 < C1 && e == hipSuccess; sg2++) {
                if (sg2 < C1 - 1 && (e = hipEventElapsedTime(&ms, repB(sg2), repE(sg2))) == hipSuccess) tr += ms * 1e-3;
                if (e == hipSuccess && (e = hipEventElapsedTime(&ms, revB(sg2), revE(sg2))) == hipSuccess) tb += ms * 1e-3;
            }
        }
        cleanup();
        if (rc) return rc;
        if (e != hipSuccess) return hip_fail(e, "pipelined reverse pass");
    } else {
        HIP_TRY(hipEventRecord(c->ev2, st), "hipEventRecord");
        for (int64_t seg = C1 - 1; seg >= 0; seg--) {
            const int64_t k0 = seg * K1, len = std::min(K1, T - k0);
            const int64_t n2 = (len + K2 - 1) / K2;
            if (seg != C1 - 1) {
                HIP_TRY(hipEventRecord(c->ev0, st), "hipEventRecord");
                HIP_TRY(hipMemcpyAsync(work, ck + (size_t)seg * ddf, ddf, hipMemcpyDeviceToDevice, st), "restore");
                if ((rc = advance_outer(work, k0, len, nullptr))) return rc;
                HIP_TRY(hipEventRecord(c->ev1, st), "hipEventRecord");
                HIP_TRY(hipEventSynchronize(c->ev1), "replay");
                tr += elapsed(c);
            }
            for (int64_t j = n2 - 1; j >= 0; j--) {
                const int64_t kk = k0 + j * K2, l2 = std::min(K2, len - j * K2);
                if (j != n2 - 1) {
                    HIP_TRY(hipEventRecord(c->ev0, st), "hipEventRecord");
                    HIP_TRY(hipMemcpyAsync(work, ick + (size_t)j * ddf, ddf, hipMemcpyDeviceToDevice, st),
                            "restore inner");
                    if ((rc = adj_forward(L, work, alpha, t0, kk, l2, o, P, rec, nullptr, ntIdx, recNt, rec19)))
                        return rc;
                    HIP_TRY(hipEventRecord(c->ev1, st), "hipEventRecord");
                    HIP_TRY(hipEventSynchronize(c->ev1), "replay");
                    tr += elapsed(c);
                }
                HIP_TRY(hipEventRecord(c->ev0, st), "hipEventRecord");
                if ((rc = reverse_steps(kk, l2, rec, recNt, rec19, st))) return rc;
                HIP_TRY(hipEventRecord(c->ev1, st), "hipEventRecord");
                HIP_TRY(hipEventSynchronize(c->ev1), "reverse steps");
                tb += elapsed(c);
            }
        }
        HIP_TRY(hipEventRecord(c->ev3, st), "hipEventRecord");
        HIP_TRY(hipEventSynchronize(c->ev3), "reverse pass");
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev2, c->ev3), "hipEventElapsedTime");
        tp = (double)ms * 1e-3;
    }
    if ((rc = reduce_to_host(L, jacc, J))) return rc;
    if (gradUin && (rc = reduce_to_host(L, gU, gradUin))) return rc;
    if (gradAlpha) HIP_TRY(hipMemcpy(gradAlpha, gA, N * 8, hipMemcpyDeviceToHost), "copy gradient");
    if (gradEps) {
        std::fill(gradEps, gradEps + N, 0.0);
        if (M) {
            std::vector<double> ge(M);
            HIP_TRY(hipMemcpy(ge.data(), gE, M * 8, hipMemcpyDeviceToHost), "copy gradient");
            for (size_t k = 0; k < M; k++) gradEps[P.ntCells[k]] = ge[k];
        }
    }
    if (info) {
        info->checkpoint_every = K2;
        info->checkpoints = (C1 - 1) + (C2 - 1);
        info->forward_seconds = tf;
        info->recompute_seconds = tr;
        info->backward_seconds = tb;
        info->workspace_bytes = W.bytes;
        info->outer_every = K1;
        info->reverse_pass_seconds = tp;
    }
    if (!std::isfinite(*J)) return fail(STRH_ERR_NOT_FINITE, "objective is not finite");
    return STRH_OK;
}

}  // namespace

extern "C" {

int strh_vjp(strh_lattice* L, const strh_objective* o, int64_t K, double* J, double* gradAlpha,
             double* gradEps, double* gradUin, strh_vjp_info* info) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!J) return fail(STRH_ERR_INVALID, "null J");
    AdjPlan P;
    if ((rc = adj_plan(L, o, P))) return rc;
    const int64_t T = P.T;
    const size_t N = L->N, S = L->S;
    const double ddf = (double)ddf_bytes(L);
    const size_t M = P.ntCells.size();
    const bool les = L->cSmago > 0.0f;
    if (K < -1) return fail(STRH_ERR_INVALID, "checkpoint_every must be >= -1");
    VjpPlan plan;
    if (K == 0) K = (int64_t)std::ceil(std::sqrt((double)T));
    if (K > 0) {
        plan.K1 = plan.K2 = std::min(K, T);
    } else {
        // fewest replays: the longest single-level segment whose recording
        // and checkpoints fit 80% of the free device memory left after the
        // fixed buffers; when none fits, two levels with outer and inner
        // segments of about T^(2/3) and T^(1/3) steps (each step replayed
        // twice, ~3 T^(1/3) images); when that does not fit either, the
        // single-level plan needing least (the allocation reports the shortfall)
        size_t freeB = 0, totalB = 0;
        HIP_TRY(hipMemGetInfo(&freeB, &totalB), "hipMemGetInfo");
        const double fixed = (double)N * (3 * 8 + 4) + 2.0 * 19 * S * 4 + 2.0 * ddf;
        const double budget = 0.8 * (double)freeB - fixed;
        // recording per step (moments and partial-cell populations, or all
        // populations under LES)
        const double perStep = les ? (double)N * 19 * L->B : (double)N * 16 + (double)M * 76;
        auto need = [&](int64_t k) { return vjp_plan_bytes(T, k, k, perStep, ddf); };
        int64_t best = 1;
        K = 0;
        for (int64_t k = std::min<int64_t>(T, budget > perStep ? (int64_t)(budget / perStep) : 1); k >= 1; k--) {
            if (need(k) <= budget) { K = k; break; }
            if (need(k) < need(best)) best = k;
        }
        if (K > 0) {
            plan.K1 = plan.K2 = K;
        } else {
            const int64_t k2 = std::max<int64_t>(1, (int64_t)std::ceil(std::cbrt((double)T)));
            const int64_t k1 = std::min<int64_t>(T, k2 * k2);
            if (vjp_plan_bytes(T, k1, k2, perStep, ddf) <= budget) {
                plan.K1 = k1;
                plan.K2 = k2;
            } else {
                for (int64_t k = 1; k <= T && (double)k * perStep <= need(best); k++)
                    if (need(k) < need(best)) best = k;
                plan.K1 = plan.K2 = best;
            }
        }
    }
    return vjp_run(L, o, P, plan, J, gradAlpha, gradEps, gradUin, info);
}

int strh_vjp_nested(strh_lattice* L, const strh_objective* o, int64_t outer_every, int64_t inner_every,
                    double* J, double* gradAlpha, double* gradEps, double* gradUin, strh_vjp_info* info) {
    int rc = check_lat(L);
    if (rc) return rc;
    if (!J) return fail(STRH_ERR_INVALID, "null J");
    if (outer_every < 1 || inner_every < 1)
        return fail(STRH_ERR_INVALID, "outer_every and inner_every must be >= 1");
    AdjPlan P;
    if ((rc = adj_plan(L, o, P))) return rc;
    VjpPlan plan;
    plan.K1 = std::min(outer_every, P.T);
    plan.K2 = std::min(inner_every, plan.K1);
    return vjp_run(L, o, P, plan, J, gradAlpha, gradEps, gradUin, info);
}

}  // extern "C"
