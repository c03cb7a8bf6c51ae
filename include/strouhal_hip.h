/*
 * strouhal_hip.h — C ABI of the MI355X (gfx950) D3Q19 lattice-Boltzmann
 * solver: the drop-in replacement for the Metal backend behind StrouhalCore's
 * `GPU` + `Simulation` (Sources/StrouhalCore/Sim.swift:60-459).
 *
 * Conventions
 *  - Every entry point returns an int status (STRH_OK = 0). Nothing aborts.
 *    On failure strh_last_error() returns a thread-local message; this replaces
 *    Swift's `throws StrouhalError` and the `precondition` traps
 *    (Sim.swift:127-136, 335, 360, 414).
 *  - Handles own their device memory; inputs are copied at call time; output
 *    buffers belong to the caller (the reference returns views of shared
 *    memory, Sim.swift:351-352 — HBM is discrete here, so the ABI copies).
 *  - One lattice is not thread-safe; several contexts may coexist in one
 *    process (the reference app runs a second `GPU()`, StrouhalApp.swift:846).
 *    Each context owns one HIP stream at default priority.
 *  - Cell index n = (z*ny + y)*nx + x; DDF slot order i*N + n
 *    (Kernels.metal:71, 243-246). D3Q19 direction order as Kernels.metal:39-47.
 */
#ifndef STROUHAL_HIP_H
#define STROUHAL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define STRH_API __attribute__((visibility("default")))
#else
#define STRH_API
#endif

typedef struct strh_ctx strh_ctx;
typedef struct strh_lattice strh_lattice;

/* Status codes. */
enum {
    STRH_OK = 0,
    STRH_ERR_INVALID = 1,     /* bad argument (null handle, size, range, tau) */
    STRH_ERR_NO_DEVICE = 2,   /* Sim.swift:127-135 StrouhalError.noDevice */
    STRH_ERR_DEVICE = 3,      /* HIP runtime / kernel failure (Sim.swift:328-330) */
    STRH_ERR_ALLOC = 4,       /* device or host allocation failed (Sim.swift:221) */
    STRH_ERR_ODD_STEPS = 5,   /* probe needs an even step count (Sim.swift:335, 360) */
    STRH_ERR_NO_FORCES = 6,   /* probe_force on a lattice created without forces */
    STRH_ERR_NOT_FINITE = 7,  /* a reduction produced a non-finite value */
    STRH_ERR_UNSUPPORTED = 8  /* the lattice uses a feature the entry point does not cover */
};

/* Storage precision. Sim.swift:5-8 has fp32 and fp16s; bf16 is new. */
enum { STRH_FP32 = 0, STRH_FP16S = 1, STRH_BF16 = 2 };

/* Cell flags, Sim.swift:46-52 / Kernels.metal:65-69 (raw values). */
enum { STRH_FLUID = 0, STRH_SOLID = 1, STRH_LID = 2, STRH_INFLOW = 3, STRH_OUTFLOW = 4 };

/* Lattice construction parameters: the arguments of Simulation.init
 * (Sim.swift:195-200). omega_minus < 0 means "same as omega" (SRT), the
 * `omegaMinus: Float? = nil` default. */
typedef struct strh_lattice_desc {
    int32_t precision;        /* STRH_FP32 / STRH_FP16S / STRH_BF16 */
    int32_t nx, ny, nz;
    float omega;
    float omega_minus;
    float lid[3];
    float uin;
    int64_t ramp_steps;
    float force[3];
    float c_smago;            /* Smagorinsky Cs^2; 0 = LES off */
    int32_t wants_forces;
} strh_lattice_desc;

/* ---- context: replaces GPU.init (Sim.swift:67-78) ----------------------- */
STRH_API int strh_ctx_create(int device_ordinal, strh_ctx** out);
STRH_API int strh_ctx_destroy(strh_ctx* ctx);
STRH_API const char* strh_last_error(void);
STRH_API int strh_device_name(strh_ctx* ctx, char* out, size_t cap);

/* ---- lattice: replaces Simulation.init (Sim.swift:195-256) --------------
 * `flags` holds nx*ny*nz raw Cell values in cell-index order; it replaces the
 * Swift `(Int, Int, Int) -> Cell` closure. The neighbour masks of
 * Sim.swift:235-255 are built by the library. The DDFs start at the shifted
 * rest state 0 (Sim.swift:225). */
STRH_API int strh_lattice_create(strh_ctx* ctx, const strh_lattice_desc* desc,
                                 const uint8_t* flags, strh_lattice** out);
STRH_API int strh_lattice_destroy(strh_lattice* lat);
STRH_API int strh_lattice_cells(strh_lattice* lat, int64_t* cells);
/* Bytes of the canonical DDF image (19 * N * storage bytes). */
STRH_API int strh_ddf_bytes(strh_lattice* lat, size_t* bytes);

/* ---- mutable public vars, Sim.swift:164-172 ------------------------------ */
STRH_API int strh_set_omega(strh_lattice* lat, float omega, float omega_minus);
STRH_API int strh_set_lid_velocity(strh_lattice* lat, float ux, float uy, float uz);
STRH_API int strh_set_uin(strh_lattice* lat, float uin_target);
STRH_API int strh_set_ramp_steps(strh_lattice* lat, int64_t ramp_steps);
STRH_API int strh_set_force(strh_lattice* lat, float fx, float fy, float fz);
STRH_API int strh_set_c_smago(strh_lattice* lat, float c_smago);
STRH_API int strh_set_inflow_uniform(strh_lattice* lat, int enabled);
/* sponge = (x0, width, tau); enabled = 0 restores `sponge = nil`. */
STRH_API int strh_set_sponge(strh_lattice* lat, int enabled, float x0, float width, float tau);
/* Noble-Torczynski solid fractions, nx*ny*nz floats (Sim.swift:413-421). */
STRH_API int strh_set_solid_fractions(strh_lattice* lat, const float* eps);
/* Brinkman porosity, nx*ny*nz floats >= 0 (no reference counterpart: the
 * reference has only the solid fractions above). Each fluid cell feels the
 * drag -alpha rho u with u = (p + F/2) / (rho (1 + alpha/2)), applied through
 * the Guo forcing terms; alpha = 0 everywhere gives the plain forced update.
 * NULL removes the field. Not combinable with solid fractions or LES, and not
 * accepted by batches (STRH_ERR_INVALID). */
STRH_API int strh_set_porosity(strh_lattice* lat, const float* alpha);

/* ---- initialisation: Simulation.initField (Sim.swift:424-439) ------------
 * mode 0 rest, 1 2-D Taylor-Green, 2 3-D Taylor-Green vortex; tau = 1/omega. */
STRH_API int strh_init_field(strh_lattice* lat, uint32_t mode, float amplitude);

/* ---- stepping: Simulation.run (Sim.swift:288-331) ------------------------ */
STRH_API int strh_run(strh_lattice* lat, int64_t steps);
/* strh_run with a HIP event pair around every step launch (no reference
 * counterpart; a measurement aid): ms[k] receives the duration of step k in
 * milliseconds (`steps` floats, steps <= 2^20). Advances stepsDone and the
 * GPU seconds like strh_run. */
STRH_API int strh_run_timed(strh_lattice* lat, int64_t steps, float* ms);
STRH_API int strh_steps_done(strh_lattice* lat, int64_t* steps);
/* Accumulated GPU time of run(), Sim.swift:176-178. */
STRH_API int strh_gpu_seconds(strh_lattice* lat, double* seconds);

/* ---- observables ----------------------------------------------------------
 * probe_moments: Sim.swift:334-353 (Kernels.metal:359-380). `out` receives
 * 4*N floats (ux, uy, uz, rho) per cell. STRH_ERR_ODD_STEPS after an odd count. */
STRH_API int strh_probe_moments(strh_lattice* lat, float* out);
/* probe_force: Sim.swift:359-372. Advances 2 steps (force on the odd pass),
 * then a fixed-order FP64 sum over z in [0,nz), y in [y0,y1], x in [x0,x1]
 * (inclusive ranges). */
STRH_API int strh_probe_force(strh_lattice* lat, int x0, int x1, int y0, int y1,
                              double out[3]);
/* stateDigest: Sim.swift:441-444 — SHA-256 of the 19*N*B raw DDF bytes in
 * slot order i*N + n. */
STRH_API int strh_state_digest(strh_lattice* lat, uint8_t out[32]);
/* massSum: Sim.swift:447-458 — FP64 sum of all 19*N shifted DDFs (equals
 * sum(rho - 1)), reduced on the GPU in a fixed order. */
STRH_API int strh_mass_sum(strh_lattice* lat, double* out);
/* New: FP64 sum over cells of |u|^2 (u as probe_moments computes it, the
 * per-cell term in FP32 as Cases.swift:1242), reduced on the GPU in a fixed
 * order. Requires an even step count. */
STRH_API int strh_kinetic_energy(strh_lattice* lat, double* out);
/* strh_kinetic_energy without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The per-cell term is
 * computed straight from the DDFs and reduced in the same fixed order, so the
 * result is the same bits as strh_kinetic_energy. The reduction works in a
 * static device array of 9,728 partial sums plus the result (77,856 bytes per
 * GPU, held until the process ends); a lattice needing more partials is
 * refused (STRH_ERR_INVALID; the plan caps at that many, so none does up to
 * the 200M-cell lattice limit). Every call on one GPU shares that array: calls
 * that overlap on one GPU (two contexts, two threads) race. */
STRH_API int strh_kinetic_energy_noalloc(strh_lattice* lat, double* out);
/* Viscous dissipation rate sum_fluid 2 nu S:S (FP64, fixed order), nu =
 * (1/omega - 1/2)/3, with the strain rate from the non-equilibrium momentum
 * flux, S = -3 Pi_neq / (2 rho tau). No reference counterpart: the reference
 * differentiates E(t) in time (Cases.swift:1254-1257). Even step count;
 * molecular viscosity only (no eddy viscosity, no body-force correction). */
STRH_API int strh_dissipation(strh_lattice* lat, double* out);

/* ---- raw state (the self-test writes fBuf directly, Cases.swift:90-99) --- */
STRH_API int strh_download_ddf(strh_lattice* lat, void* out, size_t bytes);
STRH_API int strh_upload_ddf(strh_lattice* lat, const void* in, size_t bytes);
/* Per-cell momentum-exchange field of the last force step (4*N floats). */
STRH_API int strh_download_force(strh_lattice* lat, float* out);

/* ---- checkpoint / resume (new): device-side snapshot of the DDFs and the
 * step count; restore rewinds both. One snapshot slot per lattice. */
/* State files (no reference counterpart: the reference keeps state only in
 * memory and replays to resume, Cases.swift:989-999). A 128-byte
 * little-endian header -- char magic[8] "STRHDDF1", uint32 version (1),
 * uint32 precision, int32 nx, ny, nz, uint32 bytes per value, int64 steps
 * done, uint64 image bytes, uint8 digest[32] (SHA-256, at offset 48), 48
 * reserved bytes -- followed by the canonical image (the bytes
 * strh_state_digest hashes).
 * load checks magic, precision, dimensions and the digest before it uploads
 * anything, and restores steps done. */
STRH_API int strh_save_state(strh_lattice* lat, const char* path);
STRH_API int strh_load_state(strh_lattice* lat, const char* path);
STRH_API int strh_snapshot(strh_lattice* lat);
STRH_API int strh_restore(strh_lattice* lat);

/* ---- multi-lattice batch (new) -------------------------------------------
 * Replaces the sequential rung loop of Ladder.credibility / Ladder.measure
 * (Sources/StrouhalCore/Ladder.swift:54-121): every lattice of a batch is
 * advanced by one launch per step, its own step count and probe schedule
 * carried on the device, and chunks of steps are replayed as a HIP graph.
 * Lattices must share the context and the storage precision; a batched
 * lattice evolves bit for bit as it does alone. */
typedef struct strh_batch strh_batch;
STRH_API int strh_batch_create(strh_ctx* ctx, strh_lattice* const* lattices, int count,
                               strh_batch** out);
STRH_API int strh_batch_destroy(strh_batch* batch);
/* Momentum-exchange samples of lattice `index` on the odd steps
 * first_step + j*period (j < count; first_step odd, period even), each the
 * FP64 box sum of probe_force over x in [x0,x1], y in [y0,y1], all z: the
 * schedule of Ladder.measure (Ladder.swift:56-66) with the probes kept on the
 * device. Needs wants_forces. A lattice with a schedule must not be stepped
 * outside the batch between batch runs (strh_batch_run then fails with
 * STRH_ERR_INVALID: the samples of those steps would be missing); setting the
 * schedule again restarts it. */
STRH_API int strh_batch_set_probe(strh_batch* batch, int index, int64_t first_step,
                                  int64_t period, int64_t count, int x0, int x1, int y0, int y1);
/* Advance lattice i by steps[i] steps (all in lock-step from the first). */
STRH_API int strh_batch_run(strh_batch* batch, const int64_t* steps);
/* Copy the samples taken so far (3 doubles each, Fx Fy Fz) of lattice `index`. */
STRH_API int strh_batch_probe_series(strh_batch* batch, int index, double* out,
                                     int64_t capacity, int64_t* taken);
/* Workgroup size of the batch step kernel (128, 256 or 512); results do not
 * depend on it. autotune times each on the batch and keeps the fastest (state
 * restored after). */
STRH_API int strh_batch_set_block_size(strh_batch* batch, int block_size);
STRH_API int strh_batch_autotune(strh_batch* batch, int* block_size);
/* Steps per captured HIP graph (0: plain launches). Default 256. */
STRH_API int strh_batch_set_graph_steps(strh_batch* batch, int steps);
STRH_API int strh_batch_gpu_seconds(strh_batch* batch, double* seconds);
/* Resident launches (the persistent launch of Ladder.swift:54-121's rungs):
 * lattices one cell deep (nz = 1) are cut into tiles held in LDS, one
 * workgroup per CU, and advanced up to 256 steps per launch (64 where a
 * probed lattice's ring of force buffers for 256 steps would exceed 1 GiB), trading
 * only their edge populations between steps; groups of lattices that do not
 * fit on chip together are time-sliced. Results are the same bits as the
 * streaming kernel's. mode 0: never; 1: in every run segment whose active
 * lattices all qualify and fit on chip in one launch; 2: in every segment,
 * time-sliced where needed (STRH_ERR_UNSUPPORTED when some lattice does not
 * qualify); 3 (default): as 1, and where the active lattices qualify but need
 * time-sliced groups, the first steps of the segment time both the streaming
 * and the resident graphs (one untimed and two timed replays of each) and the
 * rest of the segment runs the faster (segments of at least 8 graph lengths;
 * shorter ones stream).
 * Each qualifying lattice holds 152 bytes per cell of exchange words while it
 * is batched. A stalled exchange (which a correct build never produces)
 * returns STRH_ERR_DEVICE after about a second. */
STRH_API int strh_batch_set_resident(strh_batch* batch, int mode);
/* How long a resident launch's workgroup waits for a neighbour's exchange word
 * before the run is abandoned with STRH_ERR_DEVICE (0 < seconds <= 60; default 1).
 * After such a failure the batch's lattices are left in an undefined state. */
STRH_API int strh_batch_set_resident_timeout(strh_batch* batch, double seconds);

/* ---- discrete adjoint (no reference counterpart: the reference has no
 * gradients) ------------------------------------------------------------------
 * Objective over the T = spinup + avg steps from the current state, evaluated
 * on the pre-collision state of each step k >= spinup and averaged over avg:
 *   J = (1/avg) sum_k [ w_p (rho_in - rho_out) / 3 + w_d sum_x alpha(x) J_x(x)
 *                       + w_f sum_x ntF_x(x) ]
 * rho_in / rho_out: mean density of the fluid cells of column x = inlet_x /
 * outlet_x (-1: no column); J_x = (p_x + F_x/2) / (1 + alpha/2) is the
 * momentum the Brinkman drag acts on; ntF_x = -2 sum_i c_ix B (f_opp(i) - f_i)
 * is the x momentum exchange of a partial cell (Kernels.metal:201-227), so
 * the w_f term is the drag on a Noble-Torczynski body. The lattice may use
 * walls of every kind, Guo forcing, TRT/SRT, and porosity or partial cells
 * (not both); LES and the sponge give STRH_ERR_UNSUPPORTED. */
typedef struct strh_objective {
    int32_t inlet_x, outlet_x;
    double w_p, w_d;
    int64_t spinup, avg;
    double w_f;
} strh_objective;

typedef struct strh_vjp_info {
    int64_t checkpoint_every;   /* steps per recorded segment */
    int64_t checkpoints;        /* DDF images kept (both levels) */
    double forward_seconds;     /* first forward pass (objective, checkpoints) */
    double recompute_seconds;   /* segment replays (both levels) */
    double backward_seconds;    /* reverse steps */
    uint64_t workspace_bytes;   /* device memory of the call beyond the lattice */
    int64_t outer_every;        /* steps per outer segment (= checkpoint_every: one level) */
    double reverse_pass_seconds;   /* replays and reverse steps together, wall time on the GPU:
                                      with one level and room for a second recording, segment
                                      s - 1 is replayed on a second stream while the reverse steps
                                      of segment s run, so this is less than their sum */
} strh_vjp_info;

/* Advance T steps and return J (FP64, fixed-order reduction). */
STRH_API int strh_objective_eval(strh_lattice* lat, const strh_objective* obj, double* J);
/* Advance T steps, return J and its gradient with respect to the porosity of
 * every cell (grad_alpha: N doubles, or NULL; zero on non-fluid cells), to the
 * solid fraction of every fluid cell whose solid fraction is non-zero
 * (grad_eps: N doubles, or NULL; zero elsewhere) and to the inlet velocity
 * target uin (grad_uin, or NULL). Partial-cell lattices also record the 19
 * populations of those cells per step; Smagorinsky LES lattices record the 19
 * populations of every cell instead of the moments (the rates depend on the
 * full non-equilibrium stress), and the outlet sponge enters as per-column
 * rates. Porosity does not combine with partial cells or LES
 * (STRH_ERR_UNSUPPORTED). The state after the call
 * is the state after strh_run(lat, T), bit for bit. The forward pass keeps a
 * DDF checkpoint at the start of every segment of K = checkpoint_every steps;
 * the reverse pass replays one segment at a time, recording the four
 * pre-collision moments (rho - 1, p) per cell and step in FP32 (the
 * collision's Jacobian depends on the state only through them), and runs the
 * reverse steps over it. The last segment is recorded in the forward pass, so
 * K >= T replays nothing. K = 0: ceil(sqrt(T)); K = -1: the largest K whose
 * recording and checkpoints fit 80% of the free device memory; when none
 * fits, the two-level plan of strh_vjp_nested with outer and inner segments
 * of ceil(T^(1/3))^2 and ceil(T^(1/3)) steps if that fits (the single-level
 * K needing least memory otherwise). The adjoint state is FP32;
 * every accumulation is FP64 in a fixed order, so the result does not depend
 * on K. Device memory: (ceil(T/K) - 1) DDF images + K * N * 16 bytes (19 * N
 * stored values under LES) +
 * 2 * 19 * N floats + 3 * N doubles + one DDF image. info may be NULL. */
STRH_API int strh_vjp(strh_lattice* lat, const strh_objective* obj, int64_t checkpoint_every,
                      double* J, double* grad_alpha, double* grad_eps, double* grad_uin,
                      strh_vjp_info* info);
/* strh_vjp with two checkpoint levels, for horizons whose single-level
 * recording and checkpoints do not fit device memory (no reference
 * counterpart). The forward pass keeps the DDFs at the start of every outer
 * segment of outer_every steps; the reverse pass replays each outer segment
 * once, keeping the DDFs at the start of every inner segment of inner_every
 * steps, then replays each inner segment again, recording it, and runs its
 * reverse steps. Each step is replayed twice instead of once; device memory
 * holds (ceil(T/outer) - 1) + (ceil(outer/inner) - 1) DDF images and inner
 * recorded steps. inner_every >= outer_every is the single-level plan with
 * K = outer_every. J, the gradients and the final state are the same bits as
 * strh_vjp's for any outer_every, inner_every >= 1. */
STRH_API int strh_vjp_nested(strh_lattice* lat, const strh_objective* obj, int64_t outer_every,
                             int64_t inner_every, double* J, double* grad_alpha, double* grad_eps,
                             double* grad_uin, strh_vjp_info* info);

/* ---- host utilities (no counterpart in the reference, which hashes with
 * CryptoKit): SHA-256 of a host buffer, and a switch that forces the portable
 * block function instead of the x86 SHA extensions (returns 1 when the SHA
 * extensions are in use after the call). */
STRH_API int strh_sha256(const void* data, size_t len, uint8_t out[32]);
STRH_API int strh_sha256_force_portable(int enable);

#ifdef __cplusplus
}
#endif

#endif /* STROUHAL_HIP_H */
