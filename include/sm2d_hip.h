/* sm2d_hip.h — C ABI of the MI355X transport-step library (libsm2d_hip.so).
 *
 * The library replaces the K1–K6 transport step of SM_2D behind its boundary
 * `sm_2d::run_k1k6_pipeline_transport` (src/include/cuda/gpu_transport_wrapper.hpp:78-90,
 * implementation src/cuda/gpu_transport_wrapper.cu:207-641) and the LUT upload
 * `init_device_lut` (gpu_transport_wrapper.hpp:42, .cu:199-205).  Signatures use plain
 * pointers, sizes and versioned POD structs; no C++ or torch types cross the ABI.
 *
 * Conventions
 *  - Every entry point returns an int status (SM2D_OK = 0, negative on error) and never
 *    throws; sm2d_last_error() returns a thread-local message for the last failure.
 *  - Every struct starts with `struct_size`; callers set it to sizeof(the struct).
 *  - A context owns one HIP stream and all device memory it allocates; it is reusable across
 *    spots (no per-spot GB-scale allocation) and re-entrant (no device globals).
 *  - Pointers named d_* are device pointers; all others are host pointers.
 */
#ifndef SM2D_HIP_H
#define SM2D_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SM2D_ABI_VERSION 1

#define SM2D_OK 0
#define SM2D_ERR_ARG -1        /* invalid argument / struct_size mismatch */
#define SM2D_ERR_HIP -2        /* HIP runtime error */
#define SM2D_ERR_CAPACITY -3   /* a bounded device pool overflowed */
#define SM2D_ERR_AUDIT -4      /* K5 failed and fail_fast_on_audit was set */
#define SM2D_ERR_UNSUPPORTED -5 /* input needs a feature the GPU path does not implement */

#define SM2D_KB 32                 /* slots per cell, DEVICE_Kb (src/cuda/device/device_psic.cuh:30) */
#define SM2D_EMPTY 0xFFFFFFFFu     /* EMPTY_BLOCK_ID (src/include/core/block_encoding.hpp:30) */

/* Spatial grid, phase-space grid and local-bin layout of one transport problem.
 * The reference fixes the layout at compile time (src/include/core/local_bins.hpp:19-23:
 * 4 theta x 2 E x 8 x_sub x 4 z_sub = 256); here it is a runtime choice among the
 * layouts the library is built for: (4,2,8,4) and (8,4,4,1). */
typedef struct sm2d_grid {
    uint32_t struct_size;
    int Nx, Nz;                 /* cells in x (lateral) and z (depth) */
    float dx, dz;               /* cell size [mm] */
    float x_min, z_min;         /* grid origin [mm]; the runner uses -Nx*dx/2 and 0 */
    int N_theta;                /* global angular bins, uniform over [theta_min, theta_max] */
    float theta_min, theta_max; /* [rad]; the reference runner hard-codes +-0.35 */
    int N_E;                    /* global energy bins; edges are supplied separately */
    int n_theta_local, n_e_local, n_x_sub, n_z_sub;  /* local-bin layout of one block */
} sm2d_grid;

/* TransportConfig fields the step reads (src/include/core/incident_particle_config.hpp:174-200). */
typedef struct sm2d_transport {
    uint32_t struct_size;
    float E_fine_on, E_fine_off;  /* K1 thresholds [MeV] (defaults 10, 11) */
    float weight_active_min;      /* K1 weight threshold (1e-12) */
    float step_coarse;            /* K2 step [mm] (5) */
    int max_iterations;
    int fail_fast_on_audit;
    int step_rule;                /* 0: device rule of device_lut.cuh:168-214 (default);
                                     1: 2 % of range rule of step_control.hpp:18-66 */
    int enable_straggling;        /* K3 Vavilov straggling (pipeline default 1) */
    int enable_nuclear;           /* K3 nuclear attenuation (pipeline default 1) */
    float sigma_x_initial;        /* K1K6PipelineConfig::sigma_x_initial [mm] used by the coarse
                                     moments; sm2d_run_spot replaces it with the spot's sigma_x */
} sm2d_transport;

/* One pencil-beam spot: the beam arguments of run_k1k6_pipeline_transport. */
typedef struct sm2d_spot {
    uint32_t struct_size;
    float x0, z0, theta0, E0, W_total;
    float sigma_x, sigma_theta, sigma_E;
    int n_samples;
    uint32_t random_seed;
} sm2d_spot;

/* Energy and weight bookkeeping of one spot (whole run).  Mirrors the quantities of the
 * reference's "Energy Conservation Report" (src/cuda/k1k6_pipeline.cu:2101-2183) plus the
 * per-source-cell closure and the work counter this library adds. */
typedef struct sm2d_audit {
    uint32_t struct_size;
    int iterations;        /* iterations executed (the terminating empty one included) */
    int terminated;        /* 1 when both lists were empty before max_iterations */
    int W_fail_iters, E_fail_iters;   /* iterations whose K5 W / E check failed */
    double max_W_error, max_E_error;  /* largest K5 relative errors over iterations */
    double W_source_injected, W_source_out_of_grid, W_source_slot_drop;
    double E_source_injected, E_source_out_of_grid, E_source_slot_drop, E_representation_loss;
    double E_dep, W_cutoff, E_cutoff, W_nuclear, E_nuclear, W_boundary, E_boundary;
    double W_transport_drop, E_transport_drop, E_audit_residual;
    double W_leftover, E_leftover;    /* phase space left when max_iterations was reached */
    double max_cell_W_closure;        /* max over (iteration, processed cell) of
                                         |w_in - w_out_total| / w_in, FP64 */
    double max_cell_E_closure;        /* same for energy (K3 cells only; K2 has no closure) */
    uint64_t cell_bin_updates;        /* sum over iterations and processed cells of
                                         occupied slots x local bins */
    uint64_t processed_cells, processed_components;
    uint64_t cross_cell_emissions;    /* +z lateral-spread emissions into another listed cell's
                                         bucket (device_bucket.cuh:952-983) */
    int surrogate_active;             /* 1 when the fine transport used the learned physics */
    int surrogate_fallback;           /* 1 when a K5 drift switched it back to the plain physics */
    uint64_t fine_cell_bin_updates;   /* the part of cell_bin_updates done by the fine transport */
    uint64_t surrogate_substeps;      /* fine-transport substeps whose straggling, scattering and
                                         attenuation came from the learned-physics table */
} sm2d_audit;

/* Per-iteration K5 report: AuditReport of src/cuda/kernels/k5_audit.cuh:5-28, with the weight
 * sums kept in FP64 and reduced in a fixed order. */
typedef struct sm2d_k5_report {
    uint32_t struct_size;
    int W_pass, E_pass;
    double W_error, E_error;
    double W_in, W_out, W_cutoff, W_nuclear, W_boundary, W_source_out_of_grid,
           W_source_slot_drop, W_transport_drop;
    double E_in, E_out, E_dep, E_cutoff, E_nuclear, E_boundary, E_source_out_of_grid,
           E_source_slot_drop, E_transport_drop;
} sm2d_k5_report;

/* Kernel timing of a context, from HIP events recorded on the context's stream around each
 * stage (no counterpart in the reference, which has no kernel timing).  Times in ms. */
typedef struct sm2d_timing {
    uint32_t struct_size;
    int enabled;
    uint64_t iterations, transport_launches, gather_launches;
    double mask_ms, transport_ms, gather_ms, audit_ms;
    uint64_t cell_bin_updates;   /* updates done by the timed transport launches */
    double coarse_transport_ms, fine_transport_ms;  /* transport_ms split by phase */
    uint64_t coarse_updates, fine_updates;          /* cell_bin_updates split by phase */
    double coarse_gather_ms, fine_gather_ms;        /* gather_ms split by phase */
    uint64_t coarse_gather_values, fine_gather_values;  /* non-zero bucket values the timed
                                                          gather launches moved into psi_out */
} sm2d_timing;

typedef struct sm2d_ctx sm2d_ctx;

const char* sm2d_last_error(void);
int sm2d_abi_version(void);
/* Name of HIP device `device` (get_gpu_name_internal, src/include/cuda/gpu_transport_wrapper.hpp:95). */
int sm2d_device_name(int device, char* buf, int capacity);

/* --- host-side table builders (run on the CPU, identical to the reference's host code) --- */

/* EnergyGrid(E_min, E_max, N_E) log-spaced constructor (src/core/grids.cpp:6-23).
 * edges[N_E+1], rep[N_E]. */
int sm2d_energy_grid_log(float E_min, float E_max, int N_E, float* edges, float* rep);
/* EnergyGrid::CreatePiecewise (src/core/grids.cpp:33-96).  groups[3*n_groups] = (start, end,
 * width) triples.  Returns N_E in *n_bins; with edges/rep NULL only counts. */
int sm2d_energy_grid_piecewise(const float* groups, int n_groups, float* edges, float* rep,
                               int* n_bins);
/* AngularGrid(theta_min, theta_max, N_theta) edges (src/core/grids.cpp:121-135). */
int sm2d_angular_grid(float theta_min, float theta_max, int N_theta, float* edges);
/* GenerateRLUT(const EnergyGrid&) (src/lut/r_lut.cpp:69-129) from the NIST PSTAR water table.
 * Outputs R, S, log_E, log_R, log_S, each [N_E]. */
int sm2d_build_rlut(const float* rep, int N_E, float* R, float* S, float* log_E, float* log_R,
                    float* log_S);
/* Fermi–Eyges precompute of init_pipeline_state (src/cuda/k1k6_pipeline.cu:1061-1212):
 * A[Nz], B[Nz], sigma_total[Nz]. */
int sm2d_fe_precompute(float E0, float sigma_x0, int Nz, float dz, float* A, float* B,
                       float* sigma_total);
/* MCS post-convolution of edep rows (src/cuda/gpu_transport_wrapper.cu:551-616), in place. */
int sm2d_post_convolve(double* edep, int Nx, int Nz, float dx, const float* sigma_total,
                       float sigma_x0);

/* --- context ---------------------------------------------------------------------------- */

/* Creates a context on HIP device `device` for problems of the given grid.
 * Replaces device_psic_init x2 + init_pipeline_state (gpu_transport_wrapper.cu:282-312);
 * the context persists across spots. */
int sm2d_ctx_create(int device, const sm2d_grid* grid, sm2d_ctx** out);
int sm2d_ctx_destroy(sm2d_ctx* ctx);

/* Uploads the phase-space edges and the R-LUT.  Replaces init_device_lut
 * (gpu_transport_wrapper.cu:199-205 -> DeviceLUTWrapper::init, k3_finetransport.cuh:115-149)
 * and the edge copies of init_pipeline_state (k1k6_pipeline.cu:1012-1048). */
int sm2d_lut_upload(sm2d_ctx* ctx, const float* theta_edges, const float* E_edges,
                    const float* R, const float* S, const float* log_E, const float* log_R,
                    const float* log_S);

/* Uploads per-depth-row Fermi–Eyges moments A[Nz], B[Nz] used by the fine transport
 * (k1k6_pipeline.cu:1195-1206).  Passing NULL for both sets them to zero. */
int sm2d_fe_upload(sm2d_ctx* ctx, const float* A, const float* B);

/* Per-row material (density rho [g/cm^3] and radiation-length ratio to water) for slab
 * phantoms.  NULL restores water (rho = 1, ratio = 1), which reproduces the reference. */
int sm2d_material_upload(sm2d_ctx* ctx, const float* rho, const float* x0_ratio);

/* Full spot: injection, K1–K6 loop, EdepC extraction.  This is the body of
 * run_k1k6_pipeline_transport (gpu_transport_wrapper.cu:207-641) without the host-side MCS
 * post-convolution, which the caller applies with sm2d_post_convolve.
 * edep_out[Nz*Nx] (row-major, FP64, MeV per cell); audit may be NULL. */
int sm2d_run_spot(sm2d_ctx* ctx, const sm2d_spot* spot, const sm2d_transport* tp,
                  double* edep_out, sm2d_audit* audit);

/* --- device memory planning ------------------------------------------------------------- */

/* Device memory of a context, replacing the dense estimate and check of run_memory_preflight
 * (src/perf/memory_preflight.cpp:43-160, called at src/gpu/gpu_transport_runner.cpp:111-139).
 * The library's phase space is a sparse store with bounded capacities, so its plan lists
 * those capacities and the exact bytes sm2d_ctx_create(_batch) allocates for them (the
 * enumeration the allocation itself uses); dense_bytes is what the reference's dense store
 * would need for the same grid.  Tables and networks of the learned physics are extra. */
typedef struct sm2d_memory_plan {
    uint32_t struct_size;
    uint32_t listed_cells;    /* listed cells per phase and iteration */
    uint32_t occupied_cells;  /* occupied cells (and destinations) per iteration */
    uint32_t payload_slots;   /* payload slots of each ψ buffer and of the buckets */
    uint64_t psi_bytes;       /* both ψ buffers: slot tables, slot counts, payload pools */
    uint64_t bucket_bytes;    /* outflow buckets of the listed cells */
    uint64_t cell_bytes;      /* per-cell masks, tags and tallies */
    uint64_t other_bytes;     /* lists, audit rows, counters, grids and tables */
    uint64_t total_bytes;
    uint64_t dense_bytes;     /* estimate_dense_k1k6_memory of the reference for this grid */
    uint64_t free_bytes;      /* sm2d_memory_preflight: free device memory */
    uint64_t usable_bytes;    /* free_bytes x margin */
    int ok;                   /* total_bytes <= usable_bytes (1 from sm2d_memory_estimate) */
} sm2d_memory_plan;

/* The plan for a grid and spot count; host only (no device is touched). */
int sm2d_memory_estimate(const sm2d_grid* grid, int n_spots, sm2d_memory_plan* out);
/* The plan checked against the device's free memory times margin (0 < margin <= 1, the
 * reference's preflight_vram_margin); out->ok tells whether sm2d_ctx_create fits. */
int sm2d_memory_preflight(int device, const sm2d_grid* grid, int n_spots, float margin,
                          sm2d_memory_plan* out);
/* Device bytes a context's creation allocated (equals its plan's total_bytes). */
int sm2d_ctx_memory(sm2d_ctx* ctx, uint64_t* bytes);

/* --- batched spots ------------------------------------------------------------------------ */

/* Creates a context holding `n_spots` independent copies of the grid (spot s owns cells
 * [s*Nx*Nz, (s+1)*Nx*Nz)).  Spots of one batch are transported together: one launch per
 * stage covers every spot's listed cells.  No counterpart in the reference, whose runner
 * calls run_k1k6_pipeline_transport once per spot (src/gpu/gpu_transport_runner.cpp:28-213);
 * each spot's result is bit-identical to sm2d_run_spot on the same spot. */
int sm2d_ctx_create_batch(int device, const sm2d_grid* grid, int n_spots, sm2d_ctx** out);

/* Runs spots[0..n) (n <= the context's spot count) to completion in one batch.
 * edep_sum[Nz*Nx] (FP64, may be NULL) receives the sum over spots in spot order;
 * edep_per_spot[n*Nz*Nx] (FP32, may be NULL) each spot's EdepC; audit (may be NULL) the
 * batch totals (iterations = the batch's iteration count). */
int sm2d_run_batch(sm2d_ctx* ctx, const sm2d_spot* spots, int n, const sm2d_transport* tp,
                   double* edep_sum, float* edep_per_spot, sm2d_audit* audit);
/* Spot s's FP64 EdepC[Nz*Nx] of the last sm2d_run_batch (0 <= s < its n), before any FP32
 * rounding: the values sm2d_run_spot returns for that spot alone.  No counterpart in the
 * reference (one process per spot there, batch_run.py:200-238). */
int sm2d_batch_spot_edep(sm2d_ctx* ctx, int s, double* edep_out);

/* --- stage-level entry points (used by the parity tests) -------------------------------- */

/* Clears both phase-space buffers and all tallies. */
int sm2d_reset(sm2d_ctx* ctx);
/* Source injection only (inject_gaussian_source_kernel / inject_source_kernel,
 * k1k6_pipeline.cu:300-551) into psi_in; accounting goes to *audit (may be NULL). */
int sm2d_inject(sm2d_ctx* ctx, const sm2d_spot* spot, sm2d_audit* audit);
/* Loads psi_in from rows (cell, slot, bid, lidx, weight); rows of one (cell, slot) must
 * carry the same bid, slots of one cell must form a prefix 0..n-1, and every row's global
 * (theta, E) bin must lie on the grid (SM2D_ERR_ARG otherwise: the local bins of a partial
 * last block past N_theta or N_E have no edges). */
int sm2d_psi_load(sm2d_ctx* ctx, int n_rows, const int32_t* cell, const int32_t* slot,
                  const uint32_t* bid, const int32_t* lidx, const float* w);
/* Exports psi_in: bids[Nx*Nz*32] (may be NULL) and the rows with weight != 0 in
 * (cell, slot, lidx) order; *n_rows receives the row count (rows beyond capacity are
 * counted but not written). */
int sm2d_psi_dump(sm2d_ctx* ctx, uint32_t* bids, int capacity, int32_t* cell, int32_t* slot,
                  uint32_t* bid, int32_t* lidx, float* w, int* n_rows);
/* K1 ActiveMask + K2 CompactActive on psi_in (k1_activemask.cu:6-48,
 * k1k6_pipeline.cu:589-665).  mask_prev[N] (may be NULL = all zero), mask_out[N],
 * active/coarse lists in ascending cell order with their counts. */
int sm2d_k1k2_mask_compact(sm2d_ctx* ctx, const sm2d_transport* tp, const uint8_t* mask_prev,
                           uint8_t* mask_out, int32_t* active, int* n_active, int32_t* coarse,
                           int* n_coarse);
/* Runs `n_iter` full iterations (K1, compaction, coarse K3+K4, fine K3+K4, K5, K6) from the
 * current psi_in; stops early when both lists are empty.  tallies_out (may be NULL) receives
 * the cumulative per-cell tallies [7][N]: EdepC, W_cutoff, E_cutoff, W_nuclear, E_nuclear,
 * W_boundary, E_boundary (weights widened to FP64).  k5_last receives the last iteration's
 * K5 report. */
int sm2d_run_iterations(sm2d_ctx* ctx, const sm2d_transport* tp, int n_iter,
                        int include_source_terms, double* tallies_out, sm2d_k5_report* k5_last,
                        sm2d_audit* audit);
/* K5 on caller data (the k5_audit known-answer test, tests/gpu/test_k5_drop_channels_gpu.cu):
 * audits psi_in against psi_out with per-cell tally deltas d_tallies[7][N] (host) and the
 * drop/source scalars. */
int sm2d_k5_audit(sm2d_ctx* ctx, const double* tallies_delta, double W_src_out, double W_src_slot,
                  double E_src_out, double E_src_slot, double W_drop, double E_drop,
                  int include_source_terms, sm2d_k5_report* out);
/* K1 ActiveMask + K2 CompactActive on caller-provided device buffers, allocating nothing: no
 * hipMalloc, hipFree or allocating library call, for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).  No counterpart in
 * the reference; it takes the arguments of the reference's stage wrapper run_k1_active_mask
 * (k1k6_pipeline.cu:1261-1294) on its dense phase-space layout (DevicePsiC, device_psic.cuh:39-55):
 * d_block_ids[Nx*Nz*32] (EMPTY = 0xFFFFFFFF), d_values[Nx*Nz*32*local_bins] (local_bins 128 or
 * 256), d_mask_prev[Nx*Nz] (may be NULL), block thresholds b_E_fine_on / b_E_fine_off
 * (compute_b_E_threshold, k1k6_pipeline.cuh:419-422) and weight_active_min.  Outputs:
 * d_mask[Nx*Nz] (0/1), the active and coarse cell lists in ascending order (d_active, d_coarse:
 * Nx*Nz entries each) and their lengths d_counts[2].  Work is queued on `stream` (a hipStream_t;
 * NULL = the default stream) and the call returns without waiting.  The scratch is a fixed
 * device array of the library (4.3 MB per device, held until the process ends): grids above
 * 4 194 304 cells are rejected (SM2D_ERR_ARG).  Calls that overlap on one device race on it.
 * Results equal those of sm2d_k1k2_mask_compact on the same phase space. */
int sm2d_k1k2_mask_compact_noalloc(int Nx, int Nz, int local_bins, const uint32_t* d_block_ids,
                                   const float* d_values, const uint8_t* d_mask_prev, uint8_t* d_mask,
                                   int b_E_fine_on, int b_E_fine_off, float weight_active_min,
                                   uint32_t* d_active, uint32_t* d_coarse, uint32_t* d_counts,
                                   void* stream);

/* Copies psi_in <-> psi_out (load the second buffer for sm2d_k5_audit). */
int sm2d_swap(sm2d_ctx* ctx);

/* Transport launch shape: waves per listed cell (1, 2 or 4), or 0 for the built-in tuner, which
 * times the three shapes per phase and launch size class on the first chunk of iterations of
 * that class and then uses the fastest; lists of 2048 cells or more take one wave per cell
 * untimed (no counterpart in the reference; every shape gives the same results).  The
 * environment variable SM2D_TRANSPORT_WAVES sets the initial value. */
int sm2d_set_transport_waves(sm2d_ctx* ctx, int waves);
/* K4 gather launch shape: LDS payload rows per destination for its staged form (16, 8 or 4 KiB;
 * the staged form serves destinations whose own slot table would overflow; fewer rows let more
 * wavefronts share a compute unit, a destination needing more adds into global payloads), or 0
 * for the tuner, which times the three with the transport shapes (replaces the batch sizing role
 * of the reference's FineBatchPlanner, src/perf/fine_batch_planner.cpp:7-36; every shape gives
 * the same results). */
int sm2d_set_gather_rows(sm2d_ctx* ctx, int kib);
/* JSON summary of the tuner's measurements (per tuned launch and size class: mean launch time per
 * shape and the chosen shape), truncated to `capacity` bytes. */
int sm2d_tuner_report(sm2d_ctx* ctx, char* buf, int capacity);
/* Chunks of iterations are captured as HIP graphs and replayed when nothing they depend on has
 * changed (on by default; SM2D_GRAPHS=0 in the environment turns it off at context creation).
 * Off (0): every kernel is launched on the stream.  Chunks with stage timing or tuner
 * measurements on are always launched on the stream.  No counterpart in the reference (its
 * pipeline launches every kernel from the host, k1k6_pipeline.cu:1663-2011). */
int sm2d_set_graphs(sm2d_ctx* ctx, int on);

/* Turns stage timing on (1) or off (0) and zeroes the counters; sm2d_get_timing reads them. */
int sm2d_set_profiling(sm2d_ctx* ctx, int on);
int sm2d_get_timing(sm2d_ctx* ctx, sm2d_timing* out);

/* --- learned physics of the fine transport ------------------------------------------------
 * No counterpart in the reference.  A 4-128-128-128-33 MLP (GELU; BF16 weights, FP32
 * accumulation on the matrix cores) maps (ln E, log10 dR, rho, ln X0 ratio) to the per-substep
 * straggling width, scattering width and nuclear attenuation exponent that K3 evaluates
 * with device_energy_straggling_sigma / device_highland_sigma / device_apply_nuclear_attenuation
 * (src/cuda/device/device_physics.cuh:50-234).  Outputs: 32 logits of a joint distribution over
 * 4 x 8 log-spaced width classes (the widths are its log-moments) and ln(a + 1e-9).  It is
 * trained on the GPU on samples of the plain physics; the transport reads a table of it. */
typedef struct sm2d_mlp sm2d_mlp;

typedef struct sm2d_mlp_train_cfg {
    uint32_t struct_size;
    int n_samples;   /* training samples drawn from the plain physics (e.g. 1 << 22) */
    int batch;       /* mini-batch, a multiple of 64 */
    int steps;       /* Adam steps */
    float lr;        /* learning rate (stepped down to 0.3x and 0.1x in the last half) */
    float w_att;     /* weight of the attenuation term of the loss */
    float w_mom;     /* weight of the decoded-width (log-moment) term of the loss */
    uint32_t seed;   /* sample seed */
} sm2d_mlp_train_cfg;

int sm2d_mlp_create(int device, uint32_t seed, sm2d_mlp** out);
int sm2d_mlp_destroy(sm2d_mlp* mlp);
int sm2d_mlp_param_count(void);
/* FP32 master parameters (W0[128][4], b0, W1[128][128], b1, W2[128][128], b2, W3[33][128], b3)
 * and the width dictionaries (ln sigma_E lo/hi, ln sigma_theta lo/hi; may be NULL). */
int sm2d_mlp_get_params(sm2d_mlp* mlp, float* params, float* dict4);
int sm2d_mlp_set_params(sm2d_mlp* mlp, const float* params, const float* dict4);
/* Raw outputs z[B][33] of normalised inputs x[B][4]. */
int sm2d_mlp_forward(sm2d_mlp* mlp, const float* x, int B, float* z);
/* Decoded outputs out3[B][3] = (ln sigma_E, ln sigma_theta, a) of normalised inputs x[B][4],
 * evaluated by the same in-kernel inference the fine transport runs in mode 2 of
 * sm2d_surrogate_enable (64 samples per wavefront on the matrix cores). */
int sm2d_mlp_infer(sm2d_mlp* mlp, const float* x, int B, float* out3);
/* Mean loss and its gradient (parameter layout above) of B samples (a multiple of 64):
 * labels[B][3] = (ln sigma_E, ln sigma_theta, ln(a + 1e-9)); loss = soft-target cross-entropy
 * + w_mom * (squared errors of the decoded ln widths) + w_att * (attenuation output error)^2. */
int sm2d_mlp_grad(sm2d_mlp* mlp, const float* x, const float* labels, int B, float w_att,
                  float w_mom, float* loss, float* grad);
/* Trains on samples of the plain physics drawn with the tables of ctx; loss_hist[n_hist]
 * receives the mean loss of consecutive blocks of steps. */
int sm2d_mlp_train(sm2d_mlp* mlp, sm2d_ctx* ctx, const sm2d_mlp_train_cfg* cfg, float* loss_hist,
                   int n_hist);
/* Held-out accuracy on n fresh samples: RMS errors of ln sigma_E, ln sigma_theta and
 * ln(a + 1e-9) (a: attenuation exponent), then the largest errors of the same (err6[6]). */
int sm2d_mlp_evaluate(sm2d_mlp* mlp, sm2d_ctx* ctx, int n, uint32_t seed, float* err6);
/* Builds the context's surrogate table (the network at every material row x energy bin x
 * n_nodes log-spaced steps in [1e-4, 10] mm), validates it against the plain physics at its
 * nodes (sm2d_surrogate_check with the context tolerance, 0.25 by default) and turns the
 * learned physics on.  Fine-transport substeps outside the training range or at rejected nodes
 * use the plain physics; a failed K5 weight check or a non-finite K5 energy sum switches the
 * rest of the run back to it (sm2d_audit.surrogate_fallback). */
int sm2d_surrogate_attach(sm2d_ctx* ctx, sm2d_mlp* mlp, int n_nodes);
/* Re-validates the attached table: nodes whose learned ln sigma_E or ln sigma_theta differ from
 * the plain physics by more than tol, or whose attenuation exponent a differs by more than
 * tol x (a + 1e-4 / mm x dR), are sent back to the plain physics.  stats4 (may be null) = (nodes checked,
 * nodes rejected, largest deviation, RMS deviation); tol <= 0 only measures. */
int sm2d_surrogate_check(sm2d_ctx* ctx, float tol, float* stats4);
/* Statistics of the last validation (attach or check). */
int sm2d_surrogate_stats(sm2d_ctx* ctx, float* stats4);
/* Learned-physics mode of the fine transport: 0 the plain physics, 1 the attached table (the
 * default after sm2d_surrogate_attach), 2 the attached network evaluated inside the transport
 * for every substep (BF16 operands on the matrix cores, 64 components per wavefront; the network
 * must outlive that use).  Mode 2 applies the same range fallback and K5 drift monitor; its
 * results depend on the matrix cores' FP32 accumulation and are not compared bit for bit. */
int sm2d_surrogate_enable(sm2d_ctx* ctx, int on);
/* The attached table [n_mat][N_E][n_nodes][3] (ln sigma_E, ln sigma_theta, a; NaN outside the
 * training range) and the material index of every depth row (-1: plain physics). */
int sm2d_surrogate_export(sm2d_ctx* ctx, float* table, int capacity, int* n_mat, int* n_nodes,
                          int32_t* rowmat);
/* Installs a table in the layout above (one exported and edited, or tabulated by other means)
 * and turns the learned physics on; rowmat[Nz] in [-1, n_mat). */
int sm2d_surrogate_set_table(sm2d_ctx* ctx, const float* table, int n_mat, int n_nodes,
                             const int32_t* rowmat);

/* Stage clocks of the transport kernel (diagnostic build only, python -m sm_2d_amd.build
 * --stage-clocks; the normal build returns SM2D_ERR_UNSUPPORTED): out32[k] receives the
 * microseconds the transport workgroups spent in stage k, summed over cells (0 tags and
 * destination marks, 1 own psi_out staging, 2 count and queue, 3 physics, 4 wait for other
 * waves, 5 claims, 6 ordered adds, 7 serial apply, 8 end-of-cell barrier, 9 table flush,
 * 10 tallies, 11 accumulator batch sums), out32[15] the cells, and out32[16 + k] the parts of
 * the physics stage (0 set-up and direction, 1 step, energy loss and widths, 2 angular draw and
 * exit face, 3 spread weights, 4 emission bins, 5 lane preparation, 6 batch sums); reset = 1
 * zeroes the sums afterwards.  No counterpart in the reference. */
int sm2d_stage_clocks(int device, double* out32, int reset);

/* --- self-test of the shared elementary functions on the device -------------------------- */
/* Evaluates f(x[i]) for i < n on the GPU: which = 0 exp, 1 log, 2 sin, 3 cos, 4 erf, 5 sin and
 * 6 cos through the shared-reduction sincos. */
int sm2d_math_eval(int which, const float* x, float* y, int n);

#ifdef __cplusplus
}
#endif
#endif /* SM2D_HIP_H */
