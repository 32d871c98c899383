/* sm2d_io.h — host-side run configuration and file formats of libsm2d_hip.so.
 *
 * These entry points sit above the transport step and replace the reference's host driver
 * pieces: the INI loader (src/include/core/config_loader.hpp:281-416) with its validation
 * (src/include/core/incident_particle_config.hpp:272-391), the runner's derivation of the
 * transport inputs (src/gpu/gpu_transport_runner.cpp:28-213), the dose writers of
 * run_simulation.cpp:31-116 and the raw phase-space CSV of src/cuda/k1k6_pipeline.cu:195-291
 * (and its 6-decimal initial variant, src/cuda/gpu_transport_wrapper.cu:22-120).
 * They run on the CPU only and need no GPU.  Same conventions as sm2d_hip.h: int status,
 * sm2d_last_error() for the message, struct_size first in every struct.
 */
#ifndef SM2D_IO_H
#define SM2D_IO_H

#include <stdint.h>

#include "sm2d_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SM2D_MAX_ENERGY_GROUPS 64
#define SM2D_PATH_MAX 512

/* IncidentParticleConfig (incident_particle_config.hpp:201-244) flattened into plain fields.
 * Enumerations keep the reference's order: particle_type proton=0 electron=1 positron=2
 * alpha=3 carbon_ion=4 custom=5; beam_profile pencil=0 gaussian=1 flat=2 custom=3. */
typedef struct sm2d_run_config {
    uint32_t struct_size;
    /* [particle], [beam] */
    int particle_type;
    float particle_mass_amu, particle_charge_e;
    int beam_profile;
    float W_total;
    /* [energy] */
    float E_mean, E_sigma, E_min, E_max;
    /* [spatial] */
    float x0, z0, sigma_x, sigma_z, width_x, width_z, radius;
    /* [angular] */
    float theta0, phi0, sigma_theta, sigma_phi;
    /* [sampling] */
    int n_samples;
    uint32_t random_seed;
    int use_stratified;
    /* [grid] */
    int Nx, Nz;
    float dx, dz;
    int max_steps;
    /* [transport] */
    int N_theta, N_theta_local, N_E_local;
    int n_energy_groups;
    float energy_groups[SM2D_MAX_ENERGY_GROUPS][3]; /* (E_min, E_max, dE) MeV */
    float E_fine_on, E_fine_off, E_trigger, weight_active_min, E_coarse_max, step_coarse;
    int n_steps_per_cell, fine_batch_max_cells, fine_halo_cells;
    float preflight_vram_margin;
    int validation_mode, fail_fast_on_audit, debug_dumps, max_iterations, log_level;
    /* [output] */
    char output_dir[SM2D_PATH_MAX], dose_2d_file[SM2D_PATH_MAX], pdd_file[SM2D_PATH_MAX];
    char let_file[SM2D_PATH_MAX], format[32];
    int normalize_dose, save_2d, save_pdd, save_lat_profiles;
} sm2d_run_config;

/* The defaults of a default-constructed IncidentParticleConfig. */
int sm2d_config_defaults(sm2d_run_config* cfg);
/* load_incident_particle_config (config_loader.hpp:281-416): the same sections, keys, key
 * aliases, legacy E_trigger handling and value parsing (a value that does not parse keeps
 * the default).  SM2D_ERR_ARG when the file cannot be opened. */
int sm2d_config_load(const char* path, sm2d_run_config* cfg);
/* IncidentParticleConfig::validate (incident_particle_config.hpp:272-391): SM2D_ERR_ARG with
 * the reference's message for the first violated rule. */
int sm2d_config_validate(const sm2d_run_config* cfg);
/* What GPUTransportRunner::run derives before the transport (gpu_transport_runner.cpp:29-98,
 * 172-173): the spatial/phase-space grid (theta over +-0.35 rad, x_min = -Nx*dx/2, z_min = 0,
 * the library layout whose (N_theta_local, N_E_local) the config names), the transport
 * parameters (max_iterations 0 -> max_steps, the legacy E_trigger lift, validation mode ->
 * fail-fast) and the spot (its sigma_x is also the transport's sigma_x_initial,
 * gpu_transport_wrapper.cu:273).  grid->N_E is the piecewise energy grid's bin count. */
int sm2d_config_runtime(const sm2d_run_config* cfg, sm2d_grid* grid, sm2d_transport* tp,
                        sm2d_spot* spot);

/* save_dose_2d (run_simulation.cpp:31-76): edep[Nz*Nx] row-major. */
int sm2d_write_dose_2d(const char* path, const double* edep, int Nx, int Nz, float dx, float dz,
                       int normalize);
/* save_pdd (run_simulation.cpp:81-116) with get_depth_dose / find_bragg_peak_z
 * (src/validation/deterministic_beam.cpp:207-240). */
int sm2d_write_pdd(const char* path, const double* edep, int Nx, int Nz, float dz, int normalize);

/* Raw phase-space CSV (k1k6_pipeline.cu:195-291): rows (cell, slot, bid, lidx, w) in
 * (cell, slot, lidx) order, as sm2d_psi_dump exports them; rows with w <= 1e-12 are skipped,
 * as the reference does.  decimals = 8 (after-K4 dumps) or 6 (the initial dump).  The
 * decoded columns use the grid's layout and the given edges. */
int sm2d_write_psi_csv(const char* path, int iteration, int decimals, const sm2d_grid* grid,
                       const float* theta_edges, const float* E_edges, int n_rows,
                       const int32_t* cell, const int32_t* slot, const uint32_t* bid,
                       const int32_t* lidx, const float* w);
/* Per-cell summary CSV of the same rows (dump_nonzero_cells_to_csv, k1k6_pipeline.cu:29-190):
 * one line per cell whose FP32 weight sum exceeds 1e-12, in ascending cell order, with the
 * count of bins above 1e-12, the weighted mean and the range of their bin-centre energies and
 * angles (theta centres from the uniform bin width), 6 decimals.  Rows whose global bin lies
 * off the grid are rejected (SM2D_ERR_ARG). */
int sm2d_write_cells_csv(const char* path, const sm2d_grid* grid, const float* theta_edges,
                         const float* E_edges, int n_rows, const int32_t* cell, const int32_t* slot,
                         const uint32_t* bid, const int32_t* lidx, const float* w);
/* Reads the (cell, slot, bid, lidx, weight) columns of such a file.  *n_rows receives the row
 * count; rows beyond capacity are counted but not stored (call with capacity 0 to size). */
int sm2d_read_psi_csv(const char* path, int capacity, int32_t* cell, int32_t* slot,
                      uint32_t* bid, int32_t* lidx, float* w, int* n_rows);

#ifdef __cplusplus
}
#endif

#endif
