// sm2d_oracle.h — CPU oracle of SM_2D's K1–K6 transport step.
//
// TEST INFRASTRUCTURE ONLY.  This is a plain C++ restatement of the reference algorithm, used
// by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
// Nothing in the product library (sm_2d_amd/) may link, load or call it.
//
// Parity status of each part is recorded in the header of sm2d_oracle.cpp.
#pragma once
#include <stdint.h>
#include "../include/sm2d_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

const char* orc_last_error(void);

// Host table builders, restated from src/core/grids.cpp and src/lut/r_lut.cpp.
int orc_energy_grid_log(float E_min, float E_max, int N_E, float* edges, float* rep);
int orc_energy_grid_piecewise(const float* groups, int n_groups, float* edges, float* rep,
                              int* n_bins);
int orc_angular_grid(float theta_min, float theta_max, int N_theta, float* edges);
int orc_build_rlut(const float* rep, int N_E, float* R, float* S, float* log_E, float* log_R,
                   float* log_S);
int orc_fe_precompute(float E0, float sigma_x0, int Nz, float dz, float* A, float* B,
                      float* sigma_total);
int orc_post_convolve(double* edep, int Nx, int Nz, float dx, const float* sigma_total,
                      float sigma_x0);

// Device-rule lookups and physics (one call per value) for unit tests.
// which: 0 lookup_R, 1 lookup_S, 2 E_inverse(R), 3 max_step (device rule), 4 max_step (2 % rule),
//        5 energy_after_step(E, ds=y), 6 highland(E, ds=y), 7 straggling sigma(E, ds=y),
//        8 nuclear cross section(E), 9 find_bin(E)
int orc_eval(orc_ctx* c, int which, const float* x, const float* y, float* out, int n);
// Shared elementary functions (same source as the GPU): 0 exp 1 log 2 sin 3 cos 4 erf.
int orc_math_eval(int which, const float* x, float* y, int n);

int orc_ctx_create(const sm2d_grid* g, orc_ctx** out);
int orc_ctx_destroy(orc_ctx* c);
int orc_set_threads(orc_ctx* c, int n_threads);
int orc_lut_upload(orc_ctx* c, const float* theta_edges, const float* E_edges, const float* R,
                   const float* S, const float* log_E, const float* log_R, const float* log_S);
int orc_fe_upload(orc_ctx* c, const float* A, const float* B);
int orc_material_upload(orc_ctx* c, const float* rho, const float* x0_ratio);
int orc_surrogate_set(orc_ctx* c, const float* table, int n_mat, int n_nodes, const int32_t* rowmat, int on);
int orc_reset(orc_ctx* c);
int orc_inject(orc_ctx* c, const sm2d_spot* spot, sm2d_audit* audit);
int orc_psi_load(orc_ctx* c, int n_rows, const int32_t* cell, const int32_t* slot,
                 const uint32_t* bid, const int32_t* lidx, const float* w);
int orc_psi_dump(orc_ctx* c, uint32_t* bids, int capacity, int32_t* cell, int32_t* slot,
                 uint32_t* bid, int32_t* lidx, float* w, int* n_rows);
int orc_k1k2_mask_compact(orc_ctx* c, const sm2d_transport* tp, const uint8_t* mask_prev,
                          uint8_t* mask_out, int32_t* active, int* n_active, int32_t* coarse,
                          int* n_coarse);
int orc_run_iterations(orc_ctx* c, const sm2d_transport* tp, int n_iter,
                       int include_source_terms, double* tallies_out, sm2d_k5_report* k5_last,
                       sm2d_audit* audit);
int orc_k5_audit(orc_ctx* c, const double* tallies_delta, double W_src_out, double W_src_slot,
                 double E_src_out, double E_src_slot, double W_drop, double E_drop,
                 int include_source_terms, sm2d_k5_report* out);
int orc_swap(orc_ctx* c);
int orc_run_spot(orc_ctx* c, const sm2d_spot* spot, const sm2d_transport* tp, double* edep_out,
                 sm2d_audit* audit);

#ifdef __cplusplus
}
#endif
