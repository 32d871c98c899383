// ref_shim.cpp — C entry points over the reference's own host code, compiled from the SM_2D
// checkout (see Makefile target `ref`) into oracle/_ref/libsm2d_ref.so.  TEST INFRASTRUCTURE:
// used only to pin the oracle and the product's host table builders against the reference
// implementation itself.  No reference source is copied here; the headers and .cpp files are
// read from the checkout at build time.
#include <cstdint>
#include <cstdlib>
#include <tuple>
#include <vector>

#include <cstdio>
#include <cstring>
#include <exception>

#include "../include/sm2d_io.h"
#include "core/block_encoding.hpp"
#include "core/config_loader.hpp"
#include "core/grids.hpp"
#include "core/local_bins.hpp"
#include "lut/r_lut.hpp"
#include "physics/highland.hpp"
#include "physics/step_control.hpp"
#include "validation/deterministic_beam.hpp"

extern "C" {

int ref_energy_grid_log(float E_min, float E_max, int N_E, float* edges, float* rep) {
    EnergyGrid g(E_min, E_max, N_E);
    for (int i = 0; i <= N_E; ++i) edges[i] = g.edges[i];
    for (int i = 0; i < N_E; ++i) rep[i] = g.rep[i];
    return 0;
}

int ref_energy_grid_piecewise(const float* groups, int ng, float* edges, float* rep, int* nb) {
    std::vector<std::tuple<float, float, float>> G;
    for (int i = 0; i < ng; ++i) G.emplace_back(groups[3 * i], groups[3 * i + 1], groups[3 * i + 2]);
    EnergyGrid g = EnergyGrid::CreatePiecewise(G);
    *nb = g.N_E;
    if (edges && rep) {
        for (int i = 0; i <= g.N_E; ++i) edges[i] = g.edges[i];
        for (int i = 0; i < g.N_E; ++i) rep[i] = g.rep[i];
    }
    return 0;
}

int ref_angular_grid(float tmin, float tmax, int n, float* edges) {
    AngularGrid a(tmin, tmax, n);
    for (int i = 0; i <= n; ++i) edges[i] = a.edges[i];
    return 0;
}

// GenerateRLUT over a piecewise grid (groups) or a log grid (ng == 0: groups = {Emin, Emax, N}).
int ref_build_rlut(const float* groups, int ng, float* R, float* S, float* lE, float* lR, float* lS) {
    EnergyGrid g = (ng > 0) ? [&] {
        std::vector<std::tuple<float, float, float>> G;
        for (int i = 0; i < ng; ++i) G.emplace_back(groups[3 * i], groups[3 * i + 1], groups[3 * i + 2]);
        return EnergyGrid::CreatePiecewise(G);
    }() : EnergyGrid(groups[0], groups[1], (int)groups[2]);
    RLUT L = GenerateRLUT(g);
    for (int i = 0; i < g.N_E; ++i) {
        R[i] = L.R[i]; S[i] = L.S[i]; lE[i] = L.log_E[i]; lR[i] = L.log_R[i]; lS[i] = L.log_S[i];
    }
    return 0;
}

// Host lookups on GenerateRLUT(0.1, 300, 256), the table of the FE precompute.
int ref_log_lut_eval(int which, const float* x, const float* y, float* out, int n) {
    static RLUT L = GenerateRLUT(0.1f, 300.0f, 256);
    for (int i = 0; i < n; ++i) {
        switch (which) {
            case 0: out[i] = L.lookup_R(x[i]); break;
            case 1: out[i] = L.lookup_S(x[i]); break;
            case 2: out[i] = L.lookup_E_inverse(x[i]); break;
            case 3: out[i] = compute_energy_deposition(L, x[i], y[i]); break;
            case 4: out[i] = compute_max_step_physics(L, x[i]); break;
            case 5: out[i] = highland_sigma(x[i], y[i]); break;
            default: return -1;
        }
    }
    return 0;
}

int ref_encode_local_idx_4d(int t, int e, int x, int z) { return encode_local_idx_4d(t, e, x, z); }
int ref_decode_local_idx_4d(int lidx, int* t, int* e, int* x, int* z) {
    decode_local_idx_4d((uint16_t)lidx, *t, *e, *x, *z);
    return 0;
}
uint32_t ref_encode_block(uint32_t bt, uint32_t be) { return encode_block(bt, be); }
float ref_x_offset(int xs, float dx) { return get_x_offset_from_bin(xs, dx); }
float ref_z_offset(int zs, float dz) { return get_z_offset_from_bin(zs, dz); }
int ref_x_sub_bin(float off, float dx) { return get_x_sub_bin(off, dx); }
int ref_z_sub_bin(float off, float dz) { return get_z_sub_bin(off, dz); }
int ref_local_bins(void) { return LOCAL_BINS; }

// The reference's INI loader (config_loader.hpp:281-416) into the library's flat struct.
// Returns 0 (valid), 1 (loaded; validate() threw: msg holds its message) or 2 (load threw).
int ref_config_load(const char* path, sm2d_run_config* o, char* msg, int cap) {
    sm_2d::IncidentParticleConfig c;
    try {
        c = sm_2d::load_incident_particle_config(path);
    } catch (const std::exception& e) {
        std::snprintf(msg, cap, "%s", e.what());
        return 2;
    }
    o->particle_type = (int)c.particle_type;
    o->particle_mass_amu = c.particle_mass_amu;
    o->particle_charge_e = c.particle_charge_e;
    o->beam_profile = (int)c.beam_profile;
    o->W_total = c.W_total;
    o->E_mean = c.energy.mean_E0; o->E_sigma = c.energy.sigma_E;
    o->E_min = c.energy.min_E; o->E_max = c.energy.max_E;
    o->x0 = c.spatial.x0; o->z0 = c.spatial.z0; o->sigma_x = c.spatial.sigma_x;
    o->sigma_z = c.spatial.sigma_z; o->width_x = c.spatial.width_x; o->width_z = c.spatial.width_z;
    o->radius = c.spatial.radius;
    o->theta0 = c.angular.theta0; o->phi0 = c.angular.phi0;
    o->sigma_theta = c.angular.sigma_theta; o->sigma_phi = c.angular.sigma_phi;
    o->n_samples = c.sampling.n_samples; o->random_seed = c.sampling.random_seed;
    o->use_stratified = c.sampling.use_stratified;
    o->Nx = c.grid.Nx; o->Nz = c.grid.Nz; o->dx = c.grid.dx; o->dz = c.grid.dz;
    o->max_steps = c.grid.max_steps;
    const sm_2d::TransportConfig& t = c.transport;
    o->N_theta = t.N_theta; o->N_theta_local = t.N_theta_local; o->N_E_local = t.N_E_local;
    o->n_energy_groups = (int)t.energy_groups.size();
    for (int k = 0; k < o->n_energy_groups && k < SM2D_MAX_ENERGY_GROUPS; ++k) {
        o->energy_groups[k][0] = t.energy_groups[k].E_min_MeV;
        o->energy_groups[k][1] = t.energy_groups[k].E_max_MeV;
        o->energy_groups[k][2] = t.energy_groups[k].dE_MeV;
    }
    o->E_fine_on = t.E_fine_on; o->E_fine_off = t.E_fine_off; o->E_trigger = t.E_trigger;
    o->weight_active_min = t.weight_active_min; o->E_coarse_max = t.E_coarse_max;
    o->step_coarse = t.step_coarse; o->n_steps_per_cell = t.n_steps_per_cell;
    o->fine_batch_max_cells = t.fine_batch_max_cells; o->fine_halo_cells = t.fine_halo_cells;
    o->preflight_vram_margin = t.preflight_vram_margin; o->validation_mode = t.validation_mode;
    o->fail_fast_on_audit = t.fail_fast_on_audit; o->debug_dumps = t.debug_dumps;
    o->max_iterations = t.max_iterations; o->log_level = t.log_level;
    std::snprintf(o->output_dir, SM2D_PATH_MAX, "%s", c.output.output_dir.c_str());
    std::snprintf(o->dose_2d_file, SM2D_PATH_MAX, "%s", c.output.dose_2d_file.c_str());
    std::snprintf(o->pdd_file, SM2D_PATH_MAX, "%s", c.output.pdd_file.c_str());
    std::snprintf(o->let_file, SM2D_PATH_MAX, "%s", c.output.let_file.c_str());
    std::snprintf(o->format, sizeof o->format, "%s", c.output.format.c_str());
    o->normalize_dose = c.output.normalize_dose; o->save_2d = c.output.save_2d;
    o->save_pdd = c.output.save_pdd; o->save_lat_profiles = c.output.save_lat_profiles;
    try {
        c.validate();
    } catch (const std::exception& e) {
        std::snprintf(msg, cap, "%s", e.what());
        return 1;
    }
    return 0;
}

// get_depth_dose and find_bragg_peak_z (src/validation/deterministic_beam.cpp:207-240).
int ref_depth_dose(const double* edep, int Nx, int Nz, double* dd, int* peak) {
    SimulationResult r;
    r.Nx = Nx;
    r.Nz = Nz;
    r.dx = r.dz = 1.0f;
    r.edep.assign(Nz, std::vector<double>(Nx));
    for (int j = 0; j < Nz; ++j)
        for (int i = 0; i < Nx; ++i) r.edep[j][i] = edep[(size_t)j * Nx + i];
    std::vector<double> d = get_depth_dose(r);
    for (int j = 0; j < Nz; ++j) dd[j] = d[j];
    *peak = find_bragg_peak_z(r);
    return 0;
}

}  // extern "C"
