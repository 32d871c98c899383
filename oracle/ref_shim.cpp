// ref_shim.cpp — C entry points over the reference's own host code, compiled from the SM_2D
// checkout (see Makefile target `ref`) into oracle/_ref/libsm2d_ref.so.  TEST INFRASTRUCTURE:
// used only to pin the oracle and the product's host table builders against the reference
// implementation itself.  No reference source is copied here; the headers and .cpp files are
// read from the checkout at build time.
#include <cstdint>
#include <cstdlib>
#include <tuple>
#include <vector>

#include "core/block_encoding.hpp"
#include "core/grids.hpp"
#include "core/local_bins.hpp"
#include "lut/r_lut.hpp"
#include "physics/highland.hpp"
#include "physics/step_control.hpp"

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

}  // extern "C"
