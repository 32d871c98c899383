// sm2d_shim.cpp — the reference runner's boundary symbols on top of libsm2d_hip.so.
//
// SM_2D's runner (src/gpu/gpu_transport_runner.cpp:145-206) calls
//   sm_2d::init_device_lut(const ::RLUT&, sm_2d::DeviceLUTWrapper&)      gpu_transport_wrapper.hpp:42
//   sm_2d::run_k1k6_pipeline_transport(x0, ..., edep)                   gpu_transport_wrapper.hpp:78-90
//   sm_2d::get_gpu_name_internal()                                      gpu_transport_wrapper.hpp:95
// and owns a sm_2d::DeviceLUTWrapper (gpu_transport_wrapper.hpp:21-37).  This file defines those
// symbols with the signatures of that header, so the runner and the driver link against this
// library unchanged.  It is compiled against the reference's own headers (include path
// src/include of the SM_2D checkout): cuda/gpu_transport_wrapper.hpp, core/grids.hpp,
// core/incident_particle_config.hpp, lut/r_lut.hpp.
//
// What a call does (the body of gpu_transport_wrapper.cu:207-641 through the C ABI):
//   - the LUT wrapper keeps the host R-LUT arrays; the device copy lives in the context;
//   - one context per grid shape is kept between calls (the reference allocates and frees its
//     device phase space inside every call); a different grid or LUT recreates it;
//   - the spot runs with sm2d_run_spot, the Fermi-Eyges sigma_total of the spot comes from
//     sm2d_fe_precompute and the MCS lateral convolution from sm2d_post_convolve, as in
//     gpu_transport_wrapper.cu:551-616;
//   - errors print to stderr and return false, as the reference does (k1k6_pipeline.cuh:434-442).
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "core/grids.hpp"
#include "core/incident_particle_config.hpp"
#include "cuda/gpu_transport_wrapper.hpp"
#include "lut/r_lut.hpp"
#include "sm2d_hip.h"

// The "device LUT" the runner passes back into run_k1k6_pipeline_transport: the host tables of
// the R-LUT (the library uploads them into the context that runs the spot).
struct DeviceRLUT {
    std::vector<float> R, S, log_E, log_R, log_S;
};

namespace sm_2d {

class DeviceLUTWrapperImpl {
public:
    ::DeviceRLUT lut;
};

DeviceLUTWrapper::DeviceLUTWrapper() : p_impl(new DeviceLUTWrapperImpl), dlut_ptr(nullptr) {}

DeviceLUTWrapper::~DeviceLUTWrapper() { delete p_impl; }

bool init_device_lut(const ::RLUT& cpu_lut, DeviceLUTWrapper& wrapper) {
    const size_t n = cpu_lut.R.size();
    if (n == 0 || cpu_lut.S.size() != n || cpu_lut.log_E.size() != n || cpu_lut.log_R.size() != n ||
        cpu_lut.log_S.size() != n) {
        std::fprintf(stderr, "init_device_lut: inconsistent R-LUT arrays\n");
        return false;
    }
    ::DeviceRLUT& d = wrapper.p_impl->lut;
    d.R = cpu_lut.R;
    d.S = cpu_lut.S;
    d.log_E = cpu_lut.log_E;
    d.log_R = cpu_lut.log_R;
    d.log_S = cpu_lut.log_S;
    wrapper.dlut_ptr = &d;
    return true;
}

namespace {

struct Cached {
    sm2d_grid grid{};
    const ::DeviceRLUT* lut = nullptr;
    std::vector<float> theta_edges, E_edges;
    sm2d_ctx* ctx = nullptr;
    ~Cached() {
        if (ctx) sm2d_ctx_destroy(ctx);
    }
};

std::mutex g_mu;
std::unique_ptr<Cached> g_cache;

bool fail(const char* what) {
    std::fprintf(stderr, "run_k1k6_pipeline_transport: %s: %s\n", what, sm2d_last_error());
    return false;
}

}  // namespace

bool run_k1k6_pipeline_transport(float x0, float z0, float theta0, float E0, float W_total, float sigma_x,
                                 float sigma_theta, float sigma_E, int n_samples, unsigned int random_seed, int Nx,
                                 int Nz, float dx, float dz, float x_min, float z_min, int N_theta_local,
                                 int N_E_local, const AngularGrid& theta_grid, const EnergyGrid& E_grid,
                                 const ::DeviceRLUT& dlut, const TransportConfig& transport,
                                 std::vector<std::vector<double>>& edep) {
    const sm2d_grid grid = {
        .struct_size = sizeof(sm2d_grid),
        .Nx = Nx,
        .Nz = Nz,
        .dx = dx,
        .dz = dz,
        .x_min = x_min,
        .z_min = z_min,
        .N_theta = theta_grid.N_theta,
        .theta_min = theta_grid.theta_min,
        .theta_max = theta_grid.theta_max,
        .N_E = E_grid.N_E,
        .n_theta_local = N_theta_local,
        .n_e_local = N_E_local,
        .n_x_sub = ::N_x_sub,
        .n_z_sub = ::N_z_sub,
    };
    if ((int)dlut.R.size() != E_grid.N_E || (int)E_grid.edges.size() != E_grid.N_E + 1 ||
        (int)theta_grid.edges.size() != theta_grid.N_theta + 1) {
        std::fprintf(stderr, "run_k1k6_pipeline_transport: grid and LUT sizes disagree\n");
        return false;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    const bool same = g_cache && std::memcmp(&g_cache->grid, &grid, sizeof(grid)) == 0 && g_cache->lut == &dlut &&
                      g_cache->theta_edges == theta_grid.edges && g_cache->E_edges == E_grid.edges;
    if (!same) {
        g_cache.reset(new Cached);
        g_cache->grid = grid;
        g_cache->lut = &dlut;
        g_cache->theta_edges = theta_grid.edges;
        g_cache->E_edges = E_grid.edges;
        if (sm2d_ctx_create(0, &grid, &g_cache->ctx) != SM2D_OK) {
            g_cache.reset();
            return fail("sm2d_ctx_create");
        }
        if (sm2d_lut_upload(g_cache->ctx, theta_grid.edges.data(), E_grid.edges.data(), dlut.R.data(), dlut.S.data(),
                            dlut.log_E.data(), dlut.log_R.data(), dlut.log_S.data()) != SM2D_OK) {
            g_cache.reset();
            return fail("sm2d_lut_upload");
        }
    }
    const sm2d_transport tp = {
        .struct_size = sizeof(sm2d_transport),
        .E_fine_on = transport.E_fine_on,
        .E_fine_off = transport.E_fine_off,
        .weight_active_min = transport.weight_active_min,
        .step_coarse = transport.step_coarse,
        .max_iterations = transport.max_iterations,
        .fail_fast_on_audit = (transport.fail_fast_on_audit || transport.validation_mode) ? 1 : 0,
        .step_rule = 0,
        .enable_straggling = 1,
        .enable_nuclear = 1,
        .sigma_x_initial = sigma_x,
    };
    const sm2d_spot spot = {
        .struct_size = sizeof(sm2d_spot),
        .x0 = x0,
        .z0 = z0,
        .theta0 = theta0,
        .E0 = E0,
        .W_total = W_total,
        .sigma_x = sigma_x,
        .sigma_theta = sigma_theta,
        .sigma_E = sigma_E,
        .n_samples = n_samples,
        .random_seed = random_seed,
    };
    std::vector<double> flat((size_t)Nx * Nz);
    if (sm2d_run_spot(g_cache->ctx, &spot, &tp, flat.data(), nullptr) != SM2D_OK) return fail("sm2d_run_spot");
    std::vector<float> A(Nz), B(Nz), sigma_total(Nz);
    if (sm2d_fe_precompute(E0, sigma_x, Nz, dz, A.data(), B.data(), sigma_total.data()) != SM2D_OK)
        return fail("sm2d_fe_precompute");
    if (sm2d_post_convolve(flat.data(), Nx, Nz, dx, sigma_total.data(), sigma_x) != SM2D_OK)
        return fail("sm2d_post_convolve");
    if ((int)edep.size() != Nz) edep.assign(Nz, std::vector<double>(Nx, 0.0));
    for (int iz = 0; iz < Nz; ++iz) {
        if ((int)edep[iz].size() != Nx) edep[iz].assign(Nx, 0.0);
        for (int ix = 0; ix < Nx; ++ix) edep[iz][ix] = flat[(size_t)iz * Nx + ix];
    }
    return true;
}

std::string get_gpu_name_internal() {
    char name[256] = {0};
    if (sm2d_device_name(0, name, (int)sizeof(name)) != SM2D_OK) return "unknown";
    return name;
}

}  // namespace sm_2d
