// sm2d_run.cpp — the simulation driver over libsm2d_hip.so: `sm2d_run [config.ini]`.
//
// The flow of the reference's run_simulation (run_simulation.cpp:118-189) and
// GPUTransportRunner::run (src/gpu/gpu_transport_runner.cpp:28-213):
//   load + validate the INI, print its summary, resolve the output directory next to the
//   config file, check device memory, build the grids and the range table, run the spot
//   (injection, K1-K6 loop, MCS post-convolution), report the Bragg peak and write
//   dose_2d.txt / pdd.txt in the reference's formats.
// Extra options (none of which the reference has): --device N, and --dump-psi, which writes the
// phase-space dumps of the reference's debug build (SM2D_ENABLE_DEBUG_DUMPS, gated at run time by
// transport.debug_dumps) into the output directory: debug_ps_raw_iter_00_initial.csv after the
// injection (gpu_transport_wrapper.cu:22-120) and, after K4 of iterations 1..10, 50, 100, 150,
// 200 and 250 (k1k6_pipeline.cu:1925-1941), debug_ps_raw_iter_NN_after_K4.csv and
// debug_cells_iter_NN_after_K4.csv.  With dumps on, the loop runs in pieces that end at those
// iterations; split runs give the same results as one run.
// Errors print "Error: <message>" to stderr and exit with status 1, as the reference does.
#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sm2d_hip.h"
#include "../../include/sm2d_io.h"

namespace {

void check(int rc, const char* what) {
    if (rc != SM2D_OK) {
        const char* m = sm2d_last_error();
        throw std::runtime_error(std::string(what) + ": " + (m && *m ? m : "failed"));
    }
}

const char* particle_name(int t) {
    static const char* n[] = {"proton", "electron", "positron", "alpha", "carbon_ion", "custom"};
    return (t >= 0 && t < 6) ? n[t] : "unknown";
}
const char* profile_name(int t) {
    static const char* n[] = {"pencil", "gaussian", "flat", "custom"};
    return (t >= 0 && t < 4) ? n[t] : "pencil";
}

void summary(const sm2d_run_config& c) {
    std::printf("=== Incident Particle Configuration ===\n");
    std::printf("Particle Type: %s\nBeam Profile: %s\n", particle_name(c.particle_type),
                profile_name(c.beam_profile));
    std::printf("Energy: %g MeV", c.E_mean);
    if (c.E_sigma > 0) std::printf(" (sigma = %g MeV)", c.E_sigma);
    std::printf("\nPosition: (%g, %g) mm\nAngle: %g rad", c.x0, c.z0, c.theta0);
    if (c.sigma_theta > 0) std::printf(" (sigma = %g rad)", c.sigma_theta);
    std::printf("\nWeight: %g\n", c.W_total);
    if (c.beam_profile == 1) std::printf("Samples: %d\n", c.n_samples);
    std::printf("\nGrid: %d x %d (dx=%gmm, dz=%gmm)\n", c.Nx, c.Nz, c.dx, c.dz);
    std::printf("Transport: theta_bins=%d, E_fine_on=%g MeV, E_fine_off=%g MeV, step_coarse=%g mm, "
                "max_iterations=%d, validation_mode=%s, fail_fast_on_audit=%s\n",
                c.N_theta, c.E_fine_on, c.E_fine_off, c.step_coarse,
                c.max_iterations > 0 ? c.max_iterations : c.max_steps, c.validation_mode ? "on" : "off",
                c.fail_fast_on_audit ? "on" : "off");
    std::printf("Output Dir: %s\n======================================\n", c.output_dir);
}

bool make_dir(const std::string& p) {
    struct stat st;
    if (stat(p.c_str(), &st) == 0) return S_ISDIR(st.st_mode);
    return mkdir(p.c_str(), 0755) == 0;
}

std::string mib(uint64_t b) {
    char s[64];
    std::snprintf(s, sizeof s, "%.2f MiB", (double)b / (1024.0 * 1024.0));
    return s;
}

int run(const std::string& cfg_path, int device, bool dump_psi) {
    sm2d_run_config cfg{};
    cfg.struct_size = sizeof cfg;
    check(sm2d_config_load(cfg_path.c_str(), &cfg), "load");
    check(sm2d_config_validate(&cfg), "validate");
    summary(cfg);

    // a relative output directory is taken from the config file's directory
    std::string out_dir = cfg.output_dir;
    if (!out_dir.empty() && out_dir[0] != '/') {
        const size_t slash = cfg_path.find_last_of("/\\");
        if (slash != std::string::npos) out_dir = cfg_path.substr(0, slash) + "/" + out_dir;
    }
    if (!make_dir(out_dir)) std::fprintf(stderr, "Warning: Could not create output directory: %s\n", out_dir.c_str());

    std::printf("\n--- Running Simulation ---\n");
    sm2d_grid g{};
    sm2d_transport tp{};
    sm2d_spot sp{};
    check(sm2d_config_runtime(&cfg, &g, &tp, &sp), "runtime configuration");

    sm2d_memory_plan plan{};
    plan.struct_size = sizeof plan;
    check(sm2d_memory_preflight(device, &g, 1, cfg.preflight_vram_margin, &plan), "memory preflight");
    if (cfg.log_level >= 1)
        std::printf("GPU memory preflight: required %s, usable %s (free %s, margin %g); dense store "
                    "would need %s\n",
                    mib(plan.total_bytes).c_str(), mib(plan.usable_bytes).c_str(),
                    mib(plan.free_bytes).c_str(), cfg.preflight_vram_margin, mib(plan.dense_bytes).c_str());
    if (!plan.ok)
        throw std::runtime_error("GPU memory preflight failed: required " + mib(plan.total_bytes) +
                                 " exceeds usable " + mib(plan.usable_bytes));

    // grids and range table as the reference runner builds them
    const int NT = g.N_theta, NE = g.N_E;
    std::vector<float> te(NT + 1), ee(NE + 1), rep(NE), R(NE), S(NE), lE(NE), lR(NE), lS(NE);
    check(sm2d_angular_grid(g.theta_min, g.theta_max, NT, te.data()), "angular grid");
    int nb = 0;
    check(sm2d_energy_grid_piecewise(&cfg.energy_groups[0][0], cfg.n_energy_groups, ee.data(), rep.data(), &nb),
          "energy grid");
    check(sm2d_build_rlut(rep.data(), NE, R.data(), S.data(), lE.data(), lR.data(), lS.data()), "range table");

    sm2d_ctx* ctx = nullptr;
    check(sm2d_ctx_create(device, &g, &ctx), "context");
    struct Guard {
        sm2d_ctx* c;
        ~Guard() { sm2d_ctx_destroy(c); }
    } guard{ctx};
    check(sm2d_lut_upload(ctx, te.data(), ee.data(), R.data(), S.data(), lE.data(), lR.data(), lS.data()),
          "table upload");
    if (cfg.log_level >= 1)
        std::printf("Running K1-K6 transport: E0 %g MeV, phase space %d x %d bins, max_iterations %d\n",
                    sp.E0, NT, NE, tp.max_iterations);

    std::vector<float> A(g.Nz), B(g.Nz), st(g.Nz);
    check(sm2d_fe_precompute(sp.E0, sp.sigma_x, g.Nz, g.dz, A.data(), B.data(), st.data()), "FE moments");
    std::vector<double> edep((size_t)g.Nx * g.Nz, 0.0);
    sm2d_audit au{};
    au.struct_size = sizeof au;
    if (dump_psi && cfg.debug_dumps) {
        // psi_in after injection, then after each dumped iteration (K6 leaves the K4 output there)
        auto dump = [&](int iter) {
            int n = 0;
            check(sm2d_psi_dump(ctx, nullptr, 0, nullptr, nullptr, nullptr, nullptr, nullptr, &n), "psi dump");
            std::vector<int32_t> cell(n), slot(n), lidx(n);
            std::vector<uint32_t> bid(n);
            std::vector<float> w(n);
            check(sm2d_psi_dump(ctx, nullptr, n, cell.data(), slot.data(), bid.data(), lidx.data(), w.data(), &n),
                  "psi dump");
            char name[64];
            if (iter == 0) std::snprintf(name, sizeof name, "debug_ps_raw_iter_00_initial.csv");
            else std::snprintf(name, sizeof name, "debug_ps_raw_iter_%02d_after_K4.csv", iter);
            const std::string p = out_dir + "/" + name;
            check(sm2d_write_psi_csv(p.c_str(), iter, iter == 0 ? 6 : 8, &g, te.data(), ee.data(), n, cell.data(),
                                     slot.data(), bid.data(), lidx.data(), w.data()),
                  "psi csv");
            std::printf("  DEBUG: Wrote %d raw phase-space rows to %s\n", n, p.c_str());
            if (iter > 0) {
                std::snprintf(name, sizeof name, "debug_cells_iter_%02d_after_K4.csv", iter);
                const std::string q = out_dir + "/" + name;
                check(sm2d_write_cells_csv(q.c_str(), &g, te.data(), ee.data(), n, cell.data(), slot.data(),
                                           bid.data(), lidx.data(), w.data()),
                      "cells csv");
            }
        };
        check(sm2d_reset(ctx), "reset");
        check(sm2d_fe_upload(ctx, A.data(), B.data()), "FE upload");
        check(sm2d_inject(ctx, &sp, &au), "injection");
        dump(0);
        std::vector<double> tallies((size_t)7 * g.Nx * g.Nz);
        int done = 0;
        bool ended = false;
        while (done < tp.max_iterations && !ended) {
            // the next dumped iteration: 1..10, 50, 100, 150, 200, 250, then the rest in one piece
            int next = done < 10 ? done + 1 : (done < 250 ? (done / 50 + 1) * 50 : tp.max_iterations);
            next = std::min(next, tp.max_iterations);
            sm2d_audit piece{};
            piece.struct_size = sizeof piece;
            check(sm2d_run_iterations(ctx, &tp, next - done, done == 0 ? 1 : 0, tallies.data(), nullptr, &piece),
                  "transport");
            au.iterations += piece.iterations;
            au.W_fail_iters += piece.W_fail_iters;
            au.max_cell_W_closure = std::max(au.max_cell_W_closure, piece.max_cell_W_closure);
            ended = piece.terminated != 0;
            done += piece.iterations;
            if (!ended && done == next && (done <= 10 || (done <= 250 && done % 50 == 0))) dump(done);
        }
        for (size_t i = 0; i < edep.size(); ++i) {
            edep[i] = tallies[i];
            au.E_dep += tallies[i];
        }
    } else {
        check(sm2d_run_spot(ctx, &sp, &tp, edep.data(), &au), "transport");
    }
    // MCS lateral spread as one convolution per depth row (gpu_transport_wrapper.cu:551-616)
    check(sm2d_post_convolve(edep.data(), g.Nx, g.Nz, g.dx, st.data(), sp.sigma_x), "post-convolution");
    if (cfg.log_level >= 1)
        std::printf("GPU transport complete: %d iterations, deposited %g MeV, W closure %.3g\n",
                    au.iterations, au.E_dep, au.max_cell_W_closure);

    int peak = 0;
    double best = 0.0;
    for (int iz = 0; iz < g.Nz; ++iz) {
        double s = 0.0;
        for (int ix = 0; ix < g.Nx; ++ix) s += edep[(size_t)iz * g.Nx + ix];
        if (s > best) {
            best = s;
            peak = iz;
        }
    }
    std::printf("Simulation complete.\n  Bragg Peak: %g mm depth, %g Gy\n", (double)(peak * g.dz), best);

    std::printf("\n--- Saving Results ---\n");
    if (cfg.save_2d) {
        const std::string p = out_dir + "/" + cfg.dose_2d_file;
        check(sm2d_write_dose_2d(p.c_str(), edep.data(), g.Nx, g.Nz, g.dx, g.dz, cfg.normalize_dose), "dose_2d");
        std::printf("  2D dose saved to: %s\n", p.c_str());
    }
    if (cfg.save_pdd) {
        const std::string p = out_dir + "/" + cfg.pdd_file;
        check(sm2d_write_pdd(p.c_str(), edep.data(), g.Nx, g.Nz, g.dz, cfg.normalize_dose), "pdd");
        std::printf("  PDD saved to: %s\n", p.c_str());
    }
    std::printf("\n=== Simulation Complete ===\n");
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    std::string cfg = "sim.ini";
    int device = 0;
    bool dump_psi = false;
    for (int k = 1; k < argc; ++k) {
        if (!std::strcmp(argv[k], "--device") && k + 1 < argc) device = std::atoi(argv[++k]);
        else if (!std::strcmp(argv[k], "--dump-psi")) dump_psi = true;
        else cfg = argv[k];
    }
    std::printf("========================================\nSM_2D: Proton Therapy Simulation\n"
                "========================================\nLoading configuration from: %s\n", cfg.c_str());
    try {
        return run(cfg, device, dump_psi);
    } catch (const std::exception& e) {
        std::fflush(stdout);
        std::fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
}
