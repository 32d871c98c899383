// sm2d_io.cpp — run configuration and file formats (C ABI part of libsm2d_hip.so, CPU only).
//
//   INI loader       src/include/core/config_loader.hpp:14-209 (parser), :212-265 (energy
//                    groups), :281-416 (sections and keys)
//   validation       src/include/core/incident_particle_config.hpp:272-391
//   runner inputs    src/gpu/gpu_transport_runner.cpp:28-98, 172-206
//   dose writers     run_simulation.cpp:31-116; depth dose and Bragg peak
//                    src/validation/deterministic_beam.cpp:207-240
//   raw ψ CSV        src/cuda/k1k6_pipeline.cu:195-291 (8 decimals) and the initial dump
//                    src/cuda/gpu_transport_wrapper.cu:22-120 (6 decimals)
//
// The files these functions write are byte-compatible with the reference's: the same column
// order, separators and number formats (fixed decimals for the tables, the stream default of
// 6 significant digits for the Bragg-peak line).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sm2d_io.h"

namespace sm2d_host {
void set_error(const std::string& m);
}

namespace {

int fail(const std::string& m, int code = SM2D_ERR_ARG) {
    sm2d_host::set_error(m);
    return code;
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// One INI file: section -> key -> raw value.  Lines are trimmed; '#' or ';' in the first
// column starts a comment (there are no trailing comments: the rest of a line is the value);
// a later key overrides an earlier one; a value in matching quotes loses them.
struct Ini {
    std::map<std::string, std::map<std::string, std::string>> sec;

    bool read(const char* path) {
        std::ifstream f(path);
        if (!f.is_open()) return false;
        std::string cur = "default", line;
        while (std::getline(f, line)) {
            line = trim(line);
            if (line.empty() || line[0] == '#' || line[0] == ';') continue;
            if (line.front() == '[' && line.back() == ']') {
                cur = trim(line.substr(1, line.size() - 2));
                continue;
            }
            const size_t eq = line.find('=');
            if (eq == std::string::npos) continue;
            std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
            if (v.size() >= 2 && ((v.front() == '"' && v.back() == '"') ||
                                  (v.front() == '\'' && v.back() == '\'')))
                v = v.substr(1, v.size() - 2);
            sec[cur][k] = v;
        }
        return true;
    }
    const std::string* find(const char* s, const char* k) const {
        auto i = sec.find(s);
        if (i == sec.end()) return nullptr;
        auto j = i->second.find(k);
        return j == i->second.end() ? nullptr : &j->second;
    }
    bool has(const char* s, const char* k) const { return find(s, k) != nullptr; }
    // Numbers parse their leading part (std::stof / std::stoi); no number keeps the default.
    float f(const char* s, const char* k, float d) const {
        const std::string* v = find(s, k);
        if (!v) return d;
        try { return std::stof(*v); } catch (...) { return d; }
    }
    int i(const char* s, const char* k, int d) const {
        const std::string* v = find(s, k);
        if (!v) return d;
        try { return std::stoi(*v); } catch (...) { return d; }
    }
    // A present key is true only for true/1/yes/on (any case); anything else is false.
    int b(const char* s, const char* k, int d) const {
        const std::string* v = find(s, k);
        if (!v) return d;
        std::string t = *v;
        for (char& c : t) c = (char)std::tolower((unsigned char)c);
        return t == "true" || t == "1" || t == "yes" || t == "on";
    }
    void str(const char* s, const char* k, char* out, size_t cap) const {
        const std::string* v = find(s, k);
        if (v) std::snprintf(out, cap, "%s", v->c_str());
    }
};

int particle_type(const std::string& v) {
    static const char* names[] = {"proton", "electron", "positron", "alpha", "carbon_ion", "custom"};
    for (int k = 0; k < 6; ++k)
        if (v == names[k]) return k;
    return 0;
}
int beam_profile(const std::string& v) {
    static const char* names[] = {"pencil", "gaussian", "flat", "custom"};
    for (int k = 0; k < 4; ++k)
        if (v == names[k]) return k;
    return 0;
}

// "a:b:c, a:b:c, ..." (config_loader.hpp:228-265): any malformed group or a list with no
// group leaves the previous groups in place.
void energy_groups(const std::string& raw, sm2d_run_config* c) {
    if (raw.empty()) return;
    std::vector<float> g;
    std::stringstream ss(raw);
    std::string tok;
    while (std::getline(ss, tok, ',')) {
        tok = trim(tok);
        if (tok.empty()) continue;
        std::stringstream ts(tok);
        std::string p[3];
        for (auto& q : p)
            if (!std::getline(ts, q, ':')) return;
        try {
            for (auto& q : p) g.push_back(std::stof(trim(q)));
        } catch (...) {
            return;
        }
    }
    const int n = (int)g.size() / 3;
    if (n == 0) return;
    if (n > SM2D_MAX_ENERGY_GROUPS) throw std::invalid_argument("more than 64 energy groups");
    c->n_energy_groups = n;
    for (int k = 0; k < 3 * n; ++k) c->energy_groups[k / 3][k % 3] = g[k];
}

// std::ostream's default float format (6 significant digits, %g).
std::string dflt(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%g", v);
    return b;
}

}  // namespace

extern "C" {

int sm2d_config_defaults(sm2d_run_config* c) {
    if (!c || c->struct_size != sizeof(sm2d_run_config)) return fail("sm2d_config_defaults: struct_size");
    std::memset((char*)c + sizeof(uint32_t), 0, sizeof(*c) - sizeof(uint32_t));
    c->particle_type = 0;
    c->particle_mass_amu = 1.0f;
    c->particle_charge_e = 1.0f;
    c->beam_profile = 0;
    c->W_total = 1.0f;
    c->E_mean = 150.0f;
    c->E_sigma = 0.0f;
    c->E_min = 0.0f;
    c->E_max = 500.0f;
    c->n_samples = 1000;
    c->random_seed = 42;
    c->Nx = 100;
    c->Nz = 200;
    c->dx = 1.0f;
    c->dz = 1.0f;
    c->max_steps = 100;
    c->N_theta = 36;
    c->N_theta_local = 4;  // the reference's compile-time layout (local_bins.hpp:19-23)
    c->N_E_local = 2;
    const float groups[4][3] = {{0.1f, 2.0f, 0.1f}, {2.0f, 20.0f, 0.2f}, {20.0f, 100.0f, 0.25f},
                                {100.0f, 250.0f, 0.25f}};
    c->n_energy_groups = 4;
    std::memcpy(c->energy_groups, groups, sizeof groups);
    c->E_fine_on = 10.0f;
    c->E_fine_off = 11.0f;
    c->E_trigger = 10.0f;
    c->weight_active_min = 1e-12f;
    c->E_coarse_max = 300.0f;
    c->step_coarse = 5.0f;
    c->n_steps_per_cell = 1;
    c->fine_batch_max_cells = 0;
    c->fine_halo_cells = 1;
    c->preflight_vram_margin = 0.85f;
    c->debug_dumps = 1;
    c->max_iterations = 0;
    c->log_level = 1;
    std::snprintf(c->output_dir, SM2D_PATH_MAX, "results");
    std::snprintf(c->dose_2d_file, SM2D_PATH_MAX, "dose_2d.txt");
    std::snprintf(c->pdd_file, SM2D_PATH_MAX, "pdd.txt");
    std::snprintf(c->format, sizeof c->format, "txt");
    c->normalize_dose = 1;
    c->save_2d = 1;
    c->save_pdd = 1;
    return SM2D_OK;
}

int sm2d_config_load(const char* path, sm2d_run_config* c) {
    int rc = sm2d_config_defaults(c);
    if (rc) return rc;
    Ini ini;
    if (!path || !ini.read(path))
        return fail(std::string("Failed to load configuration file: ") + (path ? path : "(null)"));
    try {
        if (const std::string* v = ini.find("particle", "type")) c->particle_type = particle_type(*v);
        c->particle_mass_amu = ini.f("particle", "mass_amu", c->particle_mass_amu);
        c->particle_charge_e = ini.f("particle", "charge_e", c->particle_charge_e);
        if (const std::string* v = ini.find("beam", "profile")) c->beam_profile = beam_profile(*v);
        c->W_total = ini.f("beam", "weight", c->W_total);

        c->E_mean = ini.f("energy", "mean_MeV", c->E_mean);
        c->E_sigma = ini.f("energy", "sigma_MeV", c->E_sigma);
        c->E_min = ini.f("energy", "min_MeV", c->E_min);
        c->E_max = ini.f("energy", "max_MeV", c->E_max);

        c->x0 = ini.f("spatial", "x0_mm", c->x0);
        c->z0 = ini.f("spatial", "z0_mm", c->z0);
        c->sigma_x = ini.f("spatial", "sigma_x_mm", c->sigma_x);
        c->sigma_z = ini.f("spatial", "sigma_z_mm", c->sigma_z);
        c->width_x = ini.f("spatial", "width_x_mm", c->width_x);
        c->width_z = ini.f("spatial", "width_z_mm", c->width_z);
        c->radius = ini.f("spatial", "radius_mm", c->radius);

        c->theta0 = ini.f("angular", "theta0_rad", c->theta0);
        c->phi0 = ini.f("angular", "phi0_rad", c->phi0);
        c->sigma_theta = ini.f("angular", "sigma_theta_rad", c->sigma_theta);
        c->sigma_phi = ini.f("angular", "sigma_phi_rad", c->sigma_phi);

        c->n_samples = ini.i("sampling", "n_samples", c->n_samples);
        c->random_seed = (uint32_t)ini.i("sampling", "random_seed", (int)c->random_seed);
        c->use_stratified = ini.b("sampling", "use_stratified", c->use_stratified);

        c->Nx = ini.i("grid", "Nx", c->Nx);
        c->Nz = ini.i("grid", "Nz", c->Nz);
        c->dx = ini.f("grid", "dx_mm", c->dx);
        c->dz = ini.f("grid", "dz_mm", c->dz);
        c->max_steps = ini.i("grid", "max_steps", c->max_steps);

        const char* T = "transport";
        c->N_theta = ini.i(T, "N_theta", c->N_theta);
        c->N_theta_local = ini.i(T, "N_theta_local", c->N_theta_local);
        c->N_E_local = ini.i(T, "N_E_local", c->N_E_local);
        // E_fine_on / E_fine_off take either key spelling (the unit-less one wins when both
        // are given); the legacy E_trigger sets E_fine_on only when E_fine_on is absent.
        const bool on = ini.has(T, "E_fine_on_MeV") || ini.has(T, "E_fine_on");
        const bool off = ini.has(T, "E_fine_off_MeV") || ini.has(T, "E_fine_off");
        const bool trig = ini.has(T, "E_trigger_MeV") || ini.has(T, "E_trigger");
        if (on) c->E_fine_on = ini.f(T, "E_fine_on", ini.f(T, "E_fine_on_MeV", c->E_fine_on));
        if (off) c->E_fine_off = ini.f(T, "E_fine_off", ini.f(T, "E_fine_off_MeV", c->E_fine_off));
        if (trig) {
            c->E_trigger = ini.f(T, "E_trigger", ini.f(T, "E_trigger_MeV", c->E_trigger));
            if (!on) c->E_fine_on = c->E_trigger;
        } else {
            c->E_trigger = c->E_fine_on;
        }
        if (!off) c->E_fine_off = std::max(c->E_fine_off, c->E_fine_on);
        c->weight_active_min = ini.f(T, "weight_active_min", c->weight_active_min);
        c->E_coarse_max = ini.f(T, "E_coarse_max", ini.f(T, "E_coarse_max_MeV", c->E_coarse_max));
        c->step_coarse = ini.f(T, "step_coarse", ini.f(T, "step_coarse_mm", c->step_coarse));
        c->n_steps_per_cell = ini.i(T, "n_steps_per_cell", c->n_steps_per_cell);
        c->fine_batch_max_cells = ini.i(T, "fine_batch_max_cells", c->fine_batch_max_cells);
        c->fine_halo_cells = ini.i(T, "fine_halo_cells", c->fine_halo_cells);
        c->preflight_vram_margin = ini.f(T, "preflight_vram_margin", c->preflight_vram_margin);
        c->validation_mode = ini.b(T, "validation_mode", c->validation_mode);
        c->fail_fast_on_audit = ini.b(T, "fail_fast_on_audit", c->fail_fast_on_audit);
        c->debug_dumps = ini.b(T, "enable_debug_dumps", ini.b(T, "debug_dumps", c->debug_dumps));
        c->max_iterations = ini.i(T, "max_iterations", c->max_iterations);
        c->log_level = ini.i(T, "log_level", c->log_level);
        if (const std::string* v = ini.find(T, "energy_groups")) energy_groups(*v, c);

        const char* O = "output";
        ini.str(O, "output_dir", c->output_dir, SM2D_PATH_MAX);
        ini.str(O, "dose_2d_file", c->dose_2d_file, SM2D_PATH_MAX);
        ini.str(O, "pdd_file", c->pdd_file, SM2D_PATH_MAX);
        ini.str(O, "let_file", c->let_file, SM2D_PATH_MAX);
        ini.str(O, "format", c->format, sizeof c->format);
        c->normalize_dose = ini.b(O, "normalize_dose", c->normalize_dose);
        c->save_2d = ini.b(O, "save_2d", c->save_2d);
        c->save_pdd = ini.b(O, "save_pdd", c->save_pdd);
        c->save_lat_profiles = ini.b(O, "save_lat_profiles", c->save_lat_profiles);
    } catch (const std::exception& e) {
        return fail(std::string("sm2d_config_load: ") + e.what());
    }
    return SM2D_OK;
}

int sm2d_config_validate(const sm2d_run_config* c) {
    if (!c || c->struct_size != sizeof(sm2d_run_config)) return fail("sm2d_config_validate: struct_size");
    const char* P = "IncidentParticleConfig: ";
    auto bad = [&](const char* m) { return fail(std::string(P) + m); };
    if (c->E_mean <= 0.0f) return bad("mean_E0 must be positive");
    if (c->E_sigma < 0.0f) return bad("sigma_E must be non-negative");
    if (c->E_min < 0.0f) return bad("min_E must be non-negative");
    if (c->E_max <= c->E_min) return bad("max_E must be > min_E");
    if (c->sigma_x < 0.0f || c->sigma_z < 0.0f) return bad("sigma must be non-negative");
    if (c->width_x < 0.0f || c->width_z < 0.0f) return bad("width must be non-negative");
    if (c->radius < 0.0f) return bad("radius must be non-negative");
    if (c->sigma_theta < 0.0f) return bad("sigma_theta must be non-negative");
    if (c->n_samples <= 0) return bad("n_samples must be positive");
    if (c->W_total <= 0.0f) return bad("W_total must be positive");
    if (c->Nx <= 0) return bad("Nx must be positive");
    if (c->Nz <= 0) return bad("Nz must be positive");
    if (c->dx <= 0.0f) return bad("dx must be positive");
    if (c->dz <= 0.0f) return bad("dz must be positive");
    if (c->max_steps <= 0) return bad("max_steps must be positive");
    if (c->N_theta <= 0) return bad("transport.N_theta must be positive");
    if (c->N_theta_local <= 0) return bad("transport.N_theta_local must be positive");
    if (c->N_E_local <= 0) return bad("transport.N_E_local must be positive");
    if (c->E_fine_on <= 0.0f) return bad("transport.E_fine_on must be positive");
    if (c->E_fine_off < c->E_fine_on) return bad("transport.E_fine_off must be >= transport.E_fine_on");
    if (c->E_trigger <= 0.0f) return bad("transport.E_trigger must be positive");
    if (c->weight_active_min <= 0.0f) return bad("transport.weight_active_min must be positive");
    if (c->E_coarse_max <= 0.0f) return bad("transport.E_coarse_max must be positive");
    if (c->step_coarse <= 0.0f) return bad("transport.step_coarse must be positive");
    if (c->n_steps_per_cell <= 0) return bad("transport.n_steps_per_cell must be positive");
    if (c->fine_batch_max_cells < 0) return bad("transport.fine_batch_max_cells cannot be negative");
    if (c->fine_halo_cells < 0) return bad("transport.fine_halo_cells cannot be negative");
    if (c->preflight_vram_margin <= 0.0f || c->preflight_vram_margin > 1.0f)
        return bad("transport.preflight_vram_margin must be in (0,1]");
    if (c->max_iterations < 0) return bad("transport.max_iterations cannot be negative");
    if (c->log_level < 0) return bad("transport.log_level cannot be negative");
    if (c->n_energy_groups <= 0) return bad("transport.energy_groups cannot be empty");
    float last_max = c->energy_groups[0][0];
    for (int k = 0; k < c->n_energy_groups; ++k) {
        const float* g = c->energy_groups[k];
        if (g[0] <= 0.0f) return bad("transport.energy_groups E_min must be positive");
        if (g[1] <= g[0]) return bad("transport.energy_groups E_max must be > E_min");
        if (g[2] <= 0.0f) return bad("transport.energy_groups dE must be positive");
        if (k > 0 && g[0] < last_max) return bad("transport.energy_groups must be non-overlapping and ordered");
        last_max = g[1];
    }
    return SM2D_OK;
}

int sm2d_config_runtime(const sm2d_run_config* c, sm2d_grid* g, sm2d_transport* tp, sm2d_spot* sp) {
    if (!c || c->struct_size != sizeof(sm2d_run_config) || !g || !tp || !sp)
        return fail("sm2d_config_runtime: bad arguments");
    // The library's two layouts; the reference accepts only its compile-time one and throws
    // std::invalid_argument otherwise (gpu_transport_runner.cpp:29-36).
    int xs, zs;
    if (c->N_theta_local == 4 && c->N_E_local == 2) { xs = 8; zs = 4; }
    else if (c->N_theta_local == 8 && c->N_E_local == 4) { xs = 4; zs = 1; }
    else
        return fail("Transport local bin config must match a supported layout: (N_theta_local, "
                    "N_E_local) = (4, 2) or (8, 4)");
    int NE = 0;
    int rc = sm2d_energy_grid_piecewise(&c->energy_groups[0][0], c->n_energy_groups, nullptr, nullptr, &NE);
    if (rc) return rc;
    *g = sm2d_grid{};
    g->struct_size = sizeof(sm2d_grid);
    g->Nx = c->Nx;
    g->Nz = c->Nz;
    g->dx = c->dx;
    g->dz = c->dz;
    g->x_min = -(c->Nx * c->dx) / 2.0f;
    g->z_min = 0.0f;
    g->N_theta = c->N_theta;
    g->theta_min = -0.35f;  // kThetaTransportLimitRad (gpu_transport_runner.cpp:83-84)
    g->theta_max = 0.35f;
    g->N_E = NE;
    g->n_theta_local = c->N_theta_local;
    g->n_e_local = c->N_E_local;
    g->n_x_sub = xs;
    g->n_z_sub = zs;

    float on = c->E_fine_on, off = c->E_fine_off;
    // legacy compatibility: a customised E_trigger with default fine thresholds lifts them
    if (std::fabs(c->E_trigger - on) > 1e-6f && std::fabs(on - 10.0f) < 1e-6f &&
        std::fabs(off - 11.0f) < 1e-6f) {
        on = c->E_trigger;
        off = std::max(off, on);
    }
    *tp = sm2d_transport{};
    tp->struct_size = sizeof(sm2d_transport);
    tp->E_fine_on = on;
    tp->E_fine_off = off;
    tp->weight_active_min = c->weight_active_min;
    tp->step_coarse = c->step_coarse;
    tp->max_iterations = c->max_iterations == 0 ? c->max_steps : c->max_iterations;
    tp->fail_fast_on_audit = (c->fail_fast_on_audit || c->validation_mode) ? 1 : 0;
    tp->step_rule = 0;
    tp->enable_straggling = 1;  // k1k6_pipeline.cu:1385-1386
    tp->enable_nuclear = 1;
    tp->sigma_x_initial = c->sigma_x;

    *sp = sm2d_spot{};
    sp->struct_size = sizeof(sm2d_spot);
    sp->x0 = c->x0;
    sp->z0 = c->z0;
    sp->theta0 = c->theta0;
    sp->E0 = c->E_mean;
    sp->W_total = c->W_total;
    sp->sigma_x = c->sigma_x;
    sp->sigma_theta = c->sigma_theta;
    sp->sigma_E = c->E_sigma;
    sp->n_samples = c->n_samples;
    sp->random_seed = c->random_seed;
    return SM2D_OK;
}

int sm2d_write_dose_2d(const char* path, const double* edep, int Nx, int Nz, float dx, float dz,
                       int normalize) {
    if (!path || !edep || Nx <= 0 || Nz <= 0) return fail("sm2d_write_dose_2d: bad arguments");
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(std::string("Cannot open file ") + path);
    double mx = 0.0;
    if (normalize)
        for (size_t k = 0; k < (size_t)Nx * Nz; ++k) mx = std::max(mx, edep[k]);
    std::fputs("# 2D Dose Distribution\n# x(mm) z(mm) dose(Gy)", f);
    if (normalize) std::fputs(" dose_norm", f);
    std::fputs("\n", f);
    const float half = (Nx * dx) / 2.0f;
    for (int iz = 0; iz < Nz; ++iz) {
        const float z = iz * dz;
        for (int ix = 0; ix < Nx; ++ix) {
            const float x = (ix + 0.5f) * dx - half;
            double d = edep[(size_t)iz * Nx + ix];
            if (std::isnan(d) || std::isinf(d)) d = 0.0;
            std::fprintf(f, "%.4f\t%.4f\t%.4f", (double)x, (double)z, d);
            if (normalize && mx > 0) std::fprintf(f, "\t%.4f", d / mx);
            std::fputs("\n", f);
        }
        std::fputs("\n", f);
    }
    return std::fclose(f) == 0 ? SM2D_OK : fail(std::string("write failed: ") + path);
}

int sm2d_write_pdd(const char* path, const double* edep, int Nx, int Nz, float dz, int normalize) {
    if (!path || !edep || Nx <= 0 || Nz <= 0) return fail("sm2d_write_pdd: bad arguments");
    std::vector<double> dd(Nz, 0.0);
    int peak = 0;
    double best = 0.0;
    for (int iz = 0; iz < Nz; ++iz) {
        double s = 0.0;
        for (int ix = 0; ix < Nx; ++ix) s += edep[(size_t)iz * Nx + ix];
        dd[iz] = s;
        if (s > best) {  // first row of the largest row sum
            best = s;
            peak = iz;
        }
    }
    const double pd = dd[peak];
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(std::string("Cannot open file ") + path);
    std::fputs("# Depth-Dose Distribution (PDD)\n# Depth(mm) Dose(Gy)", f);
    if (normalize) std::fputs(" Dose_norm", f);
    std::fputs("\n", f);
    std::fprintf(f, "# Bragg Peak at: %s mm, %s Gy\n", dflt((double)(peak * dz)).c_str(), dflt(pd).c_str());
    for (int iz = 0; iz < Nz; ++iz) {
        const float depth = iz * dz;
        double d = dd[iz];
        if (std::isnan(d) || std::isinf(d)) d = 0.0;
        std::fprintf(f, "%.4f\t%.4f", (double)depth, d);
        if (normalize && pd > 0) std::fprintf(f, "\t%.4f", d / pd);
        std::fputs("\n", f);
    }
    return std::fclose(f) == 0 ? SM2D_OK : fail(std::string("write failed: ") + path);
}

int sm2d_write_psi_csv(const char* path, int iteration, int decimals, const sm2d_grid* g,
                       const float* te, const float* ee, int n, const int32_t* cell,
                       const int32_t* slot, const uint32_t* bid, const int32_t* lidx, const float* w) {
    if (!path || !g || g->struct_size != sizeof(sm2d_grid) || !te || !ee || n < 0 ||
        (n > 0 && (!cell || !slot || !bid || !lidx || !w)) || decimals < 0 || decimals > 12)
        return fail("sm2d_write_psi_csv: bad arguments");
    const int TL = g->n_theta_local, EL = g->n_e_local, XS = g->n_x_sub, ZS = g->n_z_sub;
    const int LB = TL * EL * XS * ZS;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (cell[a] != cell[b]) return cell[a] < cell[b];
        if (slot[a] != slot[b]) return slot[a] < slot[b];
        return lidx[a] < lidx[b];
    });
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(std::string("Failed to open ") + path + " for writing");
    std::fputs("iter,cell,ix,iz,slot,bid,lidx,weight,b_theta,b_E,theta_local,E_local,x_sub,z_sub,"
               "theta_bin,E_bin,theta_rep,E_rep,x_offset_mm,z_offset_mm\n", f);
    const int p = decimals;
    for (int r : order) {
        if (w[r] <= 1e-12f) continue;
        const int l = lidx[r];
        if (l < 0 || l >= LB || cell[r] < 0) {
            std::fclose(f);
            return fail("sm2d_write_psi_csv: row out of range");
        }
        const int bt = (int)(bid[r] & 0xFFF), bE = (int)((bid[r] >> 12) & 0xFFF);
        const int t = l % TL, e = (l / TL) % EL, x = (l / (TL * EL)) % XS, z = l / (TL * EL * XS);
        const int tb = std::max(0, std::min(bt * TL + t, g->N_theta - 1));
        const int eb = std::max(0, std::min(bE * EL + e, g->N_E - 1));
        const float trep = te[tb] + 0.5f * (te[tb + 1] - te[tb]);
        const float erep = ee[eb] + 0.5f * (ee[eb + 1] - ee[eb]);
        // sub-bin centres (local_bins.hpp:80-106); exact for power-of-two sub-bin counts
        const float xo = g->dx * ((float)(2 * x + 1 - XS) / (float)(2 * XS));
        const float zo = g->dz * ((float)(2 * z + 1 - ZS) / (float)(2 * ZS));
        std::fprintf(f, "%d,%d,%d,%d,%d,%u,%d,%.*f,%d,%d,%d,%d,%d,%d,%d,%d,%.*f,%.*f,%.*f,%.*f\n",
                     iteration, cell[r], cell[r] % g->Nx, cell[r] / g->Nx, slot[r], bid[r], l, p,
                     (double)w[r], bt, bE, t, e, x, z, tb, eb, p, (double)trep, p, (double)erep, p,
                     (double)xo, p, (double)zo);
    }
    return std::fclose(f) == 0 ? SM2D_OK : fail(std::string("write failed: ") + path);
}

int sm2d_write_cells_csv(const char* path, const sm2d_grid* g, const float* te, const float* ee, int n,
                         const int32_t* cell, const int32_t* slot, const uint32_t* bid, const int32_t* lidx,
                         const float* w) {
    if (!path || !g || g->struct_size != sizeof(sm2d_grid) || !te || !ee || n < 0 ||
        (n > 0 && (!cell || !slot || !bid || !lidx || !w)))
        return fail("sm2d_write_cells_csv: bad arguments");
    const int TL = g->n_theta_local, EL = g->n_e_local, XS = g->n_x_sub, ZS = g->n_z_sub;
    const int LB = TL * EL * XS * ZS;
    // bin centres as the dump computes them: theta from the uniform width, E from the edges
    const float t0 = te[0], dth = (te[g->N_theta] - te[0]) / (float)g->N_theta;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (cell[a] != cell[b]) return cell[a] < cell[b];
        if (slot[a] != slot[b]) return slot[a] < slot[b];
        return lidx[a] < lidx[b];
    });
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(std::string("Failed to open ") + path + " for writing");
    std::fputs("cell,ix,iz,x_mm,z_mm,total_weight,num_slots,mean_E_MeV,mean_theta_rad,min_E_MeV,max_E_MeV,"
               "min_theta_rad,max_theta_rad\n", f);
    size_t r = 0;
    while (r < order.size()) {
        const int c = cell[order[r]];
        if (c < 0) {
            std::fclose(f);
            return fail("sm2d_write_cells_csv: row out of range");
        }
        // one cell: FP32 sums in (slot, lidx) order over the bins above 1e-12
        float W = 0.f, sE = 0.f, sT = 0.f, mnE = 1e30f, mxE = -1e30f, mnT = 1e30f, mxT = -1e30f;
        int nz = 0;
        for (; r < order.size() && cell[order[r]] == c; ++r) {
            const int k = order[r];
            if (!(w[k] > 1e-12f)) continue;
            const int l = lidx[k];
            if (l < 0 || l >= LB) {
                std::fclose(f);
                return fail("sm2d_write_cells_csv: row out of range");
            }
            const int tb = (int)(bid[k] & 0xFFF) * TL + l % TL;
            const int eb = (int)((bid[k] >> 12) & 0xFFF) * EL + (l / TL) % EL;
            if (tb >= g->N_theta || eb >= g->N_E) {
                std::fclose(f);
                return fail("sm2d_write_cells_csv: bin outside the grid");
            }
            const float th = t0 + ((float)tb + 0.5f) * dth;
            const float E = ee[eb] + (ee[eb + 1] - ee[eb]) * 0.5f;
            W += w[k];
            ++nz;
            sE += E * w[k];
            sT += th * w[k];
            mnE = std::min(mnE, E);
            mxE = std::max(mxE, E);
            mnT = std::min(mnT, th);
            mxT = std::max(mxT, th);
        }
        if (!(W > 1e-12f)) continue;
        const int ix = c % g->Nx, iz = c / g->Nx;
        std::fprintf(f, "%d,%d,%d,%.6f,%.6f,%.6f,%d,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f\n", c, ix, iz,
                     (double)(((float)ix + 0.5f) * g->dx), (double)(((float)iz + 0.5f) * g->dz), (double)W, nz,
                     (double)(sE / W), (double)(sT / W), (double)mnE, (double)mxE, (double)mnT, (double)mxT);
    }
    return std::fclose(f) == 0 ? SM2D_OK : fail(std::string("write failed: ") + path);
}

int sm2d_read_psi_csv(const char* path, int cap, int32_t* cell, int32_t* slot, uint32_t* bid,
                      int32_t* lidx, float* w, int* n_rows) {
    if (!path || !n_rows || cap < 0 || (cap > 0 && (!cell || !slot || !bid || !lidx || !w)))
        return fail("sm2d_read_psi_csv: bad arguments");
    std::ifstream f(path);
    if (!f.is_open()) return fail(std::string("Failed to open ") + path);
    std::string line;
    if (!std::getline(f, line)) return fail(std::string("empty file: ") + path);
    std::vector<std::string> head;
    {
        std::stringstream ss(line);
        std::string h;
        while (std::getline(ss, h, ',')) head.push_back(trim(h));
    }
    const char* want[5] = {"cell", "slot", "bid", "lidx", "weight"};
    int col[5];
    for (int k = 0; k < 5; ++k) {
        auto it = std::find(head.begin(), head.end(), want[k]);
        if (it == head.end()) return fail(std::string("missing column ") + want[k] + " in " + path);
        col[k] = (int)(it - head.begin());
    }
    int n = 0;
    std::vector<std::string> v;
    while (std::getline(f, line)) {
        if (trim(line).empty()) continue;
        v.clear();
        std::stringstream ss(line);
        std::string tok;
        while (std::getline(ss, tok, ',')) v.push_back(tok);
        for (int k = 0; k < 5; ++k)
            if (col[k] >= (int)v.size()) return fail(std::string("short row in ") + path);
        if (n < cap) {
            char* end = nullptr;
            cell[n] = (int32_t)std::strtol(v[col[0]].c_str(), &end, 10);
            slot[n] = (int32_t)std::strtol(v[col[1]].c_str(), &end, 10);
            bid[n] = (uint32_t)std::strtoul(v[col[2]].c_str(), &end, 10);
            lidx[n] = (int32_t)std::strtol(v[col[3]].c_str(), &end, 10);
            w[n] = std::strtof(v[col[4]].c_str(), &end);
        }
        ++n;
    }
    *n_rows = n;
    return SM2D_OK;
}

}  // extern "C"
