// sm2d_host.cpp — host-side table builders of the transport step (C ABI part of
// libsm2d_hip.so).  These run once per problem on the CPU, in binary32 with the C library's
// logf/expf/sqrtf, exactly like the reference's own host code, so the tables uploaded to the
// GPU are bit-identical to the ones the reference runner builds:
//   energy grids    src/core/grids.cpp:6-96
//   angular grid    src/core/grids.cpp:121-135
//   R-LUT           src/lut/r_lut.cpp:69-129 (NIST PSTAR water, src/data/nist/pstar_water.txt)
//   FE precompute   src/cuda/k1k6_pipeline.cu:1061-1176 (host RLUT lookups r_lut.cpp:131-205,
//                   step_control.hpp:69-84, highland.hpp:59-80)
//   post-convolution src/cuda/gpu_transport_wrapper.cu:551-616
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/sm2d_hip.h"

namespace {

thread_local std::string t_err;

// NIST PSTAR liquid water: energy [MeV], electronic stopping power [MeV cm2/g],
// CSDA range [g/cm2] (the 19 rows of src/data/nist/pstar_water.txt).
struct Node { float E, S, R; };
constexpr Node kPstar[19] = {
    {0.001f, 133.7f, 0.000006319f}, {0.01f, 422.9f, 0.00003599f}, {0.1f, 814.5f, 0.0001607f},
    {0.5f, 412.8f, 0.0008869f},     {1.0f, 260.6f, 0.002458f},    {2.0f, 158.6f, 0.007555f},
    {5.0f, 79.11f, 0.03623f},       {10.0f, 45.64f, 0.1230f},     {20.0f, 26.05f, 0.4260f},
    {30.0f, 18.76f, 0.8853f},       {50.0f, 12.45f, 2.227f},      {70.0f, 9.555f, 4.080f},
    {80.0f, 8.622f, 5.184f},        {90.0f, 7.884f, 6.398f},      {100.0f, 7.286f, 7.718f},
    {150.0f, 5.443f, 15.77f},       {200.0f, 4.491f, 25.96f},     {250.0f, 3.910f, 37.94f},
    {300.0f, 3.519f, 51.45f}};

void rlut(const float* rep, int n, float* R, float* S, float* lE, float* lR, float* lS) {
    const float to_mm = 10.0f / 1.0f;  // g/cm2 -> mm of water
    for (int i = 0; i < n; ++i) {
        const float E = rep[i];
        int k = 0;
        while (k < 19 && kPstar[k].E < E) ++k;  // first node with energy >= E
        if (k == 0 || k == 19) {
            const Node& e = kPstar[k == 0 ? 0 : 18];
            R[i] = e.R * to_mm;
            S[i] = e.S;
        } else {
            const Node& a = kPstar[k - 1];
            const Node& b = kPstar[k];
            const float ra = a.R * to_mm, rb = b.R * to_mm;
            const float t = (logf(E) - logf(a.E));
            const float span = (logf(b.E) - logf(a.E));
            R[i] = expf(logf(ra) + (logf(rb) - logf(ra)) * t / span);
            S[i] = expf(logf(a.S) + (logf(b.S) - logf(a.S)) * t / span);
        }
        lE[i] = logf(rep[i]);
        lR[i] = logf(R[i]);
        lS[i] = logf(S[i]);
    }
}

void log_grid(float lo, float hi, int n, float* edges, float* rep) {
    const float a = logf(lo), b = logf(hi);
    const float step = (b - a) / n;
    for (int i = 0; i <= n; ++i) edges[i] = expf(a + i * step);
    for (int i = 0; i < n; ++i) rep[i] = sqrtf(edges[i] * edges[i + 1]);
}

// The host RLUT the FE precompute uses: GenerateRLUT(0.1, 300, 256) with the host lookups.
struct LogLut {
    int n = 256;
    float lo = 0.1f, hi = 300.0f;
    std::vector<float> edges, rep, R, S, lE, lR, lS;
    LogLut() {
        edges.resize(n + 1); rep.resize(n); R.resize(n); S.resize(n);
        lE.resize(n); lR.resize(n); lS.resize(n);
        log_grid(lo, hi, n, edges.data(), rep.data());
        rlut(rep.data(), n, R.data(), S.data(), lE.data(), lR.data(), lS.data());
    }
    int bin(float E) const {
        if (E < edges[0]) return 0;
        if (E >= edges[n]) return n - 1;
        int a = 0, b = n;
        while (a < b) {
            int m = (a + b) / 2;
            if (edges[m + 1] <= E) a = m + 1; else b = m;
        }
        return a;
    }
    float range(float E) const {
        const float Ec = fmaxf(lo, fminf(E, hi));
        const int k = bin(Ec);
        if (k == n - 1 || Ec >= hi) return R[n - 1];
        const float x = logf(Ec);
        return expf(lR[k] + (lR[k + 1] - lR[k]) * (x - lE[k]) / (lE[k + 1] - lE[k]));
    }
    float energy(float r) const {
        if (r <= 0.0f) return lo;
        if (r >= R.back()) return hi;
        if (r <= R.front()) return lo;
        int k = (int)(std::lower_bound(R.begin(), R.end(), r) - R.begin());
        k = std::max(0, std::min(k, n - 2));
        const float x = logf(r);
        const float d = lR[k + 1] - lR[k];
        if (fabsf(d) < 1e-10f) return expf(lE[k]);
        return expf(lE[k] + (lE[k + 1] - lE[k]) * (x - lR[k]) / d);
    }
    float loss(float E, float ds) const {  // 0 when the step exceeds the residual range
        const float r = range(E) - ds;
        return E - (r <= 0 ? 0.0f : energy(r));
    }
};

float highland_host(float E, float ds) {
    const float m = 938.272f;
    const float g = (E + m) / m;
    const float beta = sqrtf(fmaxf(1.0f - 1.0f / (g * g), 0.0f));
    const float p = sqrtf(fmaxf((E + m) * (E + m) - m * m, 0.0f));
    const float t = ds / 360.8f;
    if (t < 1e-6f) return 0.0f;
    const float br = fmaxf(1.0f + 0.038f * logf(t), 0.25f);
    return (13.6f * 1.0f / (beta * p)) * sqrtf(t) * br;
}

}  // namespace

namespace sm2d_host {
void set_error(const std::string& m) { t_err = m; }
const char* error() { return t_err.c_str(); }
}

extern "C" {

int sm2d_energy_grid_log(float E_min, float E_max, int N_E, float* edges, float* rep) {
    if (N_E <= 0 || !(E_min > 0.0f) || !(E_max > E_min) || !edges || !rep) {
        sm2d_host::set_error("sm2d_energy_grid_log: bad arguments");
        return SM2D_ERR_ARG;
    }
    log_grid(E_min, E_max, N_E, edges, rep);
    return SM2D_OK;
}

int sm2d_energy_grid_piecewise(const float* groups, int n_groups, float* edges, float* rep,
                               int* n_bins) {
    if (!groups || n_groups <= 0) {
        sm2d_host::set_error("sm2d_energy_grid_piecewise: no groups");
        return SM2D_ERR_ARG;
    }
    std::vector<int> counts(n_groups);
    int total = 0;
    for (int g = 0; g < n_groups; ++g) {
        const float* t = groups + 3 * g;
        counts[g] = (int)((t[1] - t[0]) / t[2] + 0.5f);
        total += counts[g];
    }
    if (n_bins) *n_bins = total;
    if (!edges || !rep) return SM2D_OK;
    int base = 0;
    for (int g = 0; g < n_groups; ++g) {
        const float* t = groups + 3 * g;
        // A group's last edge is written before the next group overwrites it with its start;
        // the representative energies of a group use the group's own end edge.
        for (int i = 0; i <= counts[g]; ++i) edges[base + i] = t[0] + i * t[2];
        for (int i = 0; i < counts[g]; ++i) rep[base + i] = sqrtf(edges[base + i] * edges[base + i + 1]);
        base += counts[g];
    }
    edges[total] = groups[3 * (n_groups - 1) + 1];
    return SM2D_OK;
}

int sm2d_angular_grid(float theta_min, float theta_max, int N_theta, float* edges) {
    if (N_theta <= 0 || !edges) {
        sm2d_host::set_error("sm2d_angular_grid: bad arguments");
        return SM2D_ERR_ARG;
    }
    const float d = (theta_max - theta_min) / N_theta;
    for (int i = 0; i <= N_theta; ++i) edges[i] = theta_min + i * d;
    return SM2D_OK;
}

int sm2d_build_rlut(const float* rep, int N_E, float* R, float* S, float* log_E, float* log_R,
                    float* log_S) {
    if (!rep || N_E <= 0 || !R || !S || !log_E || !log_R || !log_S) {
        sm2d_host::set_error("sm2d_build_rlut: bad arguments");
        return SM2D_ERR_ARG;
    }
    rlut(rep, N_E, R, S, log_E, log_R, log_S);
    return SM2D_OK;
}

int sm2d_fe_precompute(float E0, float sigma_x0, int Nz, float dz, float* A, float* B,
                       float* sigma_total) {
    if (Nz <= 0 || !A || !B || !sigma_total) {
        sm2d_host::set_error("sm2d_fe_precompute: bad arguments");
        return SM2D_ERR_ARG;
    }
    if (!(E0 > 0.0f)) {
        std::fill(A, A + Nz, 0.0f);
        std::fill(B, B + Nz, 0.0f);
        std::fill(sigma_total, sigma_total + Nz, 0.0f);
        return SM2D_OK;
    }
    static const LogLut lut;
    // Angular variance A and covariance B per row (rows past the end of range stay 0).
    std::fill(A, A + Nz, 0.0f);
    std::fill(B, B + Nz, 0.0f);
    float E = E0, a = 0.0f, b = 0.0f;
    for (int iz = 0; iz < Nz; ++iz) {
        A[iz] = a;
        B[iz] = b;
        if (E <= 0.5f) break;
        float dE = std::min(lut.loss(E, dz), E - 0.1f);
        float Em = E - 0.5f * dE;
        if (Em < 0.5f) Em = 0.5f;
        const float s = highland_host(Em, dz);
        const float T = s * s / dz;
        const float a0 = a;
        a = a0 + T * dz;
        b = b + a0 * dz + 0.5f * T * dz * dz;
        E -= dE;
        if (E < 0.1f) E = 0.1f;
    }
    // Total lateral variance C per row, held constant after the end of range.
    float C = sigma_x0 * sigma_x0;
    E = E0;
    a = 0.0f;
    b = 0.0f;
    std::fill(sigma_total, sigma_total + Nz, sigma_x0);
    for (int iz = 0; iz < Nz; ++iz) {
        sigma_total[iz] = sqrtf(std::max(0.0f, C));
        if (E <= 0.5f) {
            std::fill(sigma_total + iz + 1, sigma_total + Nz, sigma_total[iz]);
            break;
        }
        float dE = std::min(lut.loss(E, dz), E - 0.1f);
        const float Em = std::max(0.5f, E - 0.5f * dE);
        const float s = highland_host(Em, dz);
        const float T = s * s / dz;
        C += 2.0f * b * dz + a * dz * dz + T * dz * dz * dz / 3.0f;
        const float a0 = a;
        a = a0 + T * dz;
        b = b + a0 * dz + 0.5f * T * dz * dz;
        E -= dE;
        if (E < 0.1f) E = 0.1f;
    }
    return SM2D_OK;
}

int sm2d_post_convolve(double* edep, int Nx, int Nz, float dx, const float* sigma_total,
                       float sigma_x0) {
    if (!edep || !sigma_total || Nx <= 0 || Nz <= 0) {
        sm2d_host::set_error("sm2d_post_convolve: bad arguments");
        return SM2D_ERR_ARG;
    }
    const float s0 = sigma_x0 * sigma_x0;
    std::vector<double> row(Nx), ker;
    for (int iz = 0; iz < Nz; ++iz) {
        const float st = sigma_total[iz];
        const float var = st * st - s0;
        if (var <= 1e-6f) continue;
        const float cells = std::sqrt(var) / dx;
        int h = (int)std::ceil(3.0f * cells);
        h = std::max(1, std::min(h, Nx / 2));
        ker.assign(2 * h + 1, 0.0);
        double norm = 0.0;
        for (int k = -h; k <= h; ++k) {
            ker[k + h] = std::exp(-0.5 * ((double)k * k) / ((double)cells * cells));
            norm += ker[k + h];
        }
        for (double& v : ker) v /= norm;
        double* r = edep + (size_t)iz * Nx;
        std::copy(r, r + Nx, row.begin());
        for (int ix = 0; ix < Nx; ++ix) {
            double acc = 0.0;
            for (int k = -h; k <= h; ++k) {
                const int j = ix + k;
                if (j >= 0 && j < Nx) acc += row[j] * ker[k + h];
            }
            r[ix] = acc;
        }
    }
    return SM2D_OK;
}

}  // extern "C"
