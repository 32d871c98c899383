// sm2d_oracle.cpp — CPU oracle of SM_2D's K1–K6 transport step.
//
// TEST INFRASTRUCTURE ONLY (see sm2d_oracle.h).  A serial restatement of the reference
// algorithm; every function cites the reference file:line it follows.  All reference paths
// are relative to the SM_2D checkout.
//
// Arithmetic: binary32 for the phase space and the per-component physics, binary64 for the
// energy tallies, in the same expression order as the reference.  The transcendentals of the
// per-component physics come from sm_2d_amd/csrc/sm2d_math.h (the same source the GPU
// kernels compile), so that GPU and oracle agree bit for bit.  Host table builders
// (grids, R-LUT, Fermi–Eyges precompute, post-convolution) use the C library exactly as the
// reference's host code does.  Built with -DORC_FP64, the transport step runs in binary64
// instead (the FP32 results' error is stated against that build).
//
// Summation order where the reference's is a thread race (its kernels add with device atomics):
// payload adds per destination in component order; per-cell accumulators (tallies, drops,
// closure sums) one batch sum at a time, each batch of 64 components summed by a fixed
// butterfly (Batch below).  Both are the library's orders, fixed by construction and checked by
// the GPU parity tests; no reference fixture can pin them.
//
// Pinning (what checks each part; tests/ names the test):
//  - grids, R-LUT: bit-exact against the reference's own grids.cpp / r_lut.cpp compiled
//    from source into oracle/_ref (test_oracle_ref.py) and against the NIST nodes.
//  - encodings: reference KATs of tests/unit/test_local_bins.cpp, test_block_encoding.cpp.
//  - K1: truth table of tests/kernels/test_k1_activemask.cpp (test_oracle_kat.py).
//  - K5: closure KAT of tests/gpu/test_k5_drop_channels_gpu.cu:120-211 (test_oracle_kat.py).
//  - injection + one full iteration (K1, compaction, K2 or K3, K4, K5, K6): the committed
//    one-step dumps results/one_step_case_{A,B,C,D} (test_oracle_golden.py).
//  - theta split and angular draws: ten one-iteration comparisons with the reference's
//    multi-iteration dumps, the split law and the N(0,1) draw distribution (test_oracle_golden.py).
//  - physics helpers: closed forms (Highland, nuclear survival, Vavilov regimes, CSDA range
//    inversion, the device step rule) in test_oracle_kat.py; the 2 % step rule against the
//    reference's host rule compiled into oracle/_ref (test_oracle_ref.py).
//  - FE precompute: closed-form recurrence checks (A increments = T dz, B, sigma_total^2 =
//    sigma0^2 + C) and reference log-LUT lookups from oracle/_ref.
//  - post-convolution: the impulse response of every row against the discrete Gaussian of
//    sigma_MCS = sqrt(sigma_total^2 - sigma_x0^2) built independently, row sums away from the
//    edges (test_post_convolve.py).
//  - K4 slot-saturation fallback (closest-bid remap) and lateral emission into another
//    listed cell's bucket: brute force on small hand-built cases (test_oracle_cross_cell.py);
//    the reference's own order there is a thread race.
#include "sm2d_oracle.h"
#include "../sm_2d_amd/csrc/sm2d_math.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using sm2d::cosf_;
using sm2d::erff_;
using sm2d::expf_;
using sm2d::logf_;
using sm2d::sinf_;
using sm2d::sqrtf_;

namespace {

thread_local std::string g_err;
int fail(const std::string& m, int code = SM2D_ERR_ARG) {
    g_err = m;
    return code;
}

constexpr int KB = SM2D_KB;
constexpr uint32_t EMPTY = SM2D_EMPTY;

// ---------------------------------------------------------------------------------------
// NIST PSTAR water (src/data/nist/pstar_water.txt:18-36, same values as the fallback table
// of src/lut/nist_loader.cpp:26-46): energy [MeV], stopping power [MeV cm2/g], CSDA [g/cm2].
// ---------------------------------------------------------------------------------------
struct NistRow { float E, S, R; };
const NistRow kNist[] = {
    {0.001f, 133.7f, 0.000006319f}, {0.01f, 422.9f, 0.00003599f}, {0.1f, 814.5f, 0.0001607f},
    {0.5f, 412.8f, 0.0008869f},     {1.0f, 260.6f, 0.002458f},    {2.0f, 158.6f, 0.007555f},
    {5.0f, 79.11f, 0.03623f},       {10.0f, 45.64f, 0.1230f},     {20.0f, 26.05f, 0.4260f},
    {30.0f, 18.76f, 0.8853f},       {50.0f, 12.45f, 2.227f},      {70.0f, 9.555f, 4.080f},
    {80.0f, 8.622f, 5.184f},        {90.0f, 7.884f, 6.398f},      {100.0f, 7.286f, 7.718f},
    {150.0f, 5.443f, 15.77f},       {200.0f, 4.491f, 25.96f},     {250.0f, 3.910f, 37.94f},
    {300.0f, 3.519f, 51.45f}};
constexpr int kNistN = sizeof(kNist) / sizeof(kNist[0]);

// GenerateRLUT body (src/lut/r_lut.cpp:69-129), host libm.
void build_rlut(const float* rep, int N_E, float* R, float* S, float* lE, float* lR, float* lS) {
    const float g_cm2_to_mm = 10.0f / 1.0f;
    for (int i = 0; i < N_E; ++i) {
        float E = rep[i];
        int it = 0;  // lower_bound: first row with energy >= E
        while (it < kNistN && kNist[it].E < E) ++it;
        if (it == 0) {
            R[i] = kNist[0].R * g_cm2_to_mm;
            S[i] = kNist[0].S;
        } else if (it == kNistN) {
            R[i] = kNist[kNistN - 1].R * g_cm2_to_mm;
            S[i] = kNist[kNistN - 1].S;
        } else {
            float E0 = kNist[it - 1].E, E1 = kNist[it].E;
            float R0 = kNist[it - 1].R * g_cm2_to_mm, R1 = kNist[it].R * g_cm2_to_mm;
            float S0 = kNist[it - 1].S, S1 = kNist[it].S;
            float log_R = logf(R0) + (logf(R1) - logf(R0)) * (logf(E) - logf(E0)) /
                                         (logf(E1) - logf(E0));
            R[i] = expf(log_R);
            float log_S = logf(S0) + (logf(S1) - logf(S0)) * (logf(E) - logf(E0)) /
                                         (logf(E1) - logf(E0));
            S[i] = expf(log_S);
        }
        lE[i] = logf(rep[i]);
        lR[i] = logf(R[i]);
        lS[i] = logf(S[i]);
    }
}

// Host RLUT (src/lut/r_lut.cpp:131-205) over its own EnergyGrid, for the FE precompute.
struct HostLut {
    int N;
    float E_min, E_max;
    std::vector<float> edges, rep, R, S, lE, lR, lS;
    int find_bin(float E) const {  // EnergyGrid::FindBin, src/core/grids.cpp:98-112
        if (E < edges[0]) return 0;
        if (E >= edges[N]) return N - 1;
        int lo = 0, hi = N;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (edges[mid + 1] <= E) lo = mid + 1; else hi = mid;
        }
        return lo;
    }
    float lookup_R(float E) const {
        float Ec = fmaxf(E_min, fminf(E, E_max));
        int bin = find_bin(Ec);
        if (bin == N - 1 || Ec >= E_max) return R[N - 1];
        float lv = logf(Ec);
        float r = lR[bin] + (lR[bin + 1] - lR[bin]) * (lv - lE[bin]) / (lE[bin + 1] - lE[bin]);
        return expf(r);
    }
    float E_inverse(float Rin) const {
        if (Rin <= 0.0f) return E_min;
        if (Rin >= R.back()) return E_max;
        if (Rin <= R.front()) return E_min;
        int it = (int)(std::lower_bound(R.begin(), R.end(), Rin) - R.begin());
        int bin = std::max(0, std::min(it, N - 2));
        float lv = logf(Rin);
        float d = lR[bin + 1] - lR[bin];
        if (fabsf(d) < 1e-10f) return expf(lE[bin]);
        return expf(lE[bin] + (lE[bin + 1] - lE[bin]) * (lv - lR[bin]) / d);
    }
    // compute_energy_deposition, src/include/physics/step_control.hpp:69-84 (R_new<=0 -> 0).
    float deposition(float E, float ds) const {
        float Rn = lookup_R(E) - ds;
        float En = (Rn <= 0) ? 0.0f : E_inverse(Rn);
        return E - En;
    }
};

void grid_log(float E_min, float E_max, int N_E, float* edges, float* rep) {
    float lmin = logf(E_min), lmax = logf(E_max);
    float d = (lmax - lmin) / N_E;
    for (int i = 0; i <= N_E; ++i) edges[i] = expf(lmin + i * d);
    for (int i = 0; i < N_E; ++i) rep[i] = sqrtf(edges[i] * edges[i + 1]);
}

HostLut make_log_lut(float E_min, float E_max, int N) {
    HostLut L;
    L.N = N;
    L.E_min = E_min;
    L.E_max = E_max;
    L.edges.resize(N + 1);
    L.rep.resize(N);
    L.R.resize(N); L.S.resize(N); L.lE.resize(N); L.lR.resize(N); L.lS.resize(N);
    grid_log(E_min, E_max, N, L.edges.data(), L.rep.data());
    build_rlut(L.rep.data(), N, L.R.data(), L.S.data(), L.lE.data(), L.lR.data(), L.lS.data());
    return L;
}

// Host highland_sigma, src/include/physics/highland.hpp:59-80.
float host_highland(float E, float ds, float X0 = 360.8f) {
    const float m = 938.272f;
    float gamma = (E + m) / m;
    float beta = sqrtf(fmaxf(1.0f - 1.0f / (gamma * gamma), 0.0f));
    float p = sqrtf(fmaxf((E + m) * (E + m) - m * m, 0.0f));
    float t = ds / X0;
    if (t < 1e-6f) return 0.0f;
    float b = fmaxf(1.0f + 0.038f * logf(t), 0.25f);
    return (13.6f * 1.0f / (beta * p)) * sqrtf(t) * b;
}

// ---------------------------------------------------------------------------------------
// FP64 build (-DORC_FP64 -> _build/liboracle_sm2d_fp64.so).  From here to the C entry
// points every binary32 quantity of the transport step is binary64: phase-space weights,
// per-component physics, emission and spread weights, the FP32 weight tallies and the
// tables the context holds.  The transcendentals become the C library's double functions.
// Inputs stay the binary32 values the FP32 path gets: grids, range table, Fermi-Eyges
// moments, the injected phase space and the model's constants (the float literals).  So
// this build is the same algorithm evaluated in double, and the FP32 results' error is
// measured against it.
// ---------------------------------------------------------------------------------------
#ifdef ORC_FP64
#define float double
#define fmaxf fmax
#define fminf fmin
#define fabsf fabs
#define floorf floor
#define expf_(x) std::exp((double)(x))
#define logf_(x) std::log((double)(x))
#define sqrtf_(x) std::sqrt((double)(x))
#define sinf_(x) std::sin((double)(x))
#define cosf_(x) std::cos((double)(x))
#define erff_(x) std::erf((double)(x))
#endif
using real_t = float;  // the transport step's arithmetic type (double in the FP64 build)

// ---------------------------------------------------------------------------------------
// Problem description and phase-space storage
// ---------------------------------------------------------------------------------------
struct Psi {
    int N = 0, LB = 0;
    std::vector<uint32_t> bid;        // [N*KB]
    std::vector<uint8_t> nslot;       // slots 0..nslot-1 are occupied
    std::unique_ptr<float[]> val;     // [N*KB*LB], a slot's values are zeroed when claimed
    std::vector<int> occ;             // cells with nslot > 0 (for clearing)
    void init(int n, int lb) {
        N = n; LB = lb;
        bid.assign((size_t)N * KB, EMPTY);
        nslot.assign(N, 0);
        val.reset(new float[(size_t)N * KB * LB]);
        occ.clear();
    }
    float* v(int cell, int slot) { return &val[((size_t)cell * KB + slot) * LB]; }
    const float* v(int cell, int slot) const { return &val[((size_t)cell * KB + slot) * LB]; }
    uint32_t* b(int cell) { return &bid[(size_t)cell * KB]; }
    void clear() {
        for (int c : occ) {
            for (int s = 0; s < nslot[c]; ++s) bid[(size_t)c * KB + s] = EMPTY;
            nslot[c] = 0;
        }
        occ.clear();
    }
    // Claims the next empty slot of `cell` for `b_id` (the reference's CAS scan claims the
    // first EMPTY slot, so occupied slots always form a prefix).
    int claim(int cell, uint32_t b_id) {
        int s = nslot[cell];
        if (s >= KB) return -1;
        bid[(size_t)cell * KB + s] = b_id;
        std::memset(v(cell, s), 0, sizeof(float) * LB);
        if (s == 0) occ.push_back(cell);
        nslot[cell] = (uint8_t)(s + 1);
        return s;
    }
};

// device_bid_distance, src/cuda/device/device_psic.cuh:192-203.
int bid_distance(uint32_t a, uint32_t b) {
    int at = (int)(a & 0xFFF), ae = (int)((a >> 12) & 0xFFF);
    int bt = (int)(b & 0xFFF), be = (int)((b >> 12) & 0xFFF);
    return std::abs(ae - be) * 8 + std::abs(at - bt);
}

// Slot search of device_psic_find_or_allocate_slot_with_fallback (device_psic.cuh:208-260)
// and device_bucket_find_or_allocate_slot_with_fallback (device_bucket.cuh:96-128); with a
// single writer both reduce to: exact match, else first empty, else closest occupied (first
// minimum in slot order).  `claim` claims slot s for `want`.
template <class Claim>
int find_or_alloc(const uint32_t* bids, int nocc, uint32_t want, uint32_t& selected, Claim claim) {
    selected = want;
    for (int s = 0; s < nocc; ++s)
        if (bids[s] == want) return s;
    if (nocc < KB) return claim(nocc, want);
    int best = -1, bestd = 0x7fffffff;
    for (int s = 0; s < KB; ++s) {
        if (bids[s] == EMPTY) continue;
        int d = bid_distance(bids[s], want);
        if (d < bestd) { bestd = d; best = s; }
    }
    if (best >= 0) selected = bids[best];
    return best;
}

struct Bucket {  // DeviceOutflowBucket (device_bucket.cuh:48-57) without the unused fields
    uint32_t bid[KB];
    int n;
    float* val;  // [KB*LB]
};

}  // namespace

struct orc_ctx {
    sm2d_grid g{};
    int LB = 0, Tl = 0, El = 0, Xs = 0, Zs = 0, N = 0;
    std::vector<float> th_edges, E_edges, R, S, lE, lR, lS;
    std::vector<float> feA, feB, rho, x0r;
    // learned physics table of the fine transport (same layout and lookup as the library)
    std::vector<float> surr;
    std::vector<int> rowmat;
    int surr_nodes = 0, surr_on = 0, surr_active = 0;
    float sdev_max = 0;  // monitor maximum of the current iteration
    Psi psi[2];
    int in = 0;  // index of psi_in
    std::vector<uint8_t> mask, mask_prev;
    // cumulative per-cell tallies (src/cuda/k1k6_pipeline.cuh:174-180)
    std::vector<double> edep, E_cut, E_nuc, E_bnd;
    std::vector<float> W_cut, W_nuc, W_bnd;
    // source accounting (gpu_transport_wrapper.cu:339-468)
    double src_W_in = 0, src_W_out = 0, src_W_slot = 0, src_E_in = 0, src_E_out = 0,
           src_E_slot = 0, src_E_repr = 0;
    // buckets: dense per (cell, face), lazily zeroed
    std::vector<Bucket> bk;
    std::unique_ptr<float[]> bkval;
    std::vector<int> phase_of;  // 1 = listed in current phase
    int n_threads = 1;
    uint64_t cross_cell = 0;  // lateral emissions into another listed cell's +z bucket
    bool water = true;

    Psi& pin() { return psi[in]; }
    Psi& pout() { return psi[1 - in]; }

    // ---- grid helpers -------------------------------------------------------------------
    float ecenter(int b) const {  // bin centre, k3_finetransport.cu:206-209
        float lo = E_edges[b], hi = E_edges[b + 1];
        float half = (hi - lo) * 0.5f;
        return lo + 1.00f * half;
    }
    // device_find_bin_edges, src/cuda/device/device_lut.cuh:34-49
    int find_bin(float E) const {
        const float* e = E_edges.data();
        int NE = g.N_E;
        if (E <= e[0]) return 0;
        if (E >= e[NE]) return NE - 1;
        int lo = 0, hi = NE;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (e[mid + 1] <= E) lo = mid + 1; else hi = mid;
        }
        return lo;
    }
    uint16_t enc(int t, int e, int x, int z) const { return (uint16_t)(t + Tl * (e + El * (x + Xs * z))); }
    void dec(int l, int& t, int& e, int& x, int& z) const {
        t = l % Tl; int r = l / Tl; e = r % El; r /= El; x = r % Xs; z = r / Xs;
    }
    // get_x_offset_from_bin / get_z_offset_from_bin (local_bins.hpp:80-83,103-106), written
    // for a general number of sub-bins; exact (power-of-two sub-bin counts) so it equals the
    // reference's dx*(-0.4375f + 0.125f*x_sub) at N_x_sub = 8.
    float xoff(int xs) const { return g.dx * ((float)(2 * xs + 1 - Xs) / (float)(2 * Xs)); }
    float zoff(int zs) const { return g.dz * ((float)(2 * zs + 1 - Zs) / (float)(2 * Zs)); }
    int xbin(float off) const {  // get_x_sub_bin, local_bins.hpp:85-92
        float nrm = (off / g.dx) + 0.5f;
        int b = (int)(nrm * Xs);
        return b < 0 ? 0 : (b >= Xs ? Xs - 1 : b);
    }
    int zbin(float off) const {
        float nrm = (off / g.dz) + 0.5f;
        int b = (int)(nrm * Zs);
        return b < 0 ? 0 : (b >= Zs ? Zs - 1 : b);
    }

    // ---- device LUT (src/cuda/device/device_lut.cuh) ---------------------------------------
    float E_min() const { return E_edges[0]; }
    float E_max() const { return E_edges[g.N_E]; }
    float lookup_R(float E) const {  // :66-85
        int NE = g.N_E;
        float Ec = fmaxf(E_min(), fminf(E, E_max()));
        int bin = find_bin(Ec);
        if (bin == NE - 1 || Ec >= E_max()) return R[NE - 1];
        float lv = logf_(Ec);
        float r = lR[bin] + (lR[bin + 1] - lR[bin]) * (lv - lE[bin]) / (lE[bin + 1] - lE[bin]);
        return expf_(r);
    }
    float lookup_S(float E) const {  // :88-107
        int NE = g.N_E;
        float Ec = fmaxf(E_min(), fminf(E, E_max()));
        int bin = find_bin(Ec);
        if (bin == NE - 1 || Ec >= E_max()) return S[NE - 1];
        float lv = logf_(Ec);
        float r = lS[bin] + (lS[bin + 1] - lS[bin]) * (lv - lE[bin]) / (lE[bin + 1] - lE[bin]);
        return expf_(r);
    }
    float E_inverse(float Rin) const {  // :110-138 (no upper clamp, lower_bound index)
        if (Rin <= 0.0f) return E_min();
        int NE = g.N_E;
        int lo = 0, hi = NE;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (R[mid] < Rin) lo = mid + 1; else hi = mid;
        }
        int bin = std::max(0, std::min(lo, NE - 2));
        float lv = logf_(Rin);
        float d = lR[bin + 1] - lR[bin];
        if (fabsf(d) < 1e-10f) return expf_(lE[bin]);
        return expf_(lE[bin] + (lE[bin + 1] - lE[bin]) * (lv - lR[bin]) / d);
    }
    // device_compute_energy_after_step (:145-151).  In a material of density rho a geometric
    // step ds consumes rho*ds of water-equivalent range (rho = 1: the reference).
    float energy_after(float E, float ds, float rh) const {
        float Rn = lookup_R(E) - rh * ds;
        if (Rn <= 0) return E_min();
        return E_inverse(Rn);
    }
    float deposition(float E, float ds, float rh) const { return E - energy_after(E, ds, rh); }
    // device_compute_max_step (:168-214); the water range is converted to geometric mm.
    float max_step(float E, float rh, int rule) const {
        float Rw = lookup_R(E) / rh;
        if (rule == 1) {  // compute_max_step_physics, step_control.hpp:18-66
            float d = 0.02f * Rw, f = 1.0f;
            if (E < 5.0f) { f = 0.2f; d = fminf(d, 0.1f); }
            else if (E < 10.0f) { f = 0.3f; d = fminf(d, 0.2f); }
            else if (E < 20.0f) { f = 0.5f; d = fminf(d, 0.5f); }
            else if (E < 50.0f) { f = 0.7f; d = fminf(d, 0.7f); }
            d = d * f;
            return fmaxf(d, 0.05f);
        }
        float d = 0.01f * Rw;
        float fl;
        if (E < 2.0f) fl = 0.1f;
        else if (E < 20.0f) fl = 0.25f;
        else if (E < 100.0f) fl = 0.5f;
        else fl = 1.0f;
        d = fmaxf(d, fl);
        if (E < 0.5f) d = fminf(d, 0.05f);
        else if (E < 5.0f) d = fminf(d, 0.2f);
        else if (E < 20.0f) d = fminf(d, 0.5f);
        return d;
    }
};

namespace {

// Learned straggling width, scattering width and attenuation exponent from the table (the
// library's surrogate_lookup, restated).
bool surrogate_lookup(const orc_ctx* c, int ebin, int iz, float dR, float& sE, float& sth, float& a) {
    if (c->surr.empty()) return false;
    const int m = c->rowmat[iz];
    const float DR_LO = 1e-4f, DR_HI = 10.0f;
    if (m < 0 || !(dR >= DR_LO && dR <= DR_HI)) return false;
    const float lo = logf_(DR_LO), hi = logf_(DR_HI);
    const float pj = (logf_(dR) - lo) / (hi - lo) * (float)(c->surr_nodes - 1);
    const int j = std::min((int)pj, c->surr_nodes - 2);
    const float f = pj - (float)j;
    const float* t0 = c->surr.data() + ((size_t)(m * c->g.N_E + ebin) * c->surr_nodes + j) * 3;
    const float* t1 = t0 + 3;
    const float lsE = t0[0] + f * (t1[0] - t0[0]);
    const float lsT = t0[1] + f * (t1[1] - t0[1]);
    const float aa = t0[2] + f * (t1[2] - t0[2]);
    if (!(lsE == lsE && lsT == lsT && aa == aa)) return false;
    sE = expf_(lsE);
    sth = expf_(lsT);
    a = fmaxf(aa, 0.0f);
    return true;
}

// ---- device physics (src/cuda/device/device_physics.cuh) ---------------------------------
constexpr float MP = 938.272f, ME = 0.511f, X0W = 360.8f, KV = 0.307f, ZA = 0.555f;

float highland(float E, float ds, float X0) {  // :50-71
    float gamma = (E + MP) / MP;
    float beta2 = 1.0f - 1.0f / (gamma * gamma);
    float beta = sqrtf_(fmaxf(beta2, 0.0f));
    float p = sqrtf_(fmaxf((E + MP) * (E + MP) - MP * MP, 0.0f));
    float t = ds / X0;
    if (t < 1e-6f) return 0.0f;
    float bracket = fmaxf(1.0f + 0.038f * logf_(t), 0.25f);
    return (13.6f * 1.0f / (beta * p)) * sqrtf_(t) * bracket;
}
float scattering_power(float E, float ds, float X0) {  // :394-404
    float th = highland(E, ds, X0);
    if (ds < 1e-6f) return 0.0f;
    return th * th / ds;
}
float gaussian(unsigned& seed) {  // device_sample_gaussian :74-87
    seed = 1664525u * seed + 1013904223u;
    float u1 = (seed >> 16) / 65536.0f;
    seed = 1664525u * seed + 1013904223u;
    float u2 = (seed >> 16) / 65536.0f;
    u1 = fmaxf(u1, 1e-6f);
    return sqrtf_(-2.0f * logf_(u1)) * cosf_(2.0f * (float)M_PI * u2);
}
float vavilov_xi(float beta, float rh, float ds_cm) {  // :114-117
    return (KV / 2.0f) * ZA * (1.0f * 1.0f / (beta * beta)) * rh * ds_cm;
}
float straggling_sigma(float E, float ds, float rh) {  // :157-184
    float gamma = (E + MP) / MP;
    float beta2 = 1.0f - 1.0f / (gamma * gamma);
    float beta = sqrtf_(fmaxf(beta2, 0.0f));
    float ds_cm = ds / 10.0f;
    float xi = vavilov_xi(beta, rh, ds_cm);
    float m_ratio = ME / MP;  // device_vavilov_T_max :120-125
    float num = 2.0f * ME * beta * beta * gamma * gamma;
    float den = 1.0f + 2.0f * gamma * m_ratio + m_ratio * m_ratio;
    float T_max = num / fmaxf(den, 1e-10f);
    float kappa = fmaxf(xi / fmaxf(T_max, 1e-10f), 1e-6f);
    auto bohr = [&]() {  // :143-153
        float g2 = (E + MP) / MP;
        float b2 = 1.0f - 1.0f / (g2 * g2);
        float b = sqrtf_(fmaxf(b2, 0.01f));
        return 0.156f * sqrtf_(rh * ds) / b;
    };
    if (kappa > 10.0f) return bohr();
    float b = sqrtf_(fmaxf(1.0f - 1.0f / (gamma * gamma), 0.0f));
    float x = vavilov_xi(b, rh, ds / 10.0f);
    float landau = 4.0f * x / 2.355f;
    if (kappa < 0.01f) return landau;
    float sb = bohr();
    float w = 1.0f / (1.0f + kappa);
    return w * landau + (1.0f - w) * sb;
}
float nuclear_sigma(float E) {  // :201-215
    if (E < 5.0f) return 0.0f;
    if (E >= 20.0f) return 0.0012f * fmaxf(1.0f - 0.15f * logf_(E / 100.0f), 0.4f);
    return 0.0016f * ((E - 5.0f) / 15.0f);
}
float cdf(float x, float mu, float s) {  // device_gaussian_cdf :245-251
    if (s < 1e-10f) return (x >= mu) ? 1.0f : 0.0f;
    const float inv_sqrt2 = 0.70710678f;
    float arg = (x - mu) / s * inv_sqrt2;
    float c = 0.5f * (1.0f + erff_(arg));
    return fminf(1.0f, fmaxf(0.0f, c));
}
// device_gaussian_spread_weights (:264-306)
void spread_weights(float* w, float xm, float s, float dx, int N) {
    s = fmaxf(s, 1e-6f);
    float x_min = xm - 0.5f * (float)N * dx;
    float prev = cdf(x_min, xm, s);
    for (int i = 0; i < N; ++i) {
        float cur = cdf(x_min + (i + 1) * dx, xm, s);
        w[i] = cur - prev;
        prev = cur;
    }
    float sum = 0.0f;
    for (int i = 0; i < N; ++i) sum += w[i];
    if (sum > 1e-10f) { for (int i = 0; i < N; ++i) w[i] /= sum; }
    else { for (int i = 0; i < N; ++i) w[i] = (i == N / 2) ? 1.0f : 0.0f; }
}
// device_gaussian_spread_weights_subcell (:321-364)
void spread_weights_sub(float* w, float xm, float s, float dx, int N) {
    s = fmaxf(s, 1e-6f);
    float dxs = dx / N;
    float x_min = -dx * 0.5f;
    float prev = cdf(x_min, xm, s);
    for (int i = 0; i < N; ++i) {
        float cur = cdf(x_min + (i + 1) * dxs, xm, s);
        w[i] = cur - prev;
        prev = cur;
    }
    float sum = 0.0f;
    for (int i = 0; i < N; ++i) sum += w[i];
    if (sum > 1e-10f) { for (int i = 0; i < N; ++i) w[i] /= sum; }
    else { for (int i = 0; i < N; ++i) w[i] = (i == N / 2) ? 1.0f : 0.0f; }
}

enum Face { ZP = 0, ZM = 1, XP = 2, XM = 3 };

// Run-time monitor of the learned physics (the library's K5 drift check): every
// SURR_MONITOR_EVERY-th cell of a spot compares its learned widths with the plain ones; a largest
// deviation above SURR_DRIFT_TOL, a weight balance off by 1e-6 or a non-finite energy balance
// switches the following iterations to the plain physics.
constexpr int SURR_MONITOR_EVERY = 8;
constexpr float SURR_DRIFT_TOL = 0.5f;

// Deviation of learned (sigma_E, sigma_theta, a) from the plain physics of the same substep: the
// largest of the |ln sigma| differences and |a - a_plain| / (a_plain + 1e-4 / mm x rh ds), the
// criterion the library also applies to each table node before it accepts the table.
float learned_deviation(const orc_ctx* c, float E, float ds, float rh, float X0, float mdE, float sE,
                        float sth, float a) {
    const float sE_p = straggling_sigma(E, ds, rh);
    const float Em_p = fmaxf(E - 0.5f * mdE, c->E_min());
    const float sth_p = sqrtf_(fmaxf(scattering_power(Em_p, ds, X0) * ds, 0.0f));
    const float a_p = nuclear_sigma(E) * rh * ds;
    const float d0 = fabsf(logf_(fmaxf(sE, 1e-30f)) - logf_(fmaxf(sE_p, 1e-30f)));
    const float d1 = fabsf(logf_(fmaxf(sth, 1e-30f)) - logf_(fmaxf(sth_p, 1e-30f)));
    const float d2 = fabsf(a - a_p) / (a_p + 1e-4f * rh * ds);
    return fmaxf(d0, fmaxf(d1, d2));
}

// Per-cell results of one transport pass (the seven accumulators of
// k3_finetransport.cu:140-146 plus drop channels and closure sums).
struct CellAcc {
    double edep = 0, E_cut = 0, E_nuc = 0, E_bnd = 0;
    float w_cut = 0, w_nuc = 0, w_bnd = 0;
    double drop_w = 0, drop_E = 0, prune_w = 0, prune_E = 0;
    double cl_w_in = 0, cl_w_out = 0, cl_E_in = 0, cl_E_out = 0;
    uint64_t comps = 0, updates = 0, surr = 0;
    float sdev = 0;  // largest learned-vs-plain deviation of the cell's substeps (monitored cells)
};

// The library applies a cell's components in rounds of whole slots (up to BATCH_ROUND
// components) and batches of 64 consecutive components of a round; each batch adds to every
// accumulator the sum of its components' contributions, formed by a fixed butterfly over the 64
// positions (x[i] + x[i ^ o] for o = 32, 16, ..., 1; absent components contribute zero).
constexpr int BATCH_ROUND = 512;

template <class T>
T butterfly64(const T* v, int n) {
    T x[64], y[64];
    for (int i = 0; i < 64; ++i) x[i] = i < n ? v[i] : T(0);
    for (int o = 32; o > 0; o >>= 1) {
        for (int i = 0; i < 64; ++i) y[i] = x[i] + x[i ^ o];
        for (int i = 0; i < 64; ++i) x[i] = y[i];
    }
    return x[0];
}

struct Batch {
    CellAcc v[64];
    int n = 0;
    CellAcc& next() {
        v[n] = CellAcc();
        return v[n++];
    }
    template <class T>
    T sum(T CellAcc::*f) const {
        T x[64];
        for (int i = 0; i < n; ++i) x[i] = v[i].*f;
        return butterfly64(x, n);
    }
    void flush(CellAcc& a) {
        if (!n) return;
        a.edep += sum(&CellAcc::edep);
        a.E_cut += sum(&CellAcc::E_cut);
        a.E_nuc += sum(&CellAcc::E_nuc);
        a.E_bnd += sum(&CellAcc::E_bnd);
        a.w_cut += sum(&CellAcc::w_cut);
        a.w_nuc += sum(&CellAcc::w_nuc);
        a.w_bnd += sum(&CellAcc::w_bnd);
        a.drop_w += sum(&CellAcc::drop_w);
        a.drop_E += sum(&CellAcc::drop_E);
        a.cl_w_in += sum(&CellAcc::cl_w_in);
        a.cl_w_out += sum(&CellAcc::cl_w_out);
        a.cl_E_in += sum(&CellAcc::cl_E_in);
        a.cl_E_out += sum(&CellAcc::cl_E_out);
        n = 0;
    }
};

}  // namespace

// ---------------------------------------------------------------------------------------
// One transport phase (K2 coarse for `fine` = false, K3 fine for `fine` = true)
// ---------------------------------------------------------------------------------------
namespace {

// Angular split of an emitted component (device_emit_component_to_bucket_4d,
// device_bucket.cuh:452-476): the nearest theta-bin centre keeps 1 - f^2 of the weight and the
// neighbouring bin on the side of theta takes f^2, f = |theta - centre| / dtheta (none at the
// grid's edges).  The split conserves weight; its mean angle is centre + sign * f^2 * dtheta.
void theta_split(const orc_ctx* c, float theta, int tb[2], float tw[2]) {
    const float* te = c->th_edges.data();
    int NT = c->g.N_theta;
    float tmin = te[0], tmax = te[NT];
    float dth = (tmax - tmin) / NT;
    float tc = fmaxf(tmin, fminf(theta, tmax));
    float c0 = tmin + 0.5f * dth;
    float cont = (tc - c0) / dth;
    int bc = (int)floorf(cont + 0.5f);
    bc = std::max(0, std::min(bc, NT - 1));
    float tcen = c0 + (float)bc * dth;
    float delta = tc - tcen;
    tb[0] = tb[1] = bc;
    tw[0] = 1.0f;
    tw[1] = 0.0f;
    if (delta > 0.0f && bc < NT - 1) tb[1] = bc + 1;
    else if (delta < 0.0f && bc > 0) tb[1] = bc - 1;
    if (tb[1] != tb[0]) {
        float fr = fminf(1.0f, fabsf(delta) / dth);
        tw[1] = fr * fr;
        tw[0] = 1.0f - tw[1];
    }
}

// Seed of a component's angular draw (k3_finetransport.cu:355-366, k2_coarsetransport.cu:339-350).
unsigned theta_seed(int cell, int slot, int lidx, int ebin, unsigned path_seed) {
    return ((unsigned)(cell + 1) * 2654435761u) ^ ((unsigned)(slot + 1) * 2246822519u) ^
           ((unsigned)(lidx + 1) * 3266489917u) ^ (unsigned)(ebin + 17) ^ (path_seed * 747796405u);
}

struct Phase {
    orc_ctx* c;
    bool fine;
    const sm2d_transport* tp;
    std::vector<CellAcc> acc;  // per listed cell

    Bucket* bucket(int cell, int face) {
        if (cell < 0 || !c->phase_of[cell]) return nullptr;
        return &c->bk[(size_t)cell * 4 + face];
    }
    int bucket_claim(Bucket* b, int s, uint32_t want) {
        b->bid[s] = want;
        std::memset(b->val + (size_t)s * c->LB, 0, sizeof(float) * c->LB);
        b->n = s + 1;
        return s;
    }

    // device_emit_component_to_bucket_4d (device_bucket.cuh:431-533); returns dropped weight.
    float emit4d(Bucket* bk, float theta, float E, float weight, int xs, int zs, CellAcc& a) {
        if (weight <= 0.0f) return 0.0f;
        int NE = c->g.N_E;
        int tb[2];
        float tw[2];
        theta_split(c, theta, tb, tw);
        const float* ee = c->E_edges.data();
        float Ec = fmaxf(ee[0], fminf(E, ee[NE]));
        int eb = c->find_bin(Ec);
        eb = std::max(0, std::min(eb, NE - 1));
        int el = eb % c->El;
        uint32_t bE = (uint32_t)(eb / c->El);
        float dropped = 0.0f;
        for (int ti = 0; ti < 2; ++ti) {
            if (tw[ti] <= 0.0f) continue;
            float we = weight * tw[ti];
            int t = tb[ti];
            uint32_t bid = ((uint32_t)(t / c->Tl) & 0xFFF) | ((bE & 0xFFF) << 12);
            int tl = t % c->Tl;
            uint16_t lidx = c->enc(tl, el, xs, zs);
            uint32_t sel;
            int s = find_or_alloc(bk->bid, bk->n, bid, sel,
                                  [&](int sl, uint32_t w) { return bucket_claim(bk, sl, w); });
            if (s < 0) { dropped += we; continue; }
            uint16_t le = lidx;
            if (sel != bid) le = remap(lidx, bid, sel);
            bk->val[(size_t)s * c->LB + le] += we;
            a.cl_w_out += we;
            a.cl_E_out += (double)we * (double)c->ecenter(bin_E_of(sel, le));
        }
        return dropped;
    }
    int bin_E_of(uint32_t bid, int lidx) {
        int t, e, x, z;
        c->dec(lidx, t, e, x, z);
        int b = (int)((bid >> 12) & 0xFFF) * c->El + e;
        return std::max(0, std::min(b, c->g.N_E - 1));
    }
    // device_remap_lidx_to_bid (device_psic.cuh:265-292)
    uint16_t remap(uint16_t l, uint32_t src, uint32_t dst) {
        if (src == dst) return l;
        int t, e, x, z;
        c->dec(l, t, e, x, z);
        int tg = (int)(src & 0xFFF) * c->Tl + t, eg = (int)((src >> 12) & 0xFFF) * c->El + e;
        int td = tg - (int)(dst & 0xFFF) * c->Tl, ed = eg - (int)((dst >> 12) & 0xFFF) * c->El;
        td = std::max(0, std::min(td, c->Tl - 1));
        ed = std::max(0, std::min(ed, c->El - 1));
        return c->enc(td, ed, x, z);
    }

    // Lateral emissions into another listed cell's +z bucket, applied after every cell of the
    // phase has made its own emissions (apply_foreign).
    struct Foreign {
        int src_li, cell;
        float theta, E, w;
        int xs;
    };
    std::vector<Foreign> foreign;

    // device_emit_lateral_spread (device_bucket.cuh:857-987); returns dropped weight,
    // *bnd receives the out-of-domain weight.  The reference emits into a neighbour's +z bucket
    // with device atomics while that neighbour's own thread emits into it too, so the order of
    // a bucket's claims and FP32 adds is a thread race.  The order taken here, which the device
    // memory model allows (each bucket location orders its relaxed atomics independently), is:
    // a bucket's own cell first, then the emissions of other cells in ascending source-cell and
    // component order.  The emitted weight is booked to the source cell when it is emitted.
    float emit_lateral(int li, int cell, int tz, float theta, float E, float weight, float xo, float sx,
                       float* bnd, CellAcc& a) {
        if (weight <= 0.0f) return 0.0f;
        float dropped = 0.0f;
        sx = fmaxf(sx, 1e-6f);
        const int NS = 11;
        float w[NS];
        spread_weights(w, xo, sx, c->g.dx, NS);
        float sum = 0.0f;
        for (int i = 0; i < NS; ++i) { w[i] = fmaxf(0.0f, w[i]); sum += w[i]; }
        if (sum < 1e-12f) { for (int i = 0; i < NS; ++i) w[i] = 0.0f; w[NS / 2] = 1.0f; }
        else { for (int i = 0; i < NS; ++i) w[i] /= sum; }
        int Nx = c->g.Nx, Nz = c->g.Nz;
        int ix = cell % Nx, iz = cell / Nx;
        for (int i = 0; i < NS; ++i) {
            float wf = w[i];
            if (wf <= 0.0f) continue;
            int off = -(NS / 2) + i;
            int it = ix + off;
            if (it < 0 || it >= Nx || tz < 0 || tz >= Nz) { *bnd += weight * wf; continue; }
            int ec = it + iz * Nx;
            Bucket* bk = bucket(ec, ZP);
            if (!bk) { dropped += weight * wf; continue; }
            int xs = off < 0 ? c->Xs - 1 : (off > 0 ? 0 : c->Xs / 2);
            if (ec != cell) {
                c->cross_cell++;
                const float wl = weight * wf;
                int tb[2];
                float tw[2];
                theta_split(c, theta, tb, tw);
                const float* ee = c->E_edges.data();
                int eb = c->find_bin(fmaxf(ee[0], fminf(E, ee[c->g.N_E])));
                eb = std::max(0, std::min(eb, c->g.N_E - 1));
                for (int ti = 0; ti < 2; ++ti) {
                    if (tw[ti] <= 0.0f) continue;
                    const float we = wl * tw[ti];
                    a.cl_w_out += we;
                    a.cl_E_out += (double)we * (double)c->ecenter(eb);
                }
                foreign.push_back(Foreign{li, ec, theta, E, wl, xs});
                continue;
            }
            dropped += emit4d(bk, theta, E, weight * wf, xs, 0, a);
        }
        return dropped;
    }

    // The deferred lateral emissions, in production order (ascending source cell, component
    // order, target offset): each target bucket takes them after its own cell's emissions.
    void apply_foreign() {
        for (const Foreign& f : foreign) {
            CellAcc booked;  // already booked to the source cell when emitted
            float d = emit4d(bucket(f.cell, ZP), f.theta, f.E, f.w, f.xs, 0, booked);
            if (d > 0.0f) {  // a full bucket without an occupied slot: cannot happen (32 slots)
                acc[f.src_li].drop_w += (double)d;
                acc[f.src_li].drop_E += (double)(f.E * d);
            }
        }
        foreign.clear();
    }

    void run_cell(int li, int cell) {
        Psi& pi = c->pin();
        Psi& po = c->pout();
        CellAcc& a = acc[li];
        const sm2d_grid& g = c->g;
        const int Nx = g.Nx, NT = g.N_theta, NE = g.N_E;
        const float dx = g.dx, dz = g.dz;
        const float* te = c->th_edges.data();
        const float* ee = c->E_edges.data();
        const int iz = cell / Nx;
        const float rh = c->rho[iz];
        const float X0 = X0W * c->x0r[iz];
        const float half_dz = dz * 0.5f, half_dx = dx * 0.5f;
        int ns = pi.nslot[cell];
        a.updates += (uint64_t)ns * c->LB;
        Batch B;       // the accumulator contributions of the current batch, one per component
        int rpos = 0;  // the component's position in its round
        for (int slot = 0; slot < ns; ++slot) {
            uint32_t bid = pi.bid[(size_t)cell * KB + slot];
            if (bid == EMPTY) continue;
            uint32_t bT = bid & 0xFFF, bE = (bid >> 12) & 0xFFF;
            const float* vin = pi.v(cell, slot);
            int cnt = 0;
            for (int lidx = 0; lidx < c->LB; ++lidx) cnt += vin[lidx] > 0.0f ? 1 : 0;
            if (rpos + cnt > BATCH_ROUND) {  // the slot opens a new round
                B.flush(a);
                rpos = 0;
            }
            for (int lidx = 0; lidx < c->LB; ++lidx) {
                float weight = vin[lidx];
                if (!(weight > 0.0f)) continue;
                if (rpos % 64 == 0) B.flush(a);
                ++rpos;
                CellAcc& cc = B.next();
                a.comps++;
                int tl, el, xsub, zsub;
                c->dec(lidx, tl, el, xsub, zsub);
                int tbin = (int)bT * c->Tl + tl;
                int ebin = (int)bE * c->El + el;
                if (fine) {  // K3 clamps (k3_finetransport.cu:172-173); K2 does not
                    tbin = std::max(0, std::min(tbin, NT - 1));
                    ebin = std::max(0, std::min(ebin, NE - 1));
                }
                cc.cl_w_in += weight;
                float tmin = te[0], tmax = te[NT];
                float dth = (tmax - tmin) / NT;
                float E = c->ecenter(ebin);
                cc.cl_E_in += (double)weight * (double)E;
                if (fine) {
                    if (weight < 1e-15f) {  // k3_finetransport.cu:175-184
                        cc.drop_w += (double)weight;
                        cc.drop_E += (double)(E * weight);
                        cc.cl_w_out += weight;
                        cc.cl_E_out += (double)(E * weight);
                        continue;
                    }
                } else if (weight < 1e-12f) {  // k2_coarsetransport.cu:172-177
                    cc.drop_w += (double)weight;
                    cc.drop_E += (double)(E * weight);
                    cc.cl_w_out += weight;
                    cc.cl_E_out += (double)(E * weight);
                    continue;
                }
                float theta = fine ? te[tbin] + 0.5f * dth : tmin + (tbin + 0.5f) * dth;
                if (E <= 0.1f) {
                    cc.w_cut += weight;
                    cc.E_cut += E * weight;
                    cc.cl_w_out += weight;
                    cc.cl_E_out += (double)(E * weight);
                    continue;
                }
                float xc = c->xoff(xsub), zc = c->zoff(zsub);
                float mu = cosf_(theta), eta = sinf_(theta);
                float ds, dE, E_new, edep, w_new, w_rem, E_rem;
                float sigma_x, sigma_x_tr, sth;
                unsigned path_seed;
                if (fine) {
                    // K3 step, energy, nuclear, moments: k3_finetransport.cu:231-361
                    float s_zp = (mu > 0) ? (half_dz - zc) / mu : 1e30f;
                    float s_zm = (mu < 0) ? (-half_dz - zc) / mu : 1e30f;
                    float s_xp = (eta > 0) ? (half_dx - xc) / eta : 1e30f;
                    float s_xm = (eta < 0) ? (-half_dx - xc) / eta : 1e30f;
                    float sb = fminf(fminf(s_zp, s_zm), fminf(s_xp, s_xm));
                    sb = fmaxf(sb, 0.0f);
                    float sp = c->max_step(E, rh, tp->step_rule);
                    ds = fminf(sp, sb * 1.001f);
                    float mean_dE = c->deposition(E, ds, rh);
                    float s_sE = 0.0f, s_sth = 0.0f, s_a = 0.0f;
                    const bool use_s = c->surr_active && ds > 0.0f && ds / X0 >= 1e-6f &&
                                       surrogate_lookup(c, ebin, iz, rh * ds, s_sE, s_sth, s_a);
                    if (use_s) a.surr++;
                    if (use_s && cell % SURR_MONITOR_EVERY == 0)
                        a.sdev = fmaxf(a.sdev, learned_deviation(c, E, ds, rh, X0, mean_dE, s_sE, s_sth, s_a));
                    if (tp->enable_straggling) {
                        unsigned seed = (unsigned)((cell * 7 + slot * 13 + lidx * 17) ^ 0x5DEECE66DL);
                        float sE = use_s ? s_sE : straggling_sigma(E, ds, rh);
                        if (sE <= 0.0f) dE = mean_dE;
                        else dE = fmaxf(mean_dE + gaussian(seed) * sE, 0.0f);
                    } else {
                        dE = mean_dE;
                    }
                    dE = fminf(dE, E);
                    E_new = E - dE;
                    edep = dE * weight;
                    if (tp->enable_nuclear) {
                        float surv = use_s ? expf_(-s_a) : expf_(-(nuclear_sigma(E) * rh) * ds);
                        w_new = weight * surv;
                        w_rem = weight - w_new;
                        E_rem = w_rem * E;
                    } else {
                        w_new = weight; w_rem = 0.0f; E_rem = 0.0f;
                    }
                    float smm = fmaxf(ds, 0.0f);
                    float A = c->feA.empty() ? 0.0f : c->feA[iz];
                    float B = c->feB.empty() ? 0.0f : c->feB[iz];
                    float Em = fmaxf(E - 0.5f * dE, c->E_min());
                    float T = use_s ? (s_sth * s_sth) / smm : scattering_power(Em, smm, X0);
                    sth = use_s ? s_sth : sqrtf_(fmaxf(T * smm, 0.0f));
                    float dC = 2.0f * B * smm + A * smm * smm + T * smm * smm * smm / 3.0f;
                    dC = fmaxf(dC, 0.0f);
                    sigma_x = fmaxf(sqrtf_(dC), 0.01f);
                    sigma_x_tr = 0.001f;
                    path_seed = (unsigned)fminf((float)iz * dz * 1000.0f, 4294967295.0f);
                } else {
                    // K2: k2_coarsetransport.cu:135, 196-345
                    float cstep = fminf(tp->step_coarse, fminf(dx, dz));
                    float crs = cstep;
                    bool split = false;
                    if (E > tp->E_fine_on + 1e-6f) {
                        float Rc = c->lookup_R(E);
                        float Rf = c->lookup_R(fmaxf(tp->E_fine_on, c->E_min()));
                        float rtf = (Rc - Rf) / rh;
                        if (rtf > 0.0f && rtf < crs) { crs = rtf; cstep = crs; split = true; }
                    }
                    float mean_dE = c->deposition(E, crs, rh);
                    dE = fminf(mean_dE, E);
                    E_new = E - dE;
                    edep = dE * weight;
                    float surv = expf_(-(nuclear_sigma(E) * rh) * crs);
                    w_new = weight * surv;
                    w_rem = weight - w_new;
                    E_rem = w_rem * E;
                    if (split && E_new > tp->E_fine_on) {
                        float ex = E_new - tp->E_fine_on;
                        if (ex > 0.0f && w_new > 0.0f) edep += ex * w_new;
                        E_new = tp->E_fine_on;
                    }
                    float pstart = fmaxf(0.0f, (float)iz * dz + (zc + 0.5f * dz));
                    float smm = fmaxf(crs, 0.0f);
                    float st0 = 0.0f;
                    if (pstart > 0.0f) st0 = highland(E, pstart, X0);
                    float A_old = st0 * st0;
                    float B_old = 0.5f * A_old * pstart;
                    float sx0 = tp_sigma_x0;
                    float C_old = sx0 * sx0 + (A_old * pstart * pstart) / 3.0f;
                    float Em = fmaxf(E - 0.5f * dE, c->E_min());
                    float T = scattering_power(Em, smm, X0);
                    // device_fermi_eyges_step (device_physics.cuh:416-432)
                    float A_new = A_old + T * smm;
                    float C_new = C_old + 2.0f * B_old * smm + A_old * smm * smm +
                                  (1.0f / 3.0f) * T * smm * smm * smm;
                    A_new = fmaxf(A_new, 0.0f);
                    C_new = fmaxf(C_new, 0.0f);
                    sth = sqrtf_(fmaxf(A_new - A_old, 0.0f));
                    float dC = fmaxf(C_new - C_old, 0.0f);
                    float unres = fminf(0.35f, 0.04f + 0.003f * pstart);
                    float Ceff = dC + unres * C_old;
                    sigma_x = fmaxf(sqrtf_(fmaxf(Ceff, 0.0f)), 0.01f);
                    // device_lateral_spread_sigma (device_physics.cuh:374-377)
                    float lat = sinf_(fminf(sth, 1.57f)) * smm / 1.7320508f;
                    sigma_x_tr = fmaxf(lat, 0.01f);
                    path_seed = (unsigned)fminf(pstart * 1000.0f, 4294967295.0f);
                    ds = cstep;
                }
                // Angular diffusion (k3_finetransport.cu:355-366, k2_coarsetransport.cu:339-350)
                unsigned tseed = theta_seed(cell, slot, lidx, ebin, path_seed);
                float tsc = (sth <= 0.0f) ? 0.0f : gaussian(tseed) * sth;
                float th_new = theta + tsc;
                th_new = fmaxf(te[0], fminf(th_new, te[NT]));
                float mu_n = cosf_(th_new), eta_n = sinf_(th_new);
                float x_new = xc + eta_n * ds;
                float z_new = zc + mu_n * ds;
                // device_determine_exit_face (device_bucket.cuh:783-807)
                int face = -1;
                if (z_new >= half_dz - 0.001f) face = ZP;
                else if (z_new <= -half_dz + 0.001f) face = ZM;
                else if (x_new >= half_dx - 0.001f) face = XP;
                else if (x_new <= -half_dx + 0.001f) face = XM;
                x_new = fmaxf(-half_dx, fminf(x_new, half_dx));
                z_new = fmaxf(-half_dz, fminf(z_new, half_dz));

                cc.edep += edep;
                cc.w_nuc += w_rem;
                cc.E_nuc += E_rem;
                cc.cl_w_out += w_rem;
                cc.cl_E_out += (double)edep + (double)E_rem;
                if (E_new <= 0.1f) {
                    cc.w_cut += w_new;
                    cc.E_cut += E_new * w_new;
                    cc.cl_w_out += w_new;
                    cc.cl_E_out += (double)(E_new * w_new);
                    continue;
                }
                if (face >= 0) {
                    int nb = neighbor(cell, face);
                    if (nb < 0) {
                        cc.w_bnd += w_new;
                        cc.E_bnd += E_new * w_new;
                        cc.cl_w_out += w_new;
                        cc.cl_E_out += (double)(E_new * w_new);
                        continue;
                    }
                    float E_emit = c->ecenter(c->find_bin(E_new));
                    if (face == ZP) {
                        float lat_bnd = 0.0f;
                        float drop = emit_lateral(li, cell, iz + 1, th_new, E_new, w_new, x_new,
                                                  sigma_x_tr, &lat_bnd, cc);
                        if (lat_bnd > 0.0f) {
                            cc.w_bnd += lat_bnd;
                            cc.E_bnd += E_new * lat_bnd;
                            cc.cl_w_out += lat_bnd;
                            cc.cl_E_out += (double)(E_new * lat_bnd);
                        }
                        if (drop > 0.0f) {
                            cc.drop_w += (double)drop;
                            cc.drop_E += (double)(E_new * drop);
                            cc.cl_w_out += drop;
                            cc.cl_E_out += (double)(E_new * drop);
                        }
                        if (fine) {
                            float tw = fmaxf(0.0f, w_new - lat_bnd - drop);
                            if (tw > 0.0f) {
                                cc.edep += (double)((E_new - E_emit) * tw);
                                cc.cl_E_out += (double)((E_new - E_emit) * tw);
                            }
                        }
                    } else {
                        float xon = (face == XP) ? -dx * 0.5f : (face == XM ? dx * 0.5f : x_new);
                        int xsn = c->xbin(xon);
                        float zon;
                        if (face == ZM) zon = fine ? dz * 0.5f - dz * 0.125f : -dz * 0.5f + dz * 0.125f;
                        else zon = z_new;
                        int zsn = c->zbin(zon);
                        if (face == ZP) zsn = 0;
                        else if (face == ZM) zsn = c->Zs - 1;
                        Bucket* bk = bucket(cell, face);
                        float drop = emit4d(bk, th_new, E_new, w_new, xsn, zsn, cc);
                        if (drop > 0.0f) {
                            cc.drop_w += (double)drop;
                            cc.drop_E += (double)(E_new * drop);
                            cc.cl_w_out += drop;
                            cc.cl_E_out += (double)(E_new * drop);
                        }
                        if (fine) {
                            float tw = fmaxf(0.0f, w_new - drop);
                            if (tw > 0.0f) {
                                cc.edep += (double)((E_new - E_emit) * tw);
                                cc.cl_E_out += (double)((E_new - E_emit) * tw);
                            }
                        }
                    }
                    continue;
                }
                // Remain branch (k3_finetransport.cu:504-748, k2_coarsetransport.cu:460-691)
                int tbn = (int)((th_new - tmin) / dth);
                tbn = (int)fmaxf(0, fminf(tbn, NT - 1));
                int ebn = c->find_bin(E_new);
                if (!fine) ebn = (int)fmaxf(0, fminf(ebn, NE - 1));
                uint32_t bid_new = ((uint32_t)(ebn / c->El) << 12) | ((uint32_t)(tbn / c->Tl) & 0xFFF);
                uint32_t sel;
                int os = find_or_alloc(po.b(cell), po.nslot[cell], bid_new, sel,
                                       [&](int, uint32_t w) { return po.claim(cell, w); });
                if (os < 0) {
                    cc.drop_w += (double)w_new;
                    cc.drop_E += (double)(E_new * w_new);
                    cc.cl_w_out += w_new;
                    cc.cl_E_out += (double)(E_new * w_new);
                    continue;
                }
                int tbe = tbn, ebe = ebn;
                if (sel != bid_new) {
                    int bte = (int)(sel & 0xFFF), bee = (int)((sel >> 12) & 0xFFF);
                    tbe = std::max(bte * c->Tl, std::min(tbe, bte * c->Tl + c->Tl - 1));
                    ebe = std::max(bee * c->El, std::min(ebe, bee * c->El + c->El - 1));
                }
                int tln = tbe % c->Tl, eln = ebe % c->El;
                float E_emit = ee[ebe] + 0.5f * (ee[ebe + 1] - ee[ebe]);
                const int XS = c->Xs;
                float wsub[16];
                spread_weights_sub(wsub, x_new, sigma_x_tr_or_x(fine, sigma_x_tr, sigma_x), dx, XS);
                int ix = cell % Nx;
                int zst = c->zbin(z_new);
                float lb = -dx * 0.5f, rb = dx * 0.5f;
                float w_in = cdf(rb, x_new, sigma_x) - cdf(lb, x_new, sigma_x);
                float wcf = fminf(1.0f, fmaxf(0.0f, w_in));
                float wl = 0.0f, wr = 0.0f;
                float wss = 0.0f;
                for (int k = 0; k < XS; ++k) { wsub[k] = fmaxf(0.0f, wsub[k]); wss += wsub[k]; }
                if (wss < 1e-12f) {
                    for (int k = 0; k < XS; ++k) wsub[k] = 0.0f;
                    wsub[XS / 2] = 1.0f;
                    wss = 1.0f;
                }
                float* vo = po.v(cell, os);
                for (int k = 0; k < XS; ++k) {
                    float wf = (wsub[k] / wss) * wcf;
                    if (wf <= 0.0f) continue;
                    float wsp = w_new * wf;
                    vo[c->enc(tln, eln, k, zst)] += wsp;
                    cc.cl_w_out += wsp;
                    cc.cl_E_out += (double)wsp * (double)c->ecenter(ebe);
                }
                bool aborted = false;
                if (wcf < 1.0f - 1e-6f) {
                    float wlr = fmaxf(0.0f, cdf(lb, x_new, sigma_x));
                    float wrr = fmaxf(0.0f, 1.0f - cdf(rb, x_new, sigma_x));
                    float wtt = fmaxf(0.0f, 1.0f - wcf);
                    float wts = wlr + wrr;
                    if (wtt > 0.0f && wts > 1e-12f) {
                        float sc = wtt / wts;
                        wl = wlr * sc;
                        wr = wrr * sc;
                    }
                    if (wl > 0.0f) {
                        float wsl = w_new * wl;
                        if (ix > 0) {
                            Bucket* bk = bucket(cell, XM);
                            float d = emit4d(bk, th_new, E_new, wsl, XS - 1, c->zbin(z_new), cc);
                            if (d > 0.0f) {
                                cc.drop_w += (double)d;
                                cc.drop_E += (double)(E_new * d);
                                cc.cl_w_out += d;
                                cc.cl_E_out += (double)(E_new * d);
                            }
                        } else {
                            cc.w_bnd += wsl;
                            cc.E_bnd += E_new * wsl;
                            cc.cl_w_out += wsl;
                            cc.cl_E_out += (double)(E_new * wsl);
                        }
                    }
                    if (wr > 0.0f) {
                        float wsr = w_new * wr;
                        if (ix < Nx - 1) {
                            Bucket* bk = bucket(cell, XP);
                            float d = emit4d(bk, th_new, E_new, wsr, 0, c->zbin(z_new), cc);
                            if (d > 0.0f) {
                                cc.drop_w += (double)d;
                                cc.drop_E += (double)(E_new * d);
                                cc.cl_w_out += d;
                                cc.cl_E_out += (double)(E_new * d);
                            }
                        } else {
                            cc.w_bnd += wsr;
                            cc.E_bnd += E_new * wsr;
                            cc.cl_w_out += wsr;
                            cc.cl_E_out += (double)(E_new * wsr);
                        }
                    }
                }
                (void)aborted;
                if (fine) {
                    float rf = wcf;
                    if (ix > 0) rf += wl;
                    if (ix < Nx - 1) rf += wr;
                    rf = fminf(1.0f, fmaxf(0.0f, rf));
                    float rw = w_new * rf;
                    if (rw > 0.0f) {
                        cc.edep += (double)((E_new - E_emit) * rw);
                        cc.cl_E_out += (double)((E_new - E_emit) * rw);
                    }
                }
            }
        }
        B.flush(a);
    }
    static float sigma_x_tr_or_x(bool fine, float tr, float sx) { return fine ? tr : sx; }
    int neighbor(int cell, int face) {  // device_get_neighbor (device_bucket.cuh:810-829)
        int Nx = c->g.Nx, Nz = c->g.Nz;
        int ix = cell % Nx, iz = cell / Nx;
        switch (face) {
            case ZP: return iz + 1 >= Nz ? -1 : cell + Nx;
            case ZM: return iz <= 0 ? -1 : cell - Nx;
            case XP: return ix + 1 >= Nx ? -1 : cell + 1;
            case XM: return ix <= 0 ? -1 : cell - 1;
        }
        return -1;
    }
    float tp_sigma_x0 = 0.0f;
};

}  // namespace

// ---------------------------------------------------------------------------------------
// Pipeline pieces
// ---------------------------------------------------------------------------------------
namespace {

// K1_ActiveMask (src/cuda/kernels/k1_activemask.cu:6-48) over all cells.
void k1(orc_ctx* c, int b_on, int b_off, float wmin, const uint8_t* prev, uint8_t* mask) {
    Psi& p = c->pin();
    for (int cell = 0; cell < c->N; ++cell) {
        float W = 0;
        bool on = false, off = false;
        for (int s = 0; s < KB; ++s) {
            uint32_t bid = p.bid[(size_t)cell * KB + s];
            if (bid == EMPTY) continue;
            uint32_t bE = (bid >> 12) & 0xFFF;
            if (bE <= (uint32_t)b_on) on = true;
            if (bE <= (uint32_t)b_off) off = true;
            const float* v = p.v(cell, s);
            for (int l = 0; l < c->LB; ++l) W += v[l];
        }
        bool pa = prev != nullptr && prev[cell] != 0;
        bool need = on || (pa && off);
        mask[cell] = (need && W > wmin) ? 1 : 0;
    }
}

// compact_active_list / compact_coarse_list (k1k6_pipeline.cu:589-665) in ascending order.
void compact(orc_ctx* c, const uint8_t* mask, std::vector<int>& act, std::vector<int>& crs) {
    Psi& p = c->pin();
    act.clear();
    crs.clear();
    for (int cell = 0; cell < c->N; ++cell) {
        if (mask[cell] == 1) { act.push_back(cell); continue; }
        if (mask[cell] != 0) continue;
        bool has = false;
        for (int s = 0; s < KB && !has; ++s) {
            if (p.bid[(size_t)cell * KB + s] == EMPTY) continue;
            const float* v = p.v(cell, s);
            for (int l = 0; l < c->LB; ++l) if (v[l] > 0.0f) { has = true; break; }
        }
        if (has) crs.push_back(cell);
    }
}

// K4_BucketTransfer (src/cuda/kernels/k4_transfer.cu:42-201) over destination cells.
void k4(orc_ctx* c, const std::vector<int>& list, double& drop_w, double& drop_E) {
    Psi& po = c->pout();
    int Nx = c->g.Nx, Nz = c->g.Nz;
    // destinations: neighbours of listed cells, visited in ascending order
    std::vector<uint8_t> dest(c->N, 0);
    for (int cell : list) {
        int ix = cell % Nx, iz = cell / Nx;
        if (iz + 1 < Nz) dest[cell + Nx] = 1;
        if (iz > 0) dest[cell - Nx] = 1;
        if (ix + 1 < Nx) dest[cell + 1] = 1;
        if (ix > 0) dest[cell - 1] = 1;
    }
    for (int cell = 0; cell < c->N; ++cell) {
        if (!dest[cell]) continue;
        int ix = cell % Nx, iz = cell / Nx;
        for (int face = 0; face < 4; ++face) {
            int src = -1, sf = -1;
            if (face == 0 && iz > 0) { src = cell - Nx; sf = 0; }
            if (face == 1 && iz + 1 < Nz) { src = cell + Nx; sf = 1; }
            if (face == 2 && ix > 0) { src = cell - 1; sf = 2; }
            if (face == 3 && ix + 1 < Nx) { src = cell + 1; sf = 3; }
            if (src < 0 || !c->phase_of[src]) continue;
            Bucket& b = c->bk[(size_t)src * 4 + sf];
            for (int s = 0; s < b.n; ++s) {
                uint32_t bid = b.bid[s];
                if (bid == EMPTY) continue;
                uint32_t bE = (bid >> 12) & 0xFFF;
                uint32_t sel;
                int os = find_or_alloc(po.b(cell), po.nslot[cell], bid, sel,
                                       [&](int, uint32_t w) { return po.claim(cell, w); });
                const float* bv = b.val + (size_t)s * c->LB;
                if (os < 0) {
                    double dw = 0, de = 0;
                    for (int l = 0; l < c->LB; ++l) {
                        float w = bv[l];
                        if (w > 0.0f) {
                            dw += (double)w;
                            int t, e, x, z;
                            c->dec(l, t, e, x, z);
                            int eb = std::max(0, std::min((int)bE * c->El + e, c->g.N_E - 1));
                            de += (double)c->ecenter(eb) * (double)w;
                        }
                    }
                    drop_w += dw;
                    drop_E += de;
                    continue;
                }
                float* vo = po.v(cell, os);
                for (int l = 0; l < c->LB; ++l) {
                    float w = bv[l];
                    if (w > 0) {
                        int le = l;
                        if (sel != bid) {
                            int t, e, x, z;
                            c->dec(l, t, e, x, z);
                            int tg = (int)(bid & 0xFFF) * c->Tl + t, eg = (int)bE * c->El + e;
                            int td = std::max(0, std::min(tg - (int)(sel & 0xFFF) * c->Tl, c->Tl - 1));
                            int ed = std::max(0, std::min(eg - (int)((sel >> 12) & 0xFFF) * c->El, c->El - 1));
                            le = c->enc(td, ed, x, z);
                        }
                        vo[le] += w;
                    }
                }
            }
        }
    }
}

// Sum of weight and bin-centre energy over a buffer (k5_audit.cu:64-89).
void psi_totals(orc_ctx* c, Psi& p, double& W, double& E) {
    W = 0; E = 0;
    for (int cell : p.occ) {
        for (int s = 0; s < p.nslot[cell]; ++s) {
            uint32_t bid = p.bid[(size_t)cell * KB + s];
            if (bid == EMPTY) continue;
            uint32_t bE = (bid >> 12) & 0xFFF;
            const float* v = p.v(cell, s);
            for (int l = 0; l < c->LB; ++l) {
                float w = v[l];
                if (w <= 0.0f) continue;
                W += (double)w;
                int e = (l / c->Tl) % c->El;
                int eb = std::max(0, std::min((int)bE * c->El + e, c->g.N_E - 1));
                E += (double)w * (double)c->ecenter(eb);
            }
        }
    }
}

// K5_ConservationAudit (k5_audit.cu:22-167) with FP64 weight sums.
void k5(orc_ctx* c, const double* dW /*3*/, const double* dE /*4*/, double W_src_out,
        double W_src_slot, double E_src_out, double E_src_slot, double W_drop, double E_drop,
        int include_src, sm2d_k5_report* r) {
    std::memset(r, 0, sizeof(*r));
    r->struct_size = sizeof(*r);
    psi_totals(c, c->pin(), r->W_in, r->E_in);
    psi_totals(c, c->pout(), r->W_out, r->E_out);
    r->W_cutoff = dW[0]; r->W_nuclear = dW[1]; r->W_boundary = dW[2];
    r->E_dep = dE[0]; r->E_cutoff = dE[1]; r->E_nuclear = dE[2]; r->E_boundary = dE[3];
    r->W_transport_drop = std::max(0.0, W_drop);
    r->E_transport_drop = std::max(0.0, E_drop);
    if (include_src) {
        double a = std::max(0.0, W_src_out), b = std::max(0.0, W_src_slot);
        double e1 = std::max(0.0, E_src_out), e2 = std::max(0.0, E_src_slot);
        r->W_in += a + b;
        r->E_in += e1 + e2;
        r->W_source_out_of_grid = a; r->W_source_slot_drop = b;
        r->E_source_out_of_grid = e1; r->E_source_slot_drop = e2;
    }
    double Wr = r->W_out + r->W_cutoff + r->W_nuclear + r->W_boundary + r->W_transport_drop +
                r->W_source_out_of_grid + r->W_source_slot_drop;
    r->W_error = std::fabs(r->W_in - Wr) / std::max(r->W_in, 1e-20);
    r->W_pass = (std::isfinite(r->W_error) && r->W_error < 1e-6) ? 1 : 0;
    double Er = r->E_out + r->E_dep + r->E_cutoff + r->E_nuclear + r->E_boundary +
                r->E_transport_drop + r->E_source_out_of_grid + r->E_source_slot_drop;
    r->E_error = std::fabs(r->E_in - Er) / std::max(r->E_in, 1e-20);
    r->E_pass = (std::isfinite(r->E_error) && r->E_error < 1e-5) ? 1 : 0;
}

int b_threshold(orc_ctx* c, float E) {  // compute_b_E_threshold (k1k6_pipeline.cuh:419-422)
    const float* e = c->E_edges.data();
    int NE = c->g.N_E;
    int bin;
    if (E < e[0]) bin = 0;
    else if (E >= e[NE]) bin = NE - 1;
    else {
        int lo = 0, hi = NE;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (e[mid + 1] <= E) lo = mid + 1; else hi = mid;
        }
        bin = lo;
    }
    return bin / c->El;
}

// One transport phase over `list` (K2 or K3), followed by its K4.
void run_phase(orc_ctx* c, const sm2d_transport* tp, float sx0, const std::vector<int>& list,
               bool fine, sm2d_audit* au, double* it_dW, double* it_dE, double& drop_w,
               double& drop_E, double& cl_w, double& cl_E) {
    if (list.empty()) return;
    for (int cell : list) {
        c->phase_of[cell] = 1;
        for (int f = 0; f < 4; ++f) {
            Bucket& b = c->bk[(size_t)cell * 4 + f];
            b.n = 0;
            for (int s = 0; s < KB; ++s) b.bid[s] = EMPTY;
        }
    }
    Phase ph;
    ph.c = c;
    ph.fine = fine;
    ph.tp = tp;
    ph.tp_sigma_x0 = sx0;
    ph.acc.assign(list.size(), CellAcc());
    // Cells write disjoint ψ_out cells and their own buckets, so the loop parallelises; the
    // claim bookkeeping of ψ_out (occupied-cell list) is made thread-safe by a final rebuild.
    for (size_t i = 0; i < list.size(); ++i) ph.run_cell((int)i, list[i]);
    ph.apply_foreign();
    for (size_t i = 0; i < list.size(); ++i) {
        int cell = list[i];
        CellAcc& a = ph.acc[i];
        c->edep[cell] += a.edep;
        c->W_cut[cell] += a.w_cut;
        c->E_cut[cell] += a.E_cut;
        c->W_nuc[cell] += a.w_nuc;
        c->E_nuc[cell] += a.E_nuc;
        c->W_bnd[cell] += a.w_bnd;
        c->E_bnd[cell] += a.E_bnd;
        it_dW[0] += a.w_cut; it_dW[1] += a.w_nuc; it_dW[2] += a.w_bnd;
        it_dE[0] += a.edep; it_dE[1] += a.E_cut; it_dE[2] += a.E_nuc; it_dE[3] += a.E_bnd;
        drop_w += a.drop_w + a.prune_w;
        drop_E += a.drop_E + a.prune_E;
        if (a.cl_w_in > 0) cl_w = std::max(cl_w, std::fabs(a.cl_w_in - a.cl_w_out) / a.cl_w_in);
        if (fine && a.cl_E_in > 0) cl_E = std::max(cl_E, std::fabs(a.cl_E_in - a.cl_E_out) / a.cl_E_in);
        if (fine) c->sdev_max = fmaxf(c->sdev_max, a.sdev);
        if (au) {
            au->cell_bin_updates += a.updates;
            if (fine) au->fine_cell_bin_updates += a.updates;
            au->surrogate_substeps += a.surr;
            au->processed_cells += 1;
            au->processed_components += a.comps;
        }
    }
    k4(c, list, drop_w, drop_E);
    for (int cell : list) c->phase_of[cell] = 0;
}

}  // namespace

#ifdef ORC_FP64
#undef float
#undef fmaxf
#undef fminf
#undef fabsf
#undef floorf
#undef expf_
#undef logf_
#undef sqrtf_
#undef sinf_
#undef cosf_
#undef erff_
#endif

// ---------------------------------------------------------------------------------------
// C interface
// ---------------------------------------------------------------------------------------
extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_energy_grid_log(float E_min, float E_max, int N_E, float* edges, float* rep) {
    if (N_E <= 0 || !(E_min > 0) || !(E_max > E_min)) return fail("bad log grid");
    grid_log(E_min, E_max, N_E, edges, rep);
    return SM2D_OK;
}

// EnergyGrid::CreatePiecewise (src/core/grids.cpp:33-96)
int orc_energy_grid_piecewise(const float* groups, int ng, float* edges, float* rep, int* nb) {
    if (ng <= 0) return fail("no groups");
    int total = 0;
    for (int g = 0; g < ng; ++g)
        total += (int)((groups[3 * g + 1] - groups[3 * g]) / groups[3 * g + 2] + 0.5f);
    if (nb) *nb = total;
    if (!edges || !rep) return SM2D_OK;
    int off = 0;
    for (int g = 0; g < ng; ++g) {
        float s = groups[3 * g], e = groups[3 * g + 1], r = groups[3 * g + 2];
        int n = (int)((e - s) / r + 0.5f);
        for (int i = 0; i <= n; ++i) edges[off + i] = s + i * r;
        for (int i = 0; i < n; ++i) rep[off + i] = sqrtf(edges[off + i] * edges[off + i + 1]);
        off += n;
    }
    edges[total] = groups[3 * (ng - 1) + 1];
    return SM2D_OK;
}

// AngularGrid (src/core/grids.cpp:121-135)
int orc_angular_grid(float tmin, float tmax, int NT, float* edges) {
    if (NT <= 0) return fail("bad N_theta");
    float d = (tmax - tmin) / NT;
    for (int i = 0; i <= NT; ++i) edges[i] = tmin + i * d;
    return SM2D_OK;
}

int orc_build_rlut(const float* rep, int N_E, float* R, float* S, float* lE, float* lR, float* lS) {
    build_rlut(rep, N_E, R, S, lE, lR, lS);
    return SM2D_OK;
}

// init_pipeline_state FE precompute (src/cuda/k1k6_pipeline.cu:1061-1176): the first loop
// gives A and B; sigma_total comes from the corrected second pass (:1150-1176).
int orc_fe_precompute(float E0, float sx0, int Nz, float dz, float* A, float* B, float* st) {
    if (!(E0 > 0.0f)) {
        for (int i = 0; i < Nz; ++i) { A[i] = 0; B[i] = 0; st[i] = 0; }
        return SM2D_OK;
    }
    HostLut L = make_log_lut(0.1f, 300.0f, 256);
    for (int i = 0; i < Nz; ++i) { A[i] = 0.0f; B[i] = 0.0f; }
    float E = E0, Aa = 0.0f, Ba = 0.0f;
    for (int iz = 0; iz < Nz; ++iz) {
        A[iz] = Aa;
        B[iz] = Ba;
        if (E <= 0.5f) break;
        float dE = L.deposition(E, dz);
        dE = std::min(dE, E - 0.1f);
        float Em = E - 0.5f * dE;
        if (Em < 0.5f) Em = 0.5f;
        float s = host_highland(Em, dz);
        float T = s * s / dz;
        float Ao = Aa;
        Aa = Ao + T * dz;
        Ba = Ba + Ao * dz + 0.5f * T * dz * dz;
        E -= dE;
        if (E < 0.1f) E = 0.1f;
    }
    for (int i = 0; i < Nz; ++i) st[i] = sx0;
    float C = sx0 * sx0, Et = E0, At = 0.0f, Bt = 0.0f;
    for (int iz = 0; iz < Nz; ++iz) {
        st[iz] = sqrtf(std::max(0.0f, C));
        if (Et <= 0.5f) {
            for (int jz = iz + 1; jz < Nz; ++jz) st[jz] = st[iz];
            break;
        }
        float dE = L.deposition(Et, dz);
        dE = std::min(dE, Et - 0.1f);
        float Em = std::max(0.5f, Et - 0.5f * dE);
        float s = host_highland(Em, dz);
        float T = s * s / dz;
        float dC = 2.0f * Bt * dz + At * dz * dz + T * dz * dz * dz / 3.0f;
        C += dC;
        float Ao = At;
        At = Ao + T * dz;
        Bt = Bt + Ao * dz + 0.5f * T * dz * dz;
        Et -= dE;
        if (Et < 0.1f) Et = 0.1f;
    }
    return SM2D_OK;
}

// Post-processing MCS convolution (src/cuda/gpu_transport_wrapper.cu:551-616)
int orc_post_convolve(double* ed, int Nx, int Nz, float dx, const float* st, float sx0) {
    float s0 = sx0 * sx0;
    std::vector<double> row(Nx);
    for (int iz = 0; iz < Nz; ++iz) {
        float stz = st[iz];
        float sm2 = stz * stz - s0;
        if (sm2 <= 1e-6f) continue;
        float sm = std::sqrt(sm2);
        float sc = sm / dx;
        int hk = (int)std::ceil(3.0f * sc);
        hk = std::max(1, std::min(hk, Nx / 2));
        std::vector<double> k(2 * hk + 1);
        double ks = 0.0;
        for (int j = -hk; j <= hk; ++j) {
            double v = std::exp(-0.5 * ((double)j * j) / ((double)sc * sc));
            k[j + hk] = v;
            ks += v;
        }
        for (auto& v : k) v /= ks;
        for (int ix = 0; ix < Nx; ++ix) row[ix] = ed[iz * Nx + ix];
        for (int ix = 0; ix < Nx; ++ix) {
            double acc = 0.0;
            for (int j = -hk; j <= hk; ++j) {
                int jx = ix + j;
                if (jx >= 0 && jx < Nx) acc += row[jx] * k[j + hk];
            }
            ed[iz * Nx + ix] = acc;
        }
    }
    return SM2D_OK;
}

int orc_math_eval(int which, const float* x, float* y, int n) {
    for (int i = 0; i < n; ++i) {
        switch (which) {
            case 0: y[i] = expf_(x[i]); break;
            case 1: y[i] = logf_(x[i]); break;
            case 2: y[i] = sinf_(x[i]); break;
            case 3: y[i] = cosf_(x[i]); break;
            case 4: y[i] = erff_(x[i]); break;
            case 5: {  // sin, then cos, through the shared-reduction sincos
                float sn, cs;
                sm2d::sincosf_(x[i], sn, cs);
                y[i] = sn;
                break;
            }
            case 6: {
                float sn, cs;
                sm2d::sincosf_(x[i], sn, cs);
                y[i] = cs;
                break;
            }
            default: return fail("bad function");
        }
    }
    return SM2D_OK;
}

int orc_eval(orc_ctx* c, int which, const float* x, const float* y, float* out, int n) {
    for (int i = 0; i < n; ++i) {
        float a = x[i], b = y ? y[i] : 0.0f;
        switch (which) {
            case 0: out[i] = c->lookup_R(a); break;
            case 1: out[i] = c->lookup_S(a); break;
            case 2: out[i] = c->E_inverse(a); break;
            case 3: out[i] = c->max_step(a, 1.0f, 0); break;
            case 4: out[i] = c->max_step(a, 1.0f, 1); break;
            case 5: out[i] = c->energy_after(a, b, 1.0f); break;
            case 6: out[i] = highland(a, b, X0W); break;
            case 7: out[i] = straggling_sigma(a, b, 1.0f); break;
            case 8: out[i] = nuclear_sigma(a); break;
            case 9: out[i] = (float)c->find_bin(a); break;
            case 10: case 11: {  // theta split of an emission at angle a: neighbour weight, side
                int tb[2];
                real_t tw[2];
                theta_split(c, a, tb, tw);
                out[i] = which == 10 ? (float)tw[1] : (float)(tb[1] - tb[0]);
                break;
            }
            case 12: {  // the angular draw of component (cell a, slot 0, local bin b, E bin 0, path 0)
                unsigned sd = theta_seed((int)a, 0, (int)b, 0, 0u);
                out[i] = gaussian(sd);
                break;
            }
            default: return fail("bad eval id");
        }
    }
    return SM2D_OK;
}

int orc_ctx_create(const sm2d_grid* g, orc_ctx** out) {
    if (!g || g->struct_size != sizeof(sm2d_grid)) return fail("grid struct_size");
    if (g->Nx <= 0 || g->Nz <= 0 || g->N_theta <= 0 || g->N_E <= 0) return fail("bad grid");
    orc_ctx* c = new orc_ctx();
    c->g = *g;
    c->Tl = g->n_theta_local; c->El = g->n_e_local; c->Xs = g->n_x_sub; c->Zs = g->n_z_sub;
    c->LB = c->Tl * c->El * c->Xs * c->Zs;
    if (c->LB <= 0 || c->Xs > 16) { delete c; return fail("bad layout"); }
    c->N = g->Nx * g->Nz;
    c->psi[0].init(c->N, c->LB);
    c->psi[1].init(c->N, c->LB);
    c->mask.assign(c->N, 0);
    c->mask_prev.assign(c->N, 0);
    c->edep.assign(c->N, 0); c->E_cut.assign(c->N, 0); c->E_nuc.assign(c->N, 0); c->E_bnd.assign(c->N, 0);
    c->W_cut.assign(c->N, 0); c->W_nuc.assign(c->N, 0); c->W_bnd.assign(c->N, 0);
    c->bk.resize((size_t)c->N * 4);
    c->bkval.reset(new real_t[(size_t)c->N * 4 * KB * c->LB]);
    for (size_t i = 0; i < c->bk.size(); ++i) {
        c->bk[i].n = 0;
        c->bk[i].val = c->bkval.get() + i * KB * c->LB;
        for (int s = 0; s < KB; ++s) c->bk[i].bid[s] = EMPTY;
    }
    c->phase_of.assign(c->N, 0);
    c->rho.assign(g->Nz, 1.0f);
    c->x0r.assign(g->Nz, 1.0f);
    *out = c;
    return SM2D_OK;
}

int orc_ctx_destroy(orc_ctx* c) { delete c; return SM2D_OK; }
int orc_set_threads(orc_ctx* c, int n) { c->n_threads = n; return SM2D_OK; }

int orc_lut_upload(orc_ctx* c, const float* te, const float* ee, const float* R, const float* S,
                   const float* lE, const float* lR, const float* lS) {
    int NT = c->g.N_theta, NE = c->g.N_E;
    c->th_edges.assign(te, te + NT + 1);
    c->E_edges.assign(ee, ee + NE + 1);
    c->R.assign(R, R + NE); c->S.assign(S, S + NE);
    c->lE.assign(lE, lE + NE); c->lR.assign(lR, lR + NE); c->lS.assign(lS, lS + NE);
    return SM2D_OK;
}

int orc_fe_upload(orc_ctx* c, const float* A, const float* B) {
    if (!A || !B) { c->feA.assign(c->g.Nz, 0.0f); c->feB.assign(c->g.Nz, 0.0f); return SM2D_OK; }
    c->feA.assign(A, A + c->g.Nz);
    c->feB.assign(B, B + c->g.Nz);
    return SM2D_OK;
}

// The library's learned-physics table (sm2d_surrogate_export layout); on = 0 turns it off.
int orc_surrogate_set(orc_ctx* c, const float* table, int n_mat, int n_nodes, const int32_t* rowmat, int on) {
    if (on) {
        if (!table || !rowmat || n_mat <= 0 || n_nodes < 2) return fail("orc_surrogate_set: bad argument");
        c->surr.assign(table, table + (size_t)n_mat * c->g.N_E * n_nodes * 3);
        c->rowmat.assign(rowmat, rowmat + c->g.Nz);
        c->surr_nodes = n_nodes;
    }
    c->surr_on = on ? 1 : 0;
    return SM2D_OK;
}

int orc_material_upload(orc_ctx* c, const float* rho, const float* x0r) {
    if (!rho) c->rho.assign(c->g.Nz, 1.0f); else c->rho.assign(rho, rho + c->g.Nz);
    if (!x0r) c->x0r.assign(c->g.Nz, 1.0f); else c->x0r.assign(x0r, x0r + c->g.Nz);
    return SM2D_OK;
}

int orc_reset(orc_ctx* c) {
    c->psi[0].clear();
    c->psi[1].clear();
    c->in = 0;
    std::fill(c->mask.begin(), c->mask.end(), 0);
    std::fill(c->mask_prev.begin(), c->mask_prev.end(), 0);
    std::fill(c->edep.begin(), c->edep.end(), 0); std::fill(c->E_cut.begin(), c->E_cut.end(), 0);
    std::fill(c->E_nuc.begin(), c->E_nuc.end(), 0); std::fill(c->E_bnd.begin(), c->E_bnd.end(), 0);
    std::fill(c->W_cut.begin(), c->W_cut.end(), 0); std::fill(c->W_nuc.begin(), c->W_nuc.end(), 0);
    std::fill(c->W_bnd.begin(), c->W_bnd.end(), 0);
    c->src_W_in = c->src_W_out = c->src_W_slot = 0;
    c->src_E_in = c->src_E_out = c->src_E_slot = c->src_E_repr = 0;
    c->cross_cell = 0;
    return SM2D_OK;
}

// device_psic_inject_source (device_psic.cuh:347-450); returns false when no slot is free.
static bool inject_one(orc_ctx* c, int cell, float theta, float E, float w, float x, float z) {
    if (w <= 0.0f) return false;
    if (cell < 0 || cell >= c->N) return false;
    const sm2d_grid& g = c->g;
    x = fmaxf(0.0f, fminf(x, g.dx));
    z = fmaxf(0.0f, fminf(z, g.dz));
    int xs = c->xbin(x - g.dx * 0.5f), zs = c->zbin(z - g.dz * 0.5f);
    const real_t* te = c->th_edges.data();
    const real_t* ee = c->E_edges.data();
    int NT = g.N_theta, NE = g.N_E;
    float tmin = te[0], tmax = te[NT], Emin = ee[0], Emax = ee[NE];
    theta = fmaxf(tmin, fminf(theta, tmax));
    E = fmaxf(Emin, fminf(E, Emax));
    float dth = (tmax - tmin) / NT;
    float tcont = (theta - tmin) / dth;
    int tb = (int)tcont;
    float ft = tcont - tb;
    int eb = c->find_bin(E);
    float lo = ee[eb], hi = ee[eb + 1];
    float fe = (hi - lo) > 1e-10f ? (E - lo) / (hi - lo) : 0.0f;
    if (tb >= NT - 1) { tb = NT - 1; ft = 0.0f; }
    if (eb >= NE - 1) { eb = NE - 1; fe = 0.0f; }
    if (ft >= 0.5f && tb < NT - 1) tb++;
    if (fe >= 0.5f && eb < NE - 1) eb++;
    uint32_t bid = ((uint32_t)(tb / c->Tl) & 0xFFF) | (((uint32_t)(eb / c->El) & 0xFFF) << 12);
    uint16_t l = c->enc(tb % c->Tl, eb % c->El, xs, zs);
    Psi& p = c->pin();
    // device_psic_find_or_allocate_slot (device_psic.cuh:163-188): no fallback
    int s = -1;
    for (int k = 0; k < p.nslot[cell]; ++k) if (p.bid[(size_t)cell * KB + k] == bid) { s = k; break; }
    if (s < 0) s = p.claim(cell, bid);
    if (s < 0) return false;
    p.v(cell, s)[l] += w;
    return true;
}

// gaussian_sample (k1k6_pipeline.cu:433-447): u1 >= 1e-10, 2*pi*u2 evaluated in double.
static float inj_gauss(unsigned& seed) {
    seed = 1664525u * seed + 1013904223u;
    float u1 = (seed >> 16) / 65536.0f;
    seed = 1664525u * seed + 1013904223u;
    float u2 = (seed >> 16) / 65536.0f;
    float r = sqrtf_(-2.0f * logf_(fmaxf(u1, 1e-10f)));
    float th = (float)(2.0f * M_PI * u2);
    return r * cosf_(th);
}

int orc_inject(orc_ctx* c, const sm2d_spot* sp, sm2d_audit* au) {
    if (!sp || sp->struct_size != sizeof(sm2d_spot)) return fail("spot struct_size");
    const sm2d_grid& g = c->g;
    bool gauss = (sp->n_samples > 1) || (sp->sigma_x > 0) || (sp->sigma_theta > 0);
    double W_in = 0, W_out = 0, W_slot = 0, E_in = 0, E_out = 0, E_slot = 0;
    if (gauss) {  // inject_gaussian_source_kernel (k1k6_pipeline.cu:451-551), sample order
        float NEm = c->E_edges[0], NEx = c->E_edges[g.N_E];
        float tmn = c->th_edges[0], tmx = c->th_edges[g.N_theta];
        float w = sp->W_total / sp->n_samples;
        for (int i = 0; i < sp->n_samples; ++i) {
            unsigned seed = sp->random_seed + (unsigned)i;
            float xs = sp->x0 + sp->sigma_x * inj_gauss(seed);
            float ts = sp->theta0 + sp->sigma_theta * inj_gauss(seed);
            float Es = sp->E0 + sp->sigma_E * inj_gauss(seed);
            Es = fmaxf(NEm, fminf(Es, NEx));
            ts = fmaxf(tmn, fminf(ts, tmx));
            float xr = xs - g.x_min, zr = sp->z0 - g.z_min;
            int ix = (int)(xr / g.dx), iz = (int)(zr / g.dz);
            if (ix < 0 || ix >= g.Nx || iz < 0 || iz >= g.Nz) {
                W_out += w;
                E_out += (double)(Es * w);
                continue;
            }
            int cell = iz * g.Nx + ix;
            float xc = xr - ix * g.dx, zc = zr - iz * g.dz;
            if (inject_one(c, cell, ts, Es, w, xc, zc)) { W_in += w; E_in += (double)(Es * w); }
            else { W_slot += w; E_slot += (double)(Es * w); }
        }
    } else {  // inject_source_kernel pencil branch (k1k6_pipeline.cu:325-343)
        float xr = sp->x0 - g.x_min, zr = sp->z0 - g.z_min;
        int sx = (int)(xr / g.dx), sz = (int)(zr / g.dz);
        sx = sx < 0 ? 0 : (sx >= g.Nx ? g.Nx - 1 : sx);
        sz = sz < 0 ? 0 : (sz >= g.Nz ? g.Nz - 1 : sz);
        inject_one(c, sz * g.Nx + sx, sp->theta0, sp->E0, sp->W_total, xr - sx * g.dx, zr - sz * g.dz);
        W_in = sp->W_total;
        E_in = (double)sp->E0 * (double)W_in;
    }
    // represented energy (gpu_transport_wrapper.cu:125-180, :443-457)
    double Wr, Er;
    psi_totals(c, c->pin(), Wr, Er);
    double loss = E_in - Er;
    if (std::fabs(loss) < 1e-12) loss = 0.0;
    c->src_W_in = W_in; c->src_W_out = W_out; c->src_W_slot = W_slot;
    c->src_E_in = E_in; c->src_E_out = E_out; c->src_E_slot = E_slot; c->src_E_repr = loss;
    if (au) {
        au->W_source_injected = W_in; au->W_source_out_of_grid = W_out; au->W_source_slot_drop = W_slot;
        au->E_source_injected = E_in; au->E_source_out_of_grid = E_out; au->E_source_slot_drop = E_slot;
        au->E_representation_loss = loss;
    }
    return SM2D_OK;
}

int orc_psi_load(orc_ctx* c, int n, const int32_t* cell, const int32_t* slot, const uint32_t* bid,
                 const int32_t* lidx, const float* w) {
    Psi& p = c->pin();
    p.clear();
    for (int i = 0; i < n; ++i) {
        int cl = cell[i], s = slot[i];
        if (cl < 0 || cl >= c->N || s < 0 || s >= KB || lidx[i] < 0 || lidx[i] >= c->LB)
            return fail("psi row out of range");
        int t, e, x, z;
        c->dec(lidx[i], t, e, x, z);
        if ((bid[i] >> 24) != 0 || (int)(bid[i] & 0xFFF) * c->Tl + t >= c->g.N_theta ||
            (int)((bid[i] >> 12) & 0xFFF) * c->El + e >= c->g.N_E)
            return fail("psi bin outside the grid");
        while (p.nslot[cl] <= s) {
            int k = p.nslot[cl];
            p.claim(cl, EMPTY);
            (void)k;
        }
        uint32_t& b = p.bid[(size_t)cl * KB + s];
        if (b != EMPTY && b != bid[i]) return fail("rows of one slot carry different bids");
        b = bid[i];
        p.v(cl, s)[lidx[i]] = w[i];
    }
    for (int cl = 0; cl < c->N; ++cl)
        for (int s = 0; s < p.nslot[cl]; ++s)
            if (p.bid[(size_t)cl * KB + s] == EMPTY) return fail("slots of a cell must be a prefix");
    return SM2D_OK;
}

int orc_psi_dump(orc_ctx* c, uint32_t* bids, int cap, int32_t* cell, int32_t* slot, uint32_t* bid,
                 int32_t* lidx, float* w, int* nrows) {
    Psi& p = c->pin();
    if (bids) std::memcpy(bids, p.bid.data(), sizeof(uint32_t) * p.bid.size());
    int n = 0;
    for (int cl = 0; cl < c->N; ++cl) {
        for (int s = 0; s < p.nslot[cl]; ++s) {
            const real_t* v = p.v(cl, s);
            for (int l = 0; l < c->LB; ++l) {
                if (v[l] == 0.0f) continue;
                if (n < cap) {
                    cell[n] = cl; slot[n] = s; bid[n] = p.bid[(size_t)cl * KB + s]; lidx[n] = l; w[n] = v[l];
                }
                ++n;
            }
        }
    }
    *nrows = n;
    return SM2D_OK;
}

int orc_k1k2_mask_compact(orc_ctx* c, const sm2d_transport* tp, const uint8_t* prev, uint8_t* mask,
                          int32_t* act, int* na, int32_t* crs, int* nc) {
    int bon = b_threshold(c, tp->E_fine_on), boff = b_threshold(c, tp->E_fine_off);
    std::vector<uint8_t> m(c->N);
    k1(c, bon, boff, tp->weight_active_min, prev, m.data());
    std::vector<int> a, k;
    compact(c, m.data(), a, k);
    if (mask) std::memcpy(mask, m.data(), c->N);
    for (size_t i = 0; i < a.size(); ++i) act[i] = a[i];
    for (size_t i = 0; i < k.size(); ++i) crs[i] = k[i];
    *na = (int)a.size();
    *nc = (int)k.size();
    return SM2D_OK;
}

static int run_iters(orc_ctx* c, const sm2d_transport* tp, float sx0, int n_iter, int src_terms,
                     sm2d_k5_report* last, sm2d_audit* au) {
    int bon = b_threshold(c, tp->E_fine_on), boff = b_threshold(c, tp->E_fine_off);
    std::vector<int> act, crs;
    c->surr_active = c->surr_on;
    if (au) au->surrogate_active = c->surr_active;
    for (int it = 0; it < n_iter; ++it) {
        c->sdev_max = 0;
        c->mask_prev = c->mask;
        k1(c, bon, boff, tp->weight_active_min, c->mask_prev.data(), c->mask.data());
        compact(c, c->mask.data(), act, crs);
        if (au) au->iterations++;
        if (act.empty() && crs.empty()) {
            if (au) au->terminated = 1;
            return SM2D_OK;
        }
        c->pout().clear();
        double dW[3] = {0, 0, 0}, dE[4] = {0, 0, 0, 0}, drop_w = 0, drop_E = 0, clw = 0, clE = 0;
        run_phase(c, tp, sx0, crs, false, au, dW, dE, drop_w, drop_E, clw, clE);
        run_phase(c, tp, sx0, act, true, au, dW, dE, drop_w, drop_E, clw, clE);
        sm2d_k5_report r;
        bool first = src_terms && it == 0;
        k5(c, dW, dE, c->src_W_out, c->src_W_slot, c->src_E_out, c->src_E_slot, drop_w, drop_E,
           first ? 1 : 0, &r);
        if (last) *last = r;
        if (c->surr_active) {  // the drift check of K5, on the balances without the source terms
            const double Wi = r.W_in - r.W_source_out_of_grid - r.W_source_slot_drop;
            const double Wr = r.W_out + r.W_cutoff + r.W_nuclear + r.W_boundary + r.W_transport_drop;
            const double Ei = r.E_in - r.E_source_out_of_grid - r.E_source_slot_drop;
            const double Er = r.E_out + r.E_dep + r.E_cutoff + r.E_nuclear + r.E_boundary + r.E_transport_drop;
            const double werr = std::fabs(Wi - Wr) / std::max(Wi, 1e-20);
            if (!(werr < 1e-6) || !std::isfinite(Ei - Er) || !(c->sdev_max <= SURR_DRIFT_TOL)) {
                c->surr_active = 0;
                if (au) au->surrogate_fallback = 1;
            }
        }
        if (au) {
            au->W_transport_drop += drop_w;
            au->E_transport_drop += drop_E;
            double Er = r.E_out + r.E_dep + r.E_cutoff + r.E_nuclear + r.E_boundary +
                        r.E_transport_drop + r.E_source_out_of_grid + r.E_source_slot_drop;
            double res = r.E_in - Er;
            if (std::isfinite(res)) au->E_audit_residual += res;
            au->max_W_error = std::max(au->max_W_error, r.W_error);
            au->max_E_error = std::max(au->max_E_error, r.E_error);
            if (!r.W_pass) au->W_fail_iters++;
            if (!r.E_pass) au->E_fail_iters++;
            au->max_cell_W_closure = std::max(au->max_cell_W_closure, clw);
            au->max_cell_E_closure = std::max(au->max_cell_E_closure, clE);
            au->cross_cell_emissions = c->cross_cell;
        }
        if (tp->fail_fast_on_audit && (!r.W_pass || !r.E_pass)) return fail("K5 audit fail-fast", SM2D_ERR_AUDIT);
        c->in = 1 - c->in;  // K6
        c->pout().clear();
    }
    return SM2D_OK;
}

static void write_tallies(orc_ctx* c, double* t) {
    int N = c->N;
    for (int i = 0; i < N; ++i) {
        t[0 * N + i] = c->edep[i];
        t[1 * N + i] = c->W_cut[i];
        t[2 * N + i] = c->E_cut[i];
        t[3 * N + i] = c->W_nuc[i];
        t[4 * N + i] = c->E_nuc[i];
        t[5 * N + i] = c->W_bnd[i];
        t[6 * N + i] = c->E_bnd[i];
    }
}

int orc_run_iterations(orc_ctx* c, const sm2d_transport* tp, int n_iter, int src_terms,
                       double* tallies, sm2d_k5_report* last, sm2d_audit* au) {
    if (!tp || tp->struct_size != sizeof(sm2d_transport)) return fail("transport struct_size");
    if (au) { std::memset(au, 0, sizeof(*au)); au->struct_size = sizeof(*au); }
    int rc = run_iters(c, tp, tp->sigma_x_initial, n_iter, src_terms, last, au);
    if (tallies) write_tallies(c, tallies);
    return rc;
}

int orc_k5_audit(orc_ctx* c, const double* td, double a, double b, double e1, double e2, double dw,
                 double de, int src, sm2d_k5_report* out) {
    int N = c->N;
    double dW[3] = {0, 0, 0}, dE[4] = {0, 0, 0, 0};
    for (int i = 0; i < N; ++i) {
        dE[0] += std::max(0.0, td[0 * N + i]);
        dW[0] += std::max(0.0, td[1 * N + i]);
        dE[1] += std::max(0.0, td[2 * N + i]);
        dW[1] += std::max(0.0, td[3 * N + i]);
        dE[2] += std::max(0.0, td[4 * N + i]);
        dW[2] += std::max(0.0, td[5 * N + i]);
        dE[3] += std::max(0.0, td[6 * N + i]);
    }
    k5(c, dW, dE, a, b, e1, e2, dw, de, src, out);
    return SM2D_OK;
}

int orc_swap(orc_ctx* c) { c->in = 1 - c->in; return SM2D_OK; }

int orc_run_spot(orc_ctx* c, const sm2d_spot* sp, const sm2d_transport* tp, double* edep_out,
                 sm2d_audit* au) {
    if (!tp || tp->struct_size != sizeof(sm2d_transport)) return fail("transport struct_size");
    if (tp->max_iterations <= 0) return fail("max_iterations must be positive");
    sm2d_audit local;
    if (!au) au = &local;
    std::memset(au, 0, sizeof(*au));
    au->struct_size = sizeof(*au);
    orc_reset(c);
    // FE moments of this spot (init_pipeline_state, k1k6_pipeline.cu:1061-1212)
    {
        std::vector<float> st(c->g.Nz), A(c->g.Nz, 0.0f), B(c->g.Nz, 0.0f);
        int r0 = orc_fe_precompute(sp->E0, sp->sigma_x, c->g.Nz, c->g.dz, A.data(), B.data(), st.data());
        if (r0) return r0;
        c->feA.assign(A.begin(), A.end());
        c->feB.assign(B.begin(), B.end());
    }
    int rc = orc_inject(c, sp, au);
    if (rc) return rc;
    rc = run_iters(c, tp, sp->sigma_x, tp->max_iterations, 1, nullptr, au);
    // leftover when max_iterations stopped the loop (k1k6_pipeline.cu:1606, no accounting)
    if (!au->terminated) psi_totals(c, c->pin(), au->W_leftover, au->E_leftover);
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < c->N; ++i) {
        s[0] += c->edep[i]; s[1] += c->W_cut[i]; s[2] += c->E_cut[i]; s[3] += c->W_nuc[i];
        s[4] += c->E_nuc[i]; s[5] += c->W_bnd[i]; s[6] += c->E_bnd[i];
    }
    au->E_dep = s[0]; au->W_cutoff = s[1]; au->E_cutoff = s[2]; au->W_nuclear = s[3];
    au->E_nuclear = s[4]; au->W_boundary = s[5]; au->E_boundary = s[6];
    if (edep_out) std::memcpy(edep_out, c->edep.data(), sizeof(double) * c->N);
    return rc;
}

}  // extern "C"
