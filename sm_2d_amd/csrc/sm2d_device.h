// sm2d_device.h — device-side building blocks of the transport step for gfx950.
//
// Layout of one problem in HBM (one context):
//   psi (x2, ping-pong)  bid[N*32] u32   block id per (cell, slot), EMPTY = 0xFFFFFFFF
//                        ns[N] u32       occupied slots of a cell (always the prefix 0..ns-1)
//                        sref[N*32] u32  payload index of an occupied slot
//                        pool[P*LB] f32  slot payloads (LB local bins, 1 KiB at LB = 256),
//                                        bump-allocated; zeroed when a slot is claimed
//   buckets (per phase)  bid/sref[L*4*32], n[L*4]  outflow buckets of the L listed cells,
//                        indexed by list position; payloads in a second bump pool
//   per cell             epoch-tagged mask, list tags, FP64/FP32 tallies
// Nothing is cleared densely: the bid/ns tables of the previous iteration's cells are reset
// and payloads are zeroed only when claimed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sm2d_math.h"

namespace sm2d {

constexpr int KB = 32;
constexpr uint32_t EMPTY = 0xFFFFFFFFu;

// Local-bin layout lidx = t + TL*(e + EL*(x + XS*z)) (src/include/core/local_bins.hpp:32-48
// generalised to any block shape).
template <int TL_, int EL_, int XS_, int ZS_>
struct Layout {
    static constexpr int TL = TL_, EL = EL_, XS = XS_, ZS = ZS_;
    static constexpr int LB = TL * EL * XS * ZS;
    static constexpr int PL = LB / 64;  // values per lane when a wave holds one slot
    static constexpr int NL = 6144 / LB;  // LDS payload rows per transport wave (24 KiB)
    static_assert(LB % 64 == 0, "a slot must split evenly over a wavefront");
    __host__ __device__ static int enc(int t, int e, int x, int z) { return t + TL * (e + EL * (x + XS * z)); }
    __host__ __device__ static int lt(int l) { return l % TL; }
    __host__ __device__ static int le(int l) { return (l / TL) % EL; }
    __host__ __device__ static int lx(int l) { return (l / (TL * EL)) % XS; }
    __host__ __device__ static int lz(int l) { return l / (TL * EL * XS); }
};
using LayoutRef = Layout<4, 2, 8, 4>;   // the reference build (LOCAL_BINS = 256)
using LayoutCfg = Layout<8, 4, 4, 1>;   // 128 local bins: 8 theta x 4 E x 4 x_sub

// Cells of all spots of a batch are numbered cell = spot*Ns + iz*Nx + ix (Ns = Nx*Nz); the
// neighbour tests below never cross a spot.
struct Tabs {
    int Nx, Nz, N, NT, NE;
    int Ns, S;         // cells per spot, spots
    float dx, dz, x_min, z_min;
    const float* te;   // theta edges [NT+1]
    const float* ee;   // energy edges [NE+1]
    const float* R;    // CSDA range [NE] (mm of water)
    const float* lR;
    const float* lE;
    const float* feA;  // Fermi-Eyges A per spot and row [S*Nz]
    const float* feB;
    const float* sx0;  // initial beam sigma_x per spot [S] (coarse moments)
    const float* rho;  // density per row [Nz]
    const float* x0r;  // radiation-length ratio per row [Nz]
    const float* surr;    // learned-physics table [materials][NE][surr_nodes][3] or null
    const int* rowmat;    // material index of every row in that table (-1: plain physics)
    int surr_nodes;
    // the network itself, for inference inside the fine transport (null when not attached):
    // padded BF16 weights, FP32 parameters and the width dictionaries (lnE lo/hi, lnT lo/hi)
    const uint16_t* mlp_wb;
    const float* mlp_p;
    float mlp_dict[4];
};

// Device control block (u32 words).  C_DONE: 0 running, 1 ended (empty lists or a capacity
// overflow), 2 paused (the next iteration needs a phase the queued chunk left out).
// C_FNT counts the cells whose +z bucket takes lateral emissions of other cells (k_foreign
// applies them and returns it to 0).  C_DRIFT: set by the closing audit of an iteration whose K5
// row flags the learned physics (drift_check); from the next iteration on the fine transport
// uses the plain physics.
enum Ctl {
    C_NACT = 0, C_NCRS = 1, C_NDEST = 2, C_DONE = 3, C_EP = 4, C_IT = 5, C_DONE_IT = 6, C_DRIFT = 7,
    C_NOCC0 = 8, C_NOCC1 = 9, C_NDEST1 = 10, C_FNT = 13, NCTL = 16,
    // the bucket pools' shard counters (coarse, then fine phase) follow the control words in the
    // same allocation, from word C_BPC on, then the destination lists' segment counters (C_DSEG)
    C_BPC = 32, C_DSEG = C_BPC + 2 * 64 * 32, C_CTL_WORDS = C_DSEG + 2 * 64 * 32
};

// Run-time monitor of the learned physics: in every eighth cell (cell index within its spot
// divisible by 8) each fine-transport substep that takes the learned widths also evaluates the
// plain ones; the largest deviation (ln sigma_E, ln sigma_theta, and the attenuation exponent
// relative to a + 1e-4 / mm x dR, the criterion of the table validation) enters the iteration's
// K5 row, and a deviation above SURR_DRIFT_TOL switches the rest of the run to the plain physics.
constexpr int SURR_MONITOR_EVERY = 8;
constexpr float SURR_DRIFT_TOL = 0.5f;

struct Tp {
    float E_on, E_off, wmin, step_coarse;
    int b_on, b_off, rule, strag, nuc;
    int surr;    // fine transport takes the straggling, scattering and attenuation from the table
    float R_on;  // water range at max(E_on, lowest edge) (lookup_R_tab, computed on the host)
};

// Payload pools are bump allocators split into NSH shards of pcap / NSH rows, each counted on a
// 128-byte line of its own (pcount[s * SH_STRIDE]).  Atomics on one address serialize (about
// 11 ns apiece on MI355X, measured), and a batched launch reserves rows for tens of thousands of
// cells, so a workgroup takes its rows from shard (hint mod NSH) and passes to the next shard
// when that one is full.  Row positions never enter a result.
constexpr int NSH = 64, SH_STRIDE = 32;
static_assert(C_DSEG == C_BPC + 2 * NSH * SH_STRIDE && C_CTL_WORDS == C_DSEG + 2 * NSH * SH_STRIDE,
              "control-block layout of the shard counters");
constexpr uint32_t POOL_FULL = 0xFFFFFFFFu;

// First row of n consecutive rows of the pool, or POOL_FULL when no shard has n rows left.
__device__ inline uint32_t pool_take(uint32_t* pcount, uint32_t pcap, uint32_t n, uint32_t hint) {
    const uint32_t span = pcap / NSH;
    for (uint32_t k = 0; k < (uint32_t)NSH; ++k) {
        const uint32_t sh = (hint + k) % NSH;
        const uint32_t b = atomicAdd(pcount + sh * SH_STRIDE, n);
        if (b + n <= span) return sh * span + b;
    }
    return POOL_FULL;
}

struct Psi {
    uint32_t* bid;
    uint32_t* ns;
    uint32_t* sref;
    float* pool;
    uint32_t* pcount;  // NSH shard counters (stride SH_STRIDE)
    uint32_t pcap;
};

struct Bkt {
    uint32_t* bid;   // [cap*4*32]
    uint32_t* n;     // [cap*4]
    uint32_t* sref;  // [cap*4*32]
    float* pool;
    uint32_t* pcount;  // NSH shard counters (stride SH_STRIDE)
    uint32_t pcap;
    uint32_t* ftag;   // [cap] phase tag of a listed cell whose +z bucket takes other cells' lateral emissions
    uint32_t* flist;  // [cap] list positions of those cells (count in ctl[C_FNT])
};

struct Tally {  // cumulative per-cell tallies (src/cuda/k1k6_pipeline.cuh:174-180)
    double* edep;
    float* w_cut;
    double* E_cut;
    float* w_nuc;
    double* E_nuc;
    float* w_bnd;
    double* E_bnd;
};

// Per listed cell and iteration (FP64), reduced by K5 in list order.
enum Acc {
    A_EDEP, A_WCUT, A_ECUT, A_WNUC, A_ENUC, A_WBND, A_EBND, A_DROPW, A_DROPE,
    A_CLWIN, A_CLWOUT, A_CLEIN, A_CLEOUT, A_COMPS, A_UPD, A_SURR, A_SDEV, NACC
};

// Device error flags.
// F_CROSS counts lateral emissions into another listed cell's +z bucket (not an error).
enum Flag { F_PSI_FULL, F_BKT_FULL, F_CROSS, F_DEST_FULL, NFLAG };

// ---- wave helpers -----------------------------------------------------------------------
__device__ inline int lane_id() { return (int)(threadIdx.x & 63); }
__device__ inline float rlf(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ inline int rli(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ inline uint32_t rlu(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }
__device__ inline int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ inline double wave_sum(double v) {  // fixed butterfly: deterministic
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// ---- tables ---------------------------------------------------------------------------------
// device_find_bin_edges (src/cuda/device/device_lut.cuh:34-49).  With a guess g >= 0 the bins
// g, g-1, g-2 are tried first (strictly increasing edges: a bin that brackets E is the binary
// search's answer), so lookups near a known bin need no search; the result never depends on g.
SM2D_HD inline int find_bin(const float* __restrict__ e, int NE, float E, int g = -1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int b = g - k;
        if (g >= 0 && b >= 0 && b < NE && e[b] <= E && E < e[b + 1]) return b;
    }
    if (E <= e[0]) return 0;
    if (E >= e[NE]) return NE - 1;
    int lo = 0, hi = NE;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (e[mid + 1] <= E) lo = mid + 1; else hi = mid;
    }
    return lo;
}
__device__ inline float ecenter(const float* __restrict__ e, int b) {
    float lo = e[b], hi = e[b + 1];
    return lo + 1.00f * ((hi - lo) * 0.5f);
}
// device_lookup_R (device_lut.cuh:66-85) over the tables as arrays; g: bin guess (see find_bin).
// Host and device: the transport's launch constant R_on is computed on the host with it.
SM2D_HD inline float lookup_R_tab(const float* ee, const float* R, const float* lR, const float* lE, int NE, float E,
                                  int g = -1) {
    const float Emin = ee[0], Emax = ee[NE];
    float Ec = fmaxf(Emin, fminf(E, Emax));
    int b = find_bin(ee, NE, Ec, g);
    if (b == NE - 1 || Ec >= Emax) return R[NE - 1];
    float lv = logf_(Ec);
    return expf_(lR[b] + (lR[b + 1] - lR[b]) * (lv - lE[b]) / (lE[b + 1] - lE[b]));
}
__device__ inline float lookup_R(const Tabs& T, float E, int g = -1) {
    return lookup_R_tab(T.ee, T.R, T.lR, T.lE, T.NE, E, g);
}
// device_lookup_E_inverse (device_lut.cuh:110-138): lower_bound bin, no upper clamp.  With a
// guess g whose R[g] >= Rin, the lower bound is found walking down from g (at most 3 steps, R
// non-decreasing) before the binary search; the result never depends on g.
__device__ inline float E_inverse(const Tabs& T, float Rin, int g = -1) {
    if (Rin <= 0.0f) return T.ee[0];
    int lo = -1;
    if (g >= 0 && g < T.NE && T.R[g] >= Rin) {
#pragma unroll
        for (int k = 0; k < 3 && lo < 0; ++k) {
            const int c = g - k;
            if (c == 0 || T.R[c - 1] < Rin) lo = c;
        }
    }
    if (lo < 0) {
        int l0 = 0, hi = T.NE;
        while (l0 < hi) {
            int mid = (l0 + hi) >> 1;
            if (T.R[mid] < Rin) l0 = mid + 1; else hi = mid;
        }
        lo = l0;
    }
    int b = max(0, min(lo, T.NE - 2));
    float lv = logf_(Rin);
    float d = T.lR[b + 1] - T.lR[b];
    if (fabsf(d) < 1e-10f) return expf_(T.lE[b]);
    return expf_(T.lE[b] + (T.lE[b + 1] - T.lE[b]) * (lv - T.lR[b]) / d);
}
// device_compute_energy_deposition (device_lut.cuh:145-157); a geometric step ds in density
// rho consumes rho*ds of water range.
// g: bin guess of E (the inverse search starts one bin above it)
__device__ inline float deposition(const Tabs& T, float E, float ds, float rh, int g = -1) {
    float Rn = lookup_R(T, E, g) - rh * ds;
    float En = (Rn <= 0) ? T.ee[0] : E_inverse(T, Rn, g < 0 ? -1 : min(g + 1, T.NE - 1));
    return E - En;
}
// device_compute_max_step (device_lut.cuh:168-214) or the 2 % rule (step_control.hpp:18-66)
__device__ inline float max_step(const Tabs& T, float E, float rh, int rule, int g = -1) {
    float Rw = lookup_R(T, E, g) / rh;
    if (rule == 1) {
        float d = 0.02f * Rw, f = 1.0f;
        if (E < 5.0f) { f = 0.2f; d = fminf(d, 0.1f); }
        else if (E < 10.0f) { f = 0.3f; d = fminf(d, 0.2f); }
        else if (E < 20.0f) { f = 0.5f; d = fminf(d, 0.5f); }
        else if (E < 50.0f) { f = 0.7f; d = fminf(d, 0.7f); }
        return fmaxf(d * f, 0.05f);
    }
    float d = 0.01f * Rw;
    float fl = (E < 2.0f) ? 0.1f : (E < 20.0f ? 0.25f : (E < 100.0f ? 0.5f : 1.0f));
    d = fmaxf(d, fl);
    if (E < 0.5f) d = fminf(d, 0.05f);
    else if (E < 5.0f) d = fminf(d, 0.2f);
    else if (E < 20.0f) d = fminf(d, 0.5f);
    return d;
}

// ---- physics (src/cuda/device/device_physics.cuh) -----------------------------------------
constexpr float MP = 938.272f, ME = 0.511f, X0W = 360.8f;

__device__ inline float highland(float E, float ds, float X0) {  // :50-71
    float g = (E + MP) / MP;
    float beta = sqrtf_(fmaxf(1.0f - 1.0f / (g * g), 0.0f));
    float p = sqrtf_(fmaxf((E + MP) * (E + MP) - MP * MP, 0.0f));
    float t = ds / X0;
    if (t < 1e-6f) return 0.0f;
    float br = fmaxf(1.0f + 0.038f * logf_(t), 0.25f);
    return (13.6f * 1.0f / (beta * p)) * sqrtf_(t) * br;
}
__device__ inline float scat_power(float E, float ds, float X0) {  // :394-404
    float th = highland(E, ds, X0);
    if (ds < 1e-6f) return 0.0f;
    return th * th / ds;
}
__device__ inline float gauss(unsigned& s) {  // device_sample_gaussian :74-87
    s = 1664525u * s + 1013904223u;
    float u1 = (s >> 16) / 65536.0f;
    s = 1664525u * s + 1013904223u;
    float u2 = (s >> 16) / 65536.0f;
    u1 = fmaxf(u1, 1e-6f);
    return sqrtf_(-2.0f * logf_(u1)) * cosf_(2.0f * 3.14159265358979323846f * u2);
}
__device__ inline float vav_xi(float beta, float rh, float ds_cm) {
    return (0.307f / 2.0f) * 0.555f * (1.0f * 1.0f / (beta * beta)) * rh * ds_cm;
}
__device__ inline float straggling(float E, float ds, float rh) {  // :114-184
    float g = (E + MP) / MP;
    float beta = sqrtf_(fmaxf(1.0f - 1.0f / (g * g), 0.0f));
    float xi = vav_xi(beta, rh, ds / 10.0f);
    float mr = ME / MP;
    float num = 2.0f * ME * beta * beta * g * g;
    float den = 1.0f + 2.0f * g * mr + mr * mr;
    float Tm = num / fmaxf(den, 1e-10f);
    float kappa = fmaxf(xi / fmaxf(Tm, 1e-10f), 1e-6f);
    float bb = sqrtf_(fmaxf(1.0f - 1.0f / (g * g), 0.01f));
    float bohr = 0.156f * sqrtf_(rh * ds) / bb;
    if (kappa > 10.0f) return bohr;
    float landau = 4.0f * xi / 2.355f;
    if (kappa < 0.01f) return landau;
    float w = 1.0f / (1.0f + kappa);
    return w * landau + (1.0f - w) * bohr;
}
__device__ inline float nuc_sigma(float E) {  // :201-215
    if (E < 5.0f) return 0.0f;
    if (E >= 20.0f) return 0.0012f * fmaxf(1.0f - 0.15f * logf_(E / 100.0f), 0.4f);
    return 0.0016f * ((E - 5.0f) / 15.0f);
}
__device__ inline float gcdf(float x, float mu, float s) {  // :245-251
    if (s < 1e-10f) return (x >= mu) ? 1.0f : 0.0f;
    float c = 0.5f * (1.0f + erff_((x - mu) / s * 0.70710678f));
    return fminf(1.0f, fmaxf(0.0f, c));
}

}  // namespace sm2d
