// sm2d_mlp.hip — the learned part of the fine transport: a 4-128-128-128-33 MLP with GELU,
// BF16 weights and activations, FP32 accumulation on the matrix cores
// (v_mfma_f32_16x16x16_bf16), trained on the GPU with the same kernels.
//
// Inputs (normalised to [-1, 1]): ln E, log10 dR (water-equivalent step rho*ds), rho, ln of the
// radiation-length ratio to water.  Outputs: 32 logits of a joint distribution over
// 4 straggling-width classes x 8 angular-width classes (log-spaced dictionaries; the network's
// widths are the log-moments of that distribution) and one output for ln(a + 1e-9), a = the
// nuclear attenuation exponent of the substep (survival = exp(-a)).  These three quantities are
// what the per-substep Vavilov, Highland and attenuation evaluations of K3 produce
// (k3_finetransport.cu:231-361 via device_physics.cuh:50-234).
#include "sm2d_device.h"

namespace sm2d {
namespace mlp {

constexpr int NIN = 4, NH = 128, NOUT = 33, NCLS_E = 4, NCLS_T = 8, NCLS = NCLS_E * NCLS_T;
constexpr int KP0 = 16;   // first layer K padded to one MFMA K step
constexpr int NOP = 48;   // last layer N padded to three 16-wide tiles
// flat FP32 parameter layout
constexpr int OFF_W0 = 0, OFF_B0 = OFF_W0 + NH * NIN, OFF_W1 = OFF_B0 + NH, OFF_B1 = OFF_W1 + NH * NH,
              OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH, OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NOUT * NH,
              NPARAM = OFF_B3 + NOUT;
// padded BF16 weight copies used by the GEMMs: W0p[NH][KP0], W1[NH][NH], W2[NH][NH], W3p[NOP][NH]
constexpr int BOFF_W0 = 0, BOFF_W1 = BOFF_W0 + NH * KP0, BOFF_W2 = BOFF_W1 + NH * NH, BOFF_W3 = BOFF_W2 + NH * NH,
              NBW = BOFF_W3 + NOP * NH;

// Training / table ranges of the inputs.
constexpr float E_LO = 0.1f, E_HI = 250.0f, DR_LO = 1e-4f, DR_HI = 10.0f, RHO_LO = 0.2f, RHO_HI = 2.0f,
                X0R_LO = 0.5f, X0R_HI = 4.0f;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 with round to nearest even (NaN kept quiet).
SM2D_HD inline uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
SM2D_HD inline float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

SM2D_HD inline float gelu(float x) { return 0.5f * x * (1.0f + erff_(x * 0.70710678118654752f)); }
SM2D_HD inline float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff_(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf_(-0.5f * x * x);
    return cdf + x * pdf;
}

// Normalised network input of (E, dR, rho, X0 ratio); false when outside the training range.
SM2D_HD inline bool encode_input(float E, float dR, float rho, float x0r, float* x) {
    if (!(E >= E_LO && E <= E_HI && dR >= DR_LO && dR <= DR_HI && rho >= RHO_LO && rho <= RHO_HI &&
          x0r >= X0R_LO && x0r <= X0R_HI))
        return false;
    const float lE0 = logf_(E_LO), lE1 = logf_(E_HI);
    x[0] = 2.0f * (logf_(E) - lE0) / (lE1 - lE0) - 1.0f;
    x[1] = 2.0f * (logf_(dR) - logf_(DR_LO)) / (logf_(DR_HI) - logf_(DR_LO)) - 1.0f;
    x[2] = 2.0f * (rho - RHO_LO) / (RHO_HI - RHO_LO) - 1.0f;
    x[3] = 2.0f * (logf_(x0r) - logf_(X0R_LO)) / (logf_(X0R_HI) - logf_(X0R_LO)) - 1.0f;
    return true;
}

// Width dictionaries: class c of an axis stands for ln width = lo + c * (hi - lo) / (n - 1).
struct Dict {
    float lsE_lo, lsE_hi, lsT_lo, lsT_hi;
};

// Soft assignment of a ln width onto n classes (linear between the two nearest, clamped).
SM2D_HD inline void soft_assign(float v, float lo, float hi, int n, int& i0, float& f) {
    float p = (v - lo) / (hi - lo) * (float)(n - 1);
    p = fminf(fmaxf(p, 0.0f), (float)(n - 1));
    i0 = (int)p;
    if (i0 > n - 2) i0 = n - 2;
    f = p - (float)i0;
}

// ------------------------------------------------------------------------------------------
// GEMM on the matrix cores: P[z][m][n] = sum_{k in split z} A(m,k) B(k,n), A and B in BF16 with
// arbitrary strides (so transposed operands are views), FP32 partials per K split.  Block =
// 4 waves, block tile 64 x 64; wave w owns rows 16w..16w+15 and four 16-wide column tiles.
// M, N, K are multiples of 16 (the callers pad).
//   MFMA 16x16x16 BF16: lane l holds A[l%16][4(l/16)+i], B[4(l/16)+i][l%16] (i = 0..3) and
//   D[4(l/16)+i][l%16].
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gemm_bf16(const uint16_t* __restrict__ A, int sam, int sak,
                                                   const uint16_t* __restrict__ B, int sbk, int sbn, int M, int N,
                                                   int K, int ksplit, float* __restrict__ P) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m0 = blockIdx.x * 64 + w * 16, nb = blockIdx.y * 64;
    const int z = blockIdx.z;
    const int kchunk = ((K / 16 + ksplit - 1) / ksplit) * 16;
    const int k0 = z * kchunk, k1 = min(K, k0 + kchunk);
    if (m0 >= M) return;
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int am = m0 + (l & 15), kq = 4 * (l >> 4);
    for (int k = k0; k < k1; k += 16) {
        bf16x4 a;
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = (short)A[(size_t)am * sam + (size_t)(k + kq + i) * sak];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int n = nb + t * 16 + (l & 15);
            if (nb + t * 16 >= N) break;
            bf16x4 b;
#pragma unroll
            for (int i = 0; i < 4; ++i) b[i] = (short)B[(size_t)(k + kq + i) * sbk + (size_t)n * sbn];
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        if (nb + t * 16 >= N) break;
        const int n = nb + t * 16 + (l & 15);
#pragma unroll
        for (int i = 0; i < 4; ++i) P[((size_t)z * M + m0 + 4 * (l >> 4) + i) * N + n] = acc[t][i];
    }
}

// Epilogue: Z = sum_z P[z] (z ascending) + bias; mode 1 applies GELU.  Writes FP32 Z (ldz) for
// rows < mvalid and columns < nvalid, and BF16 act (lda; columns >= nvalid written as 0).
__global__ void k_epilogue(const float* __restrict__ P, int ksplit, int M, int N, int mvalid, int nvalid,
                           const float* __restrict__ bias, int mode, float* Z, int ldz, uint16_t* act, int lda) {
    const size_t total = (size_t)M * N;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int m = (int)(i / N), n = (int)(i % N);
        if (m >= mvalid) continue;
        float v = 0.0f;
        for (int z = 0; z < ksplit; ++z) v += P[(size_t)z * total + i];
        if (n >= nvalid) {
            if (act) act[(size_t)m * lda + n] = 0;
            continue;
        }
        if (bias) v += bias[n];
        if (Z) Z[(size_t)m * ldz + n] = v;
        if (act) act[(size_t)m * lda + n] = f2bf(mode == 1 ? gelu(v) : v);
    }
}

// Loss of a batch: cross-entropy of the 32-way softmax against the soft targets, the squared
// error of the decoded widths (the log-moments the transport uses; weight w_mom) and the
// squared error of the attenuation output (weight w_att), averaged over the batch; dZ (FP32 and
// BF16, [B][NOP]).  labels[b] = (ln sigma_E, ln sigma_theta, ln(a + 1e-9)).
__global__ void k_loss(const float* __restrict__ Z4, const float* __restrict__ labels, int Bn, Dict D,
                       float w_att, float w_mom, float* dZ, uint16_t* dZb, float* loss_part) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < Bn; b += gridDim.x * blockDim.x) {
        const float* z = Z4 + (size_t)b * NOP;
        int ie, it;
        float fe, ft;
        soft_assign(labels[3 * b + 0], D.lsE_lo, D.lsE_hi, NCLS_E, ie, fe);
        soft_assign(labels[3 * b + 1], D.lsT_lo, D.lsT_hi, NCLS_T, it, ft);
        float mx = z[0];
        for (int c = 1; c < NCLS; ++c) mx = fmaxf(mx, z[c]);
        float s = 0.0f;
        for (int c = 0; c < NCLS; ++c) s += expf_(z[c] - mx);
        const float lse = mx + logf_(s);
        float loss = 0.0f;
        const float inv = 1.0f / (float)Bn;
        float mE = 0.0f, mT = 0.0f;  // decoded ln widths
        for (int c = 0; c < NCLS; ++c) {
            const float q = expf_(z[c] - lse);
            mE += q * (D.lsE_lo + (D.lsE_hi - D.lsE_lo) * (float)(c / NCLS_T) / (float)(NCLS_E - 1));
            mT += q * (D.lsT_lo + (D.lsT_hi - D.lsT_lo) * (float)(c % NCLS_T) / (float)(NCLS_T - 1));
        }
        const float eE = mE - labels[3 * b + 0], eT = mT - labels[3 * b + 1];
        loss += w_mom * (eE * eE + eT * eT);
        for (int c = 0; c < NCLS; ++c) {
            const int ce = c / NCLS_T, ct = c % NCLS_T;
            const float pe = ce == ie ? 1.0f - fe : (ce == ie + 1 ? fe : 0.0f);
            const float pt = ct == it ? 1.0f - ft : (ct == it + 1 ? ft : 0.0f);
            const float tgt = pe * pt;
            const float lp = z[c] - lse;
            if (tgt > 0.0f) loss -= tgt * lp;
            const float q = expf_(lp);
            const float vE = D.lsE_lo + (D.lsE_hi - D.lsE_lo) * (float)ce / (float)(NCLS_E - 1);
            const float vT = D.lsT_lo + (D.lsT_hi - D.lsT_lo) * (float)ct / (float)(NCLS_T - 1);
            const float gm = 2.0f * w_mom * (eE * q * (vE - mE) + eT * q * (vT - mT));
            const float g = (q - tgt + gm) * inv;
            dZ[(size_t)b * NOP + c] = g;
            dZb[(size_t)b * NOP + c] = f2bf(g);
        }
        const float d = z[NCLS] - labels[3 * b + 2];
        loss += w_att * d * d;
        const float g = 2.0f * w_att * d * inv;
        dZ[(size_t)b * NOP + NCLS] = g;
        dZb[(size_t)b * NOP + NCLS] = f2bf(g);
        for (int c = NOUT; c < NOP; ++c) {
            dZ[(size_t)b * NOP + c] = 0.0f;
            dZb[(size_t)b * NOP + c] = 0;
        }
        loss_part[b] = loss;
    }
}

// dZ = dA * GELU'(Z) for the hidden layers; BF16 copy for the next GEMMs.
__global__ void k_gelu_back(const float* __restrict__ dA, const float* __restrict__ Z, size_t n, float* dZ,
                            uint16_t* dZb) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float g = dA[i] * gelu_grad(Z[i]);
        dZ[i] = g;
        dZb[i] = f2bf(g);
    }
}

// Bias gradients: db[n] = sum_b dZ[b][n] in a fixed order.  A workgroup of 16 waves takes 64
// columns (lane = column, so each row is read coalesced).  Wave w owns rows w, w+16, ...; it loads
// them 16 at a time (all in flight together) and adds them in row order into one running sum per
// lane, and the 16 waves' sums of a column are then added in wave order.
constexpr int COLSUM_WAVES = 16;
__global__ __launch_bounds__(64 * COLSUM_WAVES) void k_colsum(const float* __restrict__ dZ, int Bn, int ld,
                                                              int nvalid, float* db) {
    constexpr int U = 16;
    __shared__ float part[COLSUM_WAVES][64];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + l;
    float s = 0.0f;
    if (n < nvalid) {
        for (int b0 = w; b0 < Bn; b0 += U * COLSUM_WAVES) {
            float v[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const int b = b0 + j * COLSUM_WAVES;
                v[j] = b < Bn ? dZ[(size_t)b * ld + n] : 0.0f;
            }
#pragma unroll
            for (int j = 0; j < U; ++j) s += v[j];
        }
    }
    part[w][l] = s;
    __syncthreads();
    if (w == 0 && n < nvalid) {
        float t = 0.0f;
        for (int k = 0; k < COLSUM_WAVES; ++k) t += part[k][l];
        db[n] = t;
    }
}

// Fixed-order sum of per-sample losses (one workgroup).
__global__ void k_sum(const float* __restrict__ v, int n, float* out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += (double)v[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = (float)sh[0];
}

// Adam on the FP32 master parameters, then the padded BF16 copies.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.0f - b1) * gi;
        const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        p[i] -= lr * (a / c1) / (sqrtf_(b / c2) + eps);
    }
}

__global__ void k_pack_bf16(const float* __restrict__ p, uint16_t* __restrict__ wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NBW) return;
    float v = 0.0f;
    if (i < BOFF_W1) {
        const int n = i / KP0, k = i % KP0;
        v = k < NIN ? p[OFF_W0 + n * NIN + k] : 0.0f;
    } else if (i < BOFF_W2) {
        v = p[OFF_W1 + (i - BOFF_W1)];
    } else if (i < BOFF_W3) {
        v = p[OFF_W2 + (i - BOFF_W2)];
    } else {
        const int j = i - BOFF_W3, n = j / NH, k = j % NH;
        v = n < NOUT ? p[OFF_W3 + n * NH + k] : 0.0f;
    }
    wb[i] = f2bf(v);
}

// Counter-based uniform in [0, 1) of (seed, index, stream) (splitmix64 finaliser).
__device__ inline float hash_u01(uint32_t seed, uint64_t idx, uint32_t stream) {
    uint64_t x = idx * 0x9E3779B97F4A7C15ull ^ ((uint64_t)seed << 32) ^ ((uint64_t)stream * 0xD1B54A32D192ED03ull);
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (float)(x >> 40) * (1.0f / 16777216.0f);
}

// Training samples from the plain physics: E log-uniform in [0.1, 250] MeV, rho uniform in
// [0.2, 2], X0 ratio log-uniform in [0.5, 4], ds log-uniform between the table's smallest step
// (rho ds = 1e-4 mm) and the device step rule's step (steps cut short by cell boundaries
// included); labels from the same functions K3 calls:
// straggling width, scattering width sqrt(T(E_mid) ds) with E_mid = E - mean dE / 2, and the
// attenuation exponent sigma_n(E) rho ds.  Samples outside the input range or with a step below
// the Highland threshold (ds / X0 < 1e-6) are redrawn (at most 16 times, then replaced by a
// fixed in-range sample).
__global__ void k_gen_samples(Tabs T, uint32_t seed, int n, uint16_t* __restrict__ X, float* __restrict__ labels,
                              float* __restrict__ raw) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float E = 10.0f, rho = 1.0f, x0r = 1.0f, ds = 0.1f, x[4];
        for (int tr = 0; tr < 16; ++tr) {
            const uint64_t id = (uint64_t)i * 16 + tr;
            const float u0 = hash_u01(seed, id, 0), u1 = hash_u01(seed, id, 1), u2 = hash_u01(seed, id, 2),
                        u3 = hash_u01(seed, id, 3);
            const float Et = expf_(logf_(E_LO) + u0 * (logf_(E_HI) - logf_(E_LO)));
            const float rt = RHO_LO + u1 * (RHO_HI - RHO_LO);
            const float xt = expf_(logf_(X0R_LO) + u2 * (logf_(X0R_HI) - logf_(X0R_LO)));
            // substep log-uniform between the table's smallest step and the largest step
            const float smax = max_step(T, Et, rt, 0), smin = fminf(DR_LO / rt, smax);
            const float dt = smax * expf_(logf_(smin / smax) * (1.0f - u3));
            if (dt / (X0W * xt) < 1e-6f || !encode_input(Et, rt * dt, rt, xt, x)) continue;
            E = Et;
            rho = rt;
            x0r = xt;
            ds = dt;
            break;
        }
        encode_input(E, rho * ds, rho, x0r, x);
        const float sE = straggling(E, ds, rho);
        const float mdE = deposition(T, E, ds, rho);
        const float Em = fmaxf(E - 0.5f * mdE, T.ee[0]);
        const float sth = sqrtf_(fmaxf(scat_power(Em, ds, X0W * x0r) * ds, 0.0f));
        const float a = nuc_sigma(E) * rho * ds;
        for (int k = 0; k < KP0; ++k) X[(size_t)i * KP0 + k] = k < NIN ? f2bf(x[k]) : 0;
        labels[3 * i + 0] = logf_(fmaxf(sE, 1e-30f));
        labels[3 * i + 1] = logf_(fmaxf(sth, 1e-30f));
        labels[3 * i + 2] = logf_(a + 1e-9f);
        if (raw) {
            raw[4 * i + 0] = E;
            raw[4 * i + 1] = rho * ds;
            raw[4 * i + 2] = rho;
            raw[4 * i + 3] = x0r;
        }
    }
}

// Decoding of the network outputs: ln widths as log-moments of the class distribution and the
// attenuation exponent.  out3[b] = (ln sigma_E, ln sigma_theta, a).
SM2D_HD inline void decode_outputs(const float* z, const Dict& D, float* out3) {
    float mx = z[0];
    for (int c = 1; c < NCLS; ++c) mx = fmaxf(mx, z[c]);
    float p[NCLS], s = 0.0f;
    for (int c = 0; c < NCLS; ++c) {
        p[c] = expf_(z[c] - mx);
        s += p[c];
    }
    float lsE = 0.0f, lsT = 0.0f;
    for (int c = 0; c < NCLS; ++c) {
        const float q = p[c] / s;
        const int ce = c / NCLS_T, ct = c % NCLS_T;
        lsE += q * (D.lsE_lo + (D.lsE_hi - D.lsE_lo) * (float)ce / (float)(NCLS_E - 1));
        lsT += q * (D.lsT_lo + (D.lsT_hi - D.lsT_lo) * (float)ct / (float)(NCLS_T - 1));
    }
    out3[0] = lsE;
    out3[1] = lsT;
    out3[2] = fmaxf(expf_(z[NCLS]) - 1e-9f, 0.0f);
}

__global__ void k_decode(const float* __restrict__ Z4, int Bn, Dict D, float* out3) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < Bn; b += gridDim.x * blockDim.x)
        decode_outputs(Z4 + (size_t)b * NOP, D, out3 + 3 * (size_t)b);
}

// Inputs of the surrogate table: for material m, energy bin e and step node j the network input
// (E = bin centre, dR = node j of NJ log-spaced over [DR_LO, DR_HI]).
__global__ void k_table_inputs(Tabs T, const float* __restrict__ mrho, const float* __restrict__ mx0r, int nmat,
                               int NJ, uint16_t* X, uint8_t* valid) {
    const int total = nmat * T.NE * NJ;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int j = i % NJ, e = (i / NJ) % T.NE, m = i / (NJ * T.NE);
        const float E = ecenter(T.ee, e);
        const float dR = expf_(logf_(DR_LO) + (logf_(DR_HI) - logf_(DR_LO)) * (float)j / (float)(NJ - 1));
        float x[4] = {0, 0, 0, 0};
        const bool ok = encode_input(E, dR, mrho[m], mx0r[m], x);
        valid[i] = ok ? 1 : 0;
        for (int k = 0; k < KP0; ++k) X[(size_t)i * KP0 + k] = k < NIN ? f2bf(x[k]) : 0;
    }
}

// The plain physics at the table nodes (the labels the network is trained on): out3[i] =
// (ln sigma_E, ln sigma_theta, a) of material m, energy bin e (its centre) and step node j.
// valid[i]: the node can be reached by a fine-transport substep (in the input range, above the
// Highland threshold, and the node below it within 1.5 x the device step rule's largest step).
__global__ void k_table_plain(Tabs T, const float* __restrict__ mrho, const float* __restrict__ mx0r, int nmat,
                              int NJ, float* __restrict__ out3, uint8_t* __restrict__ valid) {
    const int total = nmat * T.NE * NJ;
    const float lo = logf_(DR_LO), hi = logf_(DR_HI);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int j = i % NJ, e = (i / NJ) % T.NE, m = i / (NJ * T.NE);
        const float E = ecenter(T.ee, e);
        const float dR = expf_(lo + (hi - lo) * (float)j / (float)(NJ - 1));
        const float dRprev = expf_(lo + (hi - lo) * (float)max(j - 1, 0) / (float)(NJ - 1));
        const float rho = mrho[m], x0r = mx0r[m], ds = dR / rho;
        float x[4];
        const bool reach = dRprev / rho <= 1.5f * max_step(T, E, rho, 0);
        valid[i] = (reach && encode_input(E, dR, rho, x0r, x) && ds / (X0W * x0r) >= 1e-6f) ? 1 : 0;
        const float sE = straggling(E, ds, rho);
        const float mdE = deposition(T, E, ds, rho);
        const float Em = fmaxf(E - 0.5f * mdE, T.ee[0]);
        const float sth = sqrtf_(fmaxf(scat_power(Em, ds, X0W * x0r) * ds, 0.0f));
        out3[3 * (size_t)i + 0] = logf_(fmaxf(sE, 1e-30f));
        out3[3 * (size_t)i + 1] = logf_(fmaxf(sth, 1e-30f));
        out3[3 * (size_t)i + 2] = nuc_sigma(E) * rho * ds;
    }
}

// ------------------------------------------------------------------------------------------
// The network inside a kernel (inference in the fine transport): the 64 lanes of a wavefront
// are 64 samples, which go through in four tiles of 16.  Transposed form Z^T = W A^T: the
// weights are the A operand (lane l: W[16t + l%16][16kk + 4(l/16) + i], four consecutive BF16 of
// a row of W[out][in]) and the activations the B operand (lane l: unit 16kk + 4(l/16) + i of
// sample 16s + l%16).  Output tile t of a layer (lane l: unit 16t + 4(l/16) + i of sample
// 16s + l%16) is then, after bias, GELU and rounding to BF16, the B operand of K step t of the
// next layer, with no data movement between layers.  Every lane of the wavefront must call it.
// ------------------------------------------------------------------------------------------
// Global-memory views of the weights (generic pointers would compile to flat loads, which also
// hold the LDS counter of the calling transport).
typedef const __attribute__((address_space(1))) uint16_t* gbf16p;
typedef const __attribute__((address_space(1))) float* gcf32p;

__device__ inline bf16x4 ld_bf16x4(gbf16p q) {  // 4 BF16 at an 8-byte aligned address
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    const u32x2 v = *reinterpret_cast<const __attribute__((address_space(1))) u32x2*>(q);
    bf16x4 r;
    r[0] = (short)(v.x & 0xFFFFu);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xFFFFu);
    r[3] = (short)(v.y >> 16);
    return r;
}

// One layer of 128 outputs over K steps of 16 inputs: out[t] = BF16(GELU(W in + b)).
template <int KSTEPS, int LD>
__device__ inline void mlp_tile_layer(gbf16p __restrict__ W, gcf32p __restrict__ b, const bf16x4* in,
                                      bf16x4* out) {
    const int l = (int)(threadIdx.x & 63), g = l >> 4, c = l & 15;
#pragma unroll
    for (int t = 0; t < NH / 16; ++t) {
        f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        const gbf16p row = W + (size_t)(16 * t + c) * LD + 4 * g;
#pragma unroll
        for (int kk = 0; kk < KSTEPS; ++kk)
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ld_bf16x4(row + 16 * kk), in[kk], acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) out[t][i] = (short)f2bf(gelu(acc[i] + b[16 * t + 4 * g + i]));
        // one tile's weight loads at a time: hoisting the next tiles' loads would hold up to 128
        // more registers
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Decoded outputs (ln sigma_E, ln sigma_theta, a) of this lane's sample, input x[4] (normalised).
__device__ __attribute__((always_inline)) inline void mlp_wave_forward(const uint16_t* __restrict__ wb_, const float* __restrict__ p_,
                                        const float* dict4, const float x[4], float& lsE_out, float& lsT_out,
                                        float& a_out) {
    const gbf16p wb = (gbf16p)wb_;
    const gcf32p p = (gcf32p)p_;
    const int l = (int)(threadIdx.x & 63), g = l >> 4, c = l & 15;
    const float eLo = dict4[0], eHi = dict4[1], tLo = dict4[2], tHi = dict4[3];
    lsE_out = lsT_out = a_out = 0.0f;
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
        // layer 0 input: features 0..3 of sample 16s + c in the lanes of group 0, zeros above
        bf16x4 x0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float v = __shfl(x[i], 16 * s + c);
            x0[i] = (short)(g == 0 ? f2bf(v) : 0);
        }
        bf16x4 h0[8], h1[8];
        mlp_tile_layer<1, KP0>(wb + BOFF_W0, p + OFF_B0, &x0, h0);
        mlp_tile_layer<NH / 16, NH>(wb + BOFF_W1, p + OFF_B1, h0, h1);
        mlp_tile_layer<NH / 16, NH>(wb + BOFF_W2, p + OFF_B2, h1, h0);
        // output layer: tiles 0, 1 are the 32 class logits, unit 32 of tile 2 is ln(a + 1e-9)
        f32x4 z[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            const gbf16p row = wb + BOFF_W3 + (size_t)(16 * t + c) * NH + 4 * g;
#pragma unroll
            for (int kk = 0; kk < NH / 16; ++kk)
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ld_bf16x4(row + 16 * kk), h0[kk], acc, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int u = 16 * t + 4 * g + i;
                z[t][i] = acc[i] + (u < NOUT ? p[OFF_B3 + u] : 0.0f);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        // class distribution of sample 16s + c: the 32 logits lie in its four lanes c + 16 g
        float mx = z[0][0];
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) mx = fmaxf(mx, z[t][i]);
        mx = fmaxf(mx, __shfl_xor(mx, 16));
        mx = fmaxf(mx, __shfl_xor(mx, 32));
        float sp = 0.0f, sE = 0.0f, sT = 0.0f;
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int u = 16 * t + 4 * g + i, ce = u / NCLS_T, ct = u % NCLS_T;
                const float q = expf_(z[t][i] - mx);
                sp += q;
                sE += q * (eLo + (eHi - eLo) * (float)ce / (float)(NCLS_E - 1));
                sT += q * (tLo + (tHi - tLo) * (float)ct / (float)(NCLS_T - 1));
            }
        sp += __shfl_xor(sp, 16);
        sE += __shfl_xor(sE, 16);
        sT += __shfl_xor(sT, 16);
        sp += __shfl_xor(sp, 32);
        sE += __shfl_xor(sE, 32);
        sT += __shfl_xor(sT, 32);
        // sample 16s + c's results sit in lane c (group 0 holds unit 32 in z[2][0])
        const float lsE = __shfl(sE / sp, l & 15), lsT = __shfl(sT / sp, l & 15), za = __shfl(z[2][0], l & 15);
        if ((l >> 4) == s) {
            lsE_out = lsE;
            lsT_out = lsT;
            a_out = fmaxf(expf_(za) - 1e-9f, 0.0f);
        }
    }
}

// Inference probe: decoded outputs of n normalised inputs x[n][4] through mlp_wave_forward (for
// the tests against the numpy restatement and the GEMM path).
__global__ __launch_bounds__(64) void k_mlp_wave_probe(const uint16_t* __restrict__ wb, const float* __restrict__ p,
                                                       Dict D, const float* __restrict__ x, int n, float* out3) {
    const int i = (int)(blockIdx.x * 64 + threadIdx.x);
    float xi[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (i < n)
        for (int k = 0; k < 4; ++k) xi[k] = x[4 * (size_t)i + k];
    const float d4[4] = {D.lsE_lo, D.lsE_hi, D.lsT_lo, D.lsT_hi};
    float a0, a1, a2;
    mlp_wave_forward(wb, p, d4, xi, a0, a1, a2);
    if (i < n) {
        out3[3 * (size_t)i + 0] = a0;
        out3[3 * (size_t)i + 1] = a1;
        out3[3 * (size_t)i + 2] = a2;
    }
}

}  // namespace mlp
}  // namespace sm2d
